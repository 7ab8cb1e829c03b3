// GATK PairHMM / PD-HMM results dumps (--pair-hmm-results-file, --pdhmm-results-file): read into
// a batch with the recorded likelihoods, or written from one (replay.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "fgkl_synthetic.hpp"

namespace fgkl {

struct ReplayData {
  Batch batch;                   // regions as reconstructed from the dump
  std::vector<double> expected;  // GATK's log10 likelihoods, region by region, [r * H + h]
  int64_t lines = 0;             // data lines read
  int64_t skipped_lines = 0;     // too few fields, unparsable numbers or flags, ragged arrays
  int64_t dropped_regions = 0;   // regions whose lines did not form reads x haplotypes
};

// Reads at most max_regions complete regions (all when <= 0).  Returns 0, or 1 with `error` set
// when the file cannot be opened.
int replay_read(const char* path, bool pd, int64_t max_regions, ReplayData& out,
                std::string& error);
// Writes every pair of `b` with expected[i] in the dump format (PD: with b's flags).
int replay_write(const Batch& b, const std::vector<double>& expected, const char* path, bool pd,
                 std::string& error);

}  // namespace fgkl
