// Aggregation of per-call regions into batches (aggregate.cpp).
#pragma once
#include <cstdint>

#include "fgkl_internal.hpp"

namespace fgkl {

// One region as a per-call entry point hands it over, validated; the pointers stay valid until
// the call returns.
struct RegionCall {
  bool pd = false;
  bool double_precision = false;
  int64_t n_reads = 0;
  const int64_t* read_off = nullptr;  // n_reads + 1
  const uint8_t *bases = nullptr, *quals = nullptr, *ins = nullptr, *del = nullptr,
                *gcp = nullptr;
  int64_t n_haps = 0;
  const int64_t* hap_off = nullptr;   // n_haps + 1
  const uint8_t* hap_bases = nullptr;
  const uint8_t* hap_flags = nullptr;  // PD only
  double* out = nullptr;
  int64_t out_len = 0;
};

// Turns aggregation on (max_regions > 1) or off; FGKL_AGGREGATE=<regions>[:<wait us>] sets it at
// the first call.
void aggregate_config(int max_regions, int max_wait_us);
bool aggregate_enabled();
void aggregate_stats(int64_t* batches, int64_t* regions);
// Frees the device memory of the batches no aggregated call is using; returns how many.
int64_t aggregate_release_idle();
// Runs the region in a batch with the calls that arrive within the wait; returns its status.
int aggregate_run(const RegionCall& call);
// Runs one region on its own (the per-call latency path); defined with the C-ABI.
int compute_region_alone(const RegionCall& call);

}  // namespace fgkl
