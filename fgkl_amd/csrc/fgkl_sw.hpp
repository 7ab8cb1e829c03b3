// The Smith-Waterman engine: GATK's SmithWatermanJavaAligner on gfx950, many pairs per launch.
#pragma once
#include <cstdint>
#include <memory>

namespace fgkl {

// Overhang strategies in the codes FgklSmithWaterman.strategyCode assigns
// (FgklSmithWaterman.java:49-57).
enum SwStrategy : int { kSwSoftClip = 0, kSwIndel = 1, kSwLeadingIndel = 2, kSwIgnore = 3 };

struct SwRunStats {
  double total_ms = 0;          // device time of the run (HIP events on the batch's stream)
  double bucket_ms[8] = {};     // each rows-per-lane bucket's launch, R = 1..8
  int64_t bucket_pairs[8] = {};
  int64_t bucket_cells[8] = {};  // reference x alternate bases of each bucket's pairs
  int64_t substring_pairs = 0;   // pairs the exact-substring shortcut resolved
};

// A batch of (reference, alternate) pairs resident on the device.  Pair p aligns
// alt[alt_off[p] .. alt_off[p+1]) to ref[ref_off[p] .. ref_off[p+1]).
class SwBatch {
 public:
  // Validates (non-empty sequences, offset tables from zero, non-decreasing) and uploads.
  static int create(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref,
                    const int64_t* alt_off, const uint8_t* alt, std::unique_ptr<SwBatch>& out);
  // Replaces the batch's pairs, reusing its device memory where they fit.
  int assign(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref, const int64_t* alt_off,
             const uint8_t* alt);
  ~SwBatch();
  // Aligns every pair with weights {match, mismatch, gap open, gap extend} and one strategy.
  // With `stats` it returns when the alignments are done; without, results() waits for them.
  int run(const int32_t params[4], int strategy, SwRunStats* stats);
  // Copies the last run's results: offsets[p], CIGAR elements of pair p in
  // cigar[starts[p] .. starts[p+1]) as BAM encodes them (length << 4 | op, op 0 M 1 I 2 D 4 S).
  int results(int32_t* offsets, int64_t* starts, uint32_t* cigar, int64_t cigar_cap);
  int64_t n_pairs() const;
  int64_t cells() const;         // reference x alternate bases over the batch
  int64_t cigar_bound() const;   // an upper bound on the CIGAR elements of the batch

 private:
  SwBatch();
  struct Impl;
  std::unique_ptr<Impl> d_;
};

}  // namespace fgkl
