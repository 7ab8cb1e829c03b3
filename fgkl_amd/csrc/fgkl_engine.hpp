// The device engine's host interface: a batch of regions resident in HBM and the ways to run it.
#pragma once

#include <cstdint>
#include <memory>

#include "fgkl_internal.hpp"

namespace fgkl {

enum RunMode : int {
  kModeFp32Rescue = 0,    // FP32 sweeps, FP64 recomputation of pairs below 1e-28 (lib.rs:377-549)
  kModeFp64 = 1,          // FP64 sweeps for every pair
  kModeFp64PerRegion = 2, // FP64, one pair per wave, one synchronous launch set per region
  kModeFp32NoRescue = 3,  // FP32 only; underflowing pairs are left NaN (double_fallback off)
};

struct RunStats {
  double total_ms = 0;     // device time of the whole run (HIP events on the batch's stream)
  double primary_ms = 0;   // the FP32 (or FP64) sweep kernels; -1 when run as one graph
  double rescue_ms = 0;    // the FP64 rescue kernels; -1 when run as one graph
  int64_t rescued_pairs = 0;
  double bucket_ms[kMaxRowsPerLane] = {};  // primary sweep launch of each rows-per-lane bucket
  int64_t bucket_cells[kMaxRowsPerLane] = {};  // DP cells (read x haplotype bases) it covers
  int64_t bucket_work[kMaxRowsPerLane] = {};   // cells it computes with prefix + suffix sharing
  double dense_rescue_ms = 0;  // of rescue_ms: dense reads rerun in FP64 with prefix sharing
};

// How the FP32 sweep of one rows-per-lane bucket is launched.  Every configuration computes
// the same bits: a read's rows, whichever lanes hold them, see the same operands in the same
// order, and identity rows reproduce DP row 0 exactly.
struct LaunchCfg {
  int rows = 0;        // kernel rows per lane, R or R + 1 for bucket R (0: R)
  int block = 256;     // threads per workgroup: 64, 128 or 256
  bool boost = false;  // the build compiled for one more wave per SIMD
  bool ug = false;     // the uniform-gap build (chosen per launch by the engine, not tuned)
};

struct TuneReport {
  LaunchCfg cfg[kMaxRowsPerLane];           // chosen per bucket (rows 0: empty bucket)
  double default_ms[kMaxRowsPerLane] = {};  // rows R, 256 threads, natural build
  double tuned_ms[kMaxRowsPerLane] = {};
  int candidates = 0;                       // configurations timed
  double seconds = 0;                       // wall time of the tuning
};

class DeviceBatch {
 public:
  ~DeviceBatch();
  // Validates and plans `in`, then copies it into device memory.
  static int create(const BatchInput& in, std::unique_ptr<DeviceBatch>& out);
  // Replaces the batch's regions with `in`, reusing its device memory, stream and events where
  // they suffice (waits for work queued on the previous inputs first).
  int assign(const BatchInput& in);
  // Computes every pair; `share` enables prefix sharing; `use_graph` replays one captured
  // hipGraph of the whole run.
  int run(int mode, int share, int use_graph, RunStats* stats);
  // Every (read, haplotype) pair in its own wave (one launch for a small batch, else one per
  // rows-per-lane bucket), then the FP64 rescue: the shortest time to results for batches too
  // small to fill the GPU.  No prefix sharing.  The results come back to pinned host memory in
  // the same stream, so results() then only copies them.  Stats are timed only when asked for.
  // Batches beyond the pair list run as run(mode, share, no graph).
  int run_latency(int mode, RunStats* stats);
  // FP32 pairs the last run_latency recomputed in FP64.
  int64_t last_rescued() const;
  // Copies the log10 likelihoods out, region by region, out[region_base + r * H + h].
  int results(double* out, int64_t n);
  int64_t n_pairs() const;
  int64_t scratch_bytes() const;
  // DP cells: nominal, with prefix sharing alone, forward and backward with suffix sharing.
  void work(int64_t out[4]) const;
  // Times every launch configuration of each FP32 bucket (rows R or R+1, 256 or 64 threads,
  // natural or occupancy-raised build), keeps the fastest for this batch and for later batches
  // of the same kind on this device, and drops captured graphs.
  int tune(int share, TuneReport* report);
  // Sets (or reads) bucket R's configuration explicitly; kInvalidArgument outside the tuned
  // space.
  int set_launch(int R, const LaunchCfg& cfg);
  // Whether prefix-sharing launches may keep a snapshot record in the wave's LDS (default on;
  // off: every record goes to an HBM slot).  Results are the same bits either way.
  void set_lds_records(bool on);
  // Whether plain FP32 main passes run the items whose reads have uniform gap qualities with the
  // uniform-gap build (default on).  Results are the same bits either way.
  void set_uniform_gaps(bool on);
  // Snapshot records of the plan per FP32 pass (HBM, LDS-eligible, restarts) and the sweep
  // launches so far that kept a record in LDS.
  void snapshots(int64_t out[4]) const;
  int get_launch(int R, LaunchCfg* cfg) const;

 private:
  DeviceBatch();
  int run_per_region(RunStats* stats);
  struct Impl;
  Impl* impl_;
};

int dpp_selftest(int32_t* out64);
int device_available();

}  // namespace fgkl
