// Host interface of the gradient engine: regions resident in HBM; probabilities in, likelihoods
// and gradients out through host memory, one synchronous call each.
#pragma once

#include <cstdint>
#include <memory>

namespace fgkl {

struct GradInput {
  int64_t n_regions = 0;
  const int64_t* region_read = nullptr;  // n_regions + 1, read index ranges
  const int64_t* region_hap = nullptr;   // n_regions + 1, haplotype index ranges
  const int64_t* read_off = nullptr;     // n_reads + 1, byte ranges into bases
  const uint8_t* bases = nullptr;
  const int64_t* hap_off = nullptr;      // n_haps + 1, byte ranges into hap_bases
  const uint8_t* hap_bases = nullptr;
};

// Probabilities per read base (n_read_bases doubles each) and the global gap-continuation
// probability.
struct GradParams {
  const double *err = nullptr, *ins = nullptr, *del = nullptr;
  double gcp = 0.0;
};

class GradBatch {
 public:
  ~GradBatch();
  static int create(const GradInput& in, std::unique_ptr<GradBatch>& out);
  // 0: FP64 forward, tape and adjoint for every pair (the default).  1: FP32 forward, tape and
  // adjoint with FP64 accumulation of the gradient terms; pairs whose FP32 likelihood falls
  // below 1e-28 are rerun in FP64.
  int set_precision(int fp32);
  // log10 L of every pair, region after region, [r * H + h] inside each.  kernel_ms (may be
  // null) receives the device time of the kernels.
  int forward(const GradParams& p, double* log10_out, double* kernel_ms);
  // Gradient of sum_p w_p log10 L_p (w = 1 when weights is null); log10_out may be null.
  int backward(const GradParams& p, const double* weights, double* log10_out, double* g_err,
               double* g_ins, double* g_del, double* g_gcp, double* kernel_ms);
  int64_t n_pairs() const;
  int64_t n_read_bases() const;
  int64_t scratch_bytes() const;

 private:
  GradBatch();
  struct Impl;
  Impl* impl_;
};

}  // namespace fgkl
