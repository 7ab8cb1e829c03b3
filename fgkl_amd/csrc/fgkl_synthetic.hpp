// Synthetic regions for tests and benchmarks (synthetic.rs restated) and the benchmark
// workload generators.
#pragma once

#include <cstdint>
#include <vector>

namespace fgkl {

// SplitMix64 (synthetic.rs:10-40).
class Rng {
 public:
  explicit Rng(uint64_t seed) : s_(seed) {}
  uint64_t next_u64();
  uint64_t below(uint64_t n);
  bool chance(double p);
  uint8_t base();
  int64_t range(int64_t lo, int64_t hi);
  double normal();

 private:
  uint64_t s_;
};

struct Read {
  std::vector<uint8_t> bases, quals, ins, del, gcp;
};

// Regions of reads and haplotypes, concatenated as the batch C-ABI takes them.
struct Batch {
  std::vector<int64_t> region_reads, region_haps;  // n_regions + 1 after end_batch()
  std::vector<int64_t> read_offsets{0}, hap_offsets{0};
  std::vector<uint8_t> bases, quals, ins, del, gcp, hap_bases, hap_flags;
  void begin_region();
  void end_batch();
  void add_read(const Read& r);
  void add_hap(const std::vector<uint8_t>& b, const std::vector<uint8_t>* flags);
  int64_t cells() const;
};

Read sample_read(Rng& rng, const std::vector<uint8_t>& source, int64_t read_len);
std::vector<uint8_t> edited(Rng& rng, const std::vector<uint8_t>& alphabet,
                            const std::vector<uint8_t>& base);
// Region::generate / PdRegion::generate: append one region (call end_batch() after the last).
void region_generate(Batch& out, uint64_t seed, int64_t num_reads, int64_t read_len,
                     int64_t num_haps, int64_t hap_len);
void pd_region_generate(Batch& out, uint64_t seed, int64_t num_reads, int64_t read_len,
                        int64_t num_haps, int64_t hap_len);
// Benchmark workloads 0-4 (one region of 32x8 FP64 parity data, the batched uneven regions,
// the underflow-heavy regions, PD regions, adjoint regions); n_regions <= 0 keeps the default.
void workload_generate(Batch& out, int config, uint64_t seed, int64_t n_regions);

}  // namespace fgkl
