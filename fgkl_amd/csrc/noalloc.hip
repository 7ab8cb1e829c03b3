// The allocation-free variant of the per-region entry point (fgkl_pairhmm_compute_native_noalloc).
//
// It computes what fgkl_pairhmm_compute_native computes, for callers that cannot allocate device
// memory (serving loops and harnesses that forbid hipMalloc): every device byte it touches beyond
// the caller's arrays lives in file-scope static device arrays, the per-call `scratch` below and
// the Phred tables.  Each GPU holds its own copy of them from module load until the process ends.
// On the first call on a device (with that device current) the arrays' device addresses are
// looked up with hipGetSymbolAddress and the tables are uploaded; after that a call allocates
// nothing, on the device or the host: it copies the caller's arrays into `scratch`, builds the
// pair list and clears the work-queue counters with a kernel, runs the same pair-per-wave sweep
// kernels as the pooled entry (FP32 with the FP64 rescue, or FP64), and copies the likelihoods
// back.  Calls on one device share its `scratch`, which keeps the last call's contents; they are
// serialised by a lock in this process, and calls from other processes or from code that enqueues
// work on the arrays' memory meanwhile would race.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>

#include "../../include/fgkl.h"
#include "fgkl_internal.hpp"
#include "sweep.hpp"

namespace fgkl {
namespace {

// The per-call scratch: the caller's read and haplotype arrays, offsets, the per-call plan
// (one region), the pair and rescue lists, the counters and the likelihoods.
constexpr size_t kScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The Phred tables in the DevTab layout (1 - e, e / 3, e, 1 - gcp, match-to-match), per precision.
constexpr int kTabQ = kMaxQual + 1;
constexpr int kTabEntries = 4 * kTabQ + kMatchToMatchSize;
static __device__ __align__(256) float tables_f32[kTabEntries];
static __device__ __align__(256) double tables_f64[kTabEntries];

constexpr int kMaxDevices = 64;
constexpr int kCounterWords = 48;  // work-queue heads and the rescue count, as the batch engine

struct DeviceSlot {
  bool ready = false;
  unsigned char* scratch = nullptr;  // device address of `scratch` on this device
  DevTab<float> tf{};
  DevTab<double> td{};
  int num_cus = 0;
  std::mutex call;  // calls on one device run one at a time
};
DeviceSlot g_slots[kMaxDevices];
std::mutex g_setup;
float g_host_f32[kTabEntries];
double g_host_f64[kTabEntries];

int fail(int code, const char* msg) {
  set_error(code, msg);
  return code;
}

#define NA_TRY(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      set_error(kInternal, std::string("HIP error: ") + hipGetErrorString(_e) + " at " + #expr); \
      return kInternal;                                                                       \
    }                                                                                         \
  } while (0)

template <typename T>
void fill_host(T* h) {
  const Tables& t = tables();
  for (int q = 0; q <= kMaxQual; ++q) {
    const double e = t.error_prob[q];
    h[q] = T(1.0 - e);              // model.rs:80-83, narrowed as kernel.rs:516-522
    h[kTabQ + q] = T(e / 3.0);
    h[2 * kTabQ + q] = T(e);
    h[3 * kTabQ + q] = T(1.0 - e);  // model.rs:73
  }
  for (int i = 0; i < kMatchToMatchSize; ++i) h[4 * kTabQ + i] = T(t.match_to_match[i]);
}

template <typename T>
DevTab<T> tab_at(T* d) {
  DevTab<T> t;
  t.pm = d;
  t.px = d + kTabQ;
  t.ep = d + 2 * kTabQ;
  t.im = d + 3 * kTabQ;
  t.mm = d + 4 * kTabQ;
  return t;
}

// Once per device, with it current: the device addresses of the static arrays and the tables.
int setup(DeviceSlot** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
    int count = 0;
    (void)hipGetDeviceCount(&count);
    if (count == 0)
      return fail(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    NA_TRY(hipGetDevice(&dev));
  }
  if (dev >= kMaxDevices) return fail(kInternal, "device index beyond the engine's table");
  std::lock_guard<std::mutex> lk(g_setup);
  DeviceSlot& s = g_slots[dev];
  if (!s.ready) {
    void* p = nullptr;
    NA_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    s.scratch = static_cast<unsigned char*>(p);
    void* pf = nullptr;
    void* pd = nullptr;
    NA_TRY(hipGetSymbolAddress(&pf, HIP_SYMBOL(tables_f32)));
    NA_TRY(hipGetSymbolAddress(&pd, HIP_SYMBOL(tables_f64)));
    fill_host(g_host_f32);
    fill_host(g_host_f64);
    NA_TRY(hipMemcpy(pf, g_host_f32, sizeof(g_host_f32), hipMemcpyHostToDevice));
    NA_TRY(hipMemcpy(pd, g_host_f64, sizeof(g_host_f64), hipMemcpyHostToDevice));
    s.tf = tab_at(static_cast<float*>(pf));
    s.td = tab_at(static_cast<double*>(pd));
    NA_TRY(hipDeviceGetAttribute(&s.num_cus, hipDeviceAttributeMultiprocessorCount, dev));
    s.ready = true;
  }
  *out = &s;
  return kOk;
}

// crates/jni/src/lib.rs:343-363 without building a copy: starts at zero, ends at the array
// length, non-decreasing.
int check_offsets(const int32_t* off, int64_t n, int64_t total) {
  if (n < 1 || !off || off[0] != 0) return fail(kInvalidArgument, "offset table must start at zero");
  if (off[n - 1] < 0 || int64_t(off[n - 1]) != total)
    return fail(kInvalidArgument, "offset table must end at the array length");
  for (int64_t i = 1; i < n; ++i)
    if (off[i] < off[i - 1]) return fail(kInvalidArgument, "offset table must be non-decreasing");
  return kOk;
}

size_t up256(size_t b) { return (b + 255) & ~size_t(255); }

// The one-region plan in scratch: 64-bit read offsets, the region tables, unsorted haplotype
// plans (no sharing: every pair from column 0), the (read, haplotype) pair list and zeroed
// counters.
__global__ void prep_kernel(const int32_t* ro32, const int32_t* ho32, int64_t n_reads,
                            int64_t n_haps, int64_t* read_off, int32_t* read_region,
                            int32_t* read_local, int64_t* region_out, int32_t* region_hap0,
                            int32_t* region_nhap, HapPlan* plan, int32_t* items,
                            int32_t* counters) {
  const int64_t pairs = n_reads * n_haps;
  int64_t total = pairs;
  if (n_reads + 1 > total) total = n_reads + 1;
  if (n_haps > total) total = n_haps;
  if (kCounterWords > total) total = kCounterWords;
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
       i += int64_t(gridDim.x) * blockDim.x) {
    if (i <= n_reads) read_off[i] = ro32[i];
    if (i < n_reads) {
      read_region[i] = 0;
      read_local[i] = int32_t(i);
    }
    if (i < n_haps) {
      HapPlan h;
      h.off = ho32[i];
      h.len = ho32[i + 1] - ho32[i];
      h.lcp = 0;
      h.src = int32_t(i);
      h.dup = 0;
      h.start_slot = -1;
      h.start_src_len = 0;
      h.snap0 = 0;
      h.nsnap = 0;
      h.cut = 0;
      h.cut_node = 0;
      plan[i] = h;
    }
    if (i < pairs) {
      items[2 * i] = int32_t(i / n_haps);
      items[2 * i + 1] = int32_t(i % n_haps);
    }
    if (i < kCounterWords) counters[i] = 0;
    if (i == 0) {
      region_out[0] = 0;
      region_out[1] = pairs;
      region_hap0[0] = 0;
      region_nhap[0] = int32_t(n_haps);
    }
  }
}

template <typename T, typename K>
int launch(K kernel, const SweepArgs& a, const DevTab<T>& tab, int64_t items, int num_cus) {
  constexpr int kBlock = 256;  // four waves, one pair each
  const size_t lds = size_t(a.lds_words) * 4 * (kBlock / 64);
  int per_cu = 0;
  NA_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, lds));
  if (per_cu < 1) per_cu = 1;
  const int64_t want = (items + kBlock / 64 - 1) / (kBlock / 64);
  const int grid = int(want < int64_t(per_cu) * num_cus ? want : int64_t(per_cu) * num_cus);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), lds, 0, a, tab);
  NA_TRY(hipGetLastError());
  return kOk;
}

}  // namespace
}  // namespace fgkl

using namespace fgkl;

extern "C" int fgkl_pairhmm_compute_native_noalloc(
    const uint8_t* bases, int64_t bases_len, const uint8_t* quals, int64_t quals_len,
    const uint8_t* ins_gop, int64_t ins_gop_len, const uint8_t* del_gop, int64_t del_gop_len,
    const uint8_t* gcp, int64_t gcp_len, const int32_t* read_offsets, int64_t read_offsets_len,
    const uint8_t* hap_bases, int64_t hap_bases_len, const int32_t* hap_offsets,
    int64_t hap_offsets_len, int double_precision, double* likelihoods, int64_t likelihoods_len) {
  // The validation of fgkl_pairhmm_compute_native (crates/jni/src/lib.rs:105-133, 325-363;
  // crates/pairhmm/src/lib.rs:54-67), in the same order and with the same messages.
  if (quals_len != bases_len || ins_gop_len != bases_len || del_gop_len != bases_len ||
      gcp_len != bases_len)
    return fail(kInvalidArgument, "read bases, qualities and penalties differ in total length");
  if (int rc = check_offsets(read_offsets, read_offsets_len, bases_len)) return rc;
  if (int rc = check_offsets(hap_offsets, hap_offsets_len, hap_bases_len)) return rc;
  const int64_t n_reads = read_offsets_len - 1, n_haps = hap_offsets_len - 1;
  int max_read = 0, max_hap = 0;
  for (int64_t i = 0; i < n_reads; ++i) {
    const int n = read_offsets[i + 1] - read_offsets[i];
    if (n == 0) {
      set_error(kInvalidArgument, "read " + std::to_string(i) + " has no bases");
      return kInvalidArgument;
    }
    max_read = n > max_read ? n : max_read;
  }
  for (int64_t i = 0; i < n_haps; ++i) {
    const int m = hap_offsets[i + 1] - hap_offsets[i];
    if (m == 0) {
      set_error(kInvalidArgument, "haplotype " + std::to_string(i) + " has no bases");
      return kInvalidArgument;
    }
    max_hap = m > max_hap ? m : max_hap;
  }
  const int64_t pairs = n_reads * n_haps;
  if (likelihoods_len != pairs) {
    set_error(kInvalidArgument, "output has " + std::to_string(likelihoods_len) +
                                    " elements but reads x haplotypes is " + std::to_string(pairs));
    return kInvalidArgument;
  }
  if (pairs == 0) return kOk;
  // This variant's own limits: the staged sweep's read and haplotype lengths, and the scratch.
  if (max_read > kMaxReadLen)
    return fail(kInvalidArgument,
                "the allocation-free entry takes reads of at most 511 bases; use "
                "fgkl_pairhmm_compute_native");
  if (max_hap > kMaxStagedHap)
    return fail(kInvalidArgument,
                "the allocation-free entry takes haplotypes of at most 2048 bases; use "
                "fgkl_pairhmm_compute_native");
  if (pairs > kMaxPairList)
    return fail(kInvalidArgument,
                "the allocation-free entry takes at most 1048576 read x haplotype pairs");
  const size_t rb = size_t(bases_len), hb = size_t(hap_bases_len);
  size_t at = 0;
  auto take = [&at](size_t bytes) {
    const size_t o = at;
    at = up256(at + (bytes ? bytes : 1));
    return o;
  };
  const size_t o_bases = take(rb), o_quals = take(rb), o_ins = take(rb), o_del = take(rb),
               o_gcp = take(rb), o_hap = take(hb);
  const size_t o_ro32 = take(4 * size_t(n_reads + 1)), o_ho32 = take(4 * size_t(n_haps + 1));
  const size_t o_read_off = take(8 * size_t(n_reads + 1));
  const size_t o_region = take(4 * size_t(n_reads)), o_local = take(4 * size_t(n_reads));
  const size_t o_rout = take(16), o_rhap0 = take(4), o_rnhap = take(4);
  const size_t o_plan = take(sizeof(HapPlan) * size_t(n_haps));
  const size_t o_items = take(8 * size_t(pairs)), o_rescue = take(8 * size_t(pairs));
  const size_t o_counters = take(4 * kCounterWords);
  const size_t o_out = take(8 * size_t(pairs));
  if (at > kScratchBytes)
    return fail(kInvalidArgument,
                "the region needs more than the allocation-free entry's 32 MiB of device "
                "scratch; use fgkl_pairhmm_compute_native");

  DeviceSlot* dev = nullptr;
  if (int rc = setup(&dev)) return rc;
  std::lock_guard<std::mutex> lk(dev->call);
  unsigned char* s = dev->scratch;  // the array's device address on this device
  const hipStream_t stream = nullptr;
  NA_TRY(hipMemcpyAsync(s + o_bases, bases, rb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_quals, quals, rb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_ins, ins_gop, rb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_del, del_gop, rb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_gcp, gcp, rb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_hap, hap_bases, hb, hipMemcpyHostToDevice, stream));
  NA_TRY(hipMemcpyAsync(s + o_ro32, read_offsets, 4 * size_t(n_reads + 1), hipMemcpyHostToDevice,
                        stream));
  NA_TRY(hipMemcpyAsync(s + o_ho32, hap_offsets, 4 * size_t(n_haps + 1), hipMemcpyHostToDevice,
                        stream));
  auto* counters = reinterpret_cast<int32_t*>(s + o_counters);
  {
    int64_t total = pairs > n_reads + 1 ? pairs : n_reads + 1;
    if (n_haps > total) total = n_haps;
    const int64_t want = (total + 255) / 256;
    const int blocks = int(want < 4096 ? (want > 0 ? want : 1) : 4096);
    hipLaunchKernelGGL(prep_kernel, dim3(blocks), dim3(256), 0, stream,
                       reinterpret_cast<const int32_t*>(s + o_ro32),
                       reinterpret_cast<const int32_t*>(s + o_ho32), n_reads, n_haps,
                       reinterpret_cast<int64_t*>(s + o_read_off),
                       reinterpret_cast<int32_t*>(s + o_region),
                       reinterpret_cast<int32_t*>(s + o_local),
                       reinterpret_cast<int64_t*>(s + o_rout),
                       reinterpret_cast<int32_t*>(s + o_rhap0),
                       reinterpret_cast<int32_t*>(s + o_rnhap),
                       reinterpret_cast<HapPlan*>(s + o_plan),
                       reinterpret_cast<int32_t*>(s + o_items), counters);
    NA_TRY(hipGetLastError());
  }
  SweepArgs a{};
  a.bases = s + o_bases;
  a.quals = s + o_quals;
  a.ins = s + o_ins;
  a.del = s + o_del;
  a.gcp = s + o_gcp;
  a.read_off = reinterpret_cast<const int64_t*>(s + o_read_off);
  a.read_region = reinterpret_cast<const int32_t*>(s + o_region);
  a.read_local = reinterpret_cast<const int32_t*>(s + o_local);
  a.region_out = reinterpret_cast<const int64_t*>(s + o_rout);
  a.region_hap0 = reinterpret_cast<const int32_t*>(s + o_rhap0);
  a.region_nhap = reinterpret_cast<const int32_t*>(s + o_rnhap);
  a.plan = reinterpret_cast<const HapPlan*>(s + o_plan);
  a.hap_bytes = s + o_hap;
  a.out = reinterpret_cast<double*>(s + o_out);
  a.pair_mode = 1;
  a.lds_words = ((max_hap + 1 + 64 + 8) + 3) & ~3;
  // Every pair in one launch at the largest rows per lane: a pair's bits do not depend on the
  // lane layout, so they equal those of the pooled entry's launches.
  const int R = rows_per_lane(max_read);
  auto* rescue = reinterpret_cast<int32_t*>(s + o_rescue);
  a.items = reinterpret_cast<const int32_t*>(s + o_items);
  a.n_items = int32_t(pairs);
  a.counter = counters;
  if (double_precision) {
    if (int rc = launch<double>(kSweepF64[0][R - 1], a, dev->td, pairs, dev->num_cus)) return rc;
  } else {
    a.rescue = 1;
    a.rescue_pairs[0] = rescue;
    a.rescue_count = counters + 16;
    if (int rc = launch<float>(kSweepF32[0][0][R - 1], a, dev->tf, pairs, dev->num_cus)) return rc;
    // The FP64 rescue of the pairs the FP32 sweep lost (crates/pairhmm/src/lib.rs:467-549).
    SweepArgs b = a;
    b.rescue = 0;
    b.items = rescue;
    b.n_items = 0;
    b.n_items_dev = counters + 16;
    b.counter = counters + 8;
    if (int rc = launch<double>(kSweepF64[0][R - 1], b, dev->td, pairs, dev->num_cus)) return rc;
  }
  NA_TRY(hipMemcpy(likelihoods, s + o_out, 8 * size_t(pairs), hipMemcpyDeviceToHost));
  return kOk;
}
