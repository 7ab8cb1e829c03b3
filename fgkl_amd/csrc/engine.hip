// The MI355X (gfx950) PairHMM / PD-HMM engine's batch runner: device state and tables, batch
// upload, the per-bucket launches (eagerly or as one captured hipGraph), the FP64 rescue, the
// per-region FP64 baseline and the launch autotuner.  The kernels are in sweep.hpp.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fgkl_engine.hpp"
#include "fgkl_internal.hpp"
#include "sweep.hpp"

namespace fgkl {

__global__ void fill_i32(int32_t* p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ void dpp_selftest_kernel(const int32_t* in, int32_t* out) {
  const int lane = threadIdx.x & 63;
  out[lane] = shr1i(-1, in[lane]);
}

// ---------------------------------------------------------------------------------------------
// Host side of the engine.

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      set_error(kInternal, std::string("HIP error: ") + hipGetErrorString(_e) + " at " +   \
                                 #expr);                                                     \
      return kInternal;                                                                      \
    }                                                                                        \
  } while (0)

namespace {

template <typename T> struct Kernels;
template <> struct Kernels<float> {
  static KernelF get(bool pd, bool boost, int rows, bool ug = false) {
    if (ug) return pd ? kSweepF32PdUniform[boost][rows - 1] : kSweepF32Uniform[boost][rows - 1];
    return kSweepF32[pd][boost][rows - 1];
  }
};
// FP64 has no occupancy-raised build; its flag selects the suffix-sharing build (plain only).
template <> struct Kernels<double> {
  static KernelD get(bool pd, bool suffix, int rows, bool ug = false) {
    if (ug) return pd ? kSweepF64PdUniform[rows - 1] : kSweepF64Uniform[suffix][rows - 1];
    return suffix && !pd ? kSweepF64Suffix[rows - 1] : kSweepF64[pd][rows - 1];
  }
};
template <typename T> struct LongKernel;
template <> struct LongKernel<float> {
  static KernelF get(bool pd) { return kLongF32[pd]; }
};
template <> struct LongKernel<double> {
  static KernelD get(bool pd) { return kLongF64[pd]; }
};

constexpr int kBlocks[3] = {64, 128, 256};
int block_index(int block) { return block == 64 ? 0 : block == 128 ? 1 : block == 256 ? 2 : -1; }

constexpr int64_t kMaxQueueItems = int64_t(1) << 24;  // 64 x (items + waves) < 2^31
constexpr int kCounters = 56;  // see Impl::counters
// Uploads above this size go straight from the caller's arrays when no staging buffer fits them.
constexpr size_t kDirectUploadBytes = size_t(64) << 20;
// Pairs a one-pair-per-wave run takes in a single launch: 256 CUs hold 8192 waves at 32 per CU.
constexpr int64_t kOneLaunchPairs = 8192;

struct DeviceState {
  int device = -1;
  int num_cus = 0;
  DevTab<float> tf{};
  DevTab<double> td{};
  // Resident blocks per CU of each kernel [fp64][pd][boost][rows-1][block], at the LDS size it
  // was queried with.
  // Resident blocks per CU of a sweep kernel at a block size and LDS bytes per block.
  std::map<std::array<int, 7>, int> occ;
  // Launch configurations the autotuner chose for FP32 buckets [pd][R-1] (rows 0: none yet).
  LaunchCfg tuned[2][kMaxRowsPerLane];
  // Resident one-wave blocks per CU of long_kernel [fp64][pd] (0: not queried yet).
  int occ_long[2][2] = {};
};

std::mutex g_dev_mutex;
DeviceState* g_dev = nullptr;

template <typename T>
int upload_tables(DevTab<T>& out) {
  const Tables& t = tables();
  std::vector<T> pm(kMaxQual + 1), px(kMaxQual + 1), ep(kMaxQual + 1), im(kMaxQual + 1),
      mm(kMatchToMatchSize);
  for (int q = 0; q <= kMaxQual; ++q) {
    const double e = t.error_prob[q];
    pm[q] = T(1.0 - e);                 // model.rs:80-83, narrowed as kernel.rs:516-522
    px[q] = T(e / 3.0);
    ep[q] = T(e);
    im[q] = T(1.0 - e);                 // model.rs:73
  }
  for (int i = 0; i < kMatchToMatchSize; ++i) mm[i] = T(t.match_to_match[i]);
  T* d = nullptr;
  const size_t n4 = size_t(kMaxQual + 1);
  HIP_TRY(hipMalloc(&d, sizeof(T) * (4 * n4 + kMatchToMatchSize)));
  HIP_TRY(hipMemcpy(d, pm.data(), sizeof(T) * n4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + n4, px.data(), sizeof(T) * n4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + 2 * n4, ep.data(), sizeof(T) * n4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + 3 * n4, im.data(), sizeof(T) * n4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(d + 4 * n4, mm.data(), sizeof(T) * kMatchToMatchSize, hipMemcpyHostToDevice));
  out.pm = d;
  out.px = d + n4;
  out.ep = d + 2 * n4;
  out.im = d + 3 * n4;
  out.mm = d + 4 * n4;
  return kOk;
}

int device_state(DeviceState** out) {
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  if (!g_dev) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) {
      set_error(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
      return kInternal;
    }
    auto* st = new DeviceState;
    HIP_TRY(hipGetDevice(&st->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, st->device));
    st->num_cus = prop.multiProcessorCount;
    if (int rc = upload_tables(st->tf)) return rc;
    if (int rc = upload_tables(st->td)) return rc;
    g_dev = st;
  }
  *out = g_dev;
  return kOk;
}

template <typename T>
int resident_blocks(DeviceState* st, bool pd, bool boost, int rows, int block, int lds_bytes,
                    int* blocks, bool ug = false) {
  const int ti = std::is_same<T, double>::value ? 1 : 0;
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  int& occ = st->occ[{ti, int(pd), int(boost), rows, block, lds_bytes, int(ug)}];
  if (occ == 0) {
    int nb = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nb, Kernels<T>::get(pd, boost, rows, ug), block, size_t(lds_bytes)));
    occ = std::max(nb, 1);
  }
  *blocks = occ;
  return kOk;
}

template <typename T> DevTab<T> dev_tab(DeviceState* st);
template <> DevTab<float> dev_tab<float>(DeviceState* st) { return st->tf; }
template <> DevTab<double> dev_tab<double>(DeviceState* st) { return st->td; }

}  // namespace

// ---------------------------------------------------------------------------------------------
struct DeviceBatch::Impl {
  DeviceState* st = nullptr;
  HostPlan plan;
  bool pd = false;
  hipStream_t stream = nullptr;
  // Buckets with little work (side[R - 1]) run their primary sweep on a second stream beside
  // the large ones, from their own part of the scratch (side_scratch_off bytes in).
  hipStream_t side_stream = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  bool side[kMaxRowsPerLane] = {};
  size_t side_scratch_off = 0;
  std::vector<void*> allocs;  // per-region mode lists, allocated on first use
  // The batch's arrays live in one device arena filled from one pinned staging buffer; both only
  // grow, so a batch re-assigned with inputs that fit reuses them (the per-call entry points do).
  void* arena = nullptr;
  size_t arena_cap = 0;
  void* staging = nullptr;
  size_t staging_cap = 0;
  // Pinned, grow-only: the rescued count (first 8 bytes) and the results of the last run, which
  // run_latency copies back in its own stream (out_on_host).
  void* readback = nullptr;
  size_t readback_cap = 0;
  bool out_on_host = false;
  int64_t last_rescued = 0;
  int reserve_readback(size_t bytes) {
    if (bytes <= readback_cap) return kOk;
    if (readback) HIP_TRY(hipHostFree(readback));
    readback = nullptr;
    readback_cap = 0;
    HIP_TRY(hipHostMalloc(&readback, bytes, hipHostMallocDefault));
    readback_cap = bytes;
    return kOk;
  }
  struct Piece {
    void** dst;
    const void* src;  // host data to upload, or null for device-only arrays
    size_t bytes, off;
  };
  std::vector<Piece> pieces;
  size_t layout_bytes = 0;
  // device copies
  uint8_t *bases = nullptr, *quals = nullptr, *ins = nullptr, *del = nullptr, *gcp = nullptr;
  int64_t* read_off = nullptr;
  int32_t *read_region = nullptr, *read_local = nullptr;
  int64_t* region_out = nullptr;
  int32_t *region_hap0 = nullptr, *region_nhap = nullptr;
  HapPlan* hplan = nullptr;
  SnapEntry* snaps = nullptr;
  uint8_t* hap_bytes = nullptr;
  uint32_t* hap_words = nullptr;
  uint8_t* region_classes = nullptr;  // PD: read-base classes from 4 on, per region
  int32_t* bucket_items[kMaxRowsPerLane] = {};
  int32_t bucket_n[kMaxRowsPerLane] = {};
  // Each bucket's items whose reads all have uniform gap qualities (and a gap-continuation
  // quality of at least kXsMinGcp) come first, bucket_ug[R - 1] of them, and run with the
  // uniform-gap build (launch_bucket).
  int32_t bucket_ug[kMaxRowsPerLane] = {};
  bool uniform_gaps = true;
  bool all_uniform = false;  // every read of the batch takes the uniform-gap build
  int64_t bucket_pairs[kMaxRowsPerLane] = {};
  int64_t bucket_cells[kMaxRowsPerLane] = {};
  int64_t bucket_work[kMaxRowsPerLane] = {};
  int32_t* rescue_pairs[kMaxRowsPerLane] = {};
  BwdSweep* bsweeps = nullptr;
  int32_t* bwrites = nullptr;
  int32_t* region_bw = nullptr;
  // Work-queue heads and counts per rows-per-lane bucket: [0,8) main, [8,16) rescue,
  // [16,24) rescue counts, [24,32) dense-read counts, [32,40) dense-rescue heads; long pairs:
  // 40 main head, 41 rescue head, 42 rescue count.
  int32_t* counters = nullptr;
  // Pairs of regions beyond the staged layout (long_kernel): (read, sorted haplotype), their
  // FP32 rescue list, and the boundary-row scratch (two buffers of (longest + 1) columns per
  // wave, sized for FP64 and the largest grid).
  int32_t* long_pairs = nullptr;
  int32_t* long_rescue = nullptr;
  int64_t n_long = 0;
  void* long_scratch = nullptr;
  size_t long_scratch_cap = 0;
  int long_grid = 0;
  uint8_t* dense_flags = nullptr;              // per read, written by every FP32 main pass
  int32_t* dense_items[kMaxRowsPerLane] = {};  // (read, -1) by the read's own rows per lane
  int64_t rows_reads[kMaxRowsPerLane] = {};    // reads of each rows per lane (plain batches)
  int64_t rows_pairs[kMaxRowsPerLane] = {};    // pairs of those reads (every batch)
  double* out = nullptr;
  void* scratch = nullptr;
  size_t scratch_bytes = 0, scratch_cap = 0;
  // Batches of up to kMaxPairList pairs (the per-call entry points, aggregated calls): every
  // (read, sorted haplotype) pair as pair-mode items grouped by the read's rows per lane (bucket
  // R from all_off[R - 1]; an item may carry two reads of a region, see pair_items), their FP32
  // rescue list, and the largest rows per lane.
  int32_t* all_pairs = nullptr;
  int32_t* all_rescue = nullptr;
  int64_t n_all_pairs = 0;  // pairs
  int64_t n_all_items = 0;  // items
  int64_t all_off[kMaxRowsPerLane + 1] = {};
  int all_rows = 0;
  bool pairs_only = false;  // planned for run_latency only
  size_t out_rb_off = 8;    // where the results start in the readback buffer
  // per-region pair lists, for the per-call mode
  int32_t* region_pairs = nullptr;
  std::vector<int64_t> region_pair_off;  // (G * 8 + 1) offsets in pairs, bucket-major per region
  int32_t* region_counters = nullptr;
  int64_t n_region_launches = 0;
  hipEvent_t ev[4] = {};
  hipEvent_t bev[2 * kMaxRowsPerLane] = {};  // around each primary bucket launch
  hipGraphExec_t graph[8] = {};  // by mode * 2 + share
  // LDS per wave, in words: the staged haplotype (the longest, plus the column a backward sweep
  // adds past its end, the ramp of the widest read and a group of four read ahead), and with
  // suffix sharing the join's FP64 row terms and chunk sums of up to 64 x rows rows, which reuse
  // the staging area once the forward sweep is done.
  int stage_hap_words = 0;
  bool join_lds = false;
  int stage_words(int rows) const {
    int w = stage_hap_words;
    if (join_lds) w = std::max(w, 2 * (64 * rows + 8 * rows + 2));
    return (w + 3) & ~3;
  }
  bool lds_records = true;  // prefix-sharing launches may keep a snapshot record in LDS
  int64_t lds_launches = 0;  // launches (eager or captured) that kept one
  LaunchCfg cfg[kMaxRowsPerLane];  // FP32 primary launch of each bucket

  void drop_graphs() {
    for (auto& g : graph) {
      if (g) (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  }

  void free_lazy() {
    for (void* p : allocs) (void)hipFree(p);
    allocs.clear();
    region_pairs = nullptr;
    region_counters = nullptr;
    region_pair_off.clear();
    n_region_launches = 0;
  }

  ~Impl() {
    if (stream) (void)hipStreamSynchronize(stream);
    drop_graphs();
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : bev)
      if (e) (void)hipEventDestroy(e);
    free_lazy();
    if (arena) (void)hipFree(arena);
    if (staging) (void)hipHostFree(staging);
    if (readback) (void)hipHostFree(readback);
    if (scratch) (void)hipFree(scratch);
    if (long_scratch) (void)hipFree(long_scratch);
    if (side_stream) (void)hipStreamSynchronize(side_stream);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    if (join_ev) (void)hipEventDestroy(join_ev);
    if (side_stream) (void)hipStreamDestroy(side_stream);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // Reserves an arena array of `count` elements (at least one), 256-byte aligned; `src` (may be
  // null) is uploaded into it by commit().
  template <typename P>
  void place(P** dst, const P* src, size_t count) {
    layout_bytes = (layout_bytes + 255) & ~size_t(255);
    const size_t bytes = sizeof(P) * (count ? count : 1);
    pieces.push_back(Piece{reinterpret_cast<void**>(dst), count ? src : nullptr, bytes, layout_bytes});
    layout_bytes += bytes;
  }
  // Grows the arena and the staging buffer to the layout, points every array into the arena and
  // uploads the host data with one copy on the batch's stream.
  int commit() {
    if (layout_bytes > arena_cap) {
      if (arena) HIP_TRY(hipFree(arena));
      arena = nullptr;
      arena_cap = 0;
      HIP_TRY(hipMalloc(&arena, layout_bytes));
      arena_cap = layout_bytes;
    }
    size_t upload_end = 0;
    for (const Piece& pc : pieces) {
      *pc.dst = static_cast<char*>(arena) + pc.off;
      if (pc.src) upload_end = pc.off + pc.bytes;
    }
    if (upload_end > kDirectUploadBytes && upload_end > staging_cap) {
      // A large upload with no staging buffer that holds it: the first writes into freshly
      // pinned pages cost more than the runtime's own staged copies from the caller's memory.
      for (const Piece& pc : pieces)
        if (pc.src)
          HIP_TRY(hipMemcpyAsync(static_cast<char*>(arena) + pc.off, pc.src, pc.bytes,
                                 hipMemcpyHostToDevice, stream));
      HIP_TRY(hipStreamSynchronize(stream));  // the caller's arrays may go once assign returns
    } else if (upload_end) {
      if (upload_end > staging_cap) {
        if (staging) HIP_TRY(hipHostFree(staging));
        staging = nullptr;
        staging_cap = 0;
        HIP_TRY(hipHostMalloc(&staging, upload_end, hipHostMallocDefault));
        staging_cap = upload_end;
      }
      // Into staging in chunks of at most 4 MiB, spread over host threads for large uploads.
      struct Chunk { char* dst; const char* src; size_t bytes; };
      std::vector<Chunk> chunks;
      constexpr size_t kChunk = size_t(4) << 20;
      for (const Piece& pc : pieces) {
        if (!pc.src) continue;
        for (size_t o = 0; o < pc.bytes; o += kChunk)
          chunks.push_back(Chunk{static_cast<char*>(staging) + pc.off + o,
                                 static_cast<const char*>(pc.src) + o, std::min(kChunk, pc.bytes - o)});
      }
      const int nt = host_threads(int64_t(upload_end), int64_t(kChunk) * 4);
      run_parallel(nt, [&](int t) {
        for (size_t i = size_t(t); i < chunks.size(); i += size_t(nt))
          std::memcpy(chunks[i].dst, chunks[i].src, chunks[i].bytes);
      });
      HIP_TRY(hipMemcpyAsync(arena, staging, upload_end, hipMemcpyHostToDevice, stream));
    }
    pieces.clear();
    layout_bytes = 0;
    return kOk;
  }

  template <typename P>
  int alloc(P** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(p), sizeof(P) * count));
    allocs.push_back(*p);
    return kOk;
  }
  template <typename P>
  int upload(P** p, const P* src, size_t count) {
    if (int rc = alloc(p, count)) return rc;
    if (count) HIP_TRY(hipMemcpyAsync(*p, src, sizeof(P) * count, hipMemcpyHostToDevice, stream));
    return kOk;
  }

  SweepArgs base_args() const {
    SweepArgs a{};
    a.bases = bases; a.quals = quals; a.ins = ins; a.del = del; a.gcp = gcp;
    a.read_off = read_off; a.read_region = read_region; a.read_local = read_local;
    a.region_out = region_out; a.region_hap0 = region_hap0; a.region_nhap = region_nhap;
    a.plan = hplan; a.snaps = snaps; a.hap_bytes = hap_bytes; a.hap_words = hap_words;
    a.region_classes = region_classes;
    a.scratch = scratch_bytes ? scratch : nullptr; a.out = out;
    a.bsweeps = bsweeps; a.bwrites = bwrites; a.region_bw = region_bw;
    return a;
  }

  // Waves a launch of bucket R's `max_items` items runs with: the resident limit, or fewer when
  // the bucket is small.  `wave_words` is the launch's LDS per wave (default: the staging only).
  template <typename T>
  int grid_waves(const LaunchCfg& c, int rows, int64_t max_items, int* grid, int wave_words = 0) {
    const int wpb = c.block / 64;
    int per_cu = 0;
    const int words = wave_words ? wave_words : stage_words(rows);
    if (int rc = resident_blocks<T>(st, pd, c.boost, rows, c.block, words * 4 * wpb, &per_cu,
                                    c.ug))
      return rc;
    const int64_t want = (max_items + wpb - 1) / wpb;
    *grid = int(std::min<int64_t>(int64_t(per_cu) * st->num_cus, want));
    return kOk;
  }

  // Words of a wave's LDS snapshot record at `rows` rows per lane: 64 lanes of snap_stride
  // values of T.
  template <typename T>
  int lds_record_words(int rows) const {
    const int nv = ((pd ? 6 * rows + 1 : 3 * rows + 1) + 1) & ~1;  // snap_stride
    return 64 * nv * int(sizeof(T) / 4);
  }
  // Gives a prefix-sharing launch the wave's LDS snapshot record when the plan puts records
  // there and the record costs no resident blocks (and fits the 64 KiB a block may hold);
  // otherwise its records all go to HBM slots.  Sets the LDS fields of `a`.
  template <typename T>
  int lds_layout(SweepArgs& a, const LaunchCfg& c, int rows) {
    const int base = stage_words(rows);
    a.lds_words = base;
    a.lds_rec = base;
    a.lds_snap = 0;
    if (!lds_records || !a.share || a.pair_mode || plan.snap_lds == 0) return kOk;
    const int wpb = c.block / 64;
    const int with = base + lds_record_words<T>(rows);
    if (int64_t(with) * 4 * wpb > 64 * 1024) return kOk;
    int b0 = 0, b1 = 0;
    if (int rc = resident_blocks<T>(st, pd, c.boost, rows, c.block, base * 4 * wpb, &b0, c.ug))
      return rc;
    if (int rc = resident_blocks<T>(st, pd, c.boost, rows, c.block, with * 4 * wpb, &b1, c.ug))
      return rc;
    if (b1 < b0) return kOk;
    a.lds_words = with;
    a.lds_snap = 1;
    ++lds_launches;
    return kOk;
  }

  // Launches bucket R's sweep with configuration `c` (FP64 launches ignore `boost` and take the
  // suffix-sharing build when `a.suffix` is set).  The scratch stride follows the kernel's rows
  // per lane.
  template <typename T>
  int launch(int R, SweepArgs a, int64_t max_items, LaunchCfg c = LaunchCfg{},
             hipStream_t on = nullptr, size_t scratch_off = 0) {
    if (max_items <= 0) return kOk;
    if (!on) on = stream;
    const int rows = c.rows ? c.rows : R;
    if (!std::is_same<T, float>::value) c.boost = a.suffix != 0;  // FP64: the suffix build
    // Pair runs, rescues and the FP64 dense pass mix the batch's reads in one item list: they
    // take the uniform-gap build when every read may.
    c.ug = uniform_gaps && (all_uniform || (c.ug && !a.pair_mode));
    if (int rc = lds_layout<T>(a, c, rows)) return rc;
    int grid = 0;
    if (int rc = grid_waves<T>(c, rows, max_items, &grid, a.lds_words)) return rc;
    if (a.scratch) {
      a.scratch_wave_stride = scratch_stride<T>(rows);
      a.node_off = snap_elems<T>(rows);
      // Every wave of the grid owns a stride of scratch: refuse a grid the batch's scratch
      // does not cover.
      const int64_t need = int64_t(grid) * (c.block / 64) * a.scratch_wave_stride * int64_t(sizeof(T));
      a.scratch = static_cast<char*>(a.scratch) + scratch_off;
      if (need + int64_t(scratch_off) > int64_t(scratch_bytes)) {
        set_error(kInternal, "a launch needs " + std::to_string(need) +
                                 " bytes of scratch, the batch holds " + std::to_string(scratch_bytes));
        return kInternal;
      }
    }
    hipLaunchKernelGGL(Kernels<T>::get(pd, c.boost, rows, c.ug), dim3(grid), dim3(c.block),
                       a.lds_words * 4 * (c.block / 64), on, a, dev_tab<T>(st));
    HIP_TRY(hipGetLastError());
    return kOk;
  }

  // Bucket R's main-pass launch over items `a` (primary_args) with configuration `c`: the items
  // of reads with uniform gap qualities with the uniform-gap build and the others (work-queue
  // counter 48 + R - 1) after them on the same stream.
  template <typename T>
  int launch_bucket(int R, const SweepArgs& a, int64_t n, LaunchCfg c, hipStream_t on,
                    size_t scratch_off) {
    const int64_t u = uniform_gaps ? bucket_ug[R - 1] : 0;
    if (u > 0) {
      SweepArgs au = a;
      au.n_items = int32_t(u);
      LaunchCfg cu = c;
      cu.ug = true;
      if (int rc = launch<T>(R, au, u, cu, on, scratch_off)) return rc;
    }
    if (u < n) {
      SweepArgs ag = a;
      ag.items = a.items + 2 * u;
      ag.n_items = int32_t(n - u);
      if (u > 0) ag.counter = counters + 48 + (R - 1);
      if (int rc = launch<T>(R, ag, n - u, c, on, scratch_off)) return rc;
    }
    return kOk;
  }

  // Scratch bytes a launch of bucket R with configuration `c` may touch.
  template <typename T>
  int64_t scratch_need(int R, LaunchCfg c, int64_t max_items = -1) {
    const int rows = c.rows ? c.rows : R;
    int grid = 0;
    if (grid_waves<T>(c, rows, max_items < 0 ? bucket_n[R - 1] : max_items, &grid)) return -1;
    return int64_t(grid) * (c.block / 64) * scratch_stride<T>(rows) * int64_t(sizeof(T));
  }
  // Per wave: the snapshot slots, then (plain) the suffix-sharing node records.
  template <typename T>
  int64_t snap_elems(int rows) const {
    const int nv = ((pd ? 6 * rows + 1 : 3 * rows + 1) + 1) & ~1;  // snap_stride
    return int64_t(plan.max_slots) * nv * 64;
  }
  template <typename T>
  int64_t scratch_stride(int rows) const {
    const bool nodes = !pd;
    const int ns = (2 * rows + 1 + 3) & ~3;  // node_stride
    return snap_elems<T>(rows) + (nodes ? int64_t(plan.max_nodes) * ns * 64 : 0);
  }

  // Elements of T per wave of long_kernel's boundary scratch: two buffers of (longest + 1)
  // columns of M, X, Y (and PD branch values).
  int64_t long_stride() const { return 2 * (int64_t(plan.long_max_hap) + 1) * (pd ? 6 : 3); }

  // Resident waves of long_kernel<T> on the device.
  template <typename T>
  int long_resident(int* waves) {
    const int ti = std::is_same<T, double>::value ? 1 : 0;
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    int& occ = st->occ_long[ti][pd];
    if (occ == 0) {
      int nb = 0;
      HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, LongKernel<T>::get(pd), 64, 0));
      occ = std::max(nb, 1);
    }
    *waves = occ * st->num_cus;
    return kOk;
  }

  // Launches long_kernel<T> over `max_items` items of `a` (pair mode, one wave per block).
  template <typename T>
  int launch_long(SweepArgs a, int64_t max_items) {
    if (max_items <= 0) return kOk;
    int resident = 0;
    if (int rc = long_resident<T>(&resident)) return rc;
    const int grid = int(std::min<int64_t>(std::min<int64_t>(resident, max_items), long_grid));
    a.scratch = long_scratch;
    a.scratch_wave_stride = long_stride();
    a.pair_mode = 1;
    hipLaunchKernelGGL(LongKernel<T>::get(pd), dim3(grid), dim3(64), 0, stream, a, dev_tab<T>(st));
    HIP_TRY(hipGetLastError());
    return kOk;
  }

  // The long pairs' pass: FP32 with their rescue list (or FP64), counters 40 and 42.
  int enqueue_long(bool fp32) {
    if (!n_long) return kOk;
    SweepArgs a = base_args();
    a.items = long_pairs;
    a.n_items = int32_t(n_long);
    a.counter = counters + 40;
    if (fp32) {
      a.rescue = 1;
      a.rescue_pairs[0] = long_rescue;
      a.rescue_count = counters + 42;
      return launch_long<float>(a, n_long);
    }
    return launch_long<double>(a, n_long);
  }
  // The FP64 rescue of the long pairs the FP32 pass lost.
  int enqueue_long_rescue() {
    if (!n_long) return kOk;
    SweepArgs a = base_args();
    a.items = long_rescue;
    a.n_items = 0;
    a.n_items_dev = counters + 42;
    a.counter = counters + 41;
    return launch_long<double>(a, n_long);
  }

  // Largest rows per lane bucket R may run at.  A read's rows, plain or PD, see the same
  // operands in the same order at R and R + 1, so both give the same bits.
  int max_rows(int R) const { return std::min(R + 1, kMaxRowsPerLane); }

  // The FP32 configurations of bucket R that the tuner times (the default first), or, with
  // `all`, every one set_launch accepts.
  std::vector<LaunchCfg> candidates(int R, bool all = false) const {
    std::vector<LaunchCfg> out;
    for (int rows = R; rows <= max_rows(R); ++rows)
      for (int boost = 0; boost < 2; ++boost)
        for (int block : {256, 64, 128})
          if (all || block != 128) out.push_back(LaunchCfg{rows, block, boost != 0});
    return out;
  }
  bool valid(int R, const LaunchCfg& c) const {
    const int rows = c.rows ? c.rows : R;
    return rows >= R && rows <= max_rows(R) && block_index(c.block) >= 0;
  }

  // Whether a main pass with prefix sharing `share` also shares suffixes (plain batches with a
  // suffix plan, both precisions).
  bool suffix_run(int share) const { return share && !pd && !plan.bsweeps.empty(); }

  SweepArgs primary_args(int R, int share, bool fp32) const {
    SweepArgs a = base_args();
    a.items = bucket_items[R - 1];
    a.n_items = bucket_n[R - 1];
    a.counter = counters + (R - 1);
    a.share = share;
    a.suffix = suffix_run(share);
    if (fp32 || a.suffix) {  // FP64: joins below 1e24 are queued for an unshared rerun
      a.rescue = 1;
      for (int q = 0; q < kMaxRowsPerLane; ++q) a.rescue_pairs[q] = rescue_pairs[q];
      a.rescue_count = counters + 16;
      a.rescue_by_rows = 1;
      if (fp32 && !pd) {  // PD batches rescue pair by pair (see enqueue)
        a.dense = dense_flags;
        for (int q = 0; q < kMaxRowsPerLane; ++q) a.dense_items[q] = dense_items[q];
        a.dense_count = counters + 24;
      }
    }
    return a;
  }

  int enqueue(int mode, int share, bool timed);
};

int DeviceBatch::Impl::enqueue(int mode, int share, bool timed) {
  // Reset the work-queue heads and rescue counts with a kernel (graph-safe).
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, stream, counters, kCounters, 0);
  HIP_TRY(hipGetLastError());
  const bool fp32 = mode == kModeFp32Rescue || mode == kModeFp32NoRescue;
  // The primary sweep, one launch per rows-per-lane bucket, each bracketed by events when the
  // run is not captured into a graph.  The small buckets run on the side stream meanwhile
  // (forked after the counters are cleared, joined before the rescue).
  bool forked = false;
  for (int R = 1; R <= kMaxRowsPerLane; ++R) {
    if (bucket_n[R - 1] && side[R - 1] && !forked) {
      HIP_TRY(hipEventRecord(fork_ev, stream));
      HIP_TRY(hipStreamWaitEvent(side_stream, fork_ev, 0));
      forked = true;
    }
  }
  for (int pass = 0; pass < 2; ++pass) {  // side buckets first, so they start at once
    for (int R = 1; R <= kMaxRowsPerLane; ++R) {
      if (!bucket_n[R - 1] || side[R - 1] != (pass == 0)) continue;
      hipStream_t on = side[R - 1] ? side_stream : stream;
      const size_t off = side[R - 1] ? side_scratch_off : 0;
      const SweepArgs a = primary_args(R, share, fp32);
      if (timed) HIP_TRY(hipEventRecord(bev[2 * (R - 1)], on));
      if (fp32) {
        if (int rc = launch_bucket<float>(R, a, bucket_n[R - 1], cfg[R - 1], on, off)) return rc;
      } else {
        if (int rc = launch_bucket<double>(R, a, bucket_n[R - 1], LaunchCfg{}, on, off)) return rc;
      }
      if (timed) HIP_TRY(hipEventRecord(bev[2 * (R - 1) + 1], on));
    }
  }
  if (forked) {
    HIP_TRY(hipEventRecord(join_ev, side_stream));
    HIP_TRY(hipStreamWaitEvent(stream, join_ev, 0));
  }
  if (int rc = enqueue_long(fp32)) return rc;
  if (timed) HIP_TRY(hipEventRecord(ev[1], stream));
  if (fp32 && mode == kModeFp32Rescue) {
    // Dense reads first: each against every haplotype of its region, FP64 with prefix sharing,
    // filling only its lost pairs.  The grid is the FP64 main pass's, which the scratch is sized
    // for; its waves take the dense reads from a queue and the ones past the count exit at once.
    // PD batches skip this: their haplotypes rarely share prefixes, and one wave walking a read
    // through every haplotype measured slower than a wave per lost pair.
    for (int R = 1; R <= kMaxRowsPerLane && !pd; ++R) {
      if (!rows_reads[R - 1]) continue;
      SweepArgs a = base_args();
      a.items = dense_items[R - 1];
      a.n_items = 0;
      a.n_items_dev = counters + 24 + (R - 1);
      a.counter = counters + 32 + (R - 1);
      a.share = 1;
      a.nan_only = 1;
      if (int rc = launch<double>(R, a, rows_reads[R - 1])) return rc;
    }
    if (timed) HIP_TRY(hipEventRecord(ev[3], stream));
  }
  // FP64 joins below 1e24 are recomputed without suffix sharing (lib.rs:467-487), as are the
  // pairs an FP32 pass lost that the dense pass did not take.
  const bool joins64 = !fp32 && suffix_run(share);
  if (fp32 || joins64) {
    // Lost pairs by their read's own rows per lane, a wave each.
    for (int R = 1; R <= kMaxRowsPerLane && (mode == kModeFp32Rescue || joins64); ++R) {
      if (!rows_pairs[R - 1]) continue;
      SweepArgs a = base_args();
      a.dense = pd || joins64 ? nullptr : dense_flags;  // pairs of dense reads are done
      a.items = rescue_pairs[R - 1];
      a.n_items = 0;
      a.n_items_dev = counters + 16 + (R - 1);
      a.counter = counters + 8 + (R - 1);
      a.pair_mode = 1;
      a.share = 0;
      a.scratch = nullptr;
      // The grid is sized for the worst case; waves beyond the rescued count exit at once.
      if (int rc = launch<double>(R, a, rows_pairs[R - 1])) return rc;
    }
    if (mode == kModeFp32Rescue)
      if (int rc = enqueue_long_rescue()) return rc;
  }
  return kOk;
}

DeviceBatch::DeviceBatch() : impl_(new Impl) {}
DeviceBatch::~DeviceBatch() { delete impl_; }

int DeviceBatch::create(const BatchInput& in, std::unique_ptr<DeviceBatch>& out) {
  std::unique_ptr<DeviceBatch> b(new DeviceBatch);
  if (int rc = b->assign(in)) return rc;
  out = std::move(b);
  return kOk;
}

// Pair-mode items of the reads whose own rows per lane is `R` (any, when `R` is 0), for a launch
// at `run_rows` rows per lane: (read, sorted haplotype k) for every haplotype of the read's region, or, for two
// reads a < b of one region whose lanes fit one wave together at run_rows, one item per
// haplotype carrying both as (a, k | (b - a) << 16 | sign bit).  Such an item is the main pass's
// two-read wave on one haplotype, and a read gets the same bits whatever shares its wave.  Reads
// pair longest with shortest among those that fit.
void pair_items(const BatchInput& in, const HostPlan& p, int R, int run_rows,
                std::vector<int32_t>& all) {
  const int64_t NR = int64_t(p.read_rows.size());
  for (int64_t g = 0; g < int64_t(p.region_nhap.size()); ++g) {
    const int32_t H = p.region_nhap[size_t(g)];
    std::vector<std::pair<int, int64_t>> reads;  // (lanes at run_rows, read)
    for (int64_t r = in.region_read[g]; r < in.region_read[g + 1] && r < NR; ++r) {
      if (R ? p.read_rows[size_t(r)] != R : p.read_rows[size_t(r)] < 1) continue;
      const int n = int(in.read_off[r + 1] - in.read_off[r]);
      reads.push_back({(n + run_rows) / run_rows, r});
    }
    std::sort(reads.begin(), reads.end(), [](const auto& x, const auto& y) {
      return x.first != y.first ? x.first > y.first : x.second < y.second;
    });
    size_t i = 0, j = reads.size();
    while (i < j) {
      int64_t a = reads[i].second, b = -1;
      if (j - i >= 2 && reads[i].first + reads[j - 1].first <= 64 && H <= 0x10000) {
        b = reads[j - 1].second;
        --j;
        if (b < a) std::swap(a, b);
        if (b - a >= 0x8000) {  // too far apart to encode: both alone
          ++j;
          b = -1;
          a = reads[i].second;
        }
      }
      ++i;
      for (int32_t k = 0; k < H; ++k) {
        all.push_back(int32_t(a));
        all.push_back(b < 0 ? k : int32_t(uint32_t(k) | uint32_t(b - a) << 16 | 0x80000000u));
      }
    }
  }
}

int DeviceBatch::assign(const BatchInput& in) {
  Impl& d = *impl_;
  // Work queued with the previous inputs finishes before their memory is reused.
  if (d.stream) HIP_TRY(hipStreamSynchronize(d.stream));
  d.out_on_host = false;
  d.drop_graphs();
  d.free_lazy();
  // A previous assign that failed validation may have left arrays placed but not committed.
  d.pieces.clear();
  d.layout_bytes = 0;
  d.plan = HostPlan{};
  if (int rc = build_plan(in, d.plan)) return rc;
  if (int rc = device_state(&d.st)) return rc;
  d.pd = in.pd;
  d.pairs_only = in.pairs_only;
  if (!d.stream) {
    HIP_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    HIP_TRY(hipStreamCreateWithFlags(&d.side_stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&d.fork_ev, hipEventDisableTiming));
    HIP_TRY(hipEventCreateWithFlags(&d.join_ev, hipEventDisableTiming));
    for (auto& e : d.ev) HIP_TRY(hipEventCreate(&e));
    for (auto& e : d.bev) HIP_TRY(hipEventCreate(&e));
  }
  const HostPlan& p = d.plan;
  const int64_t NR = p.n_reads, NH = p.n_haps, G = int64_t(p.region_hap0.size());
  const size_t rb = size_t(in.read_off[NR]);
  const size_t hb = size_t(in.hap_off[NH]);
  d.place(&d.bases, in.bases, rb);
  d.place(&d.quals, in.quals, rb);
  d.place(&d.ins, in.ins, rb);
  d.place(&d.del, in.del, rb);
  d.place(&d.gcp, in.gcp, rb);
  d.place(&d.read_off, in.read_off, size_t(NR + 1));
  d.place(&d.read_region, p.read_region.data(), size_t(NR));
  d.place(&d.read_local, p.read_local.data(), size_t(NR));
  d.place(&d.region_out, p.region_out.data(), size_t(G + 1));
  d.place(&d.region_hap0, p.region_hap0.data(), size_t(G));
  d.place(&d.region_nhap, p.region_nhap.data(), size_t(G));
  d.place(&d.hplan, p.plan.data(), size_t(NH));
  d.place(&d.snaps, p.snaps.data(), p.snaps.size());
  d.place(&d.bsweeps, p.bsweeps.data(), p.bsweeps.size());
  d.place(&d.bwrites, p.bwrites.data(), p.bwrites.size());
  d.place(&d.region_bw, p.region_bw.data(), p.region_bw.size());
  d.hap_words = nullptr;
  d.hap_bytes = nullptr;
  d.region_classes = nullptr;
  if (in.pd) {
    d.place(&d.hap_words, p.words.data(), p.words.size());
    d.place(&d.region_classes, p.region_classes.data(), p.region_classes.size());
  } else
    d.place(&d.hap_bytes, in.hap_bases, hb);
  // Reads whose insertion, deletion and gap-continuation qualities are each the same at every
  // base, with a gap-continuation quality of at least kXsMinGcp: the reads the sweep kernel's own
  // test (rows_match_first) lets scale their gap values.
  // When every read is such a read (`all_uniform`), every launch of the batch, pair runs and
  // rescues included, may take the uniform-gap build.
  std::vector<uint8_t> uniform(size_t(NR), 0);
  d.all_uniform = true;
  for (int64_t r = 0; r < NR; ++r) {
    const int64_t o0 = in.read_off[r], o1 = in.read_off[r + 1];
    bool u = o1 > o0 && in.gcp[o0] >= kXsMinGcp;
    for (int64_t i = o0 + 1; i < o1 && u; ++i)
      u = in.ins[i] == in.ins[o0] && in.del[i] == in.del[o0] && in.gcp[i] == in.gcp[o0];
    uniform[size_t(r)] = u;
    d.all_uniform = d.all_uniform && u;
  }
  if (in.pairs_only) uniform.clear();
  for (int R = 1; R <= kMaxRowsPerLane; ++R) {
    auto& items = d.plan.bucket_items[R - 1];
    d.bucket_n[R - 1] = int32_t(items.size() / 2);
    d.bucket_ug[R - 1] = 0;
    if (!uniform.empty()) {  // uniform items first, each part in its cost order
      std::vector<int32_t> ug, rest;
      for (size_t i = 0; i < items.size(); i += 2) {
        const int32_t ra = items[i], rb = items[i + 1];
        auto& to = uniform[size_t(ra)] && (rb < 0 || uniform[size_t(rb)]) ? ug : rest;
        to.push_back(ra);
        to.push_back(rb);
      }
      d.bucket_ug[R - 1] = int32_t(ug.size() / 2);
      ug.insert(ug.end(), rest.begin(), rest.end());
      items.swap(ug);
    }
    const int64_t pairs = p.bucket_pairs[R - 1];
    if (pairs > kMaxQueueItems) {  // a work queue advances by 64 per claim in an int32 counter
      set_error(kInvalidArgument, "a batch holds more than " + std::to_string(kMaxQueueItems) +
                                      " pairs of one read length class; split it");
      return kInvalidArgument;
    }
    d.bucket_pairs[R - 1] = pairs;
    d.bucket_cells[R - 1] = p.bucket_cells[R - 1];
    d.bucket_work[R - 1] = p.bucket_work[R - 1];
    d.place(&d.bucket_items[R - 1], items.data(), items.size());
  }
  // Pairs of the reads of each rows per lane: the capacity of the rescue lists.
  std::fill(std::begin(d.rows_pairs), std::end(d.rows_pairs), 0);
  if (!in.pairs_only)  // pair runs keep their own rescue list
    for (int64_t r = 0; r < NR; ++r)
      if (p.read_rows[size_t(r)] > 0)  // long-kernel reads have their own
        d.rows_pairs[p.read_rows[size_t(r)] - 1] += p.region_nhap[size_t(p.read_region[size_t(r)])];
  for (int R = 1; R <= kMaxRowsPerLane; ++R)
    d.place(&d.rescue_pairs[R - 1], static_cast<const int32_t*>(nullptr),
            size_t(2 * d.rows_pairs[R - 1]));
  // The results right after the counters, so a pair run reads back its rescued counts and its
  // results with one copy.
  d.place(&d.counters, static_cast<const int32_t*>(nullptr), kCounters);
  d.place(&d.out, static_cast<const double*>(nullptr), size_t(p.n_pairs));
  d.place(&d.dense_flags, static_cast<const uint8_t*>(nullptr), size_t(NR));
  std::fill(std::begin(d.rows_reads), std::end(d.rows_reads), 0);
  if (!in.pd && !in.pairs_only)
    for (int64_t r = 0; r < NR; ++r)
      if (p.read_rows[size_t(r)] > 0) d.rows_reads[p.read_rows[size_t(r)] - 1] += 1;
  for (int R = 1; R <= kMaxRowsPerLane; ++R)
    d.place(&d.dense_items[R - 1], static_cast<const int32_t*>(nullptr),
            size_t(2 * d.rows_reads[R - 1]));
  d.n_long = int64_t(p.long_pairs.size() / 2);
  d.place(&d.long_pairs, p.long_pairs.data(), p.long_pairs.size());
  d.place(&d.long_rescue, static_cast<const int32_t*>(nullptr), p.long_pairs.size());
  std::vector<int32_t> all;
  d.n_all_pairs = d.n_all_items = 0;
  d.all_rows = 0;
  d.all_pairs = d.all_rescue = nullptr;
  std::fill(std::begin(d.all_off), std::end(d.all_off), 0);
  if (p.n_pairs <= kMaxPairList) {
    int64_t n_pairs = 0;
    for (int64_t r = 0; r < NR; ++r)
      if (p.read_rows[size_t(r)] > 0) {
        d.all_rows = std::max(d.all_rows, int(p.read_rows[size_t(r)]));
        n_pairs += p.region_nhap[size_t(p.read_region[size_t(r)])];
      }
    d.n_all_pairs = n_pairs;
    // One launch runs every item at the largest rows per lane (run_latency); otherwise each
    // bucket runs at its own.
    const bool one_launch = n_pairs <= kOneLaunchPairs;
    all.reserve(size_t(2 * n_pairs));
    if (one_launch) {  // every item at the largest rows per lane: reads pair across buckets
      pair_items(in, p, 0, d.all_rows, all);
      for (int R = 2; R <= kMaxRowsPerLane; ++R) d.all_off[R - 1] = int64_t(all.size() / 2);
    } else {
      for (int R = 1; R <= kMaxRowsPerLane; ++R) {
        d.all_off[R - 1] = int64_t(all.size() / 2);
        pair_items(in, p, R, R, all);
      }
    }
    d.n_all_items = int64_t(all.size() / 2);
    d.all_off[kMaxRowsPerLane] = d.n_all_items;
    d.place(&d.all_pairs, all.data(), all.size());
    d.place(&d.all_rescue, static_cast<const int32_t*>(nullptr), size_t(2 * n_pairs));
  }
  if (p.max_slots > kMaxSlots || p.max_nodes > kMaxSlots) {
    set_error(kInvalidArgument, "a region holds more than " + std::to_string(kMaxSlots) +
                                    " distinct haplotype prefix columns or shared suffixes");
    return kInvalidArgument;
  }
  d.stage_hap_words = p.max_hap_len + 1 + 64 + 8;
  d.join_lds = !p.bsweeps.empty();
  if (int64_t(d.stage_words(kMaxRowsPerLane)) * 4 * (256 / 64) > 64 * 1024) {  // 4 waves a block
    set_error(kInvalidArgument, "haplotype of " + std::to_string(p.max_hap_len) +
                                    " bases exceeds the engine's LDS staging capacity");
    return kInvalidArgument;
  }
  if (int rc = d.commit()) return rc;
  {
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    for (int R = 1; R <= kMaxRowsPerLane; ++R) d.cfg[R - 1] = d.st->tuned[d.pd][R - 1];
  }
  // Side buckets: those computing under 2% of the batch's cells, which cannot fill the GPU on
  // their own (short-read and longest-read buckets of a mixed batch).
  {
    int64_t total = 0;
    for (int R = 1; R <= kMaxRowsPerLane; ++R) total += d.bucket_work[R - 1];
    int n_big = 0;
    for (int R = 1; R <= kMaxRowsPerLane; ++R) {
      d.side[R - 1] = d.bucket_n[R - 1] > 0 && d.bucket_work[R - 1] * 50 < total;
      n_big += d.bucket_n[R - 1] > 0 && !d.side[R - 1];
    }
    if (n_big == 0) std::fill(std::begin(d.side), std::end(d.side), false);
  }
  // Scratch for the largest configuration any launch may use, so tuning never reallocates: the
  // main part for launches on the batch's stream, and after it the part the side buckets use
  // while the large ones run.
  int64_t need = 0, side_need = 0;
  if (p.max_slots > 0 || p.max_nodes > 0) {
    for (int R = 1; R <= kMaxRowsPerLane; ++R) {
      // The dense FP64 rescue of the reads with R rows per lane.
      if (d.rows_reads[R - 1])
        for (bool ug : {false, true})  // the general and the uniform-gap build's grids
          need = std::max(need, d.scratch_need<double>(R, LaunchCfg{0, 256, false, ug},
                                                       d.rows_reads[R - 1]));
      if (!d.bucket_n[R - 1]) continue;
      int64_t n = 0;
      for (LaunchCfg c : d.candidates(R, true)) {
        n = std::max(n, d.scratch_need<float>(R, c));
        c.ug = true;  // the uniform-gap build's grid
        n = std::max(n, d.scratch_need<float>(R, c));
      }
      for (bool ug : {false, true}) {  // FP64 launches, general and uniform-gap builds
        n = std::max(n, d.scratch_need<double>(R, LaunchCfg{0, 256, false, ug}));
        if (d.suffix_run(1)) n = std::max(n, d.scratch_need<double>(R, LaunchCfg{0, 256, true, ug}));
      }
      need = std::max(need, n);  // tuning runs every bucket on the batch's stream
      if (d.side[R - 1]) side_need = std::max(side_need, n);
    }
  }
  if (need < 0 || side_need < 0) return kInternal;
  d.side_scratch_off = (size_t(need) + 255) & ~size_t(255);
  if (side_need > 0) need = int64_t(d.side_scratch_off) + side_need;
  if (size_t(need) > d.scratch_cap) {
    if (d.scratch) HIP_TRY(hipFree(d.scratch));
    d.scratch = nullptr;
    d.scratch_cap = 0;
    HIP_TRY(hipMalloc(&d.scratch, size_t(need)));
    d.scratch_cap = size_t(need);
  }
  d.scratch_bytes = size_t(need);
  // Boundary-row scratch of long_kernel, for the larger of its FP32 and FP64 grids.
  d.long_grid = 0;
  if (d.n_long) {
    int wf = 0, wd = 0;
    if (int rc = d.long_resident<float>(&wf)) return rc;
    if (int rc = d.long_resident<double>(&wd)) return rc;
    d.long_grid = int(std::min<int64_t>(std::max(wf, wd), d.n_long));
    const size_t bytes = size_t(d.long_grid) * size_t(d.long_stride()) * sizeof(double);
    if (bytes > d.long_scratch_cap) {
      if (d.long_scratch) HIP_TRY(hipFree(d.long_scratch));
      d.long_scratch = nullptr;
      d.long_scratch_cap = 0;
      HIP_TRY(hipMalloc(&d.long_scratch, bytes));
      d.long_scratch_cap = bytes;
    }
  }
  return kOk;
}

int DeviceBatch::run(int mode, int share, int use_graph, RunStats* stats) {
  Impl& d = *impl_;
  d.out_on_host = false;
  if (mode != kModeFp32Rescue && mode != kModeFp64 && mode != kModeFp64PerRegion &&
      mode != kModeFp32NoRescue) {
    set_error(kInvalidArgument, "unknown run mode");
    return kInvalidArgument;
  }
  if (d.plan.n_pairs == 0) {
    if (stats) *stats = RunStats{};
    return kOk;
  }
  if (d.pairs_only) {
    set_error(kInvalidArgument, "this batch was planned for one-pair-per-wave runs only");
    return kInvalidArgument;
  }
  if (mode == kModeFp64PerRegion) return run_per_region(stats);
  HIP_TRY(hipEventRecord(d.ev[0], d.stream));
  if (use_graph) {
    hipGraphExec_t& ge = d.graph[mode * 2 + (share ? 1 : 0)];
    if (!ge) {
      hipGraph_t graph = nullptr;
      HIP_TRY(hipStreamBeginCapture(d.stream, hipStreamCaptureModeThreadLocal));
      const int rc = d.enqueue(mode, share, false);
      hipGraph_t captured = nullptr;
      const hipError_t e = hipStreamEndCapture(d.stream, &captured);
      if (rc) return rc;
      HIP_TRY(e);
      graph = captured;
      HIP_TRY(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0));
      HIP_TRY(hipGraphDestroy(graph));
    }
    HIP_TRY(hipGraphLaunch(ge, d.stream));
  } else {
    if (int rc = d.enqueue(mode, share, true)) return rc;
  }
  HIP_TRY(hipEventRecord(d.ev[2], d.stream));
  HIP_TRY(hipStreamSynchronize(d.stream));
  if (stats) {
    *stats = RunStats{};
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, d.ev[0], d.ev[2]));
    stats->total_ms = total;
    int32_t counts[kCounters];
    HIP_TRY(hipMemcpy(counts, d.counters, sizeof(counts), hipMemcpyDeviceToHost));
    int64_t rescued = 0;
    if (mode == kModeFp32Rescue || mode == kModeFp32NoRescue) {
      for (int R = 0; R < kMaxRowsPerLane; ++R) rescued += counts[16 + R];
      rescued += counts[42];
    } else if (d.suffix_run(share)) {  // FP64 joins rerun unshared
      for (int R = 0; R < kMaxRowsPerLane; ++R) rescued += counts[16 + R];
    }
    stats->rescued_pairs = rescued;
    for (int R = 0; R < kMaxRowsPerLane; ++R) {
      stats->bucket_cells[R] = d.bucket_cells[R];
      stats->bucket_work[R] = d.bucket_work[R];
    }
    if (!use_graph) {
      float first = 0.f;
      HIP_TRY(hipEventElapsedTime(&first, d.ev[0], d.ev[1]));
      stats->primary_ms = first;
      stats->rescue_ms = total - first;
      if (mode == kModeFp32Rescue) {
        float dense = 0.f;
        HIP_TRY(hipEventElapsedTime(&dense, d.ev[1], d.ev[3]));
        stats->dense_rescue_ms = dense;
      }
      for (int R = 1; R <= kMaxRowsPerLane; ++R) {
        if (!d.bucket_n[R - 1]) continue;
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, d.bev[2 * (R - 1)], d.bev[2 * (R - 1) + 1]));
        stats->bucket_ms[R - 1] = ms;
      }
    } else {
      stats->primary_ms = -1.0;
      stats->rescue_ms = -1.0;
    }
  }
  return kOk;
}

int DeviceBatch::run_per_region(RunStats* stats) {
  Impl& d = *impl_;
  const HostPlan& p = d.plan;
  const int64_t G = int64_t(p.region_hap0.size());
  if (!d.region_pairs) {
    // Pair lists per region, grouped by rows per lane: the work one per-region call launches.
    std::vector<int32_t> pairs;
    pairs.reserve(size_t(2 * p.n_pairs));
    d.region_pair_off.assign(size_t(G * kMaxRowsPerLane + 1), 0);
    // Reads of a region in index order; regions are contiguous in read index.
    int64_t r = 0;
    for (int64_t g = 0; g < G; ++g) {
      const int32_t H = p.region_nhap[size_t(g)];
      int64_t r_end = r;
      while (r_end < p.n_reads && p.read_region[size_t(r_end)] == g) ++r_end;
      for (int R = 1; R <= kMaxRowsPerLane; ++R) {
        d.region_pair_off[size_t(g * kMaxRowsPerLane + R - 1)] = int64_t(pairs.size() / 2);
        for (int64_t rr = r; rr < r_end; ++rr) {
          if (p.read_rows[size_t(rr)] != R) continue;
          for (int32_t k = 0; k < H; ++k) {
            pairs.push_back(int32_t(rr));
            pairs.push_back(k);
          }
        }
      }
      r = r_end;
    }
    d.region_pair_off[size_t(G * kMaxRowsPerLane)] = int64_t(pairs.size() / 2);
    if (int rc = d.upload(&d.region_pairs, pairs.data(), pairs.size())) return rc;
    // One work-queue head per (region, rows per lane) launch, then one per region for its
    // long-kernel pairs.
    d.n_region_launches = G * (kMaxRowsPerLane + 1);
    if (int rc = d.alloc(&d.region_counters, size_t(d.n_region_launches))) return rc;
  }
  const int n = int(d.n_region_launches);
  hipLaunchKernelGGL(fill_i32, dim3((n + 255) / 256), dim3(256), 0, d.stream, d.region_counters, n, 0);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(d.stream));
  HIP_TRY(hipEventRecord(d.ev[0], d.stream));
  for (int64_t g = 0; g < G; ++g) {
    for (int R = 1; R <= kMaxRowsPerLane; ++R) {
      const int64_t i = g * kMaxRowsPerLane + R - 1;
      const int64_t lo = d.region_pair_off[size_t(i)], hi = d.region_pair_off[size_t(i + 1)];
      if (hi == lo) continue;
      SweepArgs a = d.base_args();
      a.items = d.region_pairs + 2 * lo;
      a.n_items = int32_t(hi - lo);
      a.counter = d.region_counters + i;
      a.pair_mode = 1;
      a.scratch = nullptr;
      if (int rc = d.launch<double>(R, a, hi - lo)) return rc;
    }
    const int64_t llo = p.region_long[size_t(g)], lhi = p.region_long[size_t(g + 1)];
    if (lhi > llo) {
      SweepArgs a = d.base_args();
      a.items = d.long_pairs + 2 * llo;
      a.n_items = int32_t(lhi - llo);
      a.counter = d.region_counters + G * kMaxRowsPerLane + g;
      if (int rc = d.launch_long<double>(a, lhi - llo)) return rc;
    }
    // One synchronous call per region, as HaplotypeCaller makes them.
    HIP_TRY(hipStreamSynchronize(d.stream));
  }
  HIP_TRY(hipEventRecord(d.ev[2], d.stream));
  HIP_TRY(hipStreamSynchronize(d.stream));
  if (stats) {
    *stats = RunStats{};
    float total = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, d.ev[0], d.ev[2]));
    stats->total_ms = total;
    stats->primary_ms = total;
  }
  return kOk;
}

int DeviceBatch::run_latency(int mode, RunStats* stats) {
  Impl& d = *impl_;
  d.out_on_host = false;
  d.last_rescued = 0;
  if (d.plan.n_pairs == 0) {
    if (stats) *stats = RunStats{};
    return kOk;
  }
  // Batches too large for the pair list run the batched way.
  if (!d.all_pairs) {
    RunStats st;
    const int rc = run(mode == kModeFp64PerRegion ? kModeFp64 : mode, 1, 0, &st);
    if (rc == kOk) {
      d.last_rescued = st.rescued_pairs;
      if (stats) *stats = st;
    }
    return rc;
  }
  if (mode != kModeFp32Rescue && mode != kModeFp64 && mode != kModeFp64PerRegion &&
      mode != kModeFp32NoRescue) {
    set_error(kInvalidArgument, "unknown run mode");
    return kInvalidArgument;
  }
  const bool fp32 = mode == kModeFp32Rescue || mode == kModeFp32NoRescue;
  const size_t out_bytes = sizeof(double) * size_t(d.plan.n_pairs);
  // One copy brings back counters 16.. (the rescued counts at 16 and 42) and the results, which
  // the arena places right after the counters.
  const size_t head = size_t(reinterpret_cast<const char*>(d.out) -
                             reinterpret_cast<const char*>(d.counters + 16));
  if (int rc = d.reserve_readback(head + out_bytes)) return rc;
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, d.stream, d.counters, kCounters, 0);
  HIP_TRY(hipGetLastError());
  if (stats) HIP_TRY(hipEventRecord(d.ev[0], d.stream));
  // A batch whose pairs leave the GPU partly idle runs them in one launch at the largest rows per
  // lane (launches on a stream run one after another, and each would add a pair's sweep to the
  // time to results); a larger one launches each rows-per-lane bucket at its own rows, each with
  // its own work-queue head, the FP32 launches sharing one rescue list.
  const bool one_launch = d.n_all_pairs <= kOneLaunchPairs;
  for (int R = 1; R <= kMaxRowsPerLane && d.all_rows > 0; ++R) {
    const int64_t first = one_launch ? 0 : d.all_off[R - 1];
    const int64_t n = one_launch ? (R == d.all_rows ? d.n_all_items : 0) : d.all_off[R] - first;
    if (!n) continue;
    SweepArgs a = d.base_args();
    a.items = d.all_pairs + 2 * first;
    a.n_items = int32_t(n);
    a.counter = d.counters + (R - 1);
    a.pair_mode = 1;
    a.share = 0;
    a.scratch = nullptr;
    if (fp32) {
      a.rescue = 1;
      a.rescue_pairs[0] = d.all_rescue;
      a.rescue_count = d.counters + 16;
      if (int rc = d.launch<float>(R, a, n, LaunchCfg{R, 256, false})) return rc;
    } else {
      if (int rc = d.launch<double>(R, a, n)) return rc;
    }
  }
  if (int rc = d.enqueue_long(fp32)) return rc;
  const int R = d.all_rows;
  if (stats) HIP_TRY(hipEventRecord(d.ev[1], d.stream));
  if (mode == kModeFp32Rescue && R > 0) {
    SweepArgs b = d.base_args();
    b.items = d.all_rescue;
    b.n_items = 0;
    b.n_items_dev = d.counters + 16;
    b.counter = d.counters + 8;
    b.pair_mode = 1;
    b.share = 0;
    b.scratch = nullptr;
    if (int rc = d.launch<double>(R, b, d.n_all_pairs)) return rc;
  }
  if (mode == kModeFp32Rescue)
    if (int rc = d.enqueue_long_rescue()) return rc;
  if (stats) HIP_TRY(hipEventRecord(d.ev[2], d.stream));
  auto* rb = static_cast<char*>(d.readback);
  HIP_TRY(hipMemcpyAsync(rb, d.counters + 16, head + out_bytes, hipMemcpyDeviceToHost, d.stream));
  HIP_TRY(hipStreamSynchronize(d.stream));
  int32_t rescued = 0;
  if (fp32) {  // the rescued counts of the staged pairs and of the long pairs
    int32_t c16 = 0, c42 = 0;
    std::memcpy(&c16, rb, sizeof(c16));
    std::memcpy(&c42, rb + (42 - 16) * sizeof(int32_t), sizeof(c42));
    rescued = c16 + c42;
  }
  d.last_rescued = rescued;
  d.out_on_host = true;
  d.out_rb_off = head;
  if (stats) {
    *stats = RunStats{};
    float total = 0.f, first = 0.f;
    HIP_TRY(hipEventElapsedTime(&total, d.ev[0], d.ev[2]));
    HIP_TRY(hipEventElapsedTime(&first, d.ev[0], d.ev[1]));
    stats->total_ms = total;
    stats->primary_ms = first;
    stats->rescue_ms = total - first;
    stats->rescued_pairs = rescued;
  }
  return kOk;
}

int DeviceBatch::tune(int share, TuneReport* report) {
  Impl& d = *impl_;
  d.out_on_host = false;
  const auto t0 = std::chrono::steady_clock::now();
  TuneReport rep;
  for (int R = 1; R <= kMaxRowsPerLane; ++R) {
    if (!d.bucket_n[R - 1]) continue;
    const SweepArgs a = d.primary_args(R, share, true);
    const std::vector<LaunchCfg> cands = d.candidates(R);
    double best_ms = 0;
    LaunchCfg best = cands[0];
    for (size_t i = 0; i < cands.size(); ++i) {
      double ms = 1e30;
      for (int rep_i = 0; rep_i < 2; ++rep_i) {
        hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, d.stream, d.counters, kCounters, 0);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(d.ev[0], d.stream));
        if (int rc = d.launch_bucket<float>(R, a, d.bucket_n[R - 1], cands[i], d.stream, 0))
          return rc;
        HIP_TRY(hipEventRecord(d.ev[1], d.stream));
        HIP_TRY(hipStreamSynchronize(d.stream));
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, d.ev[0], d.ev[1]));
        ms = std::min(ms, double(t));
      }
      if (i == 0) {
        rep.default_ms[R - 1] = best_ms = ms;
      } else if (ms < 0.99 * best_ms) {  // a candidate must win by 1% to replace the default
        best_ms = ms;
        best = cands[i];
      }
    }
    rep.candidates += int(cands.size());
    d.cfg[R - 1] = best;
    rep.cfg[R - 1] = best;
    rep.tuned_ms[R - 1] = best_ms;
    std::lock_guard<std::mutex> lock(g_dev_mutex);
    d.st->tuned[d.pd][R - 1] = best;
  }
  d.drop_graphs();
  rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (report) *report = rep;
  return kOk;
}

int DeviceBatch::set_launch(int R, const LaunchCfg& cfg) {
  Impl& d = *impl_;
  if (R < 1 || R > kMaxRowsPerLane || !d.valid(R, cfg)) {
    set_error(kInvalidArgument, "launch configuration outside the tuned space: bucket rows " +
                                    std::to_string(R) + ", rows " + std::to_string(cfg.rows) +
                                    ", block " + std::to_string(cfg.block));
    return kInvalidArgument;
  }
  d.cfg[R - 1] = cfg;
  d.cfg[R - 1].rows = cfg.rows ? cfg.rows : R;
  d.drop_graphs();
  return kOk;
}

void DeviceBatch::set_uniform_gaps(bool on) {
  impl_->uniform_gaps = on;
  impl_->drop_graphs();
}

void DeviceBatch::set_lds_records(bool on) {
  impl_->lds_records = on;
  impl_->drop_graphs();
}

void DeviceBatch::snapshots(int64_t out[4]) const {
  out[0] = impl_->plan.snap_hbm;
  out[1] = impl_->plan.snap_lds;
  out[2] = impl_->plan.snap_restarts;
  out[3] = impl_->lds_launches;
}

int DeviceBatch::get_launch(int R, LaunchCfg* cfg) const {
  if (R < 1 || R > kMaxRowsPerLane) {
    set_error(kInvalidArgument, "bucket rows per lane must be 1..8");
    return kInvalidArgument;
  }
  *cfg = impl_->cfg[R - 1];
  if (!cfg->rows) cfg->rows = R;
  return kOk;
}

int DeviceBatch::results(double* out, int64_t n) {
  Impl& d = *impl_;
  if (n != d.plan.n_pairs) {
    set_error(kInvalidArgument, "output has " + std::to_string(n) +
                                    " elements but the batch holds " +
                                    std::to_string(d.plan.n_pairs) + " pairs");
    return kInvalidArgument;
  }
  if (!n) return kOk;
  const size_t bytes = sizeof(double) * size_t(n);
  if (!d.out_on_host) {
    if (bytes > kDirectUploadBytes / 4 && 8 + bytes > d.readback_cap) {
      // Large results and no pinned buffer that holds them: straight into the caller's array.
      HIP_TRY(hipMemcpyAsync(out, d.out, bytes, hipMemcpyDeviceToHost, d.stream));
      HIP_TRY(hipStreamSynchronize(d.stream));
      return kOk;
    }
    // Through pinned memory in the batch's stream, after the work queued on it.
    if (int rc = d.reserve_readback(8 + bytes)) return rc;
    HIP_TRY(hipMemcpyAsync(static_cast<char*>(d.readback) + 8, d.out, bytes,
                           hipMemcpyDeviceToHost, d.stream));
    HIP_TRY(hipStreamSynchronize(d.stream));
    d.out_rb_off = 8;
  }
  std::memcpy(out, static_cast<const char*>(d.readback) + d.out_rb_off, bytes);
  return kOk;
}

int64_t DeviceBatch::last_rescued() const { return impl_->last_rescued; }

int64_t DeviceBatch::n_pairs() const { return impl_->plan.n_pairs; }
int64_t DeviceBatch::scratch_bytes() const { return int64_t(impl_->scratch_bytes); }

void DeviceBatch::work(int64_t out[4]) const {
  const HostPlan& p = impl_->plan;
  out[0] = p.work_nominal;
  out[1] = p.work_prefix;
  out[2] = p.work_forward;
  out[3] = p.work_backward;
}

int dpp_selftest(int32_t* out64) {
  DeviceState* st = nullptr;
  if (int rc = device_state(&st)) return rc;
  int32_t in[64];
  for (int i = 0; i < 64; ++i) in[i] = 1000 + i;
  int32_t *din = nullptr, *dout = nullptr;
  HIP_TRY(hipMalloc(&din, sizeof(in)));
  HIP_TRY(hipMalloc(&dout, sizeof(in)));
  HIP_TRY(hipMemcpy(din, in, sizeof(in), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(dpp_selftest_kernel, dim3(1), dim3(64), 0, nullptr, din, dout);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipMemcpy(out64, dout, sizeof(in), hipMemcpyDeviceToHost));
  HIP_TRY(hipFree(din));
  HIP_TRY(hipFree(dout));
  return kOk;
}

int device_available() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) return 0;
  return count;
}

}  // namespace fgkl

