// Smith-Waterman on gfx950: GATK's SmithWatermanJavaAligner (as fgkl restates it in
// crates/smithwaterman/src/reference.rs:10-203 and serves it in lib.rs:330-360), many pairs per
// launch.
//
// One wave aligns one pair.  The reference runs down the lanes, R consecutive rows per lane, in
// strips of 64 R rows; the alternate is swept across them one column per step, lane l working on
// column t - l + 1 at step t, so a step is one anti-diagonal of lane slabs.  A lane receives the
// score H and the vertical gap V of the row above its slab from the lane above through DPP
// (wave_shr:1); lane 0 reads them from the wave's LDS row, which holds DP row 0 for the first
// strip and the previous strip's last row after it.  Per cell the lane keeps GATK's recurrences:
//   diag = H[i-1][j-1] + (ref[i] == alt[j] ? match : mismatch)
//   V    = max(H[i-1][j] + open, V[i-1][j] + extend)   (best_gap_v, reference.rs:59-67)
//   Z    = max(H[i][j-1] + open, Z[i][j-1] + extend)   (best_gap_h, reference.rs:68-76)
//   H    = max(diag, V, Z) floored at -1e8, ties to diag, then Z, then V (reference.rs:77-88)
// and writes four trace bits: not-diagonal, insertion (Z >= V), V extends the gap above and Z
// extends the gap to the left (GATK keeps a gap length per cell; walking the extension bits back
// recovers it).  A step's R cells fill one 32-bit word per lane; words go to the wave's trace
// region step-major, 256 bytes per step.
//
// The end cell follows calculateCigar (reference.rs:102-160): the last row of the best score in
// the last column (lanes freeze at their last column, so the column is in registers when the
// strip ends), bettered by a last-row cell of higher score or of equal score nearer the diagonal
// (tracked per step in the last strip).  The traceback walks from there through the trace,
// paging 8 lanes x 64 steps of it into LDS at a time, and writes the CIGAR from its last element
// backwards.  Exact substrings short-circuit for SOFTCLIP and IGNORE (lib.rs:339-347).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fgkl_internal.hpp"
#include "fgkl_sw.hpp"
#include "wave64.hpp"

namespace fgkl {
namespace {

constexpr int32_t kCutoff = -100000000;      // MATRIX_MIN_CUTOFF (lib.rs:20)
constexpr int32_t kLowInit = INT32_MIN / 2;  // lowInitValue (lib.rs:22)
constexpr int kSwMaxRows = 8;                // rows per lane: one trace word per lane per step
constexpr int kWinLanes = 8, kWinSteps = 64, kWinWords = kWinLanes * kWinSteps;
constexpr int kMaxAlt = 16384;               // alternate bases the LDS row holds at one wave/block
constexpr uint32_t kOpM = 0, kOpI = 1, kOpD = 2, kOpS = 4;  // BAM operation codes

struct SwArgs {
  const uint8_t* ref;
  const int64_t* ref_off;
  const uint8_t* alt;
  const int64_t* alt_off;
  const int32_t* items;      // pair indices of this launch, costliest first
  int n_items;
  int* counter;              // work queue head (advances by 64 per claim)
  int* substring_pairs;      // pairs resolved by the exact-substring shortcut
  int32_t wm, wx, wo, we;    // match, mismatch, gap open, gap extend
  int strategy;
  uint32_t* trace;           // per wave: trace_stride words
  int64_t trace_stride;
  int lds_words;             // per wave: LDS row, alternate bases, traceback window
  int row_words;             // LDS words before the alternate bases (2 per column, 0..max_alt)
  int alt_words;
  uint32_t* ops;             // CIGAR elements of pair p end at ops[ops_end[p] - 1]
  const int64_t* ops_end;
  int32_t* n_ops;
  int32_t* offset;
};

// Keeps a DPP result in a register before any branch or select uses it, so the compiler cannot
// fold the DPP into an instruction that runs under a partial exec mask.
__device__ __forceinline__ int pinned(int v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <int R>
struct SwLane {
  int a[R];    // reference base of each row (0x100 below the last row: never equal to a base)
  int hl[R];   // H of each row at the lane's previous column
  int hz[R];   // best horizontal gap of each row at the previous column
  int hd;      // H of the row above the slab at the previous column
  int hb, vb;  // H and V of the slab's bottom row at its current column (read by the lane below)
  int bb;      // the alternate base of the current column
  int lb, ld, lj;  // last-row end candidate: score, distance from the diagonal, column
};

struct SwParams {
  int wm, wx, wo, we;
};

// One step of a strip.  FREEZE: some lanes are outside columns 1..m and keep their state (the
// ramp and the tail); TRACK: the last DP row is in this strip; WRITE: later strips read this
// strip's last row from the LDS row.
template <int R, bool FREEZE, bool TRACK, bool WRITE>
__device__ __forceinline__ void sw_step(SwLane<R>& s, const SwParams& w, int t, int lane, int m,
                                        int n, int rn, int2 top, int abase, int2* row,
                                        uint32_t* trace_step) {
  s.bb = pinned(shr1i(abase, s.bb));
  const int hu = pinned(shr1i(top.x, s.hb));
  const int vu = pinned(shr1i(top.y, s.vb));
  const int j = t - lane + 1;
  const bool act = !FREEZE || unsigned(t - lane) < unsigned(m);
  uint32_t word = 0;
  int up = hu, vv = vu, dg = s.hd;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int d = dg + (s.a[r] == s.bb ? w.wm : w.wx);
    const int to = up + w.wo, te = vv + w.we;
    const int v = max(to, te);
    const int ho = s.hl[r] + w.wo, he = s.hz[r] + w.we;
    const int z = max(ho, he);
    const int mx = max(v, z);
    const int h = max(max(d, mx), kCutoff);
    const uint32_t bits = uint32_t(d < mx) | (uint32_t(z >= v) << 1) | (uint32_t(to <= te) << 2) |
                          (uint32_t(ho <= he) << 3);
    word |= bits << (4 * r);
    dg = s.hl[r];
    if (FREEZE) {
      s.hl[r] = act ? h : s.hl[r];
      s.hz[r] = act ? z : s.hz[r];
    } else {
      s.hl[r] = h;
      s.hz[r] = z;
    }
    up = h;
    vv = v;
  }
  if (FREEZE) {
    s.hb = act ? up : s.hb;
    s.vb = act ? vv : s.vb;
    s.hd = act ? hu : s.hd;
  } else {
    s.hb = up;
    s.vb = vv;
    s.hd = hu;
  }
  trace_step[lane] = word;
  if (WRITE) {
    if (lane == 63 && j >= 1 && j <= m) row[j] = make_int2(up, vv);
  }
  if (TRACK) {
    int cand = s.hl[0];
#pragma unroll
    for (int r = 1; r < R; ++r) cand = rn == r ? s.hl[r] : cand;
    const int dist = abs(n - j);
    const bool better = act && (cand > s.lb || (cand == s.lb && dist < s.ld));
    s.lb = better ? cand : s.lb;
    s.ld = better ? dist : s.ld;
    s.lj = better ? j : s.lj;
  }
}

// Steps [t0, t1) of a strip, reading lane 0's inputs one step ahead.
template <int R, bool FREEZE, bool TRACK, bool WRITE>
__device__ __forceinline__ void sw_steps(SwLane<R>& s, const SwParams& w, int t0, int t1, int lane,
                                         int m, int n, int rn, int2* row, const uint8_t* lalt,
                                         uint32_t* trace) {
  if (t0 >= t1) return;
  int tc = min(t0, m - 1);
  int2 top = row[tc + 1];
  int ab = lalt[tc];
  for (int t = t0; t < t1; ++t) {
    const int tn = min(t + 1, m - 1);
    const int2 top_n = row[tn + 1];
    const int ab_n = lalt[tn];
    sw_step<R, FREEZE, TRACK, WRITE>(s, w, t, lane, m, n, rn, top, ab, row, trace + int64_t(t) * 64);
    top = top_n;
    ab = ab_n;
  }
}

// One strip: rows row0 + 1 .. row0 + 64 R of the pair.
template <int R, bool TRACK, bool WRITE>
__device__ void sw_strip(SwLane<R>& s, const SwParams& w, int lane, int m, int n, int rn,
                         int2* row, const uint8_t* lalt, uint32_t* trace) {
  const int ramp_end = 63;
  const int steady_end = max(m, 63);
  sw_steps<R, true, TRACK, WRITE>(s, w, 0, ramp_end, lane, m, n, rn, row, lalt, trace);
  sw_steps<R, false, TRACK, WRITE>(s, w, ramp_end, steady_end, lane, m, n, rn, row, lalt, trace);
  sw_steps<R, true, TRACK, WRITE>(s, w, steady_end, m + 63, lane, m, n, rn, row, lalt, trace);
}

// H of DP row i at column 0 (and of row 0 at column i): gap penalties for the indel strategies,
// zero otherwise (reference.rs:40-53).
__device__ __forceinline__ int border(bool gaps, int i, const SwParams& w) {
  return gaps && i > 0 ? w.wo + (i - 1) * w.we : 0;
}

template <int R, bool LASTROW>
__global__ void __launch_bounds__(256) sw_kernel(SwArgs a) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_all[];
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int64_t wave = int64_t(blockIdx.x) * (blockDim.x >> 6) + wib;
  uint32_t* lw = lds_all + wib * a.lds_words;
  int2* row = reinterpret_cast<int2*>(lw);
  uint8_t* lalt = reinterpret_cast<uint8_t*>(lw + a.row_words);
  uint32_t* win = lw + a.row_words + a.alt_words;
  uint32_t* trace = a.trace + wave * a.trace_stride;
  const SwParams w{a.wm, a.wx, a.wo, a.we};
  const bool gaps = a.strategy == kSwIndel || a.strategy == kSwLeadingIndel;
  constexpr int S = 64 * R;

  for (;;) {
    // Every lane adds one: a claim advances the counter by 64.
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= a.n_items) break;
    const int p = __builtin_amdgcn_readfirstlane(a.items[item]);
    const int64_t ro = a.ref_off[p], ao = a.alt_off[p];
    const int n = __builtin_amdgcn_readfirstlane(int(a.ref_off[p + 1] - ro));
    const int m = __builtin_amdgcn_readfirstlane(int(a.alt_off[p + 1] - ao));
    const uint8_t* ref = a.ref + ro;
    const uint8_t* alt = a.alt + ao;
    const int64_t oend = a.ops_end[p];

    // The exact-substring shortcut: the last occurrence of the alternate in the reference
    // (lib.rs:339-347, Utils.lastIndexOf), 64 candidate starts at a time from the right.
    if ((a.strategy == kSwSoftClip || a.strategy == kSwIgnore) && m <= n) {
      int found = -1;
      for (int base = n - m; base >= 0; base -= 64) {
        const int st = base - lane;
        bool hit = st >= 0;
        for (int k = 0; hit && k < m; ++k) hit = ref[st + k] == alt[k];
        const uint64_t mask = __ballot(hit);
        if (mask) {
          found = base - __builtin_ctzll(mask);
          break;
        }
      }
      if (found >= 0) {
        if (lane == 0) {
          a.ops[oend - 1] = (uint32_t(m) << 4) | kOpM;
          a.n_ops[p] = 1;
          a.offset[p] = found;
          atomicAdd(a.substring_pairs, 1);
        }
        continue;
      }
    }

    // Stage the alternate and DP row 0 in LDS.
    wave_sync_lds();
    for (int k = lane; k < m; k += 64) lalt[k] = alt[k];
    for (int k = lane; k <= m; k += 64) row[k] = make_int2(border(gaps, k, w), kLowInit);
    wave_sync_lds();

    const int nstrips = (n + S - 1) / S;
    int cb = INT_MIN, ci = 0;  // best score of the last column and its (last) row, this lane
    SwLane<R> s;
    s.lb = INT_MIN;
    s.ld = INT_MAX;
    s.lj = 0;
    int rn = 0, ln = 0;
    for (int strip = 0; strip < nstrips; ++strip) {
      const int row0 = strip * S;
      const int i0 = row0 + lane * R + 1;  // this lane's first DP row
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        s.a[r] = i <= n ? int(ref[i - 1]) : 0x100;
        s.hl[r] = border(gaps, i, w);
        s.hz[r] = kLowInit;
      }
      s.hd = border(gaps, i0 - 1, w);
      s.hb = 0;
      s.vb = 0;
      s.bb = 0;
      uint32_t* tr = trace + int64_t(strip) * (m + 63) * 64;
      const bool last = strip == nstrips - 1;
      if (last) {
        rn = (n - 1 - row0) % R;
        ln = (n - 1 - row0) / R;
        sw_strip<R, LASTROW, false>(s, w, lane, m, n, rn, row, lalt, tr);
      } else {
        sw_strip<R, false, true>(s, w, lane, m, n, rn, row, lalt, tr);
      }
      // Lanes froze at column m: their rows' last-column scores, later rows winning ties.
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int i = i0 + r;
        if (i <= n && s.hl[r] >= cb) {
          cb = s.hl[r];
          ci = i;
        }
      }
      wave_sync_lds();
    }
    // The trace is read back by other lanes of this wave.
    wave_sync_global();

    // The end of the alignment (reference.rs:107-141).
    int p1 = n, p2 = m, seg = 0;
    if (a.strategy != kSwIndel) {
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) {
        const int ob = __shfl_xor(cb, off), oi = __shfl_xor(ci, off);
        const bool take = ob > cb || (ob == cb && oi > ci);
        cb = take ? ob : cb;
        ci = take ? oi : ci;
      }
      p1 = __builtin_amdgcn_readfirstlane(ci);
      const int best = __builtin_amdgcn_readfirstlane(cb);
      if (LASTROW) {
        const int lb = __builtin_amdgcn_readlane(s.lb, ln);
        const int ld = __builtin_amdgcn_readlane(s.ld, ln);
        const int lj = __builtin_amdgcn_readlane(s.lj, ln);
        if (lb > best || (lb == best && ld < abs(p1 - p2))) {
          p1 = n;
          p2 = lj;
          seg = m - lj;
        }
      }
    }

    // The traceback (reference.rs:143-203), CIGAR elements written from the last backwards.
    // A CIGAR has at most n + m elements (each consumes a base); the slots hold n + m + 4.
    const int cap = n + m + 4;
    int nops = 0;
    auto push = [&](int len, uint32_t op) {
      if (lane == 0 && nops < cap) a.ops[oend - 1 - nops] = (uint32_t(len) << 4) | op;
      ++nops;
    };
    int ws = -1, wl = 0, wt = 0;  // the window: strip ws, lanes wl-7..wl, steps wt-63..wt
    auto bits = [&](int i, int j) -> uint32_t {
      const int st = (i - 1) / S;
      const int q = i - 1 - st * S;
      const int l = q / R;
      const int r = q - l * R;
      const int t = j - 1 + l;
      if (st != ws || l > wl || wl - l >= kWinLanes || t > wt || wt - t >= kWinSteps) {
        ws = st;
        wl = l;
        wt = t;
        wave_sync_lds();
        const uint32_t* tr = trace + int64_t(st) * (m + 63) * 64;
#pragma unroll
        for (int k = 0; k < kWinWords / 64; ++k) {
          const int kk = lane + 64 * k;
          const int ll = wl - (kk & (kWinLanes - 1));
          const int tt = wt - kk / kWinLanes;
          win[kk] = ll >= 0 && tt >= 0 ? tr[int64_t(tt) * 64 + ll] : 0u;
        }
        wave_sync_lds();
      }
      const uint32_t word = win[(wt - t) * kWinLanes + (wl - l)];
      return uint32_t(__builtin_amdgcn_readfirstlane(int(word >> (4 * r)))) & 15u;
    };
    if (seg > 0 && a.strategy == kSwSoftClip) {
      push(seg, kOpS);
      seg = 0;
    }
    uint32_t state = kOpM;
    do {
      const uint32_t b = bits(p1, p2);
      uint32_t next;
      int step;
      if (!(b & 1u)) {
        next = kOpM;
        step = 1;
        --p1;
        --p2;
      } else if (b & 2u) {
        step = 1;
        for (int c = p2; c > 1 && (bits(p1, c) & 8u); --c) ++step;  // column 1 never extends
        next = kOpI;
        p2 -= step;
      } else {
        step = 1;
        for (int c = p1; c > 1 && (bits(c, p2) & 4u); --c) ++step;  // row 1 never extends
        next = kOpD;
        p1 -= step;
      }
      if (next == state) {
        seg += step;
      } else {
        if (seg > 0) push(seg, state);
        seg = step;
        state = next;
      }
    } while (p1 > 0 && p2 > 0);
    int off = 0;
    if (a.strategy == kSwSoftClip) {
      push(seg, state);
      if (p2 > 0) push(p2, kOpS);
      off = p1;
    } else if (a.strategy == kSwIgnore) {
      push(seg + p2, state);
      off = p1 - p2;
    } else {
      push(seg, state);
      if (p1 > 0)
        push(p1, kOpD);
      else if (p2 > 0)
        push(p2, kOpI);
    }
    if (lane == 0) {
      a.n_ops[p] = min(nops, cap);
      a.offset[p] = off;
    }
  }
}

using SwKernelF = void (*)(SwArgs);
template <bool LASTROW>
SwKernelF sw_kernel_for(int R) {
  switch (R) {
    case 1: return sw_kernel<1, LASTROW>;
    case 2: return sw_kernel<2, LASTROW>;
    case 3: return sw_kernel<3, LASTROW>;
    case 4: return sw_kernel<4, LASTROW>;
    case 5: return sw_kernel<5, LASTROW>;
    case 6: return sw_kernel<6, LASTROW>;
    case 7: return sw_kernel<7, LASTROW>;
    default: return sw_kernel<8, LASTROW>;
  }
}
SwKernelF sw_kernel_get(int R, bool lastrow) {
  return lastrow ? sw_kernel_for<true>(R) : sw_kernel_for<false>(R);
}

__global__ void sw_fill_i32(int* p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

#define SW_TRY(expr)                                                                          \
  do {                                                                                        \
    hipError_t _e = (expr);                                                                   \
    if (_e != hipSuccess) {                                                                   \
      set_error(kInternal, std::string("HIP error: ") + hipGetErrorString(_e) + " at " #expr); \
      return kInternal;                                                                       \
    }                                                                                         \
  } while (0)

int sw_rows(int64_t n) { return int(std::min<int64_t>(kSwMaxRows, (n + 63) / 64)); }

// Device memory that only grows: a batch re-assigned with inputs that fit reuses it.
struct SwBuf {
  void* p = nullptr;
  size_t cap = 0;
  int grow(size_t bytes) {
    if (bytes <= cap) return kOk;
    if (p) SW_TRY(hipFree(p));
    p = nullptr;
    cap = 0;
    SW_TRY(hipMalloc(&p, bytes));
    cap = bytes;
    return kOk;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
};

struct SwHostBuf {  // pinned staging
  void* p = nullptr;
  size_t cap = 0;
  int grow(size_t bytes) {
    if (bytes <= cap) return kOk;
    if (p) SW_TRY(hipHostFree(p));
    p = nullptr;
    cap = 0;
    SW_TRY(hipHostMalloc(&p, bytes, hipHostMallocDefault));
    cap = bytes;
    return kOk;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
  }
};

size_t align16(size_t x) { return (x + 15) & ~size_t(15); }

}  // namespace

struct SwBatch::Impl {
  int64_t n = 0, cells = 0, ops_total = 0;
  int max_alt = 0;
  int64_t bucket_begin[kSwMaxRows + 1] = {};
  int64_t bucket_cells[kSwMaxRows] = {};
  int64_t bucket_stride[kSwMaxRows] = {};  // trace words per wave
  int num_cus = 0;
  int lds_words = 0, row_words = 0, alt_words = 0, waves_per_block = 1;
  int64_t trace_words = 0;
  // Inputs, one upload: ref_off, alt_off, ops_end (int64, n+1, n+1, n), items (int32, n), then
  // the reference and alternate bytes.  Outputs, one download: n_ops, offset (int32, n), ops.
  size_t o_alt_off = 0, o_ops_end = 0, o_items = 0, o_ref = 0, o_alt = 0, in_bytes = 0;
  size_t o_off = 0, o_ops = 0, out_bytes = 0;
  SwBuf in, out, trace, counters;
  SwHostBuf stage_in, stage_out;
  hipStream_t stream = nullptr;
  hipEvent_t ev[kSwMaxRows + 1] = {};
  bool ran = false;

  ~Impl() {
    if (stream) (void)hipStreamSynchronize(stream);
    in.release();
    out.release();
    trace.release();
    counters.release();
    stage_in.release();
    stage_out.release();
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  template <typename P>
  P* in_at(size_t off) const { return reinterpret_cast<P*>(static_cast<char*>(in.p) + off); }
  template <typename P>
  P* out_at(size_t off) const { return reinterpret_cast<P*>(static_cast<char*>(out.p) + off); }

  // Waves a bucket's launch runs: its kernel's resident waves (the larger occupancy of its two
  // variants) or enough for its pairs, in whole blocks.
  int launch_waves(int R, int64_t pairs, int64_t* result) {
    const int block = 64 * waves_per_block;
    const size_t lds = size_t(lds_words) * 4 * size_t(waves_per_block);
    int best = 1;
    for (bool lastrow : {false, true}) {
      int nb = 0;
      SW_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, sw_kernel_get(R, lastrow), block, lds));
      best = std::max(best, nb);
    }
    const int64_t resident = int64_t(best) * num_cus;
    const int64_t needed = (pairs + waves_per_block - 1) / waves_per_block;
    *result = std::min(resident, needed) * waves_per_block;
    return kOk;
  }
};

SwBatch::SwBatch() : d_(new Impl) {}
SwBatch::~SwBatch() = default;
int64_t SwBatch::n_pairs() const { return d_->n; }
int64_t SwBatch::cells() const { return d_->cells; }
int64_t SwBatch::cigar_bound() const { return d_->ops_total; }

int SwBatch::create(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref,
                    const int64_t* alt_off, const uint8_t* alt, std::unique_ptr<SwBatch>& out) {
  std::unique_ptr<SwBatch> b(new SwBatch);
  if (int rc = b->assign(n_pairs, ref_off, ref, alt_off, alt)) return rc;
  out = std::move(b);
  return kOk;
}

int SwBatch::assign(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref,
                    const int64_t* alt_off, const uint8_t* alt) {
  Impl& d = *d_;
  // The previous inputs' upload and run finish before their buffers are reused.
  if (d.stream) SW_TRY(hipStreamSynchronize(d.stream));
  d.ran = false;
  if (n_pairs < 0) {
    set_error(kInvalidArgument, "negative pair count");
    return kInvalidArgument;
  }
  if (n_pairs > (int64_t(1) << 24)) {
    set_error(kInvalidArgument, "at most 2^24 pairs per batch");
    return kInvalidArgument;
  }
  if (!ref_off || !alt_off || ref_off[0] != 0 || alt_off[0] != 0) {
    set_error(kInvalidArgument, "offset tables must start at zero");
    return kInvalidArgument;
  }
  if ((ref_off[n_pairs] > 0 && !ref) || (alt_off[n_pairs] > 0 && !alt)) {
    set_error(kInvalidArgument, "null sequence array");
    return kInvalidArgument;
  }
  const size_t np = size_t(n_pairs);
  std::vector<int> rows(np);
  std::vector<int64_t> cost(np), ops_end(np);
  int64_t cells = 0, ops_total = 0, bucket_cells[kSwMaxRows] = {}, bucket_stride[kSwMaxRows] = {};
  int max_alt = 0;
  for (int64_t p = 0; p < n_pairs; ++p) {
    const int64_t n = ref_off[p + 1] - ref_off[p], m = alt_off[p + 1] - alt_off[p];
    if (n < 0 || m < 0) {
      set_error(kInvalidArgument, "offset tables must be non-decreasing");
      return kInvalidArgument;
    }
    if (n == 0 || m == 0) {  // Error::EmptySequence (lib.rs:145-153)
      set_error(kInvalidArgument, "pair " + std::to_string(p) + ": sequences must be non-empty");
      return kInvalidArgument;
    }
    if (m > kMaxAlt || n > (int64_t(1) << 30)) {
      set_error(kInvalidArgument, "pair " + std::to_string(p) + ": alternate longer than " +
                                      std::to_string(kMaxAlt) + " bases");
      return kInvalidArgument;
    }
    const int R = sw_rows(n);
    const int64_t strips = (n + 64 * R - 1) / (64 * R);
    rows[size_t(p)] = R;
    cost[size_t(p)] = strips * (m + 63);
    cells += n * m;
    ops_total += n + m + 4;
    ops_end[size_t(p)] = ops_total;
    max_alt = std::max<int>(max_alt, int(m));
    bucket_cells[R - 1] += n * m;
    bucket_stride[R - 1] = std::max(bucket_stride[R - 1], strips * (m + 63) * 64);
  }
  // Buckets by rows per lane, each costliest first.
  std::vector<int32_t> items(np);
  for (int64_t p = 0; p < n_pairs; ++p) items[size_t(p)] = int32_t(p);
  std::stable_sort(items.begin(), items.end(), [&](int32_t x, int32_t y) {
    if (rows[size_t(x)] != rows[size_t(y)]) return rows[size_t(x)] < rows[size_t(y)];
    return cost[size_t(x)] > cost[size_t(y)];
  });
  d.n = n_pairs;
  d.cells = cells;
  d.ops_total = ops_total;
  d.max_alt = max_alt;
  for (int R = 1, k = 0; R <= kSwMaxRows; ++R) {
    d.bucket_begin[R - 1] = k;
    while (k < n_pairs && rows[size_t(items[size_t(k)])] == R) ++k;
    d.bucket_begin[R] = k;
    d.bucket_cells[R - 1] = bucket_cells[R - 1];
    d.bucket_stride[R - 1] = bucket_stride[R - 1];
  }
  // LDS per wave: the row (H, V per column 0..max_alt), the alternate bytes, the window.
  d.row_words = 2 * (max_alt + 1);
  d.alt_words = (max_alt + 3) / 4;
  d.lds_words = (d.row_words + d.alt_words + kWinWords + 1) & ~1;
  d.waves_per_block = std::max(1, std::min(4, int((64 * 1024) / (d.lds_words * 4))));

  if (!d.stream) {
    int dev = 0;
    SW_TRY(hipGetDevice(&dev));
    hipDeviceProp_t prop;
    SW_TRY(hipGetDeviceProperties(&prop, dev));
    d.num_cus = prop.multiProcessorCount;
    SW_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    for (hipEvent_t& e : d.ev) SW_TRY(hipEventCreate(&e));
    if (int rc = d.counters.grow(sizeof(int) * (kSwMaxRows + 1))) return rc;
  }
  const size_t lds_block = size_t(d.lds_words) * 4 * size_t(d.waves_per_block);
  if (lds_block > 64 * 1024) {
    for (bool lastrow : {false, true})
      for (int R = 1; R <= kSwMaxRows; ++R)
        SW_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(sw_kernel_get(R, lastrow)),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, int(lds_block)));
  }
  // The trace: each bucket's launch waves times its stride.
  d.trace_words = 0;
  for (int R = 1; R <= kSwMaxRows; ++R) {
    const int64_t pairs = d.bucket_begin[R] - d.bucket_begin[R - 1];
    if (pairs == 0) continue;
    int64_t waves = 0;
    if (int rc = d.launch_waves(R, pairs, &waves)) return rc;
    d.trace_words = std::max(d.trace_words, waves * d.bucket_stride[R - 1]);
  }
  if (int rc = d.trace.grow(sizeof(uint32_t) * size_t(std::max<int64_t>(d.trace_words, 1)))) return rc;

  const int64_t ref_len = ref_off[n_pairs], alt_len = alt_off[n_pairs];
  d.o_alt_off = align16(sizeof(int64_t) * (np + 1));
  d.o_ops_end = d.o_alt_off + align16(sizeof(int64_t) * (np + 1));
  d.o_items = d.o_ops_end + align16(sizeof(int64_t) * np);
  d.o_ref = d.o_items + align16(sizeof(int32_t) * np);
  d.o_alt = d.o_ref + align16(size_t(ref_len));
  d.in_bytes = d.o_alt + align16(size_t(alt_len));
  d.o_off = align16(sizeof(int32_t) * np);
  d.o_ops = d.o_off + align16(sizeof(int32_t) * np);
  d.out_bytes = d.o_ops + align16(sizeof(uint32_t) * size_t(ops_total));
  if (int rc = d.in.grow(d.in_bytes)) return rc;
  if (int rc = d.out.grow(d.out_bytes)) return rc;
  if (int rc = d.stage_in.grow(d.in_bytes)) return rc;
  if (int rc = d.stage_out.grow(d.out_bytes)) return rc;
  char* h = static_cast<char*>(d.stage_in.p);
  std::memcpy(h, ref_off, sizeof(int64_t) * (np + 1));
  std::memcpy(h + d.o_alt_off, alt_off, sizeof(int64_t) * (np + 1));
  if (np) {
    std::memcpy(h + d.o_ops_end, ops_end.data(), sizeof(int64_t) * np);
    std::memcpy(h + d.o_items, items.data(), sizeof(int32_t) * np);
  }
  if (ref_len) std::memcpy(h + d.o_ref, ref, size_t(ref_len));
  if (alt_len) std::memcpy(h + d.o_alt, alt, size_t(alt_len));
  SW_TRY(hipMemcpyAsync(d.in.p, h, d.in_bytes, hipMemcpyHostToDevice, d.stream));
  return kOk;
}

int SwBatch::run(const int32_t params[4], int strategy, SwRunStats* stats) {
  Impl& d = *d_;
  if (strategy < kSwSoftClip || strategy > kSwIgnore) {
    set_error(kInvalidArgument, "unknown overhang strategy code " + std::to_string(strategy));
    return kInvalidArgument;
  }
  const bool lastrow = strategy == kSwSoftClip || strategy == kSwIgnore;
  int* counters = static_cast<int*>(d.counters.p);
  sw_fill_i32<<<1, 64, 0, d.stream>>>(counters, kSwMaxRows + 1, 0);
  SW_TRY(hipGetLastError());
  if (stats) SW_TRY(hipEventRecord(d.ev[0], d.stream));
  for (int R = 1; R <= kSwMaxRows; ++R) {
    const int64_t b0 = d.bucket_begin[R - 1], pairs = d.bucket_begin[R] - b0;
    if (pairs > 0) {
      int64_t waves = 0;
      if (int rc = d.launch_waves(R, pairs, &waves)) return rc;
      if (waves * d.bucket_stride[R - 1] > d.trace_words) {
        set_error(kInternal, "Smith-Waterman trace smaller than its launch");
        return kInternal;
      }
      SwArgs a{};
      a.ref_off = d.in_at<int64_t>(0);
      a.alt_off = d.in_at<int64_t>(d.o_alt_off);
      a.ops_end = d.in_at<int64_t>(d.o_ops_end);
      a.items = d.in_at<int32_t>(d.o_items) + b0;
      a.ref = d.in_at<uint8_t>(d.o_ref);
      a.alt = d.in_at<uint8_t>(d.o_alt);
      a.n_items = int(pairs);
      a.counter = counters + (R - 1);
      a.substring_pairs = counters + kSwMaxRows;
      a.wm = params[0];
      a.wx = params[1];
      a.wo = params[2];
      a.we = params[3];
      a.strategy = strategy;
      a.trace = static_cast<uint32_t*>(d.trace.p);
      a.trace_stride = d.bucket_stride[R - 1];
      a.lds_words = d.lds_words;
      a.row_words = d.row_words;
      a.alt_words = d.alt_words;
      a.n_ops = d.out_at<int32_t>(0);
      a.offset = d.out_at<int32_t>(d.o_off);
      a.ops = d.out_at<uint32_t>(d.o_ops);
      const int blocks = int(waves / d.waves_per_block);
      const size_t lds = size_t(d.lds_words) * 4 * size_t(d.waves_per_block);
      hipLaunchKernelGGL(sw_kernel_get(R, lastrow), dim3(blocks), dim3(64 * d.waves_per_block),
                         lds, d.stream, a);
      SW_TRY(hipGetLastError());
    }
    if (stats) SW_TRY(hipEventRecord(d.ev[R], d.stream));
  }
  d.ran = true;
  if (stats) {
    SW_TRY(hipStreamSynchronize(d.stream));
    *stats = SwRunStats{};
    float ms = 0;
    SW_TRY(hipEventElapsedTime(&ms, d.ev[0], d.ev[kSwMaxRows]));
    stats->total_ms = ms;
    for (int R = 1; R <= kSwMaxRows; ++R) {
      SW_TRY(hipEventElapsedTime(&ms, d.ev[R - 1], d.ev[R]));
      stats->bucket_ms[R - 1] = ms;
      stats->bucket_pairs[R - 1] = d.bucket_begin[R] - d.bucket_begin[R - 1];
      stats->bucket_cells[R - 1] = d.bucket_cells[R - 1];
    }
    int sub = 0;
    SW_TRY(hipMemcpy(&sub, counters + kSwMaxRows, sizeof(int), hipMemcpyDeviceToHost));
    stats->substring_pairs = sub;
  }
  return kOk;
}

int SwBatch::results(int32_t* offsets, int64_t* starts, uint32_t* cigar, int64_t cigar_cap) {
  Impl& d = *d_;
  if (!d.ran) {
    set_error(kInvalidArgument, "the batch has not been run");
    return kInvalidArgument;
  }
  SW_TRY(hipMemcpyAsync(d.stage_out.p, d.out.p, d.out_bytes, hipMemcpyDeviceToHost, d.stream));
  SW_TRY(hipStreamSynchronize(d.stream));
  const char* h = static_cast<const char*>(d.stage_out.p);
  const int32_t* nops = reinterpret_cast<const int32_t*>(h);
  const int32_t* off = reinterpret_cast<const int32_t*>(h + d.o_off);
  const uint32_t* ops = reinterpret_cast<const uint32_t*>(h + d.o_ops);
  const int64_t* ops_end = reinterpret_cast<const int64_t*>(static_cast<const char*>(d.stage_in.p) + d.o_ops_end);
  int64_t total = 0;
  for (int64_t p = 0; p < d.n; ++p) total += nops[p];
  if (cigar && cigar_cap < total) {
    set_error(kInvalidArgument, "CIGAR buffer holds " + std::to_string(cigar_cap) +
                                    " elements but the batch has " + std::to_string(total));
    return kInvalidArgument;
  }
  int64_t k = 0;
  for (int64_t p = 0; p < d.n; ++p) {
    if (offsets) offsets[p] = off[p];
    if (starts) starts[p] = k;
    const int64_t c = nops[p];
    if (cigar) std::memcpy(cigar + k, ops + (ops_end[p] - c), sizeof(uint32_t) * size_t(c));
    k += c;
  }
  if (starts) starts[d.n] = k;
  return kOk;
}

}  // namespace fgkl
