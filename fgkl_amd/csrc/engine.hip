// The MI355X (gfx950) PairHMM / PD-HMM engine: device kernels and the batch runner.
//
// Kernel layout (one work item per wave64):
//  * A read of n bases is laid out down the lanes: lane l owns R consecutive rows, with
//    R = ceil((n+1)/64) (1..8) and L = ceil((n+1)/R) live lanes.  The first R*L-n >= 1 rows of
//    lane 0 are identity rows that reproduce DP row 0 exactly (prior 0, M->I 0, I->I 1, M->D 0,
//    D->D 1, left value (0, 0, C/m)), so the read's last row is always the last register of lane
//    L-1 and lane 0 never needs the row above it.
//  * The haplotype is swept along anti-diagonals: at step t lane l computes column t-l+1 for its
//    R rows.  The values of the row above a lane's first row arrive from lane l-1 with a DPP
//    wave_shr:1 move of the registers that lane computed one step earlier.  The haplotype column
//    word travels the same way, lane 0 reading it from an LDS copy of the haplotype.
//  * A work item is one read against every sorted haplotype of its region.  A haplotype starts
//    at the column where it diverges from the previous sorted haplotype, from a snapshot of that
//    column that an earlier sweep wrote into the wave's HBM scratch slot, rescaled by the ratio
//    of haplotype lengths (the recurrence is linear in C/m; kernel.rs:610-678 does the same on
//    the CPU).  Rescue and per-pair work items run a single haplotype from column 0.
//  * FP32 sweeps scale by 2^120 and flag a pair whose raw sum is below 1e-28 or NaN
//    (simd.rs:102-114, kernel.rs:897-914): it is appended to a device list that the FP64 kernel
//    drains in the same stream.  FP64 sweeps scale by 2^1020 and are never rescued.
//  * The recurrence is not a dense contraction, so MFMA is not used: every cell is 4 multiplies
//    and 4 FMAs on the vector ALU (kernel.rs:1011-1013 operation order).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fgkl_internal.hpp"
#include "fgkl_engine.hpp"
#include "wave64.hpp"

namespace fgkl {

// ---------------------------------------------------------------------------------------------
// Numeric constants (simd.rs:102-141).
template <typename T> struct Num;
template <> struct Num<float> {
  static constexpr double kScale = 0x1p120;                       // 2^120
  static constexpr double kLog10Scale = 120.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 1e-28;
};
template <> struct Num<double> {
  static constexpr double kScale = 0x1p1020;                      // 2^1020
  static constexpr double kLog10Scale = 1020.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 0.0;                     // never rescued
};

template <typename T> struct DevTab {
  const T* pm;  // 1 - e(q)
  const T* px;  // e(q) / 3
  const T* ep;  // e(q)
  const T* im;  // 1 - e(gcp)
  const T* mm;  // quantized match-to-match, triangular
};

struct SweepArgs {
  const uint8_t *bases, *quals, *ins, *del, *gcp;
  const int64_t* read_off;
  const int32_t* read_region;
  const int32_t* read_local;
  const int64_t* region_out;
  const int32_t* region_hap0;
  const int32_t* region_nhap;
  const HapPlan* plan;
  const SnapEntry* snaps;
  const uint8_t* hap_bytes;
  const uint32_t* hap_words;
  const int32_t* items;        // (read, read or -1) per item, or (read, sorted hap) in pair mode
  int32_t n_items;
  const int32_t* n_items_dev;  // when set, the item count is read from device memory
  int32_t* counter;            // work-queue head for this launch
  int32_t pair_mode;
  int32_t share;
  void* scratch;
  int64_t scratch_wave_stride;  // elements of T per wave
  double* out;
  int32_t* rescue_pairs;
  int32_t* rescue_count;
  int32_t rescue;
  int32_t lds_words;           // LDS words per wave for the staged haplotype
};

// Qualities above 254 are clamped (model.rs:53-64).
__device__ __forceinline__ int dclamp(uint32_t q) { return q > uint32_t(kMaxQual) ? kMaxQual : int(q); }

__device__ __forceinline__ uint32_t alt_bit(uint32_t b) {  // pd.rs:103-108
  switch (b) {
    case 'A': case 'a': return 1u;
    case 'C': case 'c': return 2u;
    case 'G': case 'g': return 4u;
    case 'T': case 't': return 8u;
    default: return 0u;
  }
}

// Per-lane parameters of its R rows (Workspace::prepare, kernel.rs:485-529).
template <typename T, int R, bool PD> struct Rows {
  T pm[R], px[R], mm[R], im[R], mi[R], md[R], gc[R];
  uint32_t rb[R];
  uint32_t ralt[R];
  uint32_t kind[R];  // 0: row >= 2, 1: row 1, 2: identity row (or a lane past the read)
};

template <typename T, int R, bool PD>
__device__ __forceinline__ void load_rows(const SweepArgs& a, const DevTab<T>& tab, int64_t o,
                                          int L, int pad, int lane, Rows<T, R, PD>& rw) {
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int arow = R * lane + k - pad;
    if (lane < L && arow >= 0) {
      const int64_t i = o + arow;
      const uint32_t b = a.bases[i];
      const int q = dclamp(a.quals[i]), iq = dclamp(a.ins[i]), dq = dclamp(a.del[i]);
      const int gq = dclamp(a.gcp[i]);
      rw.pm[k] = tab.pm[q];
      rw.px[k] = b == 'N' ? rw.pm[k] : tab.px[q];  // an N read base matches everything
      rw.mm[k] = tab.mm[((iq >= dq ? iq * (iq + 1) : dq * (dq + 1)) >> 1) + (iq >= dq ? dq : iq)];
      rw.mi[k] = tab.ep[iq];
      rw.md[k] = tab.ep[dq];
      rw.gc[k] = tab.ep[gq];  // insertion and deletion extension alike (model.rs:73-75)
      rw.im[k] = tab.im[gq];
      rw.rb[k] = b;
      rw.ralt[k] = alt_bit(b);
      rw.kind[k] = arow == 0 ? 1u : 0u;
    } else {
      // Identity rows: D->D 1 keeps row 0's C/m; I->I 1 is harmless because the row above an
      // identity row is zero.  Lanes past the read: all zero.
      rw.pm[k] = T(0); rw.px[k] = T(0); rw.mm[k] = T(0); rw.im[k] = T(0);
      rw.mi[k] = T(0); rw.md[k] = T(0);
      rw.gc[k] = lane < L ? T(1) : T(0);
      rw.rb[k] = 0x100u;
      rw.ralt[k] = 0u;
      rw.kind[k] = 2u;
    }
  }
}

// Number of scratch values a lane keeps per snapshot slot, and the record stride: a slot is
// 64 per-lane records of snap_stride values, so the one lane that records a column writes one
// contiguous, 16-byte-aligned block.
template <int R, bool PD> constexpr int snap_values() { return PD ? 6 * R + 7 : 3 * R + 4; }
template <int R, bool PD> constexpr int snap_stride() { return (snap_values<R, PD>() + 3) & ~3; }

// The values of one DP row in one column: match, insertion, deletion and (PD) branch values.
template <typename T> struct Cell {
  T m, x, y, bm, bx, by;
};

// What a lane carries from one anti-diagonal step to the next: its R rows' values in the column
// it computed last (the left neighbours of its next column), its running last-row sum, and the
// word of the column it is on.  The row above its first row (diagonal and up neighbours) lives
// in Cell values the sweep passes in.
template <typename T, int R, bool PD> struct Lane {
  T m[R], x[R], y[R];
  T bm[PD ? R : 1], bx[PD ? R : 1], by[PD ? R : 1];
  T acc;
  uint32_t w;
};

// Bitwise AND of a value with a per-lane 32-bit mask (both halves of a double).
__device__ __forceinline__ float and_bits(float v, uint32_t k) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & k);
}
__device__ __forceinline__ double and_bits(double v, uint32_t k) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) & ((uint64_t(k) << 32) | k));
}

// The row above a lane's first row in the column it computes next: lane l-1's last row, moved
// with DPP before either lane updates.  Lane 0 receives zeros; its first row is always an
// identity row, which ignores them.  The first lane of a second read in the wave starts with
// identity rows too; of the values above them only the insertion value would leak through (I->I
// 1 in an identity row), so it is masked off there (`keep` is zero in that lane, all ones
// elsewhere).  The mask is applied after the move by an AND on an opaque mask: a select would
// let the compiler run the move only in the lanes that keep it, and a DPP move reads a lane
// outside EXEC as zero, which would also cut the insertion value reaching the lane below.
template <typename T, int R, bool PD>
__device__ __forceinline__ void shift_up(const Lane<T, R, PD>& s, Cell<T>& u, uint32_t keep) {
  u.m = shr1z(s.m[R - 1]);
  u.x = and_bits(shr1z(s.x[R - 1]), keep);
  u.y = shr1z(s.y[R - 1]);
  if constexpr (PD) {
    u.bm = shr1z(s.bm[R - 1]);
    u.bx = shr1z(s.bx[R - 1]);
    u.by = shr1z(s.by[R - 1]);
  }
}

// The R cells of one column.  Rows are updated in place: match, deletion and branch values
// bottom-up (each row still reads the old column of the row above), then insertion values
// top-down (each row reads the new column of the row above), so no value is copied.
template <typename T, int R, bool PD>
__device__ __forceinline__ void cells(Lane<T, R, PD>& s, const Rows<T, R, PD>& rw,
                                      const Cell<T>& d, const Cell<T>& u) {
  const uint32_t w = s.w;
  const uint32_t hb = w & 0xFFu;
  const bool hN = hb == uint32_t('N');
  uint32_t st[PD ? R : 1], de[PD ? R : 1];
  if constexpr (PD) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (rw.kind[k] == 2u) {
        st[k] = kInsideDel;  // identity rows reproduce row 0, branch values included
        de[k] = 0u;
      } else {
        st[k] = rw.kind[k] == 1u ? ((w >> 12) & 3u) : ((w >> 14) & 3u);
        de[k] = (w >> 16) & 1u;
      }
    }
  }
#pragma unroll
  for (int k = R - 1; k >= 0; --k) {
    const T pdm = k ? s.m[k - 1] : d.m;
    const T pdx = k ? s.x[k - 1] : d.x;
    const T pdy = k ? s.y[k - 1] : d.y;
    bool match = hN || hb == rw.rb[k];
    if constexpr (PD) match = match || (((w >> 8) & rw.ralt[k]) != 0u);
    const T prior = match ? rw.pm[k] : rw.px[k];
    if constexpr (!PD) {
      // kernel.rs:1011-1013
      const T Y = fma(s.y[k], rw.gc[k], s.m[k] * rw.md[k]);
      s.m[k] = prior * fma(pdy, rw.im[k], fma(pdx, rw.im[k], pdm * rw.mm[k]));
      s.y[k] = Y;
    } else {
      // pd_reference.rs:62-94 for the column's deletion state.
      const bool A = st[k] == kAfterDel, I = st[k] == kInsideDel, N = st[k] == kNormal;
      const T pdbm = k ? s.bm[k - 1] : d.bm;
      const T pdbx = k ? s.bx[k - 1] : d.bx;
      const T pdby = k ? s.by[k - 1] : d.by;
      const T dme = A ? fmax(pdbm, pdm) : pdm;
      const T dxe = A ? fmax(pdbx, pdx) : pdx;
      const T dye = A ? fmax(pdby, pdy) : pdy;
      const T lme = A ? fmax(s.bm[k], s.m[k]) : s.m[k];
      const T lxe = A ? fmax(s.bx[k], s.x[k]) : s.x[k];
      const T lye = A ? fmax(s.by[k], s.y[k]) : s.y[k];
      const T Y = fma(lye, rw.gc[k], lme * rw.md[k]);
      const T BM = N ? s.m[k] : (I ? s.bm[k] : lme);
      const T BX = N ? s.x[k] : (I ? s.bx[k] : lxe);
      const T BY = N ? s.y[k] : (I ? s.by[k] : lye);
      s.m[k] = prior * fma(dye, rw.im[k], fma(dxe, rw.im[k], dme * rw.mm[k]));
      s.y[k] = Y;
      s.bm[k] = BM;
      s.bx[k] = BX;
      s.by[k] = BY;
    }
  }
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const T pum = k ? s.m[k - 1] : u.m;
    const T pux = k ? s.x[k - 1] : u.x;
    if constexpr (!PD) {
      s.x[k] = fma(pux, rw.gc[k], pum * rw.mi[k]);
    } else {
      // pd_reference.rs:95-100
      const T pubm = k ? s.bm[k - 1] : u.bm;
      const T pubx = k ? s.bx[k - 1] : u.bx;
      const T ume = de[k] ? fmax(pubm, pum) : pum;
      const T uxe = de[k] ? fmax(pubx, pux) : pux;
      s.x[k] = fma(uxe, rw.gc[k], ume * rw.mi[k]);
    }
  }
  s.acc += s.m[R - 1] + s.x[R - 1];
}

// Writes a lane's state into its snapshot record (value by value); `u` is the row above
// its first row in the column just computed.
template <typename T, int R, bool PD>
__device__ __forceinline__ void store_snapshot(const Lane<T, R, PD>& s, const Cell<T>& u, T* d) {
  constexpr int NV = snap_values<R, PD>();
#pragma unroll
  for (int k = 0; k < R; ++k) {
    d[(3 * k + 0)] = s.m[k];
    d[(3 * k + 1)] = s.x[k];
    d[(3 * k + 2)] = s.y[k];
    if constexpr (PD) {
      d[(3 * R + 3 * k + 0)] = s.bm[k];
      d[(3 * R + 3 * k + 1)] = s.bx[k];
      d[(3 * R + 3 * k + 2)] = s.by[k];
    }
  }
  constexpr int U = PD ? 6 * R : 3 * R;
  d[(U + 0)] = u.m;
  d[(U + 1)] = u.x;
  d[(U + 2)] = u.y;
  if constexpr (PD) {
    d[(U + 3)] = u.bm;
    d[(U + 4)] = u.bx;
    d[(U + 5)] = u.by;
  }
  d[(NV - 1)] = s.acc;
}

// A staged column word carries, in bits 20..31, one plus the scratch slot that must record the
// column's state (0: none).  The word travels down the lanes with the column, so each lane sees
// the mark exactly when it has just computed that column.
constexpr uint32_t kSlotShift = 20;
constexpr int kMaxSlots = 4095;

// One sweep of a read over haplotype columns start+1 .. start+W.  `lw[t]` holds the word of
// column start+1+t (zero beyond W, 16-byte aligned).  Returns the lane's raw scaled sum over the
// read's last row (meaningful in lane L-1).  `cap` receives lane cap_lane's sum right after step
// cap_t: the first read of a two-read wave finishes column W there, and its last lane would go on
// summing the columns past W while the second read's lanes finish.
template <typename T, int R, bool PD, bool SNAP>
__device__ T sweep(const Rows<T, R, PD>& rw, const uint32_t* lw, int W, int L, int lane, T init,
                   const T* snap_in, T scale, T* scratch, bool row0_above, uint32_t keep,
                   int cap_t, int cap_lane, T& cap) {
  constexpr int NV = snap_values<R, PD>();
  Lane<T, R, PD> s;
  Cell<T> P{T(0), T(0), T(0), T(0), T(0), T(0)};  // diagonal neighbours of the first row
  if (snap_in) {
    const T* p = snap_in + lane * snap_stride<R, PD>();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      s.m[k] = p[(3 * k + 0)] * scale;
      s.x[k] = p[(3 * k + 1)] * scale;
      s.y[k] = p[(3 * k + 2)] * scale;
      if constexpr (PD) {
        s.bm[k] = p[(3 * R + 3 * k + 0)] * scale;
        s.bx[k] = p[(3 * R + 3 * k + 1)] * scale;
        s.by[k] = p[(3 * R + 3 * k + 2)] * scale;
      }
    }
    constexpr int U = PD ? 6 * R : 3 * R;
    P.m = p[(U + 0)] * scale;
    P.x = p[(U + 1)] * scale;
    P.y = p[(U + 2)] * scale;
    if constexpr (PD) {
      P.bm = p[(U + 3)] * scale;
      P.bx = p[(U + 4)] * scale;
      P.by = p[(U + 5)] * scale;
    }
    s.acc = p[(NV - 1)] * scale;
  } else {
    // Column 0: zeros below row 0; identity rows carry row 0's deletion value C/m.
#pragma unroll
    for (int k = 0; k < R; ++k) {
      s.m[k] = T(0);
      s.x[k] = T(0);
      s.y[k] = (rw.kind[k] == 2u && lane < L) ? init : T(0);
      if constexpr (PD) {
        s.bm[k] = T(0);
        s.bx[k] = T(0);
        s.by[k] = T(0);
      }
    }
    s.acc = T(0);
    // When identity rows fill lane 0, read row 1 is lane 1's first row and its diagonal
    // neighbour in column 1 is DP row 0's deletion value (Y[0][0] = C/m, reference.rs:27-28).
    if (row0_above) P.y = init;
  }
  s.w = 0;

  auto record = [&](const Cell<T>& u) {
    if constexpr (SNAP) {
      const uint32_t sl = s.w >> kSlotShift;
      if (sl)
        store_snapshot<T, R, PD>(s, u, scratch + (size_t(sl - 1) * 64 + lane) * snap_stride<R, PD>());
    }
  };
  const int steps = W + L - 1;
  // Ramp: lane l starts at step l and keeps its column-`start` state until then.  Rounded up to
  // four steps so the steady loop reads aligned groups of four column words.
  const int ramp = min(steps, (L - 1 + 3) & ~3);
  int t = 0;
  Cell<T> Q;
  for (; t < ramp; ++t) {
    s.w = shr1(lw[t], s.w);
    shift_up(s, Q, keep);
    if (t >= lane) {
      cells(s, rw, P, Q);
      P = Q;
      record(P);
    }
    if (t == cap_t) cap = readlane(s.acc, cap_lane);
  }
  // Steady state: every lane is live; the diagonal and up neighbours alternate between P and Q.
  // It stops by step cap_t; the single steps after it test for the capture.
  const int steady_end = min(steps, cap_t + 1);
  uint4 q = *reinterpret_cast<const uint4*>(lw + t);
  for (; t + 4 <= steady_end; t += 4) {
    const uint4 cq = q;
    q = *reinterpret_cast<const uint4*>(lw + t + 4);
    s.w = shr1(cq.x, s.w);
    shift_up(s, Q, keep);
    cells(s, rw, P, Q);
    record(Q);
    s.w = shr1(cq.y, s.w);
    shift_up(s, P, keep);
    cells(s, rw, Q, P);
    record(P);
    s.w = shr1(cq.z, s.w);
    shift_up(s, Q, keep);
    cells(s, rw, P, Q);
    record(Q);
    s.w = shr1(cq.w, s.w);
    shift_up(s, P, keep);
    cells(s, rw, Q, P);
    record(P);
  }
  if (t - 1 == cap_t) cap = readlane(s.acc, cap_lane);
  for (; t < steps; ++t) {
    s.w = shr1(lw[t], s.w);
    shift_up(s, Q, keep);
    cells(s, rw, P, Q);
    P = Q;
    record(P);
    if (t == cap_t) cap = readlane(s.acc, cap_lane);
  }
  return s.acc;
}

// Stages the words of haplotype columns start+1 .. m into the wave's LDS, zero-padded, and marks
// the columns this sweep must record with their scratch slots.
template <bool PD>
__device__ __forceinline__ void stage_hap(const SweepArgs& a, const HapPlan& p, int start, int L,
                                          int lane, uint32_t* lw, int nsnap) {
  const int W = p.len - start;
  const int total = W + L + 8;  // the steady loop reads one group of four ahead
  wave_sync_lds();
  for (int i = lane; i < total; i += 64) {
    uint32_t v = 0;
    if (i < W) v = PD ? a.hap_words[p.off + start + i] : uint32_t(a.hap_bytes[p.off + start + i]);
    lw[i] = v;
  }
  if (nsnap) {
    wave_sync_lds();
    for (int i = lane; i < nsnap; i += 64) {
      const SnapEntry e = a.snaps[p.snap0 + i];
      lw[e.col - start - 1] |= uint32_t(e.slot + 1) << kSlotShift;
    }
  }
  wave_sync_lds();
}

template <typename T, int R, bool PD>
__global__ void __launch_bounds__(256) sweep_kernel(SweepArgs a, DevTab<T> tab) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_all[];
  const int lane = threadIdx.x & 63;
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t wave_global = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave_in_block;
  uint32_t* lw = lds_all + wave_in_block * a.lds_words;
  T* scratch = a.scratch ? static_cast<T*>(a.scratch) + wave_global * a.scratch_wave_stride : nullptr;
  const int n_items = a.n_items_dev ? *a.n_items_dev : a.n_items;
  constexpr int NV = snap_values<R, PD>();

  for (;;) {
    // Every lane adds one, so a claim advances the counter by 64 and every lane's old value lies
    // in [64 i, 64 i + 63].  A lane-dependent branch around the claim, next to the lane-0 stores
    // at the end of an item, would let the optimiser thread the loop's back edge past the claim
    // for lanes 1..63, which would then read an item without lane 0.
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= n_items) break;

    // Everything below is uniform across the wave; readfirstlane keeps it in scalar registers.
    // A main-pass item is one read, or two reads of one region sharing the wave (read A in lanes
    // [0, La), read B in [La, La+Lb)); a pair-mode item is one (read, sorted haplotype).
    int ra, rb, k_begin, k_end;
    ra = __builtin_amdgcn_readfirstlane(a.items[2 * item]);
    if (a.pair_mode) {
      rb = -1;
      k_begin = __builtin_amdgcn_readfirstlane(a.items[2 * item + 1]);
      k_end = k_begin + 1;
    } else {
      rb = __builtin_amdgcn_readfirstlane(a.items[2 * item + 1]);
      k_begin = 0;
      k_end = -1;
    }
    const bool two = rb >= 0;
    const int64_t oa = a.read_off[ra];
    const int na = __builtin_amdgcn_readfirstlane(int(a.read_off[ra + 1] - oa));
    const int64_t ob = two ? a.read_off[rb] : 0;
    const int nb = two ? __builtin_amdgcn_readfirstlane(int(a.read_off[rb + 1] - ob)) : 0;
    const int La = (na + R) / R;  // n rows plus at least one identity row
    const int Lb = two ? (nb + R) / R : 0;
    const int L = La + Lb;
    const int pada = R * La - na, padb = R * Lb - nb;
    const int g = __builtin_amdgcn_readfirstlane(a.read_region[ra]);
    const int hap0 = __builtin_amdgcn_readfirstlane(a.region_hap0[g]);
    const int H = __builtin_amdgcn_readfirstlane(a.region_nhap[g]);
    if (k_end < 0) k_end = H;
    const int64_t out_a = a.region_out[g] + int64_t(a.read_local[ra]) * H;
    const int64_t out_b = two ? a.region_out[g] + int64_t(a.read_local[rb]) * H : 0;

    // This lane's read: its offset, lanes, identity rows and lane index within the read.
    const bool in_b = lane >= La;
    Rows<T, R, PD> rw;
    load_rows<T, R, PD>(a, tab, in_b ? ob : oa, in_b ? Lb : La, in_b ? padb : pada,
                        in_b ? lane - La : lane, rw);
    // Each read's second lane starts below DP row 0 when identity rows fill its first lane.
    const bool row0_above = (lane == 1 && pada == R) || (two && lane == La + 1 && padb == R);
    // Zero in the first lane of read B, all ones elsewhere; opaque to the optimiser (see
    // shift_up).
    uint32_t keep = (two && lane == La) ? 0u : ~0u;
    asm volatile("" : "+v"(keep));

    double prev_a = 0.0, prev_b = 0.0;
    for (int k = k_begin; k < k_end; ++k) {
      HapPlan p = a.plan[hap0 + k];
      p.len = __builtin_amdgcn_readfirstlane(p.len);
      p.lcp = __builtin_amdgcn_readfirstlane(p.lcp);
      p.dup = __builtin_amdgcn_readfirstlane(p.dup);
      p.nsnap = __builtin_amdgcn_readfirstlane(p.nsnap);
      p.snap0 = __builtin_amdgcn_readfirstlane(p.snap0);
      p.start_slot = __builtin_amdgcn_readfirstlane(p.start_slot);
      double res_a, res_b;
      if (!a.pair_mode && p.dup) {
        res_a = prev_a;  // kernel.rs:237-246
        res_b = prev_b;
      } else {
        const bool shared = a.share && !a.pair_mode;
        const int start = shared ? p.lcp : 0;
        const int nsnap = shared ? p.nsnap : 0;
        const T* snap_in = nullptr;
        T scale = T(1);
        if (shared && p.start_slot >= 0) {
          snap_in = scratch + size_t(p.start_slot) * 64 * snap_stride<R, PD>();
          scale = T(double(p.start_src_len) / double(p.len));  // kernel.rs:641
        }
        const T init = T(Num<T>::kScale / double(p.len));      // kernel.rs:640
        stage_hap<PD>(a, p, start, L, lane, lw, nsnap);
        // Read A's last lane finishes column W at step W + La - 2.
        const int W = p.len - start;
        T acc_a = T(0);
        const T acc =
            nsnap ? sweep<T, R, PD, true>(rw, lw, W, L, lane, init, snap_in, scale, scratch,
                                          row0_above, keep, W + La - 2, La - 1, acc_a)
                  : sweep<T, R, PD, false>(rw, lw, W, L, lane, init, snap_in, scale, scratch,
                                           row0_above, keep, W + La - 2, La - 1, acc_a);
        // A read's last row is its last lane's last register (kernel.rs:897-914).
        auto finish = [](double raw) {
          if (Num<T>::kMinAccepted > 0.0 && !(raw >= Num<T>::kMinAccepted)) return __builtin_nan("");
          return log10(raw) - Num<T>::kLog10Scale;
        };
        res_a = finish(double(acc_a));
        res_b = finish(double(readlane(acc, L - 1)));
      }
      if (lane == 0) {
        a.out[out_a + p.src] = res_a;
        if (a.rescue && res_a != res_a) {
          const int idx = atomicAdd(a.rescue_count, 1);
          a.rescue_pairs[2 * idx] = ra;
          a.rescue_pairs[2 * idx + 1] = k;
        }
        if (two) {
          a.out[out_b + p.src] = res_b;
          if (a.rescue && res_b != res_b) {
            const int idx = atomicAdd(a.rescue_count, 1);
            a.rescue_pairs[2 * idx] = rb;
            a.rescue_pairs[2 * idx + 1] = k;
          }
        }
      }
      prev_a = res_a;
      prev_b = res_b;
    }
  }
}

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

using KernelF = void (*)(SweepArgs, DevTab<float>);
using KernelD = void (*)(SweepArgs, DevTab<double>);

template <typename T> struct KernelTable;
template <> struct KernelTable<float> {
  static constexpr KernelF plain[kMaxRowsPerLane] = {
      sweep_kernel<float, 1, false>, sweep_kernel<float, 2, false>, sweep_kernel<float, 3, false>,
      sweep_kernel<float, 4, false>, sweep_kernel<float, 5, false>, sweep_kernel<float, 6, false>,
      sweep_kernel<float, 7, false>, sweep_kernel<float, 8, false>};
  static constexpr KernelF pd[kMaxRowsPerLane] = {
      sweep_kernel<float, 1, true>, sweep_kernel<float, 2, true>, sweep_kernel<float, 3, true>,
      sweep_kernel<float, 4, true>, sweep_kernel<float, 5, true>, sweep_kernel<float, 6, true>,
      sweep_kernel<float, 7, true>, sweep_kernel<float, 8, true>};
};
template <> struct KernelTable<double> {
  static constexpr KernelD plain[kMaxRowsPerLane] = {
      sweep_kernel<double, 1, false>, sweep_kernel<double, 2, false>,
      sweep_kernel<double, 3, false>, sweep_kernel<double, 4, false>,
      sweep_kernel<double, 5, false>, sweep_kernel<double, 6, false>,
      sweep_kernel<double, 7, false>, sweep_kernel<double, 8, false>};
  static constexpr KernelD pd[kMaxRowsPerLane] = {
      sweep_kernel<double, 1, true>, sweep_kernel<double, 2, true>, sweep_kernel<double, 3, true>,
      sweep_kernel<double, 4, true>, sweep_kernel<double, 5, true>, sweep_kernel<double, 6, true>,
      sweep_kernel<double, 7, true>, sweep_kernel<double, 8, true>};
};

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / 64;
constexpr int64_t kMaxQueueItems = int64_t(1) << 24;  // 64 x (items + waves) < 2^31

struct DeviceState {
  int device = -1;
  int num_cus = 0;
  DevTab<float> tf{};
  DevTab<double> td{};
  // Resident blocks per CU for each kernel, at the LDS size it was queried with.
  int occ[2][2][kMaxRowsPerLane] = {};
  int occ_lds[2][2][kMaxRowsPerLane] = {};
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
int resident_blocks(DeviceState* st, bool pd, int R, int lds_bytes, int* blocks) {
  const int ti = std::is_same<T, double>::value ? 1 : 0;
  int& occ = st->occ[ti][pd][R - 1];
  int& occ_lds = st->occ_lds[ti][pd][R - 1];
  if (occ == 0 || occ_lds != lds_bytes) {
    int nb = 0;
    auto kern = pd ? KernelTable<T>::pd[R - 1] : KernelTable<T>::plain[R - 1];
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kBlock, size_t(lds_bytes)));
    occ = std::max(nb, 1);
    occ_lds = lds_bytes;
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
  std::vector<void*> allocs;
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
  int32_t* bucket_items[kMaxRowsPerLane] = {};
  int32_t bucket_n[kMaxRowsPerLane] = {};
  int64_t bucket_pairs[kMaxRowsPerLane] = {};
  int64_t bucket_cells[kMaxRowsPerLane] = {};
  int32_t* rescue_pairs[kMaxRowsPerLane] = {};
  int32_t* counters = nullptr;  // [0,8) main, [8,16) rescue, [16,24) rescue counts
  double* out = nullptr;
  void* scratch = nullptr;
  size_t scratch_bytes = 0;
  // per-region pair lists, for the per-call mode
  int32_t* region_pairs = nullptr;
  std::vector<int64_t> region_pair_off;  // (G * 8 + 1) offsets in pairs, bucket-major per region
  int32_t* region_counters = nullptr;
  int64_t n_region_launches = 0;
  hipEvent_t ev[4] = {};
  hipEvent_t bev[2 * kMaxRowsPerLane] = {};  // around each primary bucket launch
  hipGraphExec_t graph[8] = {};  // by mode * 2 + share
  int lds_words = 0;

  ~Impl() {
    for (auto& g : graph)
      if (g) (void)hipGraphExecDestroy(g);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : bev)
      if (e) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
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
    a.scratch = scratch; a.out = out; a.lds_words = lds_words;
    return a;
  }

  template <typename T>
  int launch(int R, SweepArgs a, int64_t max_items) {
    if (max_items <= 0) return kOk;
    const int lds_bytes = lds_words * 4 * kWavesPerBlock;
    int per_cu = 0;
    if (int rc = resident_blocks<T>(st, pd, R, lds_bytes, &per_cu)) return rc;
    const int64_t want = (max_items + kWavesPerBlock - 1) / kWavesPerBlock;
    const int grid = int(std::min<int64_t>(int64_t(per_cu) * st->num_cus, want));
    auto kern = pd ? KernelTable<T>::pd[R - 1] : KernelTable<T>::plain[R - 1];
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds_bytes, stream, a, dev_tab<T>(st));
    HIP_TRY(hipGetLastError());
    return kOk;
  }

  template <typename T>
  int64_t scratch_need(int R) {
    const int lds_bytes = lds_words * 4 * kWavesPerBlock;
    int per_cu = 0;
    if (resident_blocks<T>(st, pd, R, lds_bytes, &per_cu)) return -1;
    const int64_t waves = int64_t(per_cu) * st->num_cus * kWavesPerBlock;
    const int nv = ((pd ? 6 * R + 7 : 3 * R + 4) + 3) & ~3;  // snap_stride
    return waves * int64_t(plan.max_slots) * nv * 64 * int64_t(sizeof(T));
  }
  template <typename T>
  int64_t scratch_stride(int R) const {
    const int nv = ((pd ? 6 * R + 7 : 3 * R + 4) + 3) & ~3;  // snap_stride
    return int64_t(plan.max_slots) * nv * 64;
  }

  int enqueue(int mode, int share, bool timed);
};

int DeviceBatch::Impl::enqueue(int mode, int share, bool timed) {
  // Reset the work-queue heads and rescue counts with a kernel (graph-safe).
  hipLaunchKernelGGL(fill_i32, dim3(1), dim3(64), 0, stream, counters, 24, 0);
  HIP_TRY(hipGetLastError());
  const bool fp32 = mode == kModeFp32Rescue || mode == kModeFp32NoRescue;
  // The primary sweep, one launch per rows-per-lane bucket, each bracketed by events when the
  // run is not captured into a graph.
  for (int R = 1; R <= kMaxRowsPerLane; ++R) {
    if (!bucket_n[R - 1]) continue;
    SweepArgs a = base_args();
    a.items = bucket_items[R - 1];
    a.n_items = bucket_n[R - 1];
    a.counter = counters + (R - 1);
    a.share = share;
    if (timed) HIP_TRY(hipEventRecord(bev[2 * (R - 1)], stream));
    if (fp32) {
      a.scratch_wave_stride = scratch_stride<float>(R);
      a.rescue = 1;
      a.rescue_pairs = rescue_pairs[R - 1];
      a.rescue_count = counters + 16 + (R - 1);
      if (int rc = launch<float>(R, a, bucket_n[R - 1])) return rc;
    } else {
      a.scratch_wave_stride = scratch_stride<double>(R);
      if (int rc = launch<double>(R, a, bucket_n[R - 1])) return rc;
    }
    if (timed) HIP_TRY(hipEventRecord(bev[2 * (R - 1) + 1], stream));
  }
  if (timed) HIP_TRY(hipEventRecord(ev[1], stream));
  if (fp32) {
    for (int R = 1; R <= kMaxRowsPerLane && mode == kModeFp32Rescue; ++R) {
      if (!bucket_n[R - 1]) continue;
      SweepArgs a = base_args();
      a.items = rescue_pairs[R - 1];
      a.n_items = 0;
      a.n_items_dev = counters + 16 + (R - 1);
      a.counter = counters + 8 + (R - 1);
      a.pair_mode = 1;
      a.share = 0;
      a.scratch = nullptr;
      // The grid is sized for the worst case; waves beyond the rescued count exit at once.
      if (int rc = launch<double>(R, a, bucket_pairs[R - 1])) return rc;
    }
  }
  return kOk;
}

DeviceBatch::DeviceBatch() : impl_(new Impl) {}
DeviceBatch::~DeviceBatch() { delete impl_; }

int DeviceBatch::create(const BatchInput& in, std::unique_ptr<DeviceBatch>& out) {
  std::unique_ptr<DeviceBatch> b(new DeviceBatch);
  Impl& d = *b->impl_;
  if (int rc = build_plan(in, d.plan)) return rc;
  if (int rc = device_state(&d.st)) return rc;
  d.pd = in.pd;
  HIP_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
  for (auto& e : d.ev) HIP_TRY(hipEventCreate(&e));
  for (auto& e : d.bev) HIP_TRY(hipEventCreate(&e));
  const HostPlan& p = d.plan;
  const int64_t NR = p.n_reads, NH = p.n_haps, G = int64_t(p.region_hap0.size());
  const size_t rb = size_t(in.read_off[NR]);
  const size_t hb = size_t(in.hap_off[NH]);
  int rc = 0;
  if ((rc = d.upload(&d.bases, in.bases, rb)) || (rc = d.upload(&d.quals, in.quals, rb)) ||
      (rc = d.upload(&d.ins, in.ins, rb)) || (rc = d.upload(&d.del, in.del, rb)) ||
      (rc = d.upload(&d.gcp, in.gcp, rb)) ||
      (rc = d.upload(&d.read_off, in.read_off, size_t(NR + 1))) ||
      (rc = d.upload(&d.read_region, p.read_region.data(), size_t(NR))) ||
      (rc = d.upload(&d.read_local, p.read_local.data(), size_t(NR))) ||
      (rc = d.upload(&d.region_out, p.region_out.data(), size_t(G + 1))) ||
      (rc = d.upload(&d.region_hap0, p.region_hap0.data(), size_t(G))) ||
      (rc = d.upload(&d.region_nhap, p.region_nhap.data(), size_t(G))) ||
      (rc = d.upload(&d.hplan, p.plan.data(), size_t(NH))) ||
      (rc = d.upload(&d.snaps, p.snaps.data(), p.snaps.size())))
    return rc;
  if (in.pd) {
    if ((rc = d.upload(&d.hap_words, p.words.data(), p.words.size()))) return rc;
  } else {
    if ((rc = d.upload(&d.hap_bytes, in.hap_bases, hb))) return rc;
  }
  for (int R = 1; R <= kMaxRowsPerLane; ++R) {
    const auto& items = p.bucket_items[R - 1];
    d.bucket_n[R - 1] = int32_t(items.size() / 2);
    const int64_t pairs = p.bucket_pairs[R - 1];
    if (pairs > kMaxQueueItems) {  // a work queue advances by 64 per claim in an int32 counter
      set_error(kInvalidArgument, "a batch holds more than " + std::to_string(kMaxQueueItems) +
                                      " pairs of one read length class; split it");
      return kInvalidArgument;
    }
    d.bucket_pairs[R - 1] = pairs;
    d.bucket_cells[R - 1] = p.bucket_cells[R - 1];
    if ((rc = d.upload(&d.bucket_items[R - 1], items.data(), items.size()))) return rc;
    if ((rc = d.alloc(&d.rescue_pairs[R - 1], size_t(2 * pairs)))) return rc;
  }
  if ((rc = d.alloc(&d.counters, 32))) return rc;
  if ((rc = d.alloc(&d.out, size_t(p.n_pairs)))) return rc;
  // LDS per wave: the longest haplotype plus the ramp of the widest read, in words.
  if (p.max_slots > kMaxSlots) {
    set_error(kInvalidArgument, "a region holds more than " + std::to_string(kMaxSlots) +
                                    " distinct haplotype prefix columns");
    return kInvalidArgument;
  }
  d.lds_words = ((p.max_hap_len + 64 + 8) + 3) & ~3;
  if (int64_t(d.lds_words) * 4 * kWavesPerBlock > 64 * 1024) {
    set_error(kInvalidArgument, "haplotype of " + std::to_string(p.max_hap_len) +
                                    " bases exceeds the engine's LDS staging capacity");
    return kInvalidArgument;
  }
  int64_t need = 0;
  if (p.max_slots > 0) {
    for (int R = 1; R <= kMaxRowsPerLane; ++R) {
      if (!d.bucket_n[R - 1]) continue;
      need = std::max(need, d.scratch_need<float>(R));
      need = std::max(need, d.scratch_need<double>(R));
    }
  }
  if (need < 0) return kInternal;
  if (need > 0) {
    HIP_TRY(hipMalloc(&d.scratch, size_t(need)));
    d.allocs.push_back(d.scratch);
    d.scratch_bytes = size_t(need);
  }
  HIP_TRY(hipStreamSynchronize(d.stream));
  out = std::move(b);
  return kOk;
}

int DeviceBatch::run(int mode, int share, int use_graph, RunStats* stats) {
  Impl& d = *impl_;
  if (mode != kModeFp32Rescue && mode != kModeFp64 && mode != kModeFp64PerRegion &&
      mode != kModeFp32NoRescue) {
    set_error(kInvalidArgument, "unknown run mode");
    return kInvalidArgument;
  }
  if (d.plan.n_pairs == 0) {
    if (stats) *stats = RunStats{};
    return kOk;
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
    int32_t counts[32];
    HIP_TRY(hipMemcpy(counts, d.counters, sizeof(counts), hipMemcpyDeviceToHost));
    int64_t rescued = 0;
    if (mode == kModeFp32Rescue || mode == kModeFp32NoRescue)
      for (int R = 0; R < kMaxRowsPerLane; ++R) rescued += counts[16 + R];
    stats->rescued_pairs = rescued;
    for (int R = 0; R < kMaxRowsPerLane; ++R) stats->bucket_cells[R] = d.bucket_cells[R];
    if (!use_graph) {
      float first = 0.f;
      HIP_TRY(hipEventElapsedTime(&first, d.ev[0], d.ev[1]));
      stats->primary_ms = first;
      stats->rescue_ms = total - first;
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
    d.n_region_launches = G * kMaxRowsPerLane;
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

int DeviceBatch::results(double* out, int64_t n) {
  Impl& d = *impl_;
  if (n != d.plan.n_pairs) {
    set_error(kInvalidArgument, "output has " + std::to_string(n) +
                                    " elements but the batch holds " +
                                    std::to_string(d.plan.n_pairs) + " pairs");
    return kInvalidArgument;
  }
  if (n) HIP_TRY(hipMemcpy(out, d.out, sizeof(double) * size_t(n), hipMemcpyDeviceToHost));
  return kOk;
}

int64_t DeviceBatch::n_pairs() const { return impl_->plan.n_pairs; }
int64_t DeviceBatch::scratch_bytes() const { return int64_t(impl_->scratch_bytes); }

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
