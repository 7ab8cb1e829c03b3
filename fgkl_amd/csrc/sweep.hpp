// Device code of the MI355X (gfx950) PairHMM / PD-HMM engine: the sweep kernel template, built
// per precision in sweep_f32.hip and sweep_f64.hip and launched by engine.hip.
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
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "fgkl_internal.hpp"
#include "wave64.hpp"

namespace fgkl {

// ---------------------------------------------------------------------------------------------
// Numeric constants (simd.rs:102-141).
template <typename T> struct Num;
// kBwdScale: the weight of a path that ends, in a backward sweep (simd.rs:106-136);
// kMinAcceptedJoined: the smallest accepted scaled likelihood of a joined pair.  FP64 joins below
// 1e24 (a likelihood of about 1e-283) are rerun without suffix sharing (lib.rs:467-487).
template <> struct Num<float> {
  static constexpr double kScale = 0x1p120;                       // 2^120
  static constexpr double kLog10Scale = 120.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 1e-28;
  static constexpr double kBwdScale = 0x1p120;
  static constexpr double kMinAcceptedJoined = 1e-28;
};
template <> struct Num<double> {
  static constexpr double kScale = 0x1p1020;                      // 2^1020
  static constexpr double kLog10Scale = 1020.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 0.0;                     // never rescued unjoined
  static constexpr double kBwdScale = 1.0;
  static constexpr double kMinAcceptedJoined = 1e24;
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
  // PD: the bytes of each region's read-base classes from 4 on (kPdExtraClasses per region).
  const uint8_t* region_classes;
  const int32_t* items;        // (read, read or -1) per item, or (read, sorted hap) in pair mode
  int32_t n_items;
  const int32_t* n_items_dev;  // when set, the item count is read from device memory
  int32_t* counter;            // work-queue head for this launch
  int32_t pair_mode;
  int32_t share;
  void* scratch;
  int64_t scratch_wave_stride;  // elements of T per wave
  double* out;
  // FP32 pairs lost to underflow: with `rescue_by_rows`, listed by the read's own rows per lane
  // (rescue_pairs[R - 1], count rescue_count[R - 1]), else all in rescue_pairs[0].
  int32_t* rescue_pairs[kMaxRowsPerLane];
  int32_t* rescue_count;
  int32_t rescue_by_rows;
  int32_t rescue;
  // Dense rescue (lib.rs:488-520): a main-pass FP32 launch sets dense[read] to 1 for a read lost
  // against a third or more of its region's haplotypes (0 otherwise) and lists it in
  // dense_items; an FP64 launch with `nan_only` reruns those reads against every haplotype with
  // prefix sharing and writes only the results still NaN; a pair-mode rescue with `dense` set
  // skips their pairs.
  uint8_t* dense;
  int32_t* dense_items[kMaxRowsPerLane];  // by the read's own rows per lane
  int32_t* dense_count;                   // kMaxRowsPerLane counts
  int32_t nan_only;
  int32_t lds_words;           // LDS words per wave: the staged haplotype, then the record
  // The wave's LDS snapshot record (when lds_snap is set): lds_rec words into its LDS, one
  // snap_stride record per lane.  Snapshot entries and start slots marked kSnapLds use it.
  int32_t lds_snap;
  int32_t lds_rec;
  // Suffix sharing (plain main pass only, when `suffix` is set): the backward sweeps of
  // region g are bsweeps[region_bw[g] .. region_bw[g + 1]) with node writes in bwrites; node
  // records sit `node_off` elements into the wave's scratch, after its snapshot slots.
  const BwdSweep* bsweeps;
  const int32_t* bwrites;
  const int32_t* region_bw;
  int32_t suffix;
  int64_t node_off;
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
// UG (uniform gaps): the read's insertion, deletion and gap-continuation qualities are the same
// in every row, so mm, im, mi and md are one value per lane, kept at index 0 (g(k) is 0), and
// identity rows hold the read's values there too: their zero prior zeroes what mm and im scale,
// the rows above them hold no match value for mi and md to scale, and the one row above that
// does (read A's last row, above read B's first lane) reaches read B's identity row masked (see
// shift_up).  gc stays per row: identity rows keep D->D and I->I at 1.
//
// Scaled gap values (reads with uniform gaps and a gap continuation of at most 0.4, `xs`): a
// lane keeps x / mi in place of each insertion value x and y / md in place of each deletion value
// y of DP rows 1..n (PD: the insertion and deletion branch values alike, so the maxima of a value
// and its branch value keep the scale), so the insertion and deletion updates are single FMAs,
// fma(x_up, gc, m_up) and fma(y_left, gc, m_left), with no product (mi and md are the same in every
// row of such a read).  The match update scales the diagonal insertion value by ix = im * mi and
// the diagonal deletion value by im * md (held in im), and the last-row sum adds fma(x, ma, m) with
// ma = mi.  DP row 0 keeps its deletion value C/m unscaled (C/m / md could overflow); DP row 1,
// whose diagonal deletion value it is, divides its priors by md instead.  Elsewhere ix = im,
// ma = 1 and the priors stay, which give the unscaled values bit for bit.  Both builds keep the
// same values for a read: the uniform-gap build always scales, and the general build gives the
// rows of such reads mi = md = 1 (`scale_gaps`), since v * 1 is exact.  x / mi and y / md stay
// below m_max / (1 - gc) <= 1.7 m_max, so they cannot overflow; `yd` (md of a scaled read, else 1)
// unscales deletion values where the suffix join needs them.  The backward sweeps of suffix
// sharing run before the scaling, on the unscaled values.
template <typename T, int R, bool PD, bool UG = false> struct Rows {
  static constexpr int g(int k) { return UG ? 0 : k; }
  T pm[R], px[R], mm[UG ? 1 : R], im[UG ? 1 : R], mi[UG ? 1 : R], md[UG ? 1 : R], gc[R];
  T ix[UG ? 1 : R];
  T ma, yd;
  uint32_t rb[R];
  uint32_t ralt[R];
  // PD, staged sweep: the row's read-base class, the bit of the column word that says whether
  // the column accepts the row's base.
  uint32_t ci[PD ? R : 1];
  // Where the row's deletion-state bits sit in a PD column word: kRowShift (rows >= 2),
  // kRow1Shift (row 1) or kIdentShift (identity rows and lanes past the read).
  uint32_t sh[R];
};

// The read-base class of a read byte in a region whose other classes are `extra` (see
// pd_class_mask): A, C, G, T are 0-3, 'N' takes 0, the other bytes 4 on.
__device__ __forceinline__ uint32_t read_class(uint32_t b, const uint8_t* extra) {
  switch (b) {
    case 'A': case 'N': return 0u;
    case 'C': return 1u;
    case 'G': return 2u;
    case 'T': return 3u;
    default: break;
  }
  uint32_t c = 0u;
  if (extra)
    for (int i = 0; i < kPdExtraClasses; ++i) c = extra[i] == b && c == 0u ? 4u + i : c;
  return c;
}

// `extra` (PD, staged sweep) is the region's read-base classes from 4 on, or null.
template <typename T, int R, bool PD, bool UG = false>
__device__ __forceinline__ void load_rows(const SweepArgs& a, const DevTab<T>& tab, int64_t o,
                                          int L, int pad, int lane, Rows<T, R, PD, UG>& rw,
                                          const uint8_t* extra = nullptr) {
  if constexpr (UG) {  // the read's values, from its first base (zero in lanes past the reads)
    const bool live = lane < L;
    const int iq = dclamp(a.ins[o]), dq = dclamp(a.del[o]), gq = dclamp(a.gcp[o]);
    const T mm = tab.mm[((iq >= dq ? iq * (iq + 1) : dq * (dq + 1)) >> 1) + (iq >= dq ? dq : iq)];
    rw.mm[0] = live ? mm : T(0);
    rw.mi[0] = live ? tab.ep[iq] : T(0);
    rw.md[0] = live ? tab.ep[dq] : T(0);
    rw.im[0] = live ? tab.im[gq] : T(0);
    rw.ix[0] = rw.im[0];
  }
  rw.ma = T(1);
  rw.yd = T(1);
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
      if constexpr (!UG) {
        rw.mm[k] = tab.mm[((iq >= dq ? iq * (iq + 1) : dq * (dq + 1)) >> 1) + (iq >= dq ? dq : iq)];
        rw.mi[k] = tab.ep[iq];
        rw.md[k] = tab.ep[dq];
        rw.im[k] = tab.im[gq];
        rw.ix[k] = rw.im[k];
      }
      rw.gc[k] = tab.ep[gq];  // insertion and deletion extension alike (model.rs:73-75)
      rw.rb[k] = b;
      rw.ralt[k] = alt_bit(b);
      if constexpr (PD) rw.ci[k] = read_class(b, extra);
      rw.sh[k] = arow == 0 ? kRow1Shift : kRowShift;
    } else {
      // Identity rows: D->D 1 keeps row 0's C/m; I->I 1 is harmless because the row above an
      // identity row is zero.  Lanes past the read: all zero.
      rw.pm[k] = T(0); rw.px[k] = T(0);
      if constexpr (!UG) {
        rw.mm[k] = T(0); rw.im[k] = T(0); rw.mi[k] = T(0); rw.md[k] = T(0); rw.ix[k] = T(0);
      }
      rw.gc[k] = lane < L ? T(1) : T(0);
      rw.rb[k] = 0x100u;
      rw.ralt[k] = 0u;
      if constexpr (PD) rw.ci[k] = 0u;
      rw.sh[k] = kIdentShift;
    }
  }
}

// Whether this lane's rows of a plain read carry the insertion, deletion and gap-continuation
// quality bytes of the read's first base (true in lanes past the read).  A read scales its
// insertion values (see Rows) when every lane of it agrees and its gap-continuation quality is
// at least kXsMinGcp; the host applies the same test to pick the uniform-gap build.
template <int R>
__device__ __forceinline__ bool rows_match_first(const SweepArgs& a, int64_t o, int L, int pad,
                                                 int lane) {
  const uint32_t i0 = a.ins[o], d0 = a.del[o], g0 = a.gcp[o];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int arow = R * lane + k - pad;
    if (lane < L && arow >= 0) {
      const int64_t i = o + arow;
      ok = ok && a.ins[i] == i0 && a.del[i] == d0 && a.gcp[i] == g0;
    }
  }
  return ok;
}

// The rows of a read that scales its gap values (`xs`; always in the uniform-gap build) take
// their scaled parameters (see Rows); identity rows and lanes past the read keep theirs.  Called
// after the backward sweeps, which read the unscaled ones.
template <typename T, int R, bool PD, bool UG>
__device__ __forceinline__ void scale_gaps(Rows<T, R, PD, UG>& rw, bool xs) {
  if constexpr (UG) {
    const T md = rw.md[0];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (rw.sh[k] == kRow1Shift) {
        rw.pm[k] = rw.pm[k] / md;
        rw.px[k] = rw.px[k] / md;
      }
    }
    rw.ix[0] = rw.im[0] * rw.mi[0];
    rw.im[0] = rw.im[0] * md;
    rw.ma = rw.mi[0];
    rw.yd = md;
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      if (xs && rw.sh[k] != kIdentShift) {
        const T md = rw.md[k];
        if (rw.sh[k] == kRow1Shift) {
          rw.pm[k] = rw.pm[k] / md;
          rw.px[k] = rw.px[k] / md;
        }
        rw.ix[k] = rw.im[k] * rw.mi[k];
        rw.im[k] = rw.im[k] * md;
        rw.ma = rw.mi[k];
        rw.yd = md;
        rw.mi[k] = T(1);
        rw.md[k] = T(1);
      }
    }
  }
}

// Number of values a lane keeps per snapshot record (its rows' M, X, Y, the PD branch values and
// the running last-row sum), and the record stride: a record is 64 per-lane blocks of
// snap_stride values, so the one lane that records a column writes one contiguous block.  The
// row above a lane's first row is not kept: a restored lane takes it from the lane above with
// the same DPP move the sweep makes, which yields the value that lane recorded.
template <int R, bool PD> constexpr int snap_values() { return PD ? 6 * R + 1 : 3 * R + 1; }
template <int R, bool PD> constexpr int snap_stride() { return (snap_values<R, PD>() + 1) & ~1; }

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
  // PD: each row's diagonal maxima for its next column (see cells).
  T dm[PD ? R : 1], dx[PD ? R : 1], dy[PD ? R : 1];
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
// `keep_m` masks the match value the same way where rows take uniform gap values (Rows, UG);
// elsewhere it is all ones and folds away.
template <typename T, int R, bool PD>
__device__ __forceinline__ void shift_up(const Lane<T, R, PD>& s, Cell<T>& u, uint32_t keep,
                                         uint32_t keep_m = ~0u) {
  u.m = and_bits(shr1z(s.m[R - 1]), keep_m);
  u.x = and_bits(shr1z(s.x[R - 1]), keep);
  u.y = shr1z(s.y[R - 1]);
  if constexpr (PD) {
    // Masked like the insertion value: a DEL_END column lets the branch values above reach the
    // insertion value of the row below, identity rows included.
    u.bm = and_bits(shr1z(s.bm[R - 1]), keep);
    u.bx = and_bits(shr1z(s.bx[R - 1]), keep);
    u.by = shr1z(s.by[R - 1]);
  }
}

// Per-lane bitwise select: a where the mask is all ones, b where it is zero (v_bfi_b32).
__device__ __forceinline__ float sel_bits(uint32_t k, float a, float b) {
  const uint32_t x = __builtin_bit_cast(uint32_t, a), y = __builtin_bit_cast(uint32_t, b);
  return __builtin_bit_cast(float, (x & k) | (y & ~k));
}
__device__ __forceinline__ double sel_bits(uint32_t k, double a, double b) {
  const uint64_t m = (uint64_t(k) << 32) | k;
  const uint64_t x = __builtin_bit_cast(uint64_t, a), y = __builtin_bit_cast(uint64_t, b);
  return __builtin_bit_cast(double, (x & m) | (y & ~m));
}

// All ones if bit `b` of `w` is set, else zero.
__device__ __forceinline__ uint32_t bit_mask(uint32_t w, uint32_t b) {
  return uint32_t(__builtin_amdgcn_sbfe(int(w), b, 1));
}

// max of two non-negative, finite values.  For floats the bit patterns order as unsigned
// integers, and the integer maximum needs none of the canonicalising moves IEEE mode puts in
// front of v_max_f32 for operands the compiler cannot prove canonical.
__device__ __forceinline__ float max_nonneg(float a, float b) {
  return __builtin_bit_cast(float, max(__builtin_bit_cast(uint32_t, a), __builtin_bit_cast(uint32_t, b)));
}
__device__ __forceinline__ double max_nonneg(double a, double b) { return fmax(a, b); }

// The R cells of one column.  Rows are updated in place: match, deletion and branch values
// bottom-up (each row still reads the old column of the row above), then insertion values
// top-down (each row reads the new column of the row above), so no value is copied.
// CARRY (PD): the state after a DEL_END column is AFTER_DEL for every read row, so column j+1's
// after-deletion mask is column j's DEL_END mask for every row.  A row's diagonal maxima in
// column j+1 are then the up maxima it forms in column j, and the left maxima of the row above
// it are the same values: both are kept (s.dm, with Y included) from one step to the next, and
// only the lane's last row forms its left maxima afresh.  Identity rows and lanes past the read
// read other mask bits but hold zero branch values (or have a zero prior), where the masks do
// not matter; so are all values of column 0, where a sweep from the first column starts.  The
// sweep forms the diagonal maxima of its first column afresh.
// CLS (PD): the column word holds read-base classes (pd_class_mask) and a row's prior is chosen
// by its class bit; otherwise it holds the haplotype byte and alternates and bytes are compared.
template <typename T, int R, bool PD, bool CARRY = false, bool CLS = false, bool UG = false>
__device__ __forceinline__ void cells(Lane<T, R, PD>& s, const Rows<T, R, PD, UG>& rw,
                                      const Cell<T>& d, const Cell<T>& u) {
  const uint32_t w = s.w;
  const uint32_t hb = w & 0xFFu;
  const bool hN = hb == uint32_t('N');
  // PD: the column's DEL_END flag as a mask, and per row its deletion state as two masks
  // (inside, after; both zero in NORMAL columns), read at the row's shift; identity rows read
  // the word's constant bits and hold their values as inside a deletion (row 0, branch values
  // included).  All values are non-negative, so max(v, b & after) is v outside AFTER_DEL and
  // max(v, b) inside it, and the states need no compares or selects.
  const uint32_t de = PD ? bit_mask(w, kDelEndBit) : 0u;
#pragma unroll
  for (int k = R - 1; k >= 0; --k) {
    const T pdm = k ? s.m[k - 1] : d.m;
    const T pdx = k ? s.x[k - 1] : d.x;
    const T pdy = k ? s.y[k - 1] : d.y;
    // Bitwise ORs, no short circuit: a branch around the alternate-base test (a load) would turn
    // the prior below into a phi of two loads, which the optimiser folds into one load at a
    // selected offset of the row table, keeping the table in private memory.
    T prior;
    if constexpr (PD && CLS) {
      prior = sel_bits(bit_mask(w, rw.ci[k]), rw.pm[k], rw.px[k]);
    } else {
      bool match = hN | (hb == rw.rb[k]);
      if constexpr (PD) match = match | (((w >> 8) & rw.ralt[k]) != 0u);
      prior = match ? rw.pm[k] : rw.px[k];
    }
    if constexpr (!PD) {
      // kernel.rs:1011-1013 on scaled gap values (see Rows)
      T Y;
      if constexpr (UG)
        Y = fma(s.y[k], rw.gc[k], s.m[k]);
      else
        Y = fma(s.y[k], rw.gc[k], s.m[k] * rw.md[k]);
      s.m[k] = prior * fma(pdy, rw.im[rw.g(k)], fma(pdx, rw.ix[rw.g(k)], pdm * rw.mm[rw.g(k)]));
      s.y[k] = Y;
    } else {
      // pd_reference.rs:62-94 for the column's deletion state.
      const uint32_t I = bit_mask(w, rw.sh[k]), A = bit_mask(w, rw.sh[k] + 1);
      const T pdbm = k ? s.bm[k - 1] : d.bm;
      const T pdbx = k ? s.bx[k - 1] : d.bx;
      const T pdby = k ? s.by[k - 1] : d.by;
      T dme, dxe, dye;
      if constexpr (CARRY) {
        dme = s.dm[k];
        dxe = s.dx[k];
        dye = s.dy[k];
      } else {
        dme = max_nonneg(pdm, and_bits(pdbm, A));
        dxe = max_nonneg(pdx, and_bits(pdbx, A));
        dye = max_nonneg(pdy, and_bits(pdby, A));
      }
      T lme, lxe, lye;
      if (!CARRY || k == R - 1) {
        lme = max_nonneg(s.m[k], and_bits(s.bm[k], A));
        lxe = max_nonneg(s.x[k], and_bits(s.bx[k], A));
        lye = max_nonneg(s.y[k], and_bits(s.by[k], A));
      } else {  // the row below's diagonal maxima, formed from the same values last step
        lme = s.dm[k + 1];
        lxe = s.dx[k + 1];
        lye = s.dy[k + 1];
      }
      T Y;  // on scaled gap values (see Rows): the maxima keep the scale of their operands
      if constexpr (UG)
        Y = fma(lye, rw.gc[k], lme);
      else
        Y = fma(lye, rw.gc[k], lme * rw.md[k]);
      // NORMAL copies the row's values (lme is m there), AFTER_DEL takes the maxima, INSIDE_DEL
      // holds the branch.
      // Opaque to the optimiser: otherwise it distributes the next step's `branch & after` over
      // the select and keeps both halves of each select live as well.
      T BM = sel_bits(I, s.bm[k], lme);
      T BX = sel_bits(I, s.bx[k], lxe);
      T BY = sel_bits(I, s.by[k], lye);
      asm volatile("" : "+v"(BM), "+v"(BX), "+v"(BY));
      s.m[k] = prior * fma(dye, rw.im[rw.g(k)], fma(dxe, rw.ix[rw.g(k)], dme * rw.mm[rw.g(k)]));
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
      if constexpr (UG)
        s.x[k] = fma(pux, rw.gc[k], pum);  // scaled gap values (see Rows)
      else
        s.x[k] = fma(pux, rw.gc[k], pum * rw.mi[k]);
    } else {
      // pd_reference.rs:95-100
      const T pubm = k ? s.bm[k - 1] : u.bm;
      const T pubx = k ? s.bx[k - 1] : u.bx;
      const T ume = max_nonneg(pum, and_bits(pubm, de));
      const T uxe = max_nonneg(pux, and_bits(pubx, de));
      if constexpr (UG)
        s.x[k] = fma(uxe, rw.gc[k], ume);
      else
        s.x[k] = fma(uxe, rw.gc[k], ume * rw.mi[k]);
      if constexpr (CARRY) {
        const T puy = k ? s.y[k - 1] : u.y;
        const T puby = k ? s.by[k - 1] : u.by;
        s.dm[k] = ume;
        s.dx[k] = uxe;
        s.dy[k] = max_nonneg(puy, and_bits(puby, de));
      }
    }
  }
  s.acc += fma(s.x[R - 1], rw.ma, s.m[R - 1]);  // m + x, unscaling x (see Rows)
}

// Writes a lane's state into its snapshot record (value by value), in HBM or in LDS.
template <typename T, int R, bool PD>
__device__ __forceinline__ void store_snapshot(const Lane<T, R, PD>& s, T* d) {
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
  d[(NV - 1)] = s.acc;
}

// A staged column word carries, in bits 20..31, the record the column's state goes to: 0 none,
// kLdsField the wave's LDS record, else one plus the HBM scratch slot.  The word travels down
// the lanes with the column, so each lane sees the mark exactly when it has just computed that
// column.
constexpr uint32_t kSlotShift = 20;
constexpr uint32_t kLdsField = 4095;
constexpr int kMaxSlots = 4094;

// The column word a lane works on next: lane l-1's word moved down with DPP, or, in the first
// lane of each read, the word from LDS (lane 0 gets it as the move's old operand; `head` is all
// ones in the first lane of a second read).  Merged by bitwise ops on an opaque mask rather than a
// select, so the move always runs in every lane (see shift_up).
__device__ __forceinline__ uint32_t next_word(uint32_t from_lds, uint32_t w, uint32_t head) {
  return (shr1(from_lds, w) & ~head) | (from_lds & head);
}

// One sweep of a wave's reads over haplotype columns start+1 .. start+W.  `lw[t]` holds the word
// of column start+1+t (zero beyond W, 16-byte aligned).  Each read runs its own anti-diagonal:
// the lane `pos` lanes below the read's first lane computes column t - pos + 1 at step t, so a
// second read in the wave starts with the first instead of behind its last lane.  Every lane stops
// once it has computed column W, so it ends holding its rows of column W (a cut haplotype joins
// from there) and s.acc of a read's last lane is the raw scaled sum over the read's last row.
// `span` is the lane count of the wave's longest read.
// `snap_in` (HBM) or `lds_in` (the wave's LDS record) holds the start column's state when the
// sweep resumes from one; `lrec` is the wave's LDS record that columns marked kLdsField go to.
template <typename T, int R, bool PD, bool SNAP, bool UG = false>
__device__ void sweep(const Rows<T, R, PD, UG>& rw, const uint32_t* lw, int W, int span, int L,
                      int pos, int lane, uint32_t head, T init, const T* snap_in,
                      const T* lds_in, T scale, T* scratch, T* lrec, bool row0_above,
                      uint32_t keep, Lane<T, R, PD>& s) {
  constexpr int NV = snap_values<R, PD>();
  const uint32_t keep_m = UG ? keep : ~0u;  // see shift_up
  Cell<T> P{T(0), T(0), T(0), T(0), T(0), T(0)};  // diagonal neighbours of the first row
  if (snap_in || lds_in) {
    const T* p = (lds_in ? lds_in : snap_in) + lane * snap_stride<R, PD>();
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
    s.acc = p[(NV - 1)] * scale;
    // The row above the first row at the start column: the restored last row of the lane above,
    // moved and masked as the sweep moves it, which is the value the lane saw when it recorded.
    shift_up(s, P, keep, keep_m);
  } else {
    // Column 0: zeros below row 0; identity rows carry row 0's deletion value C/m.
#pragma unroll
    for (int k = 0; k < R; ++k) {
      s.m[k] = T(0);
      s.x[k] = T(0);
      s.y[k] = (rw.sh[k] == kIdentShift && lane < L) ? init : T(0);
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
  if constexpr (PD) {
    // The diagonal maxima of the first column, under that column's after-deletion masks.
    const uint32_t w0 = lw[0];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      const uint32_t A = bit_mask(w0, rw.sh[k] + 1);
      const T pm = k ? s.m[k - 1] : P.m, px = k ? s.x[k - 1] : P.x, py = k ? s.y[k - 1] : P.y;
      const T pbm = k ? s.bm[k - 1] : P.bm, pbx = k ? s.bx[k - 1] : P.bx;
      const T pby = k ? s.by[k - 1] : P.by;
      s.dm[k] = max_nonneg(pm, and_bits(pbm, A));
      s.dx[k] = max_nonneg(px, and_bits(pbx, A));
      s.dy[k] = max_nonneg(py, and_bits(pby, A));
    }
  }

  auto record = [&](const Cell<T>&) {
    if constexpr (SNAP) {
      const uint32_t sl = s.w >> kSlotShift;
      if (sl) {
        if (sl == kLdsField)
          store_snapshot<T, R, PD>(s, lrec + lane * snap_stride<R, PD>());
        else
          store_snapshot<T, R, PD>(s, scratch + (size_t(sl - 1) * 64 + lane) * snap_stride<R, PD>());
      }
    }
  };
  const int steps = W + span - 1;
  // Ramp: a lane starts at step `pos` and keeps its column-`start` state until then.  Rounded up
  // to four steps so the steady loop reads aligned groups of four column words.
  const int ramp = min(steps, (span - 1 + 3) & ~3);
  int t = 0;
  Cell<T> Q;
  for (; t < ramp; ++t) {
    s.w = next_word(lw[t], s.w, head);
    shift_up(s, Q, keep, keep_m);
    if (t >= pos && t - pos < W) {
      cells<T, R, PD, PD, PD>(s, rw, P, Q);
      P = Q;
      record(P);
    }
  }
  // Steady state: every lane of the reads is inside columns 1..W until the lanes at `pos` 0 pass
  // column W at step W; the diagonal and up neighbours alternate between P and Q.
  uint4 q = *reinterpret_cast<const uint4*>(lw + t);
  for (; t + 4 <= W; t += 4) {
    const uint4 cq = q;
    q = *reinterpret_cast<const uint4*>(lw + t + 4);
    s.w = next_word(cq.x, s.w, head);
    shift_up(s, Q, keep, keep_m);
    cells<T, R, PD, PD, PD>(s, rw, P, Q);
    record(Q);
    s.w = next_word(cq.y, s.w, head);
    shift_up(s, P, keep, keep_m);
    cells<T, R, PD, PD, PD>(s, rw, Q, P);
    record(P);
    s.w = next_word(cq.z, s.w, head);
    shift_up(s, Q, keep, keep_m);
    cells<T, R, PD, PD, PD>(s, rw, P, Q);
    record(Q);
    s.w = next_word(cq.w, s.w, head);
    shift_up(s, P, keep, keep_m);
    cells<T, R, PD, PD, PD>(s, rw, Q, P);
    record(P);
  }
  for (; t < steps; ++t) {
    s.w = next_word(lw[t], s.w, head);
    shift_up(s, Q, keep, keep_m);
    if (t >= pos && t - pos < W) {
      cells<T, R, PD, PD, PD>(s, rw, P, Q);
      P = Q;
      record(P);
    }
  }
}

// Stages the words of haplotype columns start+1 .. m into the wave's LDS, zero-padded, and marks
// the columns this sweep must record with their record (the wave's LDS record or an HBM slot).
template <bool PD>
__device__ __forceinline__ void stage_hap(const SweepArgs& a, const HapPlan& p, int start, int W,
                                          int L, int lane, uint32_t* lw, int nsnap) {
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
      const uint32_t field = (a.lds_snap && (e.slot & kSnapLds))
                                 ? kLdsField
                                 : uint32_t(e.slot & (kSnapLds - 1)) + 1;
      if (e.col - start <= W) lw[e.col - start - 1] |= field << kSlotShift;
    }
  }
  wave_sync_lds();
}

// ---------------------------------------------------------------------------------------------
// Suffix sharing (plain, FP32 and FP64): backward sweeps and the join (kernel.rs:579-602,
// 744-891, 1029-1126).  B_M, B_X, B_Y of a cell are the scaled mass of the paths from that cell to
// the end of the read (2^120 per path that ends in FP32, 1 in FP64: simd.rs:109, 134).  The backward recurrence mirrors the
// forward one: row i at column j reads row i+1 at column j+1 (match, through the prior and
// transitions of read position i+1), row i+1 at column j (insertion) and row i at column j+1
// (deletion).  In the lane layout rows run bottom-up inside a lane and the row below a lane's
// last row arrives from lane l+1 by DPP wave_shl:1; columns run right to left, each read's last
// lane first.  The identity row that stands for DP row 0 keeps D->D 1, so its deletion value
// becomes the running sum S of the weights of paths that start in row 0 at or after the column
// (kernel.rs:1104-1126); identity rows above it stay zero.

// A node record: per lane its rows' B_M and B_Y at the node's column and B_M of the row below
// the lane's last row there (the next column's diagonal), padded to four values.
template <int R> constexpr int node_stride() { return (2 * R + 1 + 3) & ~3; }

// Parameters of the row below a lane's last row (lane l+1's first row), zero for a read's last
// lane (nothing lies below its last row).
template <typename T> struct EdgeRow {
  T pm, px, mm, im, mi, gc;
  uint32_t rb;
};

// Stages the words of backward sweep `b` into the wave's LDS: word t is for column right-1-t and
// holds the haplotype base of column right-t (the column its match moves into; none for the
// last column) and, at the columns the sweep records, the node record index plus one.
__device__ __forceinline__ void stage_bwd(const SweepArgs& a, const BwdSweep& b, const HapPlan& h,
                                          int L, int lane, uint32_t* lw) {
  const int Wb = b.right - b.left;
  const int total = Wb + L + 8;
  wave_sync_lds();
  for (int i = lane; i < total; i += 64) {
    uint32_t v = 0x100u;  // no base: matches nothing
    const int j = b.right - 1 - i;
    if (i < Wb && j < h.len) v = uint32_t(a.hap_bytes[h.off + j]);
    lw[i] = v;
  }
  wave_sync_lds();
  for (int i = lane; i < b.nw; i += 64) {
    const int col = a.bwrites[2 * (b.w0 + i)], node = a.bwrites[2 * (b.w0 + i) + 1];
    lw[b.right - 1 - col] |= uint32_t(node + 1) << kSlotShift;
  }
  wave_sync_lds();
}

// One backward sweep of Wb columns.  A lane starts `delay` steps after its read's last lane (the
// `bottom` lanes read the words from LDS; the others take them from the lane below).  `sink` is
// 2^120 in a read's last lane (whose last register holds the read's last row), zero elsewhere.
// Starts from node record `rec_in` (this lane's values at column `right`) or from zeros.
// `from_lds` is all ones in the bottom lanes, zero elsewhere, and opaque to the optimiser: words
// are merged with bit operations rather than a select, which the compiler could turn into a DPP
// move run only in the lanes that keep its result (a DPP move reads a lane outside EXEC as zero).
// The state a lane carries through a backward sweep: its rows' B_M, B_X, B_Y at the column it
// computed last, B_M of the row below its last row one column further right, and the word of
// the column it is on.
template <typename T, int R> struct BLane {
  T bm[R], bx[R], by[R];
  T P;
  uint32_t w;
};

// One backward column of a lane's rows, bottom-up: row k reads row k+1 at this column (insertion,
// new values) and one column to the right (match, the old value), and itself one column to the
// right (deletion); the row below the last row is lane l+1's first row (`qm`, `qx` this column,
// P the previous one), with parameters `eg`.  Records a node when the word carries one.
template <typename T, int R, bool UG>
__device__ __forceinline__ void bcells(BLane<T, R>& s, const Rows<T, R, false, UG>& rw,
                                       const EdgeRow<T>& eg, T qm, T qx, T sink, T* nodes,
                                       int lane) {
  const uint32_t hb = s.w & 0x1FFu;
  const bool hN = hb == uint32_t('N');
  T below_x = qx, diag = s.P;
#pragma unroll
  for (int k = R - 1; k >= 0; --k) {
    const bool edge = k == R - 1;
    const int kb = edge ? k : k + 1;
    const T pmb = edge ? eg.pm : rw.pm[kb];
    const T pxb = edge ? eg.px : rw.px[kb];
    const uint32_t rbb = edge ? eg.rb : rw.rb[kb];
    const T mmb = edge ? eg.mm : rw.mm[rw.g(kb)];
    const T imb = edge ? eg.im : rw.im[rw.g(kb)];
    const T mib = edge ? eg.mi : rw.mi[rw.g(kb)];
    const T iib = edge ? eg.gc : rw.gc[kb];
    const bool match = hN | (hb == rbb);
    const T into = (match ? pmb : pxb) * diag;  // kernel.rs:1071-1082
    const T ind = into * imb;
    T nbx = fma(below_x, iib, ind);
    const T nby = fma(s.by[k], rw.gc[k], ind);
    T nbm = fma(s.by[k], rw.md[rw.g(k)], fma(below_x, mib, into * mmb));
    if (edge) {
      nbm += sink;
      nbx += sink;
    }
    diag = s.bm[k];  // row k one column to the right: the diagonal of row k-1
    s.bm[k] = nbm;
    s.bx[k] = nbx;
    s.by[k] = nby;
    below_x = nbx;
  }
  s.P = qm;
  const uint32_t node = s.w >> kSlotShift;
  if (node) {
    T* r = nodes + (size_t(node - 1) * 64 + lane) * node_stride<R>();
#pragma unroll
    for (int k = 0; k < R; ++k) {
      r[k] = s.bm[k];
      r[R + k] = s.by[k];
    }
    r[2 * R] = s.P;
  }
}

// One backward sweep of Wb columns.  A lane starts `delay` steps after its read's last lane: the
// read's last lanes take the words from LDS (`from_lds` all ones there, zero elsewhere, opaque to
// the optimiser: words are merged with bit operations rather than a select, which the compiler
// could turn into a DPP move run only in the lanes that keep its result, and a DPP move reads a
// lane outside EXEC as zero); the others take them from the lane below.  `sink` is 2^120 in a
// read's last lane (whose last register holds the read's last row), zero elsewhere.  Starts from
// node record `rec_in` (this lane's values at column `right`) or from zeros.  Lanes past the
// reads (lane >= L) start from zeros: their records were never written, and the upward DPP moves
// lane L's values into lane L-1, where a stale NaN times the zero parameters below a read's last
// row would still be NaN.  `full` is the step from which every lane of the item is inside the
// sweep (the largest delay).
template <typename T, int R, bool UG>
__device__ void bsweep(const Rows<T, R, false, UG>& rw, const EdgeRow<T>& eg, const uint32_t* lw,
                       int Wb, int steps, int full, int lane, int L, int delay, uint32_t from_lds,
                       T sink, const T* rec_in, T* nodes) {
  constexpr int NS = node_stride<R>();
  BLane<T, R> s;
  s.P = T(0);
  s.w = 0;
  if (rec_in) {
    const T* r = rec_in + lane * NS;
    const bool live = lane < L;
#pragma unroll
    for (int k = 0; k < R; ++k) {
      s.bm[k] = live ? r[k] : T(0);
      s.by[k] = live ? r[R + k] : T(0);
      s.bx[k] = T(0);
    }
    s.P = live ? r[2 * R] : T(0);
  } else {
#pragma unroll
    for (int k = 0; k < R; ++k) s.bm[k] = s.bx[k] = s.by[k] = T(0);
  }
  auto step = [&](uint32_t word, bool live) {
    s.w = (shl1z(s.w) & ~from_lds) | (word & from_lds);
    const T qm = shl1z(s.bm[0]), qx = shl1z(s.bx[0]);  // the row below, this column
    if (live) bcells<T, R>(s, rw, eg, qm, qx, sink, nodes, lane);
  };
  // Ramp: lanes start one after the other; rounded up to four steps so the steady loop reads
  // aligned groups of four words.
  const int ramp = min(steps, (full + 3) & ~3);
  int t = 0;
  for (; t < ramp; ++t) step(lw[t], t >= delay && t - delay < Wb);
  // Steady state: every lane of the item is inside the sweep until its last lanes pass column
  // `left` (they have delay 0).
  for (; t + 4 <= Wb; t += 4) {
    const uint4 q = *reinterpret_cast<const uint4*>(lw + t);
    step(q.x, true);
    step(q.y, true);
    step(q.z, true);
    step(q.w, true);
  }
  for (; t < steps; ++t) step(lw[t], t >= delay && t - delay < Wb);
}

// Waves per SIMD the occupancy-raised FP32 variants ask the compiler for: one more than the
// register use of the plain build allows (plain R=1..8 use 69..164 VGPRs, 7..3 waves; PD
// 87..213, 5..2 waves, of 512 per SIMD lane).  The raised variant may spill; the autotuner
// times both and keeps the faster per bucket.
constexpr int boost_waves(int R, bool PD) {
  return PD ? (R <= 1 ? 6 : R <= 3 ? 5 : R <= 5 ? 4 : 3)
            : (R <= 1 ? 8 : R <= 2 ? 7 : R <= 3 ? 6 : R <= 5 ? 5 : 4);
}

// BOOST selects the occupancy-raised variant (FP32 only).  SFX compiles the suffix-sharing path
// (backward sweeps and joins); FP64 has a build without it, whose registers are not raised by
// that path, for the runs that share no suffixes (the rescue of FP32 pairs among them).
template <typename T, int R, bool PD, bool BOOST, bool SFX, bool UG = false>
__global__ void __launch_bounds__(256, BOOST ? boost_waves(R, PD) : 1)
    sweep_kernel(SweepArgs a, DevTab<T> tab) {
  extern __shared__ __attribute__((aligned(16))) uint32_t lds_all[];
  const int lane = threadIdx.x & 63;
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t wave_global = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave_in_block;
  uint32_t* lw = lds_all + wave_in_block * a.lds_words;
  T* lrec = reinterpret_cast<T*>(lw + a.lds_rec);
  T* scratch = a.scratch ? static_cast<T*>(a.scratch) + wave_global * a.scratch_wave_stride : nullptr;
  const int n_items = a.n_items_dev ? *a.n_items_dev : a.n_items;

  for (;;) {
    // Every lane adds one, so a claim advances the counter by 64 and every lane's old value lies
    // in [64 i, 64 i + 63].  A lane-dependent branch around the claim, next to the lane-0 stores
    // at the end of an item, would let the optimiser thread the loop's back edge past the claim
    // for lanes 1..63, which would then read an item without lane 0.
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= n_items) break;

    // Everything below is uniform across the wave; readfirstlane keeps it in scalar registers.
    // A main-pass item is one read, or two reads of one region sharing the wave (read A in lanes
    // [0, La), read B in [La, La+Lb)); a pair-mode item is one (read, sorted haplotype k), or,
    // with the sign bit set, two reads a < b of one region on haplotype k, encoded
    // (a, k | (b - a) << 16 | sign).
    int ra, rb, k_begin, k_end;
    ra = __builtin_amdgcn_readfirstlane(a.items[2 * item]);
    if (a.pair_mode) {
      const int v = __builtin_amdgcn_readfirstlane(a.items[2 * item + 1]);
      rb = v < 0 ? ra + ((v >> 16) & 0x7FFF) : -1;
      k_begin = v < 0 ? (v & 0xFFFF) : v;
      k_end = k_begin + 1;
      if (a.dense && __builtin_amdgcn_readfirstlane(a.dense[ra])) continue;
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
    Rows<T, R, PD, UG> rw;
    load_rows<T, R, PD, UG>(a, tab, in_b ? ob : oa, in_b ? Lb : La, in_b ? padb : pada,
                        in_b ? lane - La : lane, rw,
                        PD ? a.region_classes + int64_t(g) * kPdExtraClasses : nullptr);
    // Each read's second lane starts below DP row 0 when identity rows fill its first lane.
    const bool row0_above = (lane == 1 && pada == R) || (two && lane == La + 1 && padb == R);
    // Zero in the first lane of read B, all ones elsewhere; opaque to the optimiser (see
    // shift_up).
    uint32_t keep = (two && lane == La) ? 0u : ~0u;
    asm volatile("" : "+v"(keep));
    // Each read runs its own anti-diagonal (see sweep): a lane's offset below its read's first
    // lane, the first lane of read B taking the column word from LDS, and the longest read's lanes.
    const int pos = in_b ? lane - La : lane;
    const uint32_t head = ~keep;
    const int span = two ? max(La, Lb) : La;

    // Suffix sharing: the region's backward sweeps, once per item, before its forward sweeps.
    constexpr bool kSuffix = SFX && !PD;
    const bool suffix = kSuffix && a.suffix && !a.pair_mode;
    T* nodes = scratch ? scratch + a.node_off : nullptr;
    if constexpr (kSuffix) {
      if (suffix) {
        const bool bottom = lane == La - 1 || (two && lane == L - 1);
        const int delay = lane < La ? La - 1 - lane : (lane < L ? L - 1 - lane : 1 << 20);
        const T sink = bottom ? T(Num<T>::kBwdScale) : T(0);
        // All ones in the read's last lanes; opaque (see shift_up and bsweep).
        uint32_t from_lds = bottom ? ~0u : 0u;
        asm volatile("" : "+v"(from_lds));
        EdgeRow<T> eg;  // lane l+1's first row, zero below a read's last lane
        eg.pm = and_bits(shl1z(rw.pm[0]), ~from_lds);
        eg.px = and_bits(shl1z(rw.px[0]), ~from_lds);
        eg.mm = and_bits(shl1z(rw.mm[0]), ~from_lds);
        eg.im = and_bits(shl1z(rw.im[0]), ~from_lds);
        eg.mi = and_bits(shl1z(rw.mi[0]), ~from_lds);
        eg.gc = and_bits(shl1z(rw.gc[0]), ~from_lds);
        eg.rb = shl1z(rw.rb[0]) & ~from_lds;
        const int steps_extra = max(La, Lb) - 1;
        const int w1 = __builtin_amdgcn_readfirstlane(a.region_bw[g + 1]);
        for (int w = __builtin_amdgcn_readfirstlane(a.region_bw[g]); w < w1; ++w) {
          BwdSweep b = a.bsweeps[w];
          b.hap = __builtin_amdgcn_readfirstlane(b.hap);
          b.right = __builtin_amdgcn_readfirstlane(b.right);
          b.left = __builtin_amdgcn_readfirstlane(b.left);
          b.start_node = __builtin_amdgcn_readfirstlane(b.start_node);
          b.w0 = __builtin_amdgcn_readfirstlane(b.w0);
          b.nw = __builtin_amdgcn_readfirstlane(b.nw);
          HapPlan h = a.plan[b.hap];
          h.len = __builtin_amdgcn_readfirstlane(h.len);
          stage_bwd(a, b, h, L, lane, lw);
          const int Wb = b.right - b.left;
          bsweep<T, R>(rw, eg, lw, Wb, Wb + steps_extra, steps_extra, lane, L, delay, from_lds, sink,
                    b.start_node >= 0 ? nodes + size_t(b.start_node) * 64 * node_stride<R>()
                                      : nullptr,
                    nodes);
        }
      }
    }

    // Reads that scale their gap values (see Rows; every read of the uniform-gap build) take
    // their scaled parameters, after the backward sweeps, which use the unscaled ones.
    {
      bool xs = true;
      if constexpr (!UG) {
        const bool ok = rows_match_first<R>(a, in_b ? ob : oa, in_b ? Lb : La,
                                            in_b ? padb : pada, in_b ? lane - La : lane);
        const uint64_t bad = __ballot(!ok);
        const uint64_t lanes_a = La >= 64 ? ~0ull : (1ull << La) - 1;
        const bool xs_a = !(bad & lanes_a) && a.gcp[oa] >= kXsMinGcp;
        const bool xs_b = two && !(bad & ~lanes_a) && a.gcp[ob] >= kXsMinGcp;
        xs = in_b ? xs_b : xs_a;
      }
      scale_gaps(rw, xs);
    }

    double prev_a = 0.0, prev_b = 0.0;
    int lost_a = 0, lost_b = 0;
    for (int k = k_begin; k < k_end; ++k) {
      HapPlan p = a.plan[hap0 + k];
      p.len = __builtin_amdgcn_readfirstlane(p.len);
      p.lcp = __builtin_amdgcn_readfirstlane(p.lcp);
      p.dup = __builtin_amdgcn_readfirstlane(p.dup);
      p.nsnap = __builtin_amdgcn_readfirstlane(p.nsnap);
      p.snap0 = __builtin_amdgcn_readfirstlane(p.snap0);
      p.start_slot = __builtin_amdgcn_readfirstlane(p.start_slot);
      p.cut = __builtin_amdgcn_readfirstlane(p.cut);
      p.cut_node = __builtin_amdgcn_readfirstlane(p.cut_node);
      double res_a, res_b;
      if (!a.pair_mode && p.dup) {
        res_a = prev_a;  // kernel.rs:237-246
        res_b = prev_b;
      } else {
        const bool shared = a.share && !a.pair_mode;
        const int start = shared ? p.lcp : 0;
        const int nsnap = shared ? p.nsnap : 0;
        const T* snap_in = nullptr;
        const T* lds_in = nullptr;
        T scale = T(1);
        if (shared && p.start_slot >= 0) {
          if (a.lds_snap && (p.start_slot & kSnapLds))
            lds_in = lrec;
          else
            snap_in = scratch + size_t(p.start_slot & (kSnapLds - 1)) * 64 * snap_stride<R, PD>();
          scale = T(double(p.start_src_len) / double(p.len));  // kernel.rs:641
        }
        const T init = T(Num<T>::kScale / double(p.len));      // kernel.rs:640
        // A cut haplotype sweeps up to the column before its cut (kernel.rs:251-252).
        const bool cut = suffix && p.cut > 0;
        const int W = (cut ? p.cut - 1 : p.len) - start;
        stage_hap<PD>(a, p, start, W, L, lane, lw, nsnap);
        Lane<T, R, PD> st;
        if (nsnap)
          sweep<T, R, PD, true>(rw, lw, W, span, L, pos, lane, head, init, snap_in, lds_in, scale,
                                scratch, lrec, row0_above, keep, st);
        else
          sweep<T, R, PD, false>(rw, lw, W, span, L, pos, lane, head, init, snap_in, lds_in,
                                 scale, scratch, lrec, row0_above, keep, st);
        // A read's last row is its last lane's last register (kernel.rs:897-914).
        // A joined pair is held to kMinAcceptedJoined (kernel.rs:261-275).
        auto finish = [cut](double raw) {
          const double lo = cut ? Num<T>::kMinAcceptedJoined : Num<T>::kMinAccepted;
          if (lo > 0.0 && !(raw >= lo)) return __builtin_nan("");
          return log10(raw) - Num<T>::kLog10Scale;
        };
        double raw_a = double(readlane(st.acc, La - 1));
        double raw_b = double(readlane(st.acc, L - 1));
        if constexpr (kSuffix) {
          if (cut) {
            // The join (kernel.rs:744-791): the values entering the cut column from column W
            // (match from the diagonal, deletion from the left), times the cut node's backward
            // values, summed in FP64 per read; DP row 0's identity row contributes C/m times
            // the start-weight sum S through its deletion value.
            // Each DP row's term goes to LDS at its row index (read A's rows 0..na, then read
            // B's), is summed in chunks of eight rows, and the chunks in order: the grouping
            // depends on row numbers only, not on how the rows sit in lanes, so a read gets the
            // same bits whatever it shares its wave with and at any rows per lane.
            const T um = shr1z(st.m[R - 1]), ux = shr1z(st.x[R - 1]), uy = shr1z(st.y[R - 1]);
            const uint32_t hc = uint32_t(a.hap_bytes[p.off + p.cut - 1]);
            const T* rec = nodes + (size_t(p.cut_node) * 64 + lane) * node_stride<R>();
            double* rows = reinterpret_cast<double*>(lw);   // [na + 1 + nb + 1]
            double* chunks = rows + 64 * R;  // after the 64 x R rows a wave holds
            const int base = in_b ? na + 1 : 0;
            const int first = R * (in_b ? lane - La : lane) - (in_b ? padb : pada);  // row - 1
            // Opaque per join, so the optimiser keeps no per-row factors live across the sweeps.
            T yd = rw.yd;
            asm volatile("" : "+v"(yd));
            wave_sync_lds();
#pragma unroll
            for (int k = 0; k < R; ++k) {
              const T dm = k ? st.m[k - 1] : um;
              const T dx = k ? st.x[k - 1] : ux;
              const T dy = k ? st.y[k - 1] : uy;
              const bool match = (hc == uint32_t('N')) | (hc == rw.rb[k]);
              const T jm = (match ? rw.pm[k] : rw.px[k]) *
                           fma(dy, rw.im[rw.g(k)], fma(dx, rw.ix[rw.g(k)], dm * rw.mm[rw.g(k)]));
              // The deletion value entering the cut column, unscaled by yd in scaled rows (DP
              // row 0 keeps C/m unscaled; see Rows).
              const T jy = UG ? fma(st.y[k], rw.gc[k], st.m[k])
                              : fma(st.y[k], rw.gc[k], st.m[k] * rw.md[rw.g(k)]);
              const double f = rw.sh[k] == kIdentShift ? 1.0 : double(yd);
              const double term =
                  fma(double(jm), double(rec[k]), double(jy) * f * double(rec[R + k]));
              const int row = first + k + 1;  // DP row; rows above DP row 0 hold nothing
              if (lane < L && row >= 0) rows[base + row] = term;
            }
            wave_sync_lds();
            const int ca = (na + 1 + 7) >> 3, cb = two ? (nb + 1 + 7) >> 3 : 0;
            for (int c = lane; c < ca + cb; c += 64) {
              const int r0 = c < ca ? 8 * c : na + 1 + 8 * (c - ca);
              const int r1 = c < ca ? min(8 * c + 8, na + 1) : min(r0 + 8, na + 1 + nb + 1);
              double sum = 0.0;
              for (int r = r0; r < r1; ++r) sum += rows[r];
              chunks[c] = sum;
            }
            wave_sync_lds();
            double sa = 0.0, sb = 0.0;
            for (int c = 0; c < ca; ++c) sa += chunks[c];
            for (int c = ca; c < ca + cb; ++c) sb += chunks[c];
            wave_sync_lds();
            raw_a += sa * (1.0 / Num<T>::kBwdScale);
            raw_b += sb * (1.0 / Num<T>::kBwdScale);
          }
        }
        res_a = finish(raw_a);
        res_b = finish(raw_b);
      }
      lost_a += res_a != res_a;
      lost_b += res_b != res_b;
      if (lane == 0) {
        double* oa = a.out + out_a + p.src;
        if (!a.nan_only || *oa != *oa) *oa = res_a;
        if (a.rescue && res_a != res_a) {
          const int q = a.rescue_by_rows ? (na + 64) / 64 - 1 : 0;
          const int idx = atomicAdd(a.rescue_count + q, 1);
          a.rescue_pairs[q][2 * idx] = ra;
          a.rescue_pairs[q][2 * idx + 1] = k;
        }
        if (two) {
          double* ob = a.out + out_b + p.src;
          if (!a.nan_only || *ob != *ob) *ob = res_b;
          if (a.rescue && res_b != res_b) {
            const int q = a.rescue_by_rows ? (nb + 64) / 64 - 1 : 0;
            const int idx = atomicAdd(a.rescue_count + q, 1);
            a.rescue_pairs[q][2 * idx] = rb;
            a.rescue_pairs[q][2 * idx + 1] = k;
          }
        }
      }
      prev_a = res_a;
      prev_b = res_b;
    }
    // A main-pass item has run its reads against every haplotype: mark the dense ones, listed
    // by their own rows per lane (a read paired into a wider bucket reruns alone at its own).
    if (a.dense && !a.pair_mode && lane == 0) {
      const bool da = lost_a > 0 && 3 * lost_a >= H;
      a.dense[ra] = da;
      if (da) {
        const int q = (na + 64) / 64 - 1;
        const int idx = atomicAdd(a.dense_count + q, 1);
        a.dense_items[q][2 * idx] = ra;
        a.dense_items[q][2 * idx + 1] = -1;
      }
      if (two) {
        const bool db = lost_b > 0 && 3 * lost_b >= H;
        a.dense[rb] = db;
        if (db) {
          const int q = (nb + 64) / 64 - 1;
          const int idx = atomicAdd(a.dense_count + q, 1);
          a.dense_items[q][2 * idx] = rb;
          a.dense_items[q][2 * idx + 1] = -1;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------------------------
// Pairs beyond the staged layout: reads of more than 511 bases, haplotypes longer than the LDS
// staging holds, and every other pair of their regions (which then share nothing).  One
// (read, haplotype) pair per wave, from column 0.  The read's rows are laid out as in sweep_kernel
// at 8 rows per lane over V = ceil((n + 1) / 8) virtual lanes (identity rows at the top of
// virtual lane 0), and the wave sweeps them 64 virtual lanes (one strip) at a time, top strip
// first.  A strip's last lane stores its last row's values at every column into the wave's
// boundary buffer in HBM; the next strip's lane 0 takes them as the row above its first row.
// Lane 0's inputs (the haplotype column word, the boundary row) are loaded 64 columns at a time,
// one column per lane, and handed to lane 0 with readlane, so neither length is bounded by LDS.

// Values of one boundary column: M, X, Y and (PD) the branch values.
template <bool PD> constexpr int bnd_values() { return PD ? 6 : 3; }

// The row above lane 0's first row comes from `above` (uniform); every other lane's from the lane
// before it, as shift_up moves it.
template <typename T, int R, bool PD>
__device__ __forceinline__ void shift_up_from(const Lane<T, R, PD>& s, Cell<T>& u,
                                              const Cell<T>& above) {
  u.m = shr1(above.m, s.m[R - 1]);
  u.x = shr1(above.x, s.x[R - 1]);
  u.y = shr1(above.y, s.y[R - 1]);
  if constexpr (PD) {
    u.bm = shr1(above.bm, s.bm[R - 1]);
    u.bx = shr1(above.bx, s.bx[R - 1]);
    u.by = shr1(above.by, s.by[R - 1]);
  }
}

template <typename T, bool PD>
__global__ void __launch_bounds__(64) long_kernel(SweepArgs a, DevTab<T> tab) {
  constexpr int R = kMaxRowsPerLane;
  constexpr int NB = bnd_values<PD>();
  const int lane = threadIdx.x & 63;
  const int64_t wave_global = int64_t(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  T* const bnd = static_cast<T*>(a.scratch) + wave_global * a.scratch_wave_stride;
  const int64_t bnd_half = a.scratch_wave_stride / 2;  // one boundary buffer: (W + 1) x NB
  const int n_items = a.n_items_dev ? *a.n_items_dev : a.n_items;

  for (;;) {
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= n_items) break;
    const int ra = __builtin_amdgcn_readfirstlane(a.items[2 * item]);
    const int k = __builtin_amdgcn_readfirstlane(a.items[2 * item + 1]);
    const int64_t o = a.read_off[ra];
    const int n = __builtin_amdgcn_readfirstlane(int(a.read_off[ra + 1] - o));
    const int g = __builtin_amdgcn_readfirstlane(a.read_region[ra]);
    const int hap0 = __builtin_amdgcn_readfirstlane(a.region_hap0[g]);
    const int H = __builtin_amdgcn_readfirstlane(a.region_nhap[g]);
    const int64_t out_idx = a.region_out[g] + int64_t(a.read_local[ra]) * H;
    HapPlan p = a.plan[hap0 + k];
    const int W = __builtin_amdgcn_readfirstlane(p.len);
    const int64_t hoff = p.off;
    const int src = __builtin_amdgcn_readfirstlane(p.src);
    const int V = (n + R) / R;        // virtual lanes: n rows plus at least one identity row
    const int pad = R * V - n;        // identity rows at the top of virtual lane 0
    const int S = (V + 63) / 64;      // strips
    const T init = T(Num<T>::kScale / double(W));  // kernel.rs:640

    T* bin = bnd;
    T* bout = bnd + bnd_half;
    double raw = 0.0;
    for (int st = 0; st < S; ++st) {
      const int vl = 64 * st + lane;
      Rows<T, R, PD> rw;
      load_rows<T, R, PD>(a, tab, o, V, pad, vl, rw);
      const int L = min(64, V - 64 * st);  // live lanes of this strip
      const bool last_strip = st == S - 1;
      Lane<T, R, PD> s;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        s.m[r] = T(0);
        s.x[r] = T(0);
        // Column 0: identity rows carry row 0's deletion value C/m; real rows are zero.
        s.y[r] = (rw.sh[r] == kIdentShift && vl < V) ? init : T(0);
        if constexpr (PD) {
          s.bm[r] = T(0);
          s.bx[r] = T(0);
          s.by[r] = T(0);
        }
      }
      s.acc = T(0);
      s.w = 0;
      Cell<T> P{T(0), T(0), T(0), T(0), T(0), T(0)};
      // Identity rows fill virtual lane 0: read row 1's diagonal in column 1 is DP row 0's C/m.
      if (st == 0 && lane == 1 && pad == R) P.y = init;
      const int steps = W + L - 1;
      // Lane l's loads for a chunk starting at step cb: the word of column cb + l + 1 and the
      // boundary row at that column (zero past column W and in the top strip).
      auto load_word = [&](int cb) -> uint32_t {
        const int j = cb + lane;  // 0-based column index
        if (j >= W) return 0u;
        return PD ? a.hap_words[hoff + j] : uint32_t(a.hap_bytes[hoff + j]);
      };
      auto load_bnd = [&](int cb, T* v) {
        const int c = cb + lane + 1;
#pragma unroll
        for (int q = 0; q < NB; ++q) v[q] = (st > 0 && c <= W) ? bin[int64_t(c) * NB + q] : T(0);
      };
      uint32_t wv_next = load_word(0);
      T bv_next[NB];
      load_bnd(0, bv_next);
      for (int cb = 0; cb < steps; cb += 64) {
        const uint32_t wv = wv_next;
        T bv[NB];
#pragma unroll
        for (int q = 0; q < NB; ++q) bv[q] = bv_next[q];
        wv_next = load_word(cb + 64);
        load_bnd(cb + 64, bv_next);
        const int nsteps = min(64, steps - cb);
        for (int i = 0; i < nsteps; ++i) {
          const int t = cb + i;
          s.w = shr1(uint32_t(__builtin_amdgcn_readlane(int(wv), i)), s.w);
          Cell<T> above{T(0), T(0), T(0), T(0), T(0), T(0)};
          above.m = readlane(bv[0], i);
          above.x = readlane(bv[1], i);
          above.y = readlane(bv[2], i);
          if constexpr (PD) {
            above.bm = readlane(bv[3], i);
            above.bx = readlane(bv[4], i);
            above.by = readlane(bv[5], i);
          }
          Cell<T> Q;
          shift_up_from(s, Q, above);
          const int c = t - lane + 1;  // the column this lane computes at step t
          if (c >= 1 && c <= W) {
            cells(s, rw, P, Q);
            P = Q;
            // A strip's last lane hands its last row at column c to the next strip.
            if (!last_strip && lane == 63) {
              T* d = bout + int64_t(c) * NB;
              d[0] = s.m[R - 1];
              d[1] = s.x[R - 1];
              d[2] = s.y[R - 1];
              if constexpr (PD) {
                d[3] = s.bm[R - 1];
                d[4] = s.bx[R - 1];
                d[5] = s.by[R - 1];
              }
            }
          }
        }
      }
      if (last_strip) raw = double(readlane(s.acc, L - 1));
      // The next strip reads what this one's last lane stored.
      wave_sync_global();
      T* tmp = bin;
      bin = bout;
      bout = tmp;
    }
    double res;
    if (Num<T>::kMinAccepted > 0.0 && !(raw >= Num<T>::kMinAccepted))
      res = __builtin_nan("");  // kernel.rs:897-914
    else
      res = log10(raw) - Num<T>::kLog10Scale;
    if (lane == 0) {
      a.out[out_idx + src] = res;
      if (a.rescue && res != res) {
        const int idx = atomicAdd(a.rescue_count, 1);
        a.rescue_pairs[0][2 * idx] = ra;
        a.rescue_pairs[0][2 * idx + 1] = k;
      }
    }
  }
}

using KernelF = void (*)(SweepArgs, DevTab<float>);
using KernelD = void (*)(SweepArgs, DevTab<double>);

// Every instantiation the host launches: [PD][BOOST][R-1] for FP32 (sweep_f32.hip), [PD][R-1]
// for FP64 without suffix sharing and [R-1] for plain FP64 with it (sweep_f64.hip); the
// uniform-gap builds by [BOOST][R-1] (FP32 plain and PD), [SFX][R-1] (FP64 plain) and [R-1]
// (FP64 PD).
extern const KernelF kSweepF32[2][2][kMaxRowsPerLane];
extern const KernelF kSweepF32Uniform[2][kMaxRowsPerLane];
extern const KernelF kSweepF32PdUniform[2][kMaxRowsPerLane];
extern const KernelD kSweepF64PdUniform[kMaxRowsPerLane];
extern const KernelD kSweepF64[2][kMaxRowsPerLane];
extern const KernelD kSweepF64Suffix[kMaxRowsPerLane];
extern const KernelD kSweepF64Uniform[2][kMaxRowsPerLane];
// long_kernel by [PD].
extern const KernelF kLongF32[2];
extern const KernelD kLongF64[2];

}  // namespace fgkl
