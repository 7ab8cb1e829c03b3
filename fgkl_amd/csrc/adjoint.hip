// The adjoint of the PairHMM forward algorithm on gfx950: log10 likelihoods and the gradient of
// sum_p w_p log10 L_p with respect to each read base's error, insertion and deletion
// probabilities and the global gap-continuation probability.  The reference has no gradient;
// the model is its FP64 forward recurrence (reference.rs:20-64) in a continuous
// parameterisation: priors (1 - e, e / 3) with N matching everything, M->M = 1 - d_ins - d_del,
// M->I = d_ins, M->D = d_del, I->I = D->D = g, I->M = D->M = 1 - g, row 0 deletions C / m.
//
// Layout (one work item per wave64, a read against every haplotype of its region):
//  * Read rows 1..n are cut into strips of 64; lane l of strip s owns row 64 s + l + 1.  One row
//    per lane keeps the adjoint's working set in registers.  DP row 0 is stored as the boundary
//    row above strip 0.
//  * Pass F sweeps each strip along anti-diagonals (at step t lane l is on column t - l + 1,
//    the row above arriving by DPP wave_shr:1 from lane l-1; lane 0 takes it from the previous
//    strip's last row, which lane 63 stored) and stores every lane's state once every K steps:
//    the checkpoints.  It returns L.
//  * Pass B walks the strips bottom-up and, inside a strip, the K-step blocks right to left: a
//    block is recomputed from its checkpoint into a register tape, then swept in reverse.  The
//    adjoint of a cell needs the row below's values one and two steps later, which arrive by
//    DPP wave_shl:1 from lane l+1 (lane 63 reads the lower strip's top row, stored by its lane
//    0).  Gradient terms are summed per block in the working precision and accumulated in FP64
//    registers per row, written once per strip.
//  * Two precisions.  FP64: C = 2^1020, K = 8, every pair.  FP32 (the forward, the tape and the
//    adjoint): C = 2^120, K = 16, the adjoint seeded with w / L so that it is d ln L / d cell
//    (its values then stay within FP32's normal range from the last row to the first); a pair
//    whose FP32 likelihood falls below 1e-28 (kernel.rs:897-914's rescue threshold) is flagged,
//    and the FP64 kernel reruns the flagged pairs of each read and adds their terms.
//  * Checkpoint memory is n/64 strips x (m+63)/K blocks x 6 values per lane; the tape is K+1
//    cells per lane in registers, and the K rows above that the recompute moved in (the reverse
//    sweep's diagonal and up values, read back instead of moved again).  Work: forward twice
//    plus the reverse sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "fgkl_grad.hpp"
#include "fgkl_internal.hpp"
#include "wave64.hpp"

namespace fgkl {
namespace {

// Per precision: the row-0 scale C (reference.rs:12 for FP64, simd.rs:106 for FP32), the
// smallest accepted scaled likelihood, and the checkpoint block of K steps.
template <typename T> struct GT;
template <> struct GT<double> {
  static constexpr double kInit = 0x1p1020;
  static constexpr double kLog10Init = 1020.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 0.0;
  static constexpr int kTape = 8;
  static constexpr bool kPrefetch = false;  // its registers would spill
  static constexpr bool kTapeUp = false;    // likewise: the reverse sweep moves the rows again
};
template <> struct GT<float> {
  static constexpr double kInit = 0x1p120;
  static constexpr double kLog10Init = 120.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 1e-28;
  static constexpr int kTape = 16;  // about sqrt(m) for haplotypes of a few hundred bases; the
                                    // fastest with the row-above tape (measured 10..20)
  static constexpr bool kPrefetch = true;   // the next block's inputs load during this block
  static constexpr bool kTapeUp = true;     // the recompute keeps the rows above (see fblock)
};

constexpr double kInvLn10 = 0.434294481903251827651128918916605082;
constexpr int kHapPad = 80;      // zero bytes before the staged haplotype (columns down to -79)
constexpr int kSlot = 32;        // bytes of one dummy slot (holds any staged record)

struct GradArgs {
  const uint8_t* bases;
  const double *err, *ins, *del, *gcp;
  const int64_t* read_off;
  const int32_t* read_region;
  const int32_t* read_local;
  const int64_t* region_out;
  const int32_t* region_hap0;
  const int32_t* region_nhap;
  const int64_t* hap_off;
  const uint8_t* hap_bases;
  // Items of two read indices, most work first: (ra, -1) alone, or (ra, rb) sharing a strip
  // sequence (rescuing: the reads with lost pairs, alone).
  const int32_t* items;
  int32_t n_items;
  const int32_t* n_items_dev;  // when set, the item count is read from device memory
  int32_t* counter;
  const double* weights;     // per pair, or null for all ones
  double* log10_out;         // per pair, or null
  double *g_err, *g_ins, *g_del;
  double* dg_part;           // per read
  // FP32: lost[pair] is set to 1 for a pair below the threshold (0 otherwise), and each read
  // with one is appended to lost_items.  Rescuing (FP64): only pairs with lost[pair] run, and
  // their terms are added to the gradients already written.
  uint8_t* lost;
  int32_t* lost_items;
  int32_t* lost_count;
  int32_t rescue;
  char* scratch;
  int64_t wave_stride;       // bytes of scratch per wave
  int64_t ck_off, bnd_off, abnd_off;  // sections of a wave's scratch, in bytes
  int64_t dummy_off;         // 64 slots that absorb the stores of lanes with none to make
  int32_t nb_max;            // checkpoint blocks per strip
  int32_t bnd_cols;          // columns of a stored boundary row
  int32_t lds_bytes;         // per wave
};

// A lane's read row: its priors and gap-open parameters, the row below's (the adjoint reads
// them), and the global gap continuation g and 1 - g, which every row shares.
template <typename T> struct GRow {
  T pm, px, mm, mi, md, mmB, miB;
  T g, ig;
  uint32_t rb;
};

template <typename T> struct St {
  T m, x, y;
};
template <typename T> struct alignas(4 * sizeof(T)) St4 {
  T m, x, y, pad;
};
template <typename T> struct alignas(2 * sizeof(T)) Adj2 {
  T ax, q;
};

// Row v (1-based) of the read at `o`; rows past the read are zero and stay zero.  Every lane
// loads (from a clamped, valid index) and selects, so no lane-dependent branch is taken.
template <typename T>
__device__ __forceinline__ void load_grow(const GradArgs& a, int64_t o, int n, int v, double g,
                                          GRow<T>& rw) {
  const bool real = v <= n, below = v < n;
  const int64_t i = o + (real ? v : n) - 1;
  const int64_t j = below ? i + 1 : i;
  const double e = a.err[i], di = a.ins[i], dd = a.del[i];
  const double di2 = a.ins[j], dd2 = a.del[j];
  const uint32_t b = a.bases[i];
  rw.pm = T(real ? 1.0 - e : 0.0);
  rw.px = T(real ? e / 3.0 : 0.0);
  rw.mm = T(real ? 1.0 - di - dd : 0.0);
  rw.mi = T(real ? di : 0.0);
  rw.md = T(real ? dd : 0.0);
  rw.mmB = T(below ? 1.0 - di2 - dd2 : 0.0);
  rw.miB = T(below ? di2 : 0.0);
  rw.g = T(g);
  rw.ig = T(1.0 - g);
  rw.rb = real ? b : 0x100u;
}

__device__ __forceinline__ bool is_match(uint32_t hb, uint32_t rb) {
  return hb == rb || hb == uint32_t('N') || rb == uint32_t('N');
}

template <typename T>
__device__ __forceinline__ St<T> shr_st(const St4<T>& old, const St<T>& src) {
  return St<T>{shr1(old.m, src.m), shr1(old.x, src.x), shr1(old.y, src.y)};
}

// Two reads in one strip sequence: read B's rows follow read A's, B's first row in the lane below
// A's last row (lane la + 1 of A's last strip).  That lane takes B's DP row 0, (0, 0, C/m), as
// the row above, where the DPP move brings A's last row (`keep_f` is zero there, all ones
// elsewhere), and A's last row takes zero adjoints from below (`keep_r` zero in lane la).  The
// masks are applied by bitwise ops after the move, so the move runs in every lane.
__device__ __forceinline__ float and_bits(float v, uint32_t k) {
  return __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, v) & k);
}
__device__ __forceinline__ double and_bits(double v, uint32_t k) {
  return __builtin_bit_cast(double, __builtin_bit_cast(uint64_t, v) & ((uint64_t(k) << 32) | k));
}
__device__ __forceinline__ float sel_bits(uint32_t k, float a, float b) {
  const uint32_t x = __builtin_bit_cast(uint32_t, a), y = __builtin_bit_cast(uint32_t, b);
  return __builtin_bit_cast(float, (x & k) | (y & ~k));
}
__device__ __forceinline__ double sel_bits(uint32_t k, double a, double b) {
  const uint64_t m = (uint64_t(k) << 32) | k;
  const uint64_t x = __builtin_bit_cast(uint64_t, a), y = __builtin_bit_cast(uint64_t, b);
  return __builtin_bit_cast(double, (x & m) | (y & ~m));
}
template <typename T>
__device__ __forceinline__ void row0_above(St<T>& q, uint32_t keep_f, T init) {
  q.m = and_bits(q.m, keep_f);
  q.x = and_bits(q.x, keep_f);
  q.y = sel_bits(keep_f, q.y, init);
}

// A wave's LDS: the staged boundary rows and haplotype, and the per-block staging of the values
// one lane hands to the neighbouring strip.  The lane that owns a boundary writes its staging
// slot each step; every other lane writes a private dummy slot, so the writes need no branch.
template <typename T> struct Lds {
  St4<T>* stage;   // K + 1 columns of the row above the strip (lane 0's neighbour)
  Adj2<T>* astage; // K columns of the adjoint of the row below the strip (lane 63's)
  St4<T>* fout;    // K columns of lane 63's state, for the strip below
  Adj2<T>* aout;   // K columns of lane 0's adjoint, for the strip above
  uint8_t* dummy;  // 64 + K slots of kSlot bytes
  uint8_t* hap;    // kHapPad zeros, the haplotype, zeros
};

template <typename T>
constexpr int lds_fixed() {
  constexpr int K = GT<T>::kTape;
  return (K + 1) * int(sizeof(St4<T>)) + K * int(sizeof(Adj2<T>)) + K * int(sizeof(St4<T>)) +
         K * int(sizeof(Adj2<T>)) + (64 + K) * kSlot;
}

template <typename T>
__device__ __forceinline__ Lds<T> lds_parts(uint8_t* base) {
  constexpr int K = GT<T>::kTape;
  Lds<T> l;
  l.stage = reinterpret_cast<St4<T>*>(base);
  base += (K + 1) * sizeof(St4<T>);
  l.astage = reinterpret_cast<Adj2<T>*>(base);
  base += K * sizeof(Adj2<T>);
  l.fout = reinterpret_cast<St4<T>*>(base);
  base += K * sizeof(St4<T>);
  l.aout = reinterpret_cast<Adj2<T>*>(base);
  base += K * sizeof(Adj2<T>);
  l.dummy = base;
  base += (64 + K) * kSlot;
  l.hap = base;
  return l;
}

template <typename T>
__device__ __forceinline__ T prior_of(const Lds<T>& l, int c, const GRow<T>& rw) {
  return is_match(l.hap[kHapPad + c - 1], rw.rb) ? rw.pm : rw.px;
}

// K forward steps from `st` (the state before step t0) with diagonal `P`.  FM freezes lanes
// that have not reached column 1; ACC adds M + X of columns 1..m; OUT stages lane 63's state.
// With TAPE, tp[k] receives the state before step k and, when GT<T>::kTapeUp, tq[k] the row
// above at step k (the next step's diagonal), which the reverse sweep then reads instead of
// moving them again.
// MID: the strip holds the first row of a second read (see row0_above).
template <typename T, bool FM, bool ACC, bool OUT, bool TAPE, bool MID>
__device__ __forceinline__ void fblock(St<T>& st, St<T>& P, T& acc, St<T>* tp, St<T>* tq,
                                       const GRow<T>& rw, const Lds<T>& l, St4<T>* out, int t0,
                                       int lane, int m, uint32_t keep_f, T init) {
  constexpr int K = GT<T>::kTape;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if constexpr (TAPE) tp[k] = st;
    const int c = t0 + k - lane + 1;
    St<T> Q = shr_st(l.stage[k + 1], st);
    if constexpr (MID) row0_above(Q, keep_f, init);
    if constexpr (TAPE && GT<T>::kTapeUp) tq[k] = Q;
    const T prior = prior_of(l, c, rw);
    const T Y = fma(st.y, rw.g, st.m * rw.md);
    const T M = prior * fma(P.y, rw.ig, fma(P.x, rw.ig, P.m * rw.mm));
    const T X = fma(Q.x, rw.g, Q.m * rw.mi);
    if constexpr (FM) {
      const bool live = c >= 1;
      st.m = live ? M : st.m;
      st.x = live ? X : st.x;
      st.y = live ? Y : st.y;
    } else {
      st = St<T>{M, X, Y};
    }
    P = Q;
    if constexpr (ACC) {
      if constexpr (FM)
        acc += (c >= 1 && c <= m) ? st.m + st.x : T(0);
      else
        acc += (c <= m) ? st.m + st.x : T(0);
    }
    if constexpr (OUT) out[k] = St4<T>{st.m, st.x, st.y, T(0)};
  }
  if constexpr (TAPE) tp[K] = st;
}

// Column t0 + j (clamped to 0..m) of a stored boundary row, the value lane j <= K stages for
// lane 0 (lanes past K load column t0 + K and stage it into the dummy area).  Loaded one block
// ahead of its use, so the global load's latency overlaps a block of compute.
template <typename T>
__device__ __forceinline__ St4<T> load_above(const T* bnd, int t0, int m, int lane) {
  constexpr int K = GT<T>::kTape;
  const int col = min(t0 + min(lane, K), m);
  return *reinterpret_cast<const St4<T>*>(bnd + size_t(col) * 4);
}

template <typename T>
__device__ __forceinline__ void put_above(const Lds<T>& l, const St4<T>& w, int lane) {
  constexpr int K = GT<T>::kTape;
  St4<T>* dst = lane <= K ? &l.stage[lane] : reinterpret_cast<St4<T>*>(l.dummy + lane * kSlot);
  *dst = St4<T>{w.m, w.x, w.y, T(0)};
}

// Pass F over strip s: returns the lane's sum of M + X over columns 1..m when ACC (row n's lane
// is the one that matters).
template <typename T, bool CK, bool ACC, bool MID>
__device__ T forward_strip(const GradArgs& a, const Lds<T>& l, const GRow<T>& rw, int lane, int s,
                           int m, int L, int nb, bool has_next, char* wscr, uint32_t keep_f,
                           T init) {
  constexpr int K = GT<T>::kTape;
  const T* bnd_in = reinterpret_cast<const T*>(wscr + a.bnd_off) + size_t(s) * a.bnd_cols * 4;
  T* bnd_out = reinterpret_cast<T*>(wscr + a.bnd_off) + size_t(s + 1) * a.bnd_cols * 4;
  T* ck = reinterpret_cast<T*>(wscr + a.ck_off) + size_t(s) * a.nb_max * 6 * 64;
  St4<T>* out = lane == 63 ? l.fout : reinterpret_cast<St4<T>*>(l.dummy + lane * kSlot);
  St<T> st{T(0), T(0), T(0)};  // column 0 below DP row 0
  St<T> P{T(0), T(0), T(0)};
  T acc = T(0);
  St4<T>* gdummy = reinterpret_cast<St4<T>*>(wscr + a.dummy_off + lane * kSlot);
  if (has_next) *reinterpret_cast<St4<T>*>(bnd_out) = St4<T>{T(0), T(0), T(0), T(0)};
  St4<T> above = load_above<T>(bnd_in, 0, m, lane);
  for (int b = 0; b < nb; ++b) {
    const int t0 = b * K;
    if constexpr (!GT<T>::kPrefetch) above = load_above<T>(bnd_in, t0, m, lane);
    wave_sync_lds();
    put_above(l, above, lane);
    if constexpr (GT<T>::kPrefetch)
      above = load_above<T>(bnd_in, t0 + K, m, lane);  // the next block's (clamped: in bounds)
    wave_sync_lds();
    if (b == 0) {  // the diagonal of step 0: column 0 of the row above
      P = shr_st(l.stage[0], st);
      if constexpr (MID) row0_above(P, keep_f, init);
    }
    if constexpr (CK) {
      T* c = ck + size_t(b) * 6 * 64 + lane;
      c[0 * 64] = st.m;
      c[1 * 64] = st.x;
      c[2 * 64] = st.y;
      c[3 * 64] = P.m;
      c[4 * 64] = P.x;
      c[5 * 64] = P.y;
    }
    const bool fm = t0 < L - 1;
    if (has_next) {
      if (fm)
        fblock<T, true, ACC, true, false, MID>(st, P, acc, nullptr, nullptr, rw, l, out,
                                               t0, lane, m, keep_f, init);
      else
        fblock<T, false, ACC, true, false, MID>(st, P, acc, nullptr, nullptr, rw, l, out,
                                                t0, lane, m, keep_f, init);
      wave_sync_lds();
      // Lanes 0..K-1 store lane 63's columns t0+lane-62 that lie in 1..m.
      const int c = t0 + lane - 62;
      const bool ok = lane < K && c >= 1 && c <= m;
      const St4<T> w = l.fout[min(lane, K - 1)];
      St4<T>* dst = ok ? reinterpret_cast<St4<T>*>(bnd_out + size_t(max(c, 0)) * 4) : gdummy;
      *dst = St4<T>{w.m, w.x, w.y, T(0)};
    } else {
      if (fm)
        fblock<T, true, ACC, false, false, MID>(st, P, acc, nullptr, nullptr, rw, l, out,
                                                t0, lane, m, keep_f, init);
      else
        fblock<T, false, ACC, false, false, MID>(st, P, acc, nullptr, nullptr, rw, l, out,
                                                 t0, lane, m, keep_f, init);
    }
  }
  return acc;
}

// Gradient terms of one block, summed in the working precision (added to the FP64 accumulators
// once per block).
template <typename T> struct GSum {
  T ge, gmm, gim, gmi, gmd, gg;
};

// The reverse sweep of one block over the tape.  BM zeroes the adjoint outside columns 1..m;
// SEED adds `seedv` (non-zero only in the read's last row), which only the bottom strip holds.
template <typename T, bool BM, bool SEED, bool MID>
__device__ __forceinline__ void rblock(const St<T>* tp, const St<T>* tq, const St<T>& P0,
                                       const GRow<T>& rw, const Lds<T>& l, Adj2<T>* aout, int t0,
                                       int lane, int m, T seedv, T& AX, T& AY, T& Q, T& Bq,
                                       GSum<T>& gs, uint32_t keep_f, uint32_t keep_r, T init) {
  constexpr int K = GT<T>::kTape;
  // The row above at step k: from the tape, or moved again from the state before step k.
  auto up_at = [&](int k) {
    if constexpr (GT<T>::kTapeUp) {
      return tq[k];
    } else {
      St<T> u = shr_st(l.stage[k + 1], tp[k]);
      if constexpr (MID) row0_above(u, keep_f, init);
      return u;
    }
  };
  St<T> U = up_at(K - 1);  // up values of the last step
#pragma unroll
  for (int k = K - 1; k >= 0; --k) {
    const int col = t0 + k - lane + 1;
    const bool mt = is_match(l.hap[kHapPad + col - 1], rw.rb);
    const T prior = mt ? rw.pm : rw.px;
    const T ds = mt ? T(-1) : T(1.0 / 3.0);  // d prior / d e
    // The row below: its insertion adjoint in this column and Q one column to the right.
    const Adj2<T> below = l.astage[k];
    T AXb = shl1(below.ax, AX);
    T Qn = shl1(below.q, Q);
    if constexpr (MID) {
      AXb = and_bits(AXb, keep_r);
      Qn = and_bits(Qn, keep_r);
    }
    const T Qd = Bq;
    Bq = Qn;
    const St<T> D = k ? up_at(k - 1) : P0;  // the diagonal: the row above one step earlier
    const St<T>& Lf = tp[k];
    const T pr = Qd * rw.ig;
    T ay = fma(AY, rw.g, pr);
    const T amp = fma(Qd, rw.mmB, AY * rw.md);
    T ax = fma(AXb, rw.g, pr);
    T am = fma(AXb, rw.miB, amp);
    if constexpr (SEED) {
      ax += seedv;
      am += seedv;
    }
    if constexpr (BM) {
      const bool live = col >= 1 && col <= m;
      ay = live ? ay : T(0);
      ax = live ? ax : T(0);
      am = live ? am : T(0);
    }
    const T q = am * prior;
    const T xy = D.x + D.y;
    const T spre = fma(D.m, rw.mm, xy * rw.ig);  // M before the prior
    gs.ge = fma(am * spre, ds, gs.ge);
    gs.gmm = fma(q, D.m, gs.gmm);
    gs.gim = fma(q, xy, gs.gim);
    gs.gmi = fma(ax, U.m, gs.gmi);
    gs.gmd = fma(ay, Lf.m, gs.gmd);
    gs.gg = fma(ax, U.x, fma(ay, Lf.y, gs.gg));
    AY = ay;
    AX = ax;
    Q = q;
    U = D;
    aout[k] = Adj2<T>{ax, q};
  }
}

// Pass B over strip s.  The lane's row has gradient index `gidx` (-1: no row); `dgl_a` and
// `dgl_b` collect its gap-continuation terms for the first and the second read of the item.
template <typename T, bool SEED, bool MID>
__device__ void backward_strip(const GradArgs& a, const Lds<T>& l, const GRow<T>& rw, int lane,
                               int s, int m, int L, int nb, T seedv, int64_t gidx, bool in_b,
                               char* wscr, double& dgl_a, double& dgl_b, uint32_t keep_f,
                               uint32_t keep_r, T init) {
  constexpr int K = GT<T>::kTape;
  const T* bnd_in = reinterpret_cast<const T*>(wscr + a.bnd_off) + size_t(s) * a.bnd_cols * 4;
  const T* ck = reinterpret_cast<const T*>(wscr + a.ck_off) + size_t(s) * a.nb_max * 6 * 64;
  const int abnd_cols = a.bnd_cols + 1;
  Adj2<T>* abnd = reinterpret_cast<Adj2<T>*>(wscr + a.abnd_off);
  const Adj2<T>* ab_in = abnd + size_t((s + 1) & 1) * abnd_cols;
  Adj2<T>* ab_out = abnd + size_t(s & 1) * abnd_cols;
  Adj2<T>* aout = lane == 0 ? l.aout : reinterpret_cast<Adj2<T>*>(l.dummy + lane * kSlot);
  char* gdummy = wscr + a.dummy_off + lane * kSlot;
  ab_out[0] = Adj2<T>{T(0), T(0)};
  ab_out[m + 1] = Adj2<T>{T(0), T(0)};
  T AX = T(0), AY = T(0), Q = T(0), Bq = T(0);  // columns beyond m: zero adjoints
  // d/d e, d/d d_ins (M->I minus M->M), d/d d_del (M->D minus M->M), d/d g (I->I, D->D minus
  // I->M, D->M): each block's terms combine in the working precision, then add in FP64.
  double ge = 0.0, gi = 0.0, gd = 0.0, gq = 0.0;
  // A block's inputs, loaded one block ahead: the row above (columns t0..t0+K, for lane 0), the
  // adjoint row below (columns t0-62..t0-63+K, for lane 63) and the lane's checkpoint.
  struct In {
    St4<T> above;
    Adj2<T> below;
    T c[6];
  };
  auto load = [&](int b) {
    const int t0 = b * K;
    In in;
    in.above = load_above<T>(bnd_in, t0, m, lane);
    in.below = ab_in[min(max(t0 + lane - 62, 0), m + 1)];
    const T* c = ck + size_t(b) * 6 * 64 + lane;
#pragma unroll
    for (int i = 0; i < 6; ++i) in.c[i] = c[i * 64];
    return in;
  };
  In next = load(nb - 1);
  for (int b = nb - 1; b >= 0; --b) {
    const int t0 = b * K;
    In cur;
    if constexpr (GT<T>::kPrefetch) {
      cur = next;
      next = load(max(b - 1, 0));
    } else {
      cur = load(b);
    }
    wave_sync_lds();
    *(lane < K ? &l.astage[lane] : reinterpret_cast<Adj2<T>*>(l.dummy + lane * kSlot)) = cur.below;
    put_above(l, cur.above, lane);
    wave_sync_lds();
    // Recompute the block from its checkpoint into the tape: tp[k] is the state before step k.
    St<T> tp[K + 1], tq[GT<T>::kTapeUp ? K : 1];
    St<T> st{cur.c[0], cur.c[1], cur.c[2]};
    const St<T> P0{cur.c[3], cur.c[4], cur.c[5]};
    St<T> P = P0;
    T unused = T(0);
    if (t0 < L - 1)
      fblock<T, true, false, false, true, MID>(st, P, unused, tp, tq, rw, l, nullptr, t0, lane,
                                               m, keep_f, init);
    else
      fblock<T, false, false, false, true, MID>(st, P, unused, tp, tq, rw, l, nullptr, t0, lane,
                                                m, keep_f, init);
    GSum<T> gs{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (t0 < L - 1 || t0 + K > m)
      rblock<T, true, SEED, MID>(tp, tq, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq, gs,
                                 keep_f, keep_r, init);
    else
      rblock<T, false, SEED, MID>(tp, tq, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq,
                                  gs, keep_f, keep_r, init);
    ge += double(gs.ge);
    gi += double(gs.gmi - gs.gmm);
    gd += double(gs.gmd - gs.gmm);
    gq += double(gs.gg - gs.gim);
    wave_sync_lds();
    {
      const int col = t0 + lane + 1;  // lane 0's column at step t0 + lane
      const bool ok = lane < K && col <= m;
      const Adj2<T> w = l.aout[min(lane, K - 1)];
      Adj2<T>* dst = ok ? &ab_out[col] : reinterpret_cast<Adj2<T>*>(gdummy);
      *dst = w;
    }
  }
  const bool real = gidx >= 0;
  const int64_t i = real ? gidx : 0;
  double* pe = real ? a.g_err + i : reinterpret_cast<double*>(gdummy);
  double* pi = real ? a.g_ins + i : reinterpret_cast<double*>(gdummy) + 1;
  double* pd = real ? a.g_del + i : reinterpret_cast<double*>(gdummy) + 2;
  *pe += ge * kInvLn10;
  *pi += gi * kInvLn10;  // M->M = 1 - d_ins - d_del
  *pd += gd * kInvLn10;
  dgl_a += real && !in_b ? gq : 0.0;  // I->M = D->M = 1 - g
  dgl_b += real && in_b ? gq : 0.0;
}

// FP32: three waves per SIMD (at most 168 VGPRs) measured faster than two without spills.
template <typename T, bool BWD>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(std::is_same<T, float>::value ? 3 : 1)))
grad_kernel(GradArgs a) {
  constexpr int K = GT<T>::kTape;
  constexpr bool kFp32 = std::is_same<T, float>::value;
  extern __shared__ __attribute__((aligned(32))) uint8_t glds[];
  const int lane = threadIdx.x & 63;
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t wave_global = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave_in_block;
  const Lds<T> l = lds_parts<T>(glds + wave_in_block * a.lds_bytes);
  char* wscr = a.scratch + wave_global * a.wave_stride;
  char* gslot = wscr + a.dummy_off + lane * kSlot;
  const double g = a.gcp[0];
  const int n_items = a.n_items_dev ? *a.n_items_dev : a.n_items;

  for (;;) {
    // Every lane adds one, so a claim advances the counter by 64 and every lane's old value lies
    // in [64 i, 64 i + 63]; no lane-dependent branch reaches the readfirstlane.
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= n_items) break;
    // An item is one read, or two reads of one region (ra, then rb) in one sequence of strips:
    // rb's rows follow ra's, its first row in the lane below ra's last row, inside ra's last strip.
    // FP64 runs single reads only: its registers have no room for the shared strip's code.
    const int ra = __builtin_amdgcn_readfirstlane(a.items[2 * item]);
    const int rb = kFp32 ? __builtin_amdgcn_readfirstlane(a.items[2 * item + 1]) : -1;
    const bool two = kFp32 && rb >= 0;
    const int64_t oa = a.read_off[ra];
    const int na = __builtin_amdgcn_readfirstlane(int(a.read_off[ra + 1] - oa));
    const int64_t ob = two ? a.read_off[rb] : 0;
    const int nr_b = two ? __builtin_amdgcn_readfirstlane(int(a.read_off[rb + 1] - ob)) : 0;
    const int N = na + nr_b;  // rows of the item
    const int gi = __builtin_amdgcn_readfirstlane(a.read_region[ra]);
    const int hap0 = __builtin_amdgcn_readfirstlane(a.region_hap0[gi]);
    const int H = __builtin_amdgcn_readfirstlane(a.region_nhap[gi]);
    const int64_t base_a = a.region_out[gi] + int64_t(a.read_local[ra]) * H;
    const int64_t base_b = two ? a.region_out[gi] + int64_t(a.read_local[rb]) * H : 0;
    const int S = (N + 63) / 64;   // strips of 64 rows
    const int sa = (na - 1) / 64;  // the strip of ra's last row (and of rb's first)
    const int la = (na - 1) % 64;  // ra's last row's lane there
    const int lb = (N - 1) % 64;   // rb's last row's lane in strip S - 1
    uint32_t keep_f = two && lane == la + 1 ? 0u : ~0u;
    uint32_t keep_r = two && lane == la ? 0u : ~0u;
    asm volatile("" : "+v"(keep_f), "+v"(keep_r));
    // The lane's row in strip s: its read's offset, length and row number (1-based).
    struct RowOf {
      int64_t o;
      int n, v;
      bool in_b;
    };
    auto row_of = [&](int s) {
      const int v = 64 * s + lane + 1;
      const bool in_b = two && v > na;
      return RowOf{in_b ? ob : oa, in_b ? nr_b : na, in_b ? v - na : v, in_b};
    };
    if (BWD && !a.rescue) {  // a rescue adds to what the FP32 pass wrote
      for (int s = 0; s < S; ++s) {
        const RowOf w = row_of(s);
        const bool real = w.v <= w.n;
        const int64_t i = w.o + (real ? w.v : w.n) - 1;
        *(real ? a.g_err + i : reinterpret_cast<double*>(gslot)) = 0.0;
        *(real ? a.g_ins + i : reinterpret_cast<double*>(gslot) + 1) = 0.0;
        *(real ? a.g_del + i : reinterpret_cast<double*>(gslot) + 2) = 0.0;
      }
    }
    double dgl_a = 0.0, dgl_b = 0.0;
    int any_lost_a = 0, any_lost_b = 0;
    for (int h = 0; h < H; ++h) {
      // Rescue items are single reads.
      if (a.rescue && !__builtin_amdgcn_readfirstlane(int(a.lost[base_a + h]))) continue;
      const int64_t hoff = a.hap_off[hap0 + h];
      const int m = __builtin_amdgcn_readfirstlane(int(a.hap_off[hap0 + h + 1] - hoff));
      wave_sync_lds();
      const int hap_total = kHapPad + m + 64 + K + kHapPad;
      for (int i0 = 0; i0 < hap_total; i0 += 64) {
        const int i = i0 + lane;
        const int c = i - kHapPad;
        const bool in = c >= 0 && c < m;
        const uint8_t hb = a.hap_bases[hoff + (in ? c : 0)];
        *(i < hap_total ? &l.hap[i] : l.dummy + lane) = in ? hb : uint8_t(0);
      }
      // The row above strip 0 is DP row 0: no match or insertion mass, deletions C / m in every
      // column (reference.rs:27-28).  rb's DP row 0 is the same row, given by row0_above.
      const T init = T(GT<T>::kInit / double(m));
      T* bnd0 = reinterpret_cast<T*>(wscr + a.bnd_off);
      for (int c0 = 0; c0 <= m; c0 += 64) {
        const int c = c0 + lane;
        St4<T>* dst = c <= m ? reinterpret_cast<St4<T>*>(bnd0 + size_t(c) * 4)
                             : reinterpret_cast<St4<T>*>(gslot);
        *dst = St4<T>{T(0), T(0), init, T(0)};
      }
      wave_sync_lds();
      wave_sync_global();
      const double wa = BWD ? (a.weights ? a.weights[base_a + h] : 1.0) : 0.0;
      const double wb = BWD && two ? (a.weights ? a.weights[base_b + h] : 1.0) : 0.0;
      const bool grad = BWD && (wa != 0.0 || wb != 0.0);
      // Pass F.  The strips holding a read's last row sum its M + X (ACC).
      double total_a = 0.0, total_b = 0.0;
      for (int s = 0; s < S; ++s) {
        const RowOf w = row_of(s);
        const int L = min(64, N - 64 * s);
        const int nb = (m + L - 1 + K - 1) / K;
        GRow<T> rw;
        load_grow(a, w.o, w.n, w.v, g, rw);
        const bool next = s + 1 < S;
        T acc = T(0);
        if (two && s == sa) {  // the strip where rb begins; it also ends ra
          acc = grad ? forward_strip<T, true, true, true>(a, l, rw, lane, s, m, L, nb, next, wscr,
                                                         keep_f, init)
                     : forward_strip<T, false, true, true>(a, l, rw, lane, s, m, L, nb, next,
                                                          wscr, keep_f, init);
        } else if (s == sa || s == S - 1) {
          acc = grad ? forward_strip<T, true, true, false>(a, l, rw, lane, s, m, L, nb, next,
                                                          wscr, keep_f, init)
                     : forward_strip<T, false, true, false>(a, l, rw, lane, s, m, L, nb, next,
                                                           wscr, keep_f, init);
        } else if (grad) {
          forward_strip<T, true, false, false>(a, l, rw, lane, s, m, L, nb, next, wscr, keep_f,
                                               init);
        } else {
          forward_strip<T, false, false, false>(a, l, rw, lane, s, m, L, nb, next, wscr, keep_f,
                                                init);
        }
        if (s == sa) total_a = double(readlane(acc, la));
        if (two && s == S - 1) total_b = double(readlane(acc, lb));
        wave_sync_global();
      }
      // FP32: a pair below the rescue threshold (or NaN) is left to the FP64 pass.
      const bool lost_a = kFp32 && !(total_a >= GT<T>::kMinAccepted);
      const bool lost_b = kFp32 && two && !(total_b >= GT<T>::kMinAccepted);
      if (kFp32 && lane == 0) {
        a.lost[base_a + h] = lost_a ? 1 : 0;
        if (two) a.lost[base_b + h] = lost_b ? 1 : 0;
      }
      any_lost_a |= lost_a;
      any_lost_b |= lost_b;
      if (a.log10_out && lane == 0) {
        a.log10_out[base_a + h] = lost_a ? __builtin_nan("") : log10(total_a) - GT<T>::kLog10Init;
        if (two)
          a.log10_out[base_b + h] =
              lost_b ? __builtin_nan("") : log10(total_b) - GT<T>::kLog10Init;
      }
      auto usable = [](double t) { return t > 0.0 && t <= 1.7976931348623157e308; };
      const bool ga = grad && wa != 0.0 && !lost_a && usable(total_a);
      const bool gb = grad && two && wb != 0.0 && !lost_b && usable(total_b);
      if (!ga && !gb) continue;
      // Pass B: the bottom strip has nothing below it.
      {
        const int abnd_cols = a.bnd_cols + 1;
        Adj2<T>* z = reinterpret_cast<Adj2<T>*>(wscr + a.abnd_off) + size_t(S & 1) * abnd_cols;
        for (int c0 = 0; c0 <= m + 1; c0 += 64) {
          const int c = c0 + lane;
          *(c <= m + 1 ? &z[c] : reinterpret_cast<Adj2<T>*>(gslot)) = Adj2<T>{T(0), T(0)};
        }
        wave_sync_global();
      }
      // Every adjoint is w d ln L / d cell: seeded with w / L at each read's last row (zero for
      // a read that contributes nothing, whose rows then take zero adjoints).
      const T seed_a = ga ? T(wa / total_a) : T(0);
      const T seed_b = gb ? T(wb / total_b) : T(0);
      for (int s = S - 1; s >= 0; --s) {
        const RowOf w = row_of(s);
        const int L = min(64, N - 64 * s);
        const int nb = (m + L - 1 + K - 1) / K;
        GRow<T> rw;
        load_grow(a, w.o, w.n, w.v, g, rw);
        const int64_t gidx = w.v <= w.n ? w.o + w.v - 1 : -1;
        const T seedv = !w.in_b && w.v == na ? seed_a : (w.in_b && w.v == nr_b ? seed_b : T(0));
        if (two && s == sa)
          backward_strip<T, true, true>(a, l, rw, lane, s, m, L, nb, seedv, gidx, w.in_b, wscr,
                                        dgl_a, dgl_b, keep_f, keep_r, init);
        else if (s == sa || s == S - 1)  // a read's last row lies in this strip
          backward_strip<T, true, false>(a, l, rw, lane, s, m, L, nb, seedv, gidx, w.in_b, wscr,
                                         dgl_a, dgl_b, keep_f, keep_r, init);
        else
          backward_strip<T, false, false>(a, l, rw, lane, s, m, L, nb, T(0), gidx, w.in_b,
                                          wscr, dgl_a, dgl_b, keep_f, keep_r, init);
        wave_sync_global();
      }
    }
    if (kFp32 && lane == 0) {
      if (any_lost_a) {
        const int idx = atomicAdd(a.lost_count, 1);
        a.lost_items[2 * idx] = ra;
        a.lost_items[2 * idx + 1] = -1;
      }
      if (any_lost_b) {
        const int idx = atomicAdd(a.lost_count, 1);
        a.lost_items[2 * idx] = rb;
        a.lost_items[2 * idx + 1] = -1;
      }
    }
    if constexpr (BWD) {
      // Sum over lanes in lane order through LDS; every lane stores the same value.
      double* red = reinterpret_cast<double*>(l.dummy);
      for (int q = 0; q < (two ? 2 : 1); ++q) {
        wave_sync_lds();
        red[lane] = q ? dgl_b : dgl_a;
        wave_sync_lds();
        double tot = 0.0;
#pragma unroll 4
        for (int i = 0; i < 64; ++i) tot += red[i];
        const int r = q ? rb : ra;
        if (a.rescue)
          a.dg_part[r] += tot * kInvLn10;
        else
          a.dg_part[r] = tot * kInvLn10;
      }
    }
  }
}

__global__ void grad_fill(int32_t* p, int n, int32_t v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// Sums the per-read gap-continuation terms in a fixed order (one block of 256 threads).
__global__ void __launch_bounds__(256) grad_reduce(const double* part, int64_t n, double* out) {
  __shared__ double buf[256];
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
  buf[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) buf[threadIdx.x] += buf[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = buf[0];
}

__global__ void dpp_shl_selftest_kernel(const int32_t* in, int32_t* out) {
  const int lane = threadIdx.x & 63;
  out[lane] = shl1i(-1, in[lane]);
}

#define GRAD_TRY(expr)                                                                       \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess) {                                                                  \
      set_error(kInternal, std::string("HIP error: ") + hipGetErrorString(_e) + " at " +   \
                                 #expr);                                                     \
      return kInternal;                                                                      \
    }                                                                                        \
  } while (0)

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

}  // namespace

struct GradBatch::Impl {
  std::vector<void*> allocs;
  int num_cus = 0;
  int64_t n_reads = 0, n_pairs = 0, n_read_bases = 0;
  uint8_t* bases = nullptr;
  int64_t* read_off = nullptr;
  int32_t *read_region = nullptr, *read_local = nullptr;
  int64_t* region_out = nullptr;
  int32_t *region_hap0 = nullptr, *region_nhap = nullptr;
  int64_t* hap_off = nullptr;
  uint8_t* hap_bases = nullptr;
  int32_t* items = nullptr;         // FP32: reads alone or in pairs
  int32_t* items_single = nullptr;  // FP64: reads alone
  int32_t n_single = 0;
  // Work-queue heads and counts: 0 forward, 1 backward, 2 reads with lost pairs, 3 the rescue.
  int32_t* counters = nullptr;
  double* dg_part = nullptr;
  uint8_t* lost = nullptr;        // per pair (FP32 forward)
  int32_t* lost_items = nullptr;  // reads with lost pairs, as items (read, -1)
  char* scratch = nullptr;
  int64_t scratch_waves = 0;
  GradArgs proto[2] = {};         // [fp64]: precision-specific scratch and LDS layout
  int occ[2][2] = {};             // resident blocks per CU [fp64][bwd]
  int fp32 = 0;                   // precision of the forward, the tape and the adjoint
  hipStream_t stream = nullptr;
  hipEvent_t ev[2] = {};
  double *d_err = nullptr, *d_ins = nullptr, *d_del = nullptr, *d_gcp = nullptr;
  double *d_w = nullptr, *d_log10 = nullptr;
  double *d_ge = nullptr, *d_gi = nullptr, *d_gd = nullptr, *d_gg = nullptr;

  ~Impl() {
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (void* p : allocs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // Copies the probabilities in; the caller's arrays must hold n_read_bases values.
  int put_params(const GradParams& p) {
    const size_t nb = sizeof(double) * size_t(n_read_bases);
    if (nb && (!p.err || !p.ins || !p.del)) {
      set_error(kInvalidArgument, "probability arrays are null");
      return kInvalidArgument;
    }
    if (nb) {
      GRAD_TRY(hipMemcpyAsync(d_err, p.err, nb, hipMemcpyHostToDevice, stream));
      GRAD_TRY(hipMemcpyAsync(d_ins, p.ins, nb, hipMemcpyHostToDevice, stream));
      GRAD_TRY(hipMemcpyAsync(d_del, p.del, nb, hipMemcpyHostToDevice, stream));
    }
    GRAD_TRY(hipMemcpyAsync(d_gcp, &p.gcp, sizeof(double), hipMemcpyHostToDevice, stream));
    return kOk;
  }
  int finish(double* kernel_ms) {
    GRAD_TRY(hipStreamSynchronize(stream));
    if (kernel_ms) {
      float ms = 0.f;
      GRAD_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
      *kernel_ms = ms;
    }
    return kOk;
  }
  template <typename P>
  int alloc(P** p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    GRAD_TRY(hipMalloc(reinterpret_cast<void**>(p), sizeof(P) * count));
    allocs.push_back(*p);
    return kOk;
  }
  template <typename P>
  int upload(P** p, const P* src, size_t count) {
    if (int rc = alloc(p, count)) return rc;
    if (count) GRAD_TRY(hipMemcpy(*p, src, sizeof(P) * count, hipMemcpyHostToDevice));
    return kOk;
  }
  template <typename T>
  int launch_one(bool bwd, GradArgs a, int64_t max_items) {
    if (max_items <= 0) return kOk;
    const int ti = std::is_same<T, double>::value ? 1 : 0;
    const int64_t want = (max_items + kWaves - 1) / kWaves;
    const int64_t resident = int64_t(occ[ti][bwd ? 1 : 0]) * num_cus;
    const int grid = int(std::min({want, resident, scratch_waves / kWaves}));
    if (bwd)
      hipLaunchKernelGGL((grad_kernel<T, true>), dim3(grid), dim3(kBlock),
                         size_t(a.lds_bytes) * kWaves, stream, a);
    else
      hipLaunchKernelGGL((grad_kernel<T, false>), dim3(grid), dim3(kBlock),
                         size_t(a.lds_bytes) * kWaves, stream, a);
    GRAD_TRY(hipGetLastError());
    return kOk;
  }
  // FP64: one pass over every read.  FP32: one pass, then the FP64 pass over the reads with
  // pairs below the threshold, which reruns only those pairs and adds their terms.
  int launch(bool bwd, const GradArgs& base) {
    hipLaunchKernelGGL(grad_fill, dim3(1), dim3(64), 0, stream, counters, 4, 0);
    GRAD_TRY(hipGetLastError());
    if (base.n_items <= 0) return kOk;
    if (!fp32) {
      GradArgs a = base;
      a.items = items_single;
      a.n_items = n_single;
      a.counter = counters + (bwd ? 1 : 0);
      return launch_one<double>(bwd, a, a.n_items);
    }
    GradArgs a = base;
    const GradArgs& lf = proto[0];
    a.ck_off = lf.ck_off; a.bnd_off = lf.bnd_off; a.abnd_off = lf.abnd_off;
    a.dummy_off = lf.dummy_off; a.nb_max = lf.nb_max; a.lds_bytes = lf.lds_bytes;
    a.counter = counters + (bwd ? 1 : 0);
    a.lost = lost;
    a.lost_items = lost_items;
    a.lost_count = counters + 2;
    if (int rc = launch_one<float>(bwd, a, a.n_items)) return rc;
    GradArgs r = base;  // the FP64 layout
    r.items = lost_items;
    r.n_items = 0;
    r.n_items_dev = counters + 2;
    r.counter = counters + 3;
    r.lost = lost;
    r.rescue = 1;
    return launch_one<double>(bwd, r, n_reads);
  }
};

GradBatch::GradBatch() : impl_(new Impl) {}
GradBatch::~GradBatch() { delete impl_; }

namespace {
// Per-wave scratch of one precision, in bytes: checkpoints, forward boundary rows of every strip
// (row 0 and each strip's last row), the two adjoint boundary rows, the dummy slots.
template <typename T>
void layout(GradArgs& a, int S_max, int max_m) {
  constexpr int K = GT<T>::kTape;
  a.nb_max = (max_m + 63 + K - 1) / K;
  a.bnd_cols = max_m + 1;
  a.lds_bytes = (lds_fixed<T>() + 2 * kHapPad + max_m + 64 + K + 31) & ~31;
  const int64_t sz = int64_t(sizeof(T));
  a.ck_off = 0;
  a.bnd_off = (int64_t(S_max) * a.nb_max * 6 * 64 * sz + 255) & ~int64_t(255);
  a.abnd_off = (a.bnd_off + int64_t(S_max + 1) * a.bnd_cols * 4 * sz + 255) & ~int64_t(255);
  a.dummy_off = (a.abnd_off + 2 * int64_t(a.bnd_cols + 1) * 2 * sz + 255) & ~int64_t(255);
  a.wave_stride = (a.dummy_off + 64 * kSlot + 255) & ~int64_t(255);
}
}  // namespace

int GradBatch::create(const GradInput& in, std::unique_ptr<GradBatch>& out) {
  if (in.n_regions < 0 || (in.n_regions > 0 && (!in.region_read || !in.region_hap))) {
    set_error(kInvalidArgument, "region tables are missing");
    return kInvalidArgument;
  }
  const int64_t G = in.n_regions;
  const int64_t NR = G ? in.region_read[G] : 0, NH = G ? in.region_hap[G] : 0;
  if (G && (in.region_read[0] != 0 || in.region_hap[0] != 0)) {
    set_error(kInvalidArgument, "region tables must start at zero");
    return kInvalidArgument;
  }
  for (int64_t g = 0; g < G; ++g)
    if (in.region_read[g + 1] < in.region_read[g] || in.region_hap[g + 1] < in.region_hap[g]) {
      set_error(kInvalidArgument, "region tables must be non-decreasing");
      return kInvalidArgument;
    }
  if (NR > (int64_t(1) << 24) || NH > 0x7fffffff) {  // the work queue advances by 64 per claim
    set_error(kInvalidArgument, "too many reads or haplotypes for one batch");
    return kInvalidArgument;
  }
  int max_m = 0;
  for (int64_t r = 0; r < NR; ++r) {
    const int64_t len = in.read_off[r + 1] - in.read_off[r];
    if (len <= 0) {
      set_error(kInvalidArgument, "read " + std::to_string(r) + " has no bases");
      return kInvalidArgument;
    }
    if (len > kMaxReadLen) {
      set_error(kInvalidArgument, "read " + std::to_string(r) + " has " + std::to_string(len) +
                                      " bases; the engine supports at most " +
                                      std::to_string(kMaxReadLen));
      return kInvalidArgument;
    }
  }
  for (int64_t h = 0; h < NH; ++h) {
    const int64_t len = in.hap_off[h + 1] - in.hap_off[h];
    if (len <= 0) {
      set_error(kInvalidArgument, "haplotype " + std::to_string(h) + " has no bases");
      return kInvalidArgument;
    }
    if (len > 16384) {
      set_error(kInvalidArgument, "haplotype " + std::to_string(h) + " is longer than 16384 bases");
      return kInvalidArgument;
    }
    max_m = std::max(max_m, int(len));
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_error(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    return kInternal;
  }
  std::unique_ptr<GradBatch> b(new GradBatch);
  Impl& d = *b->impl_;
  int dev = 0;
  GRAD_TRY(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  GRAD_TRY(hipGetDeviceProperties(&prop, dev));
  d.num_cus = prop.multiProcessorCount;
  d.n_reads = NR;
  d.n_read_bases = NR ? in.read_off[NR] : 0;
  // Region tables and the work list, longest work first.
  const size_t nr = static_cast<size_t>(NR), ng = static_cast<size_t>(G);
  std::vector<int32_t> read_region(nr), read_local(nr), hap0(ng), nhap(ng);
  std::vector<int64_t> region_out(ng + 1, 0);
  // Work items: a read alone, or two reads of a region whose last, partial strips fit in one
  // (both lengths not multiples of 64, their remainders summing to at most 64), which then run
  // one strip fewer.  Per region the reads go by remainder, largest first, each with the unpaired
  // read of the largest remainder that fits.  Costliest items first.
  std::vector<int32_t> items, singles;
  std::vector<int64_t> item_cost, single_cost;
  items.reserve(2 * nr);
  singles.reserve(2 * nr);
  int S_max = 1;
  for (int64_t g = 0; g < G; ++g) {
    const int64_t R = in.region_read[g + 1] - in.region_read[g];
    const int64_t H = in.region_hap[g + 1] - in.region_hap[g];
    hap0[size_t(g)] = int32_t(in.region_hap[g]);
    nhap[size_t(g)] = int32_t(H);
    region_out[size_t(g + 1)] = region_out[size_t(g)] + R * H;
    const int64_t hap_steps =
        in.hap_off[in.region_hap[g + 1]] - in.hap_off[in.region_hap[g]] + 63 * H;
    auto len = [&](int64_t r) { return int(in.read_off[r + 1] - in.read_off[r]); };
    std::vector<std::pair<int, int32_t>> by_rest;  // (length mod 64, read)
    for (int64_t r = in.region_read[g]; r < in.region_read[g + 1]; ++r) {
      read_region[size_t(r)] = int32_t(g);
      read_local[size_t(r)] = int32_t(r - in.region_read[g]);
      by_rest.emplace_back(len(r) % 64, int32_t(r));
      const int S = (len(r) + 63) / 64;
      S_max = std::max(S_max, S);
      singles.push_back(int32_t(r));
      singles.push_back(-1);
      single_cost.push_back(int64_t(S) * hap_steps);
    }
    std::stable_sort(by_rest.begin(), by_rest.end(),
                     [](const std::pair<int, int32_t>& x, const std::pair<int, int32_t>& y) {
                       return x.first > y.first;
                     });
    std::vector<char> used(by_rest.size(), 0);
    for (size_t i = 0; i < by_rest.size(); ++i) {
      if (used[i]) continue;
      used[i] = 1;
      const int32_t ra = by_rest[i].second;
      int32_t rb = -1;
      if (by_rest[i].first > 0)
        for (size_t j = i + 1; j < by_rest.size(); ++j)
          if (!used[j] && by_rest[j].first > 0 && by_rest[i].first + by_rest[j].first <= 64) {
            used[j] = 1;
            rb = by_rest[j].second;
            break;
          }
      const int rows = len(ra) + (rb >= 0 ? len(rb) : 0);
      const int S = (rows + 63) / 64;
      S_max = std::max(S_max, S);
      items.push_back(ra);
      items.push_back(rb);
      item_cost.push_back(int64_t(S) * hap_steps);
    }
  }
  d.n_pairs = region_out[size_t(G)];
  auto by_cost = [](std::vector<int32_t>& list, const std::vector<int64_t>& cost) {
    const size_t n = cost.size();
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](size_t x, size_t y) { return cost[x] > cost[y]; });
    std::vector<int32_t> sorted(list.size());
    for (size_t i = 0; i < n; ++i) {
      sorted[2 * i] = list[2 * order[i]];
      sorted[2 * i + 1] = list[2 * order[i] + 1];
    }
    list.swap(sorted);
  };
  by_cost(items, item_cost);
  by_cost(singles, single_cost);
  const int64_t n_items = int64_t(item_cost.size());
  const int64_t nb = NR ? in.hap_off[NH] : 0;
  int rc = 0;
  GRAD_TRY(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
  for (auto& e : d.ev) GRAD_TRY(hipEventCreate(&e));
  const size_t nrb = size_t(d.n_read_bases), npair = size_t(d.n_pairs);
  if ((rc = d.alloc(&d.d_err, nrb)) || (rc = d.alloc(&d.d_ins, nrb)) ||
      (rc = d.alloc(&d.d_del, nrb)) || (rc = d.alloc(&d.d_gcp, 1)) ||
      (rc = d.alloc(&d.d_w, npair)) || (rc = d.alloc(&d.d_log10, npair)) ||
      (rc = d.alloc(&d.d_ge, nrb)) || (rc = d.alloc(&d.d_gi, nrb)) ||
      (rc = d.alloc(&d.d_gd, nrb)) || (rc = d.alloc(&d.d_gg, 1)) ||
      (rc = d.alloc(&d.lost, npair)) || (rc = d.alloc(&d.lost_items, 2 * nr)))
    return rc;
  if ((rc = d.upload(&d.bases, in.bases, size_t(d.n_read_bases))) ||
      (rc = d.upload(&d.read_off, in.read_off, size_t(NR + 1))) ||
      (rc = d.upload(&d.read_region, read_region.data(), size_t(NR))) ||
      (rc = d.upload(&d.read_local, read_local.data(), size_t(NR))) ||
      (rc = d.upload(&d.region_out, region_out.data(), size_t(G + 1))) ||
      (rc = d.upload(&d.region_hap0, hap0.data(), size_t(G))) ||
      (rc = d.upload(&d.region_nhap, nhap.data(), size_t(G))) ||
      (rc = d.upload(&d.hap_off, in.hap_off, size_t(NH + 1))) ||
      (rc = d.upload(&d.hap_bases, in.hap_bases, size_t(nb))) ||
      (rc = d.upload(&d.items, items.data(), items.size())) ||
      (rc = d.upload(&d.items_single, singles.data(), singles.size())) ||
      (rc = d.alloc(&d.counters, 4)) || (rc = d.alloc(&d.dg_part, size_t(NR))))
    return rc;
  // Per-wave LDS and scratch, per precision; both passes share the larger per-wave stride.
  layout<float>(d.proto[0], S_max, max_m);
  layout<double>(d.proto[1], S_max, max_m);
  for (const GradArgs& a : d.proto)
    if (a.lds_bytes * kWaves > 64 * 1024) {
      set_error(kInvalidArgument, "haplotypes are too long for the gradient engine's LDS staging");
      return kInvalidArgument;
    }
  const int64_t stride = std::max(d.proto[0].wave_stride, d.proto[1].wave_stride);
  for (int t = 0; t < 2; ++t)
    for (int k = 0; k < 2; ++k) {
      int nbk = 0;
      const void* fn = t ? (k ? reinterpret_cast<const void*>(grad_kernel<double, true>)
                              : reinterpret_cast<const void*>(grad_kernel<double, false>))
                         : (k ? reinterpret_cast<const void*>(grad_kernel<float, true>)
                              : reinterpret_cast<const void*>(grad_kernel<float, false>));
      GRAD_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nbk, fn, kBlock,
                                                           size_t(d.proto[t].lds_bytes) * kWaves));
      d.occ[t][k] = std::max(nbk, 1);
    }
  int occ_max = 0;
  for (auto& row : d.occ)
    for (int o : row) occ_max = std::max(occ_max, o);
  d.n_single = int32_t(NR);
  const int64_t want_waves = std::max<int64_t>((NR + kWaves - 1) / kWaves * kWaves, kWaves);
  d.scratch_waves = std::min<int64_t>(int64_t(occ_max) * d.num_cus * kWaves, want_waves);
  if ((rc = d.alloc(&d.scratch, size_t(d.scratch_waves * stride)))) return rc;
  for (GradArgs& a : d.proto) {
    a.wave_stride = stride;
    a.bases = d.bases;
    a.read_off = d.read_off;
    a.read_region = d.read_region;
    a.read_local = d.read_local;
    a.region_out = d.region_out;
    a.region_hap0 = d.region_hap0;
    a.region_nhap = d.region_nhap;
    a.hap_off = d.hap_off;
    a.hap_bases = d.hap_bases;
    a.items = d.items;
    a.n_items = int32_t(n_items);
    a.dg_part = d.dg_part;
    a.scratch = d.scratch;
    a.err = d.d_err;
    a.ins = d.d_ins;
    a.del = d.d_del;
    a.gcp = d.d_gcp;
    a.g_err = d.d_ge;
    a.g_ins = d.d_gi;
    a.g_del = d.d_gd;
  }
  out = std::move(b);
  return kOk;
}

int GradBatch::set_precision(int fp32) {
  if (fp32 != 0 && fp32 != 1) {
    set_error(kInvalidArgument, "precision must be 0 (FP64) or 1 (FP32 with FP64 rescue)");
    return kInvalidArgument;
  }
  impl_->fp32 = fp32;
  return kOk;
}

int GradBatch::forward(const GradParams& p, double* log10_out, double* kernel_ms) {
  Impl& d = *impl_;
  if (d.n_pairs && !log10_out) {
    set_error(kInvalidArgument, "log10 likelihood output is null");
    return kInvalidArgument;
  }
  if (int rc = d.put_params(p)) return rc;
  GRAD_TRY(hipEventRecord(d.ev[0], d.stream));
  if (d.n_pairs) {
    GradArgs a = d.proto[1];
    a.log10_out = d.d_log10;
    if (int rc = d.launch(false, a)) return rc;
  }
  GRAD_TRY(hipEventRecord(d.ev[1], d.stream));
  if (d.n_pairs)
    GRAD_TRY(hipMemcpyAsync(log10_out, d.d_log10, sizeof(double) * size_t(d.n_pairs),
                            hipMemcpyDeviceToHost, d.stream));
  return d.finish(kernel_ms);
}

int GradBatch::backward(const GradParams& p, const double* weights, double* log10_out,
                        double* g_err, double* g_ins, double* g_del, double* g_gcp,
                        double* kernel_ms) {
  Impl& d = *impl_;
  if ((d.n_read_bases && (!g_err || !g_ins || !g_del)) || !g_gcp) {
    set_error(kInvalidArgument, "gradient outputs are null");
    return kInvalidArgument;
  }
  if (int rc = d.put_params(p)) return rc;
  GradArgs a = d.proto[1];
  if (weights && d.n_pairs) {
    GRAD_TRY(hipMemcpyAsync(d.d_w, weights, sizeof(double) * size_t(d.n_pairs),
                            hipMemcpyHostToDevice, d.stream));
    a.weights = d.d_w;
  }
  a.log10_out = d.d_log10;
  GRAD_TRY(hipEventRecord(d.ev[0], d.stream));
  if (d.n_reads) {
    if (int rc = d.launch(true, a)) return rc;
  }
  hipLaunchKernelGGL(grad_reduce, dim3(1), dim3(kBlock), 0, d.stream, d.dg_part, d.n_reads,
                     d.d_gg);
  GRAD_TRY(hipGetLastError());
  GRAD_TRY(hipEventRecord(d.ev[1], d.stream));
  const size_t nb = sizeof(double) * size_t(d.n_read_bases);
  if (nb) {
    GRAD_TRY(hipMemcpyAsync(g_err, d.d_ge, nb, hipMemcpyDeviceToHost, d.stream));
    GRAD_TRY(hipMemcpyAsync(g_ins, d.d_gi, nb, hipMemcpyDeviceToHost, d.stream));
    GRAD_TRY(hipMemcpyAsync(g_del, d.d_gd, nb, hipMemcpyDeviceToHost, d.stream));
  }
  GRAD_TRY(hipMemcpyAsync(g_gcp, d.d_gg, sizeof(double), hipMemcpyDeviceToHost, d.stream));
  if (log10_out && d.n_pairs)
    GRAD_TRY(hipMemcpyAsync(log10_out, d.d_log10, sizeof(double) * size_t(d.n_pairs),
                            hipMemcpyDeviceToHost, d.stream));
  return d.finish(kernel_ms);
}

int64_t GradBatch::n_pairs() const { return impl_->n_pairs; }
int64_t GradBatch::n_read_bases() const { return impl_->n_read_bases; }
int64_t GradBatch::scratch_bytes() const {
  return impl_->scratch_waves * impl_->proto[1].wave_stride;
}

int dpp_shl_selftest(int32_t* out64) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    set_error(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    return kInternal;
  }
  int32_t in[64];
  for (int i = 0; i < 64; ++i) in[i] = 1000 + i;
  int32_t *din = nullptr, *dout = nullptr;
  GRAD_TRY(hipMalloc(&din, sizeof(in)));
  GRAD_TRY(hipMalloc(&dout, sizeof(in)));
  GRAD_TRY(hipMemcpy(din, in, sizeof(in), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(dpp_shl_selftest_kernel, dim3(1), dim3(64), 0, nullptr, din, dout);
  GRAD_TRY(hipGetLastError());
  GRAD_TRY(hipMemcpy(out64, dout, sizeof(in), hipMemcpyDeviceToHost));
  GRAD_TRY(hipFree(din));
  GRAD_TRY(hipFree(dout));
  return kOk;
}

}  // namespace fgkl
