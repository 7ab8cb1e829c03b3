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
//    cells per lane in registers.  Work: forward twice plus the reverse sweep.
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
};
template <> struct GT<float> {
  static constexpr double kInit = 0x1p120;
  static constexpr double kLog10Init = 120.0 * 0.301029995663981195213738894724493027;
  static constexpr double kMinAccepted = 1e-28;
  static constexpr int kTape = 16;
  static constexpr bool kPrefetch = true;   // the next block's inputs load during this block
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
  const int32_t* items;      // reads, most work first (or, rescuing, the reads with lost pairs)
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
// tp[k] receives the state before step k (TAPE).
template <typename T, bool FM, bool ACC, bool OUT, bool TAPE>
__device__ __forceinline__ void fblock(St<T>& st, St<T>& P, T& acc, St<T>* tp, const GRow<T>& rw,
                                       const Lds<T>& l, St4<T>* out, int t0, int lane, int m) {
  constexpr int K = GT<T>::kTape;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if constexpr (TAPE) tp[k] = st;
    const int c = t0 + k - lane + 1;
    const St<T> Q = shr_st(l.stage[k + 1], st);
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
template <typename T, bool CK, bool ACC>
__device__ T forward_strip(const GradArgs& a, const Lds<T>& l, const GRow<T>& rw, int lane, int s,
                           int m, int L, int nb, bool has_next, char* wscr) {
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
    if (b == 0) P = shr_st(l.stage[0], st);  // the diagonal of step 0: column 0 of the row above
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
        fblock<T, true, ACC, true, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      else
        fblock<T, false, ACC, true, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      wave_sync_lds();
      // Lanes 0..K-1 store lane 63's columns t0+lane-62 that lie in 1..m.
      const int c = t0 + lane - 62;
      const bool ok = lane < K && c >= 1 && c <= m;
      const St4<T> w = l.fout[min(lane, K - 1)];
      St4<T>* dst = ok ? reinterpret_cast<St4<T>*>(bnd_out + size_t(max(c, 0)) * 4) : gdummy;
      *dst = St4<T>{w.m, w.x, w.y, T(0)};
    } else {
      if (fm)
        fblock<T, true, ACC, false, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      else
        fblock<T, false, ACC, false, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
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
template <typename T, bool BM, bool SEED>
__device__ __forceinline__ void rblock(const St<T>* tp, const St<T>& P0, const GRow<T>& rw,
                                       const Lds<T>& l, Adj2<T>* aout, int t0, int lane, int m,
                                       T seedv, T& AX, T& AY, T& Q, T& Bq, GSum<T>& gs) {
  constexpr int K = GT<T>::kTape;
  St<T> U = shr_st(l.stage[K], tp[K - 1]);  // up values of the last step
#pragma unroll
  for (int k = K - 1; k >= 0; --k) {
    const int col = t0 + k - lane + 1;
    const bool mt = is_match(l.hap[kHapPad + col - 1], rw.rb);
    const T prior = mt ? rw.pm : rw.px;
    const T ds = mt ? T(-1) : T(1.0 / 3.0);  // d prior / d e
    // The row below: its insertion adjoint in this column and Q one column to the right.
    const Adj2<T> below = l.astage[k];
    const T AXb = shl1(below.ax, AX);
    const T Qn = shl1(below.q, Q);
    const T Qd = Bq;
    Bq = Qn;
    const St<T> D = k ? shr_st(l.stage[k], tp[k - 1]) : P0;
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

// Pass B over strip s.  `dgl` collects the lane's gap-continuation terms.
template <typename T, bool SEED>
__device__ void backward_strip(const GradArgs& a, const Lds<T>& l, const GRow<T>& rw, int lane,
                               int s, int v, int n, int m, int L, int nb, T seedv, int64_t o,
                               char* wscr, double& dgl) {
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
    St<T> tp[K + 1];
    St<T> st{cur.c[0], cur.c[1], cur.c[2]};
    const St<T> P0{cur.c[3], cur.c[4], cur.c[5]};
    St<T> P = P0;
    T unused = T(0);
    if (t0 < L - 1)
      fblock<T, true, false, false, true>(st, P, unused, tp, rw, l, nullptr, t0, lane, m);
    else
      fblock<T, false, false, false, true>(st, P, unused, tp, rw, l, nullptr, t0, lane, m);
    GSum<T> gs{T(0), T(0), T(0), T(0), T(0), T(0)};
    if (t0 < L - 1 || t0 + K > m)
      rblock<T, true, SEED>(tp, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq, gs);
    else
      rblock<T, false, SEED>(tp, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq, gs);
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
  const bool real = v <= n;
  const int64_t i = o + (real ? v : n) - 1;
  double* pe = real ? a.g_err + i : reinterpret_cast<double*>(gdummy);
  double* pi = real ? a.g_ins + i : reinterpret_cast<double*>(gdummy) + 1;
  double* pd = real ? a.g_del + i : reinterpret_cast<double*>(gdummy) + 2;
  *pe += ge * kInvLn10;
  *pi += gi * kInvLn10;  // M->M = 1 - d_ins - d_del
  *pd += gd * kInvLn10;
  dgl += real ? gq : 0.0;  // I->M = D->M = 1 - g
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
    const int r = __builtin_amdgcn_readfirstlane(a.items[item]);
    const int64_t o = a.read_off[r];
    const int n = __builtin_amdgcn_readfirstlane(int(a.read_off[r + 1] - o));
    const int gi = __builtin_amdgcn_readfirstlane(a.read_region[r]);
    const int hap0 = __builtin_amdgcn_readfirstlane(a.region_hap0[gi]);
    const int H = __builtin_amdgcn_readfirstlane(a.region_nhap[gi]);
    const int64_t out_base = a.region_out[gi] + int64_t(a.read_local[r]) * H;
    const int S = (n + 63) / 64;  // strips of 64 read rows
    if (BWD && !a.rescue) {  // a rescue adds to what the FP32 pass wrote
      for (int s = 0; s < S; ++s) {
        const int v = 64 * s + lane + 1;
        const bool real = v <= n;
        const int64_t i = o + (real ? v : n) - 1;
        *(real ? a.g_err + i : reinterpret_cast<double*>(gslot)) = 0.0;
        *(real ? a.g_ins + i : reinterpret_cast<double*>(gslot) + 1) = 0.0;
        *(real ? a.g_del + i : reinterpret_cast<double*>(gslot) + 2) = 0.0;
      }
    }
    double dgl = 0.0;
    int any_lost = 0;
    for (int h = 0; h < H; ++h) {
      if (a.rescue && !__builtin_amdgcn_readfirstlane(int(a.lost[out_base + h]))) continue;
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
      // column (reference.rs:27-28).
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
      const double w = BWD ? (a.weights ? a.weights[out_base + h] : 1.0) : 0.0;
      const bool grad = BWD && w != 0.0;
      // Pass F.
      double total = 0.0;
      for (int s = 0; s < S; ++s) {
        const int v = 64 * s + lane + 1;
        const int L = min(64, n - 64 * s);
        const int nb = (m + L - 1 + K - 1) / K;
        GRow<T> rw;
        load_grow(a, o, n, v, g, rw);
        if (s + 1 < S) {
          if (grad)
            forward_strip<T, true, false>(a, l, rw, lane, s, m, L, nb, true, wscr);
          else
            forward_strip<T, false, false>(a, l, rw, lane, s, m, L, nb, true, wscr);
        } else {
          const T acc = grad ? forward_strip<T, true, true>(a, l, rw, lane, s, m, L, nb, false, wscr)
                             : forward_strip<T, false, true>(a, l, rw, lane, s, m, L, nb, false, wscr);
          total = double(readlane(acc, n - 1 - 64 * s));
        }
        wave_sync_global();
      }
      // FP32: a pair below the rescue threshold (or NaN) is left to the FP64 pass.
      const bool lost = kFp32 && !(total >= GT<T>::kMinAccepted);
      if (kFp32 && lane == 0) a.lost[out_base + h] = lost ? 1 : 0;
      any_lost |= lost;
      if (a.log10_out && lane == 0)
        a.log10_out[out_base + h] = lost ? __builtin_nan("") : log10(total) - GT<T>::kLog10Init;
      if (lost || !grad || !(total > 0.0) || !(total <= 1.7976931348623157e308)) continue;
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
      // Every adjoint is w d ln L / d cell: seeded with w / L at the last row.
      const T seed = T(w / total);
      for (int s = S - 1; s >= 0; --s) {
        const int v = 64 * s + lane + 1;
        const int L = min(64, n - 64 * s);
        const int nb = (m + L - 1 + K - 1) / K;
        GRow<T> rw;
        load_grow(a, o, n, v, g, rw);
        if (s == S - 1)  // the read's last row lies in its bottom strip
          backward_strip<T, true>(a, l, rw, lane, s, v, n, m, L, nb, v == n ? seed : T(0), o,
                                  wscr, dgl);
        else
          backward_strip<T, false>(a, l, rw, lane, s, v, n, m, L, nb, T(0), o, wscr, dgl);
        wave_sync_global();
      }
    }
    if (kFp32 && any_lost && lane == 0) {
      const int idx = atomicAdd(a.lost_count, 1);
      a.lost_items[idx] = r;
    }
    if constexpr (BWD) {
      // Sum over lanes in lane order through LDS; every lane stores the same value.
      double* red = reinterpret_cast<double*>(l.dummy);
      wave_sync_lds();
      red[lane] = dgl;
      wave_sync_lds();
      double tot = 0.0;
#pragma unroll 4
      for (int i = 0; i < 64; ++i) tot += red[i];
      if (a.rescue)
        a.dg_part[r] += tot * kInvLn10;
      else
        a.dg_part[r] = tot * kInvLn10;
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
  int32_t* items = nullptr;
  // Work-queue heads and counts: 0 forward, 1 backward, 2 reads with lost pairs, 3 the rescue.
  int32_t* counters = nullptr;
  double* dg_part = nullptr;
  uint8_t* lost = nullptr;        // per pair (FP32 forward)
  int32_t* lost_items = nullptr;  // reads with lost pairs
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
  std::vector<int64_t> cost(nr);
  int max_n = 0;
  for (int64_t g = 0; g < G; ++g) {
    const int64_t R = in.region_read[g + 1] - in.region_read[g];
    const int64_t H = in.region_hap[g + 1] - in.region_hap[g];
    hap0[size_t(g)] = int32_t(in.region_hap[g]);
    nhap[size_t(g)] = int32_t(H);
    region_out[size_t(g + 1)] = region_out[size_t(g)] + R * H;
    const int64_t hap_sum = in.hap_off[in.region_hap[g + 1]] - in.hap_off[in.region_hap[g]];
    for (int64_t r = in.region_read[g]; r < in.region_read[g + 1]; ++r) {
      read_region[size_t(r)] = int32_t(g);
      read_local[size_t(r)] = int32_t(r - in.region_read[g]);
      const int64_t n = in.read_off[r + 1] - in.read_off[r];
      max_n = std::max(max_n, int(n));
      cost[size_t(r)] = ((n + 64) / 64) * (hap_sum + 64 * H);
    }
  }
  d.n_pairs = region_out[size_t(G)];
  std::vector<int32_t> items(nr);
  for (int64_t r = 0; r < NR; ++r) items[size_t(r)] = int32_t(r);
  std::stable_sort(items.begin(), items.end(),
                   [&](int32_t x, int32_t y) { return cost[size_t(x)] > cost[size_t(y)]; });
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
      (rc = d.alloc(&d.lost, npair)) || (rc = d.alloc(&d.lost_items, nr)))
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
      (rc = d.upload(&d.items, items.data(), size_t(NR))) ||
      (rc = d.alloc(&d.counters, 4)) || (rc = d.alloc(&d.dg_part, size_t(NR))))
    return rc;
  // Per-wave LDS and scratch, per precision; both passes share the larger per-wave stride.
  const int S_max = (max_n + 64) / 64;
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
    a.n_items = int32_t(NR);
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
