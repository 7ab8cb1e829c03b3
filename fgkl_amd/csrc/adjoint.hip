// The adjoint of the PairHMM forward algorithm on gfx950: log10 likelihoods and the gradient of
// sum_p w_p log10 L_p with respect to each read base's error, insertion and deletion
// probabilities and the global gap-continuation probability.  The reference has no gradient;
// the model is its FP64 forward recurrence (reference.rs:20-64) in a continuous
// parameterisation: priors (1 - e, e / 3) with N matching everything, M->M = 1 - d_ins - d_del,
// M->I = d_ins, M->D = d_del, I->I = D->D = g, I->M = D->M = 1 - g, row 0 deletions 2^1020 / m.
//
// Layout (one work item per wave64, a read against every haplotype of its region):
//  * Read rows 1..n are cut into strips of 64; lane l of strip s owns row 64 s + l + 1.  One row
//    per lane keeps the FP64 working set of the adjoint in registers.  DP row 0 is stored as the
//    boundary row above strip 0.
//  * Pass F sweeps each strip along anti-diagonals (at step t lane l is on column t - l + 1,
//    the row above arriving by DPP wave_shr:1 from lane l-1; lane 0 takes it from the previous
//    strip's last row, which lane 63 stored) and stores every lane's state once every K steps:
//    the checkpoints.  It returns L.
//  * Pass B walks the strips bottom-up and, inside a strip, the K-step blocks right to left: a
//    block is recomputed from its checkpoint into a register tape, then swept in reverse.  The
//    adjoint of a cell needs the row below's values one and two steps later, which arrive by
//    DPP wave_shl:1 from lane l+1 (lane 63 reads the lower strip's top row, stored by its lane
//    0).  Gradient terms accumulate in FP64 registers per row and are written once per strip.
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

constexpr int kTape = 8;         // steps per checkpoint block, recomputed into registers
constexpr double kInitial = 0x1p1020;  // reference.rs:12
constexpr double kLog10Initial = 1020.0 * 0.301029995663981195213738894724493027;
constexpr double kInvLn10 = 0.434294481903251827651128918916605082;
constexpr int kHapPad = 80;      // zero bytes before the staged haplotype (columns down to -79)

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
  const int32_t* items;      // reads, most work first
  int32_t n_items;
  int32_t* counter;
  const double* weights;     // per pair, or null for all ones
  double* log10_out;         // per pair, or null
  double *g_err, *g_ins, *g_del;
  double* dg_part;           // per read
  double* scratch;
  int64_t wave_stride;       // doubles of scratch per wave
  int64_t ck_off, bnd_off, abnd_off;  // sections of a wave's scratch, in doubles
  int64_t dummy_off;         // 64 x 4 doubles that absorb the stores of lanes with none to make
  int32_t nb_max;            // checkpoint blocks per strip
  int32_t bnd_cols;          // columns of a stored boundary row
  int32_t lds_bytes;         // per wave
};

// A lane's read row: its priors and gap-open parameters, the row below's (the adjoint reads
// them), and the global gap continuation g and 1 - g, which every row shares.
struct GRow {
  double pm, px, mm, mi, md, mmB, miB;
  double g, ig;
  uint32_t rb;
};

struct St {
  double m, x, y;
};
struct alignas(32) St4 {
  double m, x, y, pad;
};
struct alignas(16) Adj2 {
  double ax, q;
};

// Row v (1-based) of the read at `o`; rows past the read are zero and stay zero.  Every lane
// loads (from a clamped, valid index) and selects, so no lane-dependent branch is taken.
__device__ __forceinline__ void load_grow(const GradArgs& a, int64_t o, int n, int v, double g,
                                          GRow& rw) {
  const bool real = v <= n, below = v < n;
  const int64_t i = o + (real ? v : n) - 1;
  const int64_t j = below ? i + 1 : i;
  const double e = a.err[i], di = a.ins[i], dd = a.del[i];
  const double di2 = a.ins[j], dd2 = a.del[j];
  const uint32_t b = a.bases[i];
  rw.pm = real ? 1.0 - e : 0.0;
  rw.px = real ? e / 3.0 : 0.0;
  rw.mm = real ? 1.0 - di - dd : 0.0;
  rw.mi = real ? di : 0.0;
  rw.md = real ? dd : 0.0;
  rw.mmB = below ? 1.0 - di2 - dd2 : 0.0;
  rw.miB = below ? di2 : 0.0;
  rw.g = g;
  rw.ig = 1.0 - g;
  rw.rb = real ? b : 0x100u;
}

__device__ __forceinline__ bool is_match(uint32_t hb, uint32_t rb) {
  return hb == rb || hb == uint32_t('N') || rb == uint32_t('N');
}

__device__ __forceinline__ St shr_st(const St4& old, const St& src) {
  return St{shr1(old.m, src.m), shr1(old.x, src.x), shr1(old.y, src.y)};
}

// A wave's LDS: the staged boundary rows and haplotype, and the per-block staging of the values
// one lane hands to the neighbouring strip.  The lane that owns a boundary writes its staging
// slot each step; every other lane writes a private dummy slot, so the writes need no branch.
struct Lds {
  St4* stage;     // kTape + 1 columns of the row above the strip (lane 0's neighbour)
  Adj2* astage;   // kTape columns of the adjoint of the row below the strip (lane 63's)
  St4* fout;      // kTape columns of lane 63's state, for the strip below
  Adj2* aout;     // kTape columns of lane 0's adjoint, for the strip above
  uint8_t* dummy; // 64 + kTape slots of 32 bytes
  uint8_t* hap;   // kHapPad zeros, the haplotype, zeros
};

constexpr int kLdsFixed = (kTape + 1) * 32 + kTape * 16 + kTape * 32 + kTape * 16 + (64 + kTape) * 32;

__device__ __forceinline__ Lds lds_parts(uint8_t* base) {
  Lds l;
  l.stage = reinterpret_cast<St4*>(base);
  base += (kTape + 1) * 32;
  l.astage = reinterpret_cast<Adj2*>(base);
  base += kTape * 16;
  l.fout = reinterpret_cast<St4*>(base);
  base += kTape * 32;
  l.aout = reinterpret_cast<Adj2*>(base);
  base += kTape * 16;
  l.dummy = base;
  base += (64 + kTape) * 32;
  l.hap = base;
  return l;
}

__device__ __forceinline__ double prior_of(const Lds& l, int c, const GRow& rw) {
  return is_match(l.hap[kHapPad + c - 1], rw.rb) ? rw.pm : rw.px;
}

// kTape forward steps from `st` (the state before step t0) with diagonal `P`.  FM freezes lanes
// that have not reached column 1; ACC adds M + X of columns 1..m; OUT stages lane 63's state.
// tp[k] receives the state before step k (TAPE).
template <bool FM, bool ACC, bool OUT, bool TAPE>
__device__ __forceinline__ void fblock(St& st, St& P, double& acc, St* tp, const GRow& rw,
                                       const Lds& l, St4* out, int t0, int lane, int m) {
#pragma unroll
  for (int k = 0; k < kTape; ++k) {
    if constexpr (TAPE) tp[k] = st;
    const int c = t0 + k - lane + 1;
    const St Q = shr_st(l.stage[k + 1], st);
    const double prior = prior_of(l, c, rw);
    const double Y = fma(st.y, rw.g, st.m * rw.md);
    const double M = prior * fma(P.y, rw.ig, fma(P.x, rw.ig, P.m * rw.mm));
    const double X = fma(Q.x, rw.g, Q.m * rw.mi);
    if constexpr (FM) {
      const bool live = c >= 1;
      st.m = live ? M : st.m;
      st.x = live ? X : st.x;
      st.y = live ? Y : st.y;
    } else {
      st = St{M, X, Y};
    }
    P = Q;
    if constexpr (ACC) {
      if constexpr (FM)
        acc += (c >= 1 && c <= m) ? st.m + st.x : 0.0;
      else
        acc += (c <= m) ? st.m + st.x : 0.0;
    }
    if constexpr (OUT) out[k] = St4{st.m, st.x, st.y, 0.0};
  }
  if constexpr (TAPE) tp[kTape] = st;
}

// Stages columns t0 .. t0+kTape of a stored boundary row (clamped to 0..m) for lane 0; lanes
// past kTape write their copy into the dummy area.
__device__ __forceinline__ void stage_above(const Lds& l, const double* bnd, int t0, int m,
                                            int lane) {
  wave_sync_lds();
  const int j = min(lane, kTape);
  const int col = min(t0 + j, m);
  const double4 w = *reinterpret_cast<const double4*>(bnd + size_t(col) * 4);
  St4* dst = lane <= kTape ? &l.stage[j] : reinterpret_cast<St4*>(l.dummy + lane * 32);
  *dst = St4{w.x, w.y, w.z, 0.0};
  wave_sync_lds();
}

// Pass F over strip s: returns the lane's sum of M + X over columns 1..m when ACC (row n's lane
// is the one that matters).
template <bool CK, bool ACC>
__device__ double forward_strip(const GradArgs& a, const Lds& l, const GRow& rw, int lane, int s,
                                int v, int m, int L, int nb, bool has_next, double* wscr) {
  const double* bnd_in = wscr + a.bnd_off + size_t(s) * a.bnd_cols * 4;
  double* bnd_out = wscr + a.bnd_off + size_t(s + 1) * a.bnd_cols * 4;
  double* ck = wscr + a.ck_off + size_t(s) * a.nb_max * 6 * 64;
  St4* out = lane == 63 ? l.fout : reinterpret_cast<St4*>(l.dummy + lane * 32);
  St st{0.0, 0.0, 0.0};  // column 0 below DP row 0
  St P{0.0, 0.0, 0.0};
  double acc = 0.0;
  double* gdummy = wscr + a.dummy_off + lane * 4;
  if (has_next) *reinterpret_cast<double4*>(bnd_out) = make_double4(0.0, 0.0, 0.0, 0.0);
  for (int b = 0; b < nb; ++b) {
    const int t0 = b * kTape;
    stage_above(l, bnd_in, t0, m, lane);
    if (b == 0) P = shr_st(l.stage[0], st);  // the diagonal of step 0: column 0 of the row above
    if constexpr (CK) {
      double* c = ck + size_t(b) * 6 * 64 + lane;
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
        fblock<true, ACC, true, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      else
        fblock<false, ACC, true, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      wave_sync_lds();
      // Lanes 0..kTape-1 store lane 63's columns t0+lane-62 that lie in 1..m.
      const int c = t0 + lane - 62;
      const bool ok = lane < kTape && c >= 1 && c <= m;
      const St4 w = l.fout[min(lane, kTape - 1)];
      double* dst = ok ? bnd_out + size_t(max(c, 0)) * 4 : gdummy;
      *reinterpret_cast<double4*>(dst) = make_double4(w.m, w.x, w.y, 0.0);
    } else {
      if (fm)
        fblock<true, ACC, false, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
      else
        fblock<false, ACC, false, false>(st, P, acc, nullptr, rw, l, out, t0, lane, m);
    }
  }
  return acc;
}

// The reverse sweep of one block over the tape.  BM zeroes the adjoint outside columns 1..m.
template <bool BM>
__device__ __forceinline__ void rblock(const St* tp, const St& P0, const GRow& rw, const Lds& l,
                                       Adj2* aout, int t0, int lane, int m, double seedv,
                                       double& AX, double& AY, double& Q, double& Bq, double& ge,
                                       double& gmm, double& gim, double& gmi, double& gmd,
                                       double& gg) {
  St U = shr_st(l.stage[kTape], tp[kTape - 1]);  // up values of the last step
#pragma unroll
  for (int k = kTape - 1; k >= 0; --k) {
    const int col = t0 + k - lane + 1;
    const bool mt = is_match(l.hap[kHapPad + col - 1], rw.rb);
    const double prior = mt ? rw.pm : rw.px;
    const double ds = mt ? -1.0 : (1.0 / 3.0);  // d prior / d e
    // The row below: its insertion adjoint in this column and Q one column to the right.
    const Adj2 below = l.astage[k];
    const double AXb = shl1(below.ax, AX);
    const double Qn = shl1(below.q, Q);
    const double Qd = Bq;
    Bq = Qn;
    const St D = k ? shr_st(l.stage[k], tp[k - 1]) : P0;
    const St& Lf = tp[k];
    const double pr = Qd * rw.ig;
    double ay = fma(AY, rw.g, pr);
    const double amp = fma(Qd, rw.mmB, AY * rw.md);
    double ax = fma(AXb, rw.g, pr) + seedv;
    double am = fma(AXb, rw.miB, amp) + seedv;
    if constexpr (BM) {
      const bool live = col >= 1 && col <= m;
      ay = live ? ay : 0.0;
      ax = live ? ax : 0.0;
      am = live ? am : 0.0;
    }
    const double q = am * prior;
    const double xy = D.x + D.y;
    const double spre = fma(D.m, rw.mm, xy * rw.ig);  // M before the prior
    ge = fma(am * spre, ds, ge);
    gmm = fma(q, D.m, gmm);
    gim = fma(q, xy, gim);
    gmi = fma(ax, U.m, gmi);
    gmd = fma(ay, Lf.m, gmd);
    gg = fma(ax, U.x, fma(ay, Lf.y, gg));
    AY = ay;
    AX = ax;
    Q = q;
    U = D;
    aout[k] = Adj2{ax, q};
  }
}

// Pass B over strip s.  `dgl` collects the lane's gap-continuation terms.
__device__ void backward_strip(const GradArgs& a, const Lds& l, const GRow& rw, int lane, int s,
                               int v, int n, int m, int L, int nb, double seedv, int64_t o,
                               double* wscr, double& dgl) {
  const double* bnd_in = wscr + a.bnd_off + size_t(s) * a.bnd_cols * 4;
  const double* ck = wscr + a.ck_off + size_t(s) * a.nb_max * 6 * 64;
  const int abnd_cols = a.bnd_cols + 1;
  const Adj2* ab_in = reinterpret_cast<const Adj2*>(wscr + a.abnd_off) + size_t((s + 1) & 1) * abnd_cols;
  Adj2* ab_out = reinterpret_cast<Adj2*>(wscr + a.abnd_off) + size_t(s & 1) * abnd_cols;
  Adj2* aout = lane == 0 ? l.aout : reinterpret_cast<Adj2*>(l.dummy + lane * 32);
  double* gdummy = wscr + a.dummy_off + lane * 4;
  ab_out[0] = Adj2{0.0, 0.0};
  ab_out[m + 1] = Adj2{0.0, 0.0};
  double AX = 0.0, AY = 0.0, Q = 0.0, Bq = 0.0;  // columns beyond m: zero adjoints
  double ge = 0.0, gmm = 0.0, gim = 0.0, gmi = 0.0, gmd = 0.0, gg = 0.0;
  for (int b = nb - 1; b >= 0; --b) {
    const int t0 = b * kTape;
    {
      const Adj2 w = ab_in[min(max(t0 + lane - 62, 0), m + 1)];
      Adj2* dst = lane < kTape ? &l.astage[lane] : reinterpret_cast<Adj2*>(l.dummy + lane * 32);
      wave_sync_lds();
      *dst = w;
    }
    stage_above(l, bnd_in, t0, m, lane);
    // Recompute the block from its checkpoint into the tape: tp[k] is the state before step k.
    const double* c = ck + size_t(b) * 6 * 64 + lane;
    St tp[kTape + 1];
    St st{c[0 * 64], c[1 * 64], c[2 * 64]};
    const St P0{c[3 * 64], c[4 * 64], c[5 * 64]};
    St P = P0;
    double unused = 0.0;
    if (t0 < L - 1)
      fblock<true, false, false, true>(st, P, unused, tp, rw, l, nullptr, t0, lane, m);
    else
      fblock<false, false, false, true>(st, P, unused, tp, rw, l, nullptr, t0, lane, m);
    if (t0 < L - 1 || t0 + kTape > m)
      rblock<true>(tp, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq, ge, gmm, gim, gmi, gmd, gg);
    else
      rblock<false>(tp, P0, rw, l, aout, t0, lane, m, seedv, AX, AY, Q, Bq, ge, gmm, gim, gmi, gmd, gg);
    wave_sync_lds();
    {
      const int col = t0 + lane + 1;  // lane 0's column at step t0 + lane
      const bool ok = lane < kTape && col <= m;
      const Adj2 w = l.aout[min(lane, kTape - 1)];
      Adj2* dst = ok ? &ab_out[col] : reinterpret_cast<Adj2*>(gdummy);
      *dst = w;
    }
  }
  const bool real = v <= n;
  const int64_t i = o + (real ? v : n) - 1;
  double* pe = real ? a.g_err + i : gdummy;
  double* pi = real ? a.g_ins + i : gdummy + 1;
  double* pd = real ? a.g_del + i : gdummy + 2;
  *pe += ge * kInvLn10;
  *pi += (gmi - gmm) * kInvLn10;  // M->M = 1 - d_ins - d_del
  *pd += (gmd - gmm) * kInvLn10;
  dgl += real ? gg - gim : 0.0;   // I->M = D->M = 1 - g
}

template <bool BWD>
__global__ void __launch_bounds__(256) grad_kernel(GradArgs a) {
  extern __shared__ __attribute__((aligned(32))) uint8_t glds[];
  const int lane = threadIdx.x & 63;
  const int wave_in_block = threadIdx.x >> 6;
  const int64_t wave_global = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave_in_block;
  const Lds l = lds_parts(glds + wave_in_block * a.lds_bytes);
  double* wscr = a.scratch + wave_global * a.wave_stride;
  const double g = a.gcp[0];

  for (;;) {
    // Every lane adds one, so a claim advances the counter by 64 and every lane's old value lies
    // in [64 i, 64 i + 63]; no lane-dependent branch reaches the readfirstlane.
    const int item = __builtin_amdgcn_readfirstlane(atomicAdd(a.counter, 1)) >> 6;
    if (item >= a.n_items) break;
    const int r = __builtin_amdgcn_readfirstlane(a.items[item]);
    const int64_t o = a.read_off[r];
    const int n = __builtin_amdgcn_readfirstlane(int(a.read_off[r + 1] - o));
    const int gi = __builtin_amdgcn_readfirstlane(a.read_region[r]);
    const int hap0 = __builtin_amdgcn_readfirstlane(a.region_hap0[gi]);
    const int H = __builtin_amdgcn_readfirstlane(a.region_nhap[gi]);
    const int64_t out_base = a.region_out[gi] + int64_t(a.read_local[r]) * H;
    const int S = (n + 63) / 64;  // strips of 64 read rows
    if constexpr (BWD) {
      double* gdummy = wscr + a.dummy_off + lane * 4;
      for (int s = 0; s < S; ++s) {
        const int v = 64 * s + lane + 1;
        const bool real = v <= n;
        const int64_t i = o + (real ? v : n) - 1;
        *(real ? a.g_err + i : gdummy) = 0.0;
        *(real ? a.g_ins + i : gdummy + 1) = 0.0;
        *(real ? a.g_del + i : gdummy + 2) = 0.0;
      }
    }
    double dgl = 0.0;
    for (int h = 0; h < H; ++h) {
      const int64_t hoff = a.hap_off[hap0 + h];
      const int m = __builtin_amdgcn_readfirstlane(int(a.hap_off[hap0 + h + 1] - hoff));
      wave_sync_lds();
      const int hap_total = kHapPad + m + 64 + kTape + kHapPad;
      for (int i0 = 0; i0 < hap_total; i0 += 64) {
        const int i = i0 + lane;
        const int c = i - kHapPad;
        const bool in = c >= 0 && c < m;
        const uint8_t hb = a.hap_bases[hoff + (in ? c : 0)];
        *(i < hap_total ? &l.hap[i] : l.dummy + lane) = in ? hb : uint8_t(0);
      }
      // The row above strip 0 is DP row 0: no match or insertion mass, deletions 2^1020 / m in
      // every column (reference.rs:27-28).
      const double init = kInitial / double(m);
      for (int c0 = 0; c0 <= m; c0 += 64) {
        const int c = c0 + lane;
        double* dst = c <= m ? wscr + a.bnd_off + size_t(c) * 4 : wscr + a.dummy_off + lane * 4;
        *reinterpret_cast<double4*>(dst) = make_double4(0.0, 0.0, init, 0.0);
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
        const int nb = (m + L - 1 + kTape - 1) / kTape;
        GRow rw;
        load_grow(a, o, n, v, g, rw);
        if (s + 1 < S) {
          if (grad)
            forward_strip<true, false>(a, l, rw, lane, s, v, m, L, nb, true, wscr);
          else
            forward_strip<false, false>(a, l, rw, lane, s, v, m, L, nb, true, wscr);
        } else {
          const double acc =
              grad ? forward_strip<true, true>(a, l, rw, lane, s, v, m, L, nb, false, wscr)
                   : forward_strip<false, true>(a, l, rw, lane, s, v, m, L, nb, false, wscr);
          total = readlane(acc, n - 1 - 64 * s);
        }
        wave_sync_global();
      }
      if (a.log10_out) a.log10_out[out_base + h] = log10(total) - kLog10Initial;  // every lane
      if (!grad || !(total > 0.0) || !(total <= 1.7976931348623157e308)) continue;
      // Pass B: the bottom strip has nothing below it.
      {
        const int abnd_cols = a.bnd_cols + 1;
        Adj2* z = reinterpret_cast<Adj2*>(wscr + a.abnd_off) + size_t(S & 1) * abnd_cols;
        for (int c0 = 0; c0 <= m + 1; c0 += 64) {
          const int c = c0 + lane;
          *(c <= m + 1 ? &z[c] : reinterpret_cast<Adj2*>(wscr + a.dummy_off + lane * 4)) =
              Adj2{0.0, 0.0};
        }
        wave_sync_global();
      }
      const double seed = w / total;  // every adjoint is w d ln L / d cell
      for (int s = S - 1; s >= 0; --s) {
        const int v = 64 * s + lane + 1;
        const int L = min(64, n - 64 * s);
        const int nb = (m + L - 1 + kTape - 1) / kTape;
        GRow rw;
        load_grow(a, o, n, v, g, rw);
        backward_strip(a, l, rw, lane, s, v, n, m, L, nb, v == n ? seed : 0.0, o, wscr, dgl);
        wave_sync_global();
      }
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
  int32_t* counters = nullptr;
  double* dg_part = nullptr;
  double* scratch = nullptr;
  int64_t scratch_waves = 0;
  GradArgs proto{};
  int occ[2] = {0, 0};
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
  int launch(bool bwd, GradArgs a, hipStream_t stream) {
    hipLaunchKernelGGL(grad_fill, dim3(1), dim3(64), 0, stream, counters, 2, 0);
    GRAD_TRY(hipGetLastError());
    if (a.n_items > 0) {
      const int64_t want = (int64_t(a.n_items) + kWaves - 1) / kWaves;
      const int64_t resident = int64_t(occ[bwd ? 1 : 0]) * num_cus;
      const int grid = int(std::min({want, resident, scratch_waves / kWaves}));
      a.counter = counters + (bwd ? 1 : 0);
      if (bwd)
        hipLaunchKernelGGL(grad_kernel<true>, dim3(grid), dim3(kBlock), size_t(a.lds_bytes) * kWaves,
                           stream, a);
      else
        hipLaunchKernelGGL(grad_kernel<false>, dim3(grid), dim3(kBlock),
                           size_t(a.lds_bytes) * kWaves, stream, a);
      GRAD_TRY(hipGetLastError());
    }
    return kOk;
  }
};

GradBatch::GradBatch() : impl_(new Impl) {}
GradBatch::~GradBatch() { delete impl_; }

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
      (rc = d.alloc(&d.d_gd, nrb)) || (rc = d.alloc(&d.d_gg, 1)))
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
      (rc = d.alloc(&d.counters, 2)) || (rc = d.alloc(&d.dg_part, size_t(NR))))
    return rc;
  // Per-wave LDS and scratch.
  const int lds = (kLdsFixed + 2 * kHapPad + max_m + 64 + kTape + 31) & ~31;
  if (lds * kWaves > 64 * 1024) {
    set_error(kInvalidArgument, "haplotypes are too long for the gradient engine's LDS staging");
    return kInvalidArgument;
  }
  const int S_max = (max_n + 64) / 64;
  const int nb_max = (max_m + 63 + kTape - 1) / kTape;
  const int bnd_cols = max_m + 1;
  GradArgs& a = d.proto;
  a.nb_max = nb_max;
  a.bnd_cols = bnd_cols;
  a.lds_bytes = lds;
  a.ck_off = 0;
  a.bnd_off = int64_t(S_max) * nb_max * 6 * 64;
  a.abnd_off = a.bnd_off + int64_t(S_max + 1) * bnd_cols * 4;
  a.dummy_off = (a.abnd_off + 2 * int64_t(bnd_cols + 1) * 2 + 3) & ~int64_t(3);
  a.wave_stride = (a.dummy_off + 64 * 4 + 63) & ~int64_t(63);
  for (int k = 0; k < 2; ++k) {
    int nbk = 0;
    GRAD_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nbk, k ? grad_kernel<true> : grad_kernel<false>, kBlock, size_t(lds) * kWaves));
    d.occ[k] = std::max(nbk, 1);
  }
  const int64_t want_waves = std::max<int64_t>((NR + kWaves - 1) / kWaves * kWaves, kWaves);
  d.scratch_waves = std::min<int64_t>(int64_t(std::max(d.occ[0], d.occ[1])) * d.num_cus * kWaves,
                                      want_waves);
  if ((rc = d.alloc(&d.scratch, size_t(d.scratch_waves * a.wave_stride)))) return rc;
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
  out = std::move(b);
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
    GradArgs a = d.proto;
    a.log10_out = d.d_log10;
    if (int rc = d.launch(false, a, d.stream)) return rc;
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
  GradArgs a = d.proto;
  if (weights && d.n_pairs) {
    GRAD_TRY(hipMemcpyAsync(d.d_w, weights, sizeof(double) * size_t(d.n_pairs),
                            hipMemcpyHostToDevice, d.stream));
    a.weights = d.d_w;
  }
  a.log10_out = d.d_log10;
  GRAD_TRY(hipEventRecord(d.ev[0], d.stream));
  if (d.n_reads) {
    if (int rc = d.launch(true, a, d.stream)) return rc;
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
  return impl_->scratch_waves * impl_->proto.wave_stride * int64_t(sizeof(double));
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
