// Host side of the engine: the Phred tables, input validation and the per-batch plan (sorted
// haplotypes, common prefixes, duplicates, snapshot columns, PD column words, work buckets).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <set>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>

#include "fgkl_internal.hpp"

namespace fgkl {

namespace {
thread_local std::string g_last_error;
}

void set_error(int code, const std::string& msg) {
  (void)code;
  g_last_error = msg;
}
const char* last_error() { return g_last_error.c_str(); }

// ---- model.rs:36-106: GATK's tables, in f64 --------------------------------------------------
namespace {
constexpr double kJacobianMaxTolerance = 8.0;   // model.rs:23
constexpr double kJacobianTableStep = 0.0001;   // model.rs:24
const double kJacobianInvStep = 1.0 / kJacobianTableStep;                 // model.rs:25
const double kInvLn10 = 1.0 / 2.302585092994045684017991454684364208;     // model.rs:26

int64_t fast_round(double d) { return d > 0.0 ? int64_t(d + 0.5) : int64_t(d - 0.5); }

double jacobian(double diff) {
  const int64_t k = fast_round(diff * kJacobianInvStep);
  return std::log10(1.0 + std::pow(10.0, -double(k) * kJacobianTableStep));
}

double approximate_log10_sum_log10(double a, double b) {
  if (a > b) return approximate_log10_sum_log10(b, a);
  if (a == -HUGE_VAL) return b;
  const double diff = b - a;
  return diff < kJacobianMaxTolerance ? b + jacobian(diff) : b;
}

Tables* build_tables() {
  auto* t = new Tables;
  for (int q = 0; q <= kMaxQual; ++q) t->error_prob[q] = std::pow(10.0, double(q) / -10.0);
  int offset = 0;
  for (int i = 0; i <= kMaxQual; ++i) {
    for (int j = 0; j <= i; ++j) {
      const double log10_sum = approximate_log10_sum_log10(-0.1 * double(i), -0.1 * double(j));
      const double p = std::min(std::pow(10.0, log10_sum), 1.0);
      const double m2m_log10 = std::log1p(-p) * kInvLn10;
      t->match_to_match[offset + j] = std::pow(10.0, m2m_log10);
    }
    offset += i + 1;
  }
  return t;
}
}  // namespace

const Tables& tables() {
  static const Tables* t = build_tables();
  return *t;
}

// ---- validation (lib.rs:54-67, 272-289; pd.rs:75-92) and planning -------------------------
namespace {

// segments() of pd_kernel.rs:723-743 reduced to the state before each column and the state
// carried out of the row.
uint32_t column_states(const uint8_t* flags, int64_t m, uint32_t state, uint32_t* st) {
  for (int64_t j = 0; j < m; ++j) {
    const uint8_t f = flags[j];
    if (st) st[j] = state;
    if (state == kAfterDel) state = kNormal;
    if (f & kDelStart) state = kInsideDel;
    if (f & kDelEnd) state = kAfterDel;
  }
  return state;
}

// ---- plan.rs:55-230: suffix sharing ----------------------------------------------------------
// A distinct sorted haplotype whose longest prefix shared with another is p and longest suffix
// shared with another is s stops its forward sweep at column len - d - 1 with
// d = min(s, len - p - 1) (when d >= 4), and joins there the backward values of depth d, which
// backward sweeps compute once per shared suffix: in reversed-bases order, each resuming from
// the deepest depth computed along its suffix so far.
struct SuffixPlan {
  std::vector<int32_t> depth;  // per sorted haplotype (0: no cut)
  std::vector<int32_t> node;   // per sorted haplotype: the node its cut joins
  struct Sweep {
    int32_t hap, resume, need, start;
    std::map<int32_t, int32_t> node_of;  // depth -> node, for the depths it records
  };
  std::vector<Sweep> sweeps;
  int32_t nodes = 1;  // node 0: depth 0 (never stored)
};

SuffixPlan plan_suffixes(const std::vector<const uint8_t*>& bases, const std::vector<int32_t>& len,
                         const std::vector<int64_t>& lcp, const std::vector<char>& dup) {
  constexpr int32_t kMinSaving = 4;  // plan.rs:9
  const int32_t H = int32_t(bases.size());
  SuffixPlan plan;
  plan.depth.assign(size_t(H), 0);
  plan.node.assign(size_t(H), 0);
  std::vector<int32_t> distinct;
  for (int32_t k = 0; k < H; ++k)
    if (!dup[size_t(k)]) distinct.push_back(k);
  auto common_suffix = [&](int32_t a, int32_t b) {
    int32_t n = 0;
    while (n < len[size_t(a)] && n < len[size_t(b)] &&
           bases[size_t(a)][len[size_t(a)] - 1 - n] == bases[size_t(b)][len[size_t(b)] - 1 - n])
      ++n;
    return n;
  };
  std::vector<int32_t> order = distinct;  // by reversed bases, then sorted index
  std::sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
    const int32_t n = std::min(len[size_t(a)], len[size_t(b)]);
    for (int32_t i = 0; i < n; ++i) {
      const uint8_t x = bases[size_t(a)][len[size_t(a)] - 1 - i];
      const uint8_t y = bases[size_t(b)][len[size_t(b)] - 1 - i];
      if (x != y) return x < y;
    }
    return len[size_t(a)] != len[size_t(b)] ? len[size_t(a)] < len[size_t(b)] : a < b;
  });
  const int32_t D = int32_t(order.size());
  std::vector<int32_t> lcs(size_t(D), 0), pos(size_t(H), 0), sweep_at(size_t(D), -1);
  for (int32_t r = 0; r < D; ++r) {
    if (r) lcs[size_t(r)] = common_suffix(order[size_t(r - 1)], order[size_t(r)]);
    pos[size_t(order[size_t(r)])] = r;
  }
  for (size_t t = 0; t < distinct.size(); ++t) {
    const int32_t k = distinct[t];
    const int64_t next = t + 1 < distinct.size() ? lcp[size_t(distinct[t + 1])] : 0;
    const int32_t shared_prefix = int32_t(std::max<int64_t>(lcp[size_t(k)], next));
    const int32_t L = len[size_t(k)];
    if (shared_prefix >= L) continue;
    const int32_t r = pos[size_t(k)];
    const int32_t s = std::max(lcs[size_t(r)], r + 1 < D ? lcs[size_t(r + 1)] : 0);
    const int32_t d = std::min(s, L - shared_prefix - 1);
    if (d >= kMinSaving) plan.depth[size_t(k)] = d;
  }
  int32_t reach = 0;
  for (int32_t r = 0; r < D; ++r) {
    const int32_t k = order[size_t(r)];
    const int32_t resume = std::min(lcs[size_t(r)], reach);
    if (plan.depth[size_t(k)] > resume) {
      sweep_at[size_t(r)] = int32_t(plan.sweeps.size());
      plan.sweeps.push_back(SuffixPlan::Sweep{k, resume, plan.depth[size_t(k)], 0, {}});
    }
    reach = std::max(resume, plan.depth[size_t(k)]);
  }
  // The node holding depth d of the suffix at position r: the one the latest sweep at or before
  // r wrote for a suffix r shares at least that deep (plan.rs:205-230).
  auto node_for = [&](int32_t r, int32_t d) -> int32_t {
    if (d == 0) return 0;
    int32_t shared = INT32_MAX;
    for (int32_t q = r; q >= 0; --q) {
      const int32_t si = sweep_at[size_t(q)];
      if (si >= 0 && plan.sweeps[size_t(si)].resume < d && d <= plan.sweeps[size_t(si)].need) {
        auto ins = plan.sweeps[size_t(si)].node_of.emplace(d, plan.nodes);
        if (ins.second) ++plan.nodes;
        return ins.first->second;
      }
      shared = std::min(shared, lcs[size_t(q)]);
      if (shared < d) break;
    }
    return -1;  // unreachable for a consistent plan
  };
  for (auto& sw : plan.sweeps) sw.start = node_for(pos[size_t(sw.hap)], sw.resume);
  for (int32_t k : distinct)
    if (plan.depth[size_t(k)] > 0) plan.node[size_t(k)] = node_for(pos[size_t(k)], plan.depth[size_t(k)]);
  return plan;
}

// lanes_for(n, R) for every read length and rows per lane, without a division.
struct LaneTable {
  int16_t t[kMaxRowsPerLane + 1][kMaxReadLen + 1];
  LaneTable() {
    for (int R = 1; R <= kMaxRowsPerLane; ++R)
      for (int n = 0; n <= kMaxReadLen; ++n) t[R][n] = int16_t(lanes_for(n, R));
  }
  int64_t of(int64_t n, int R) const { return t[R][n]; }
};
const LaneTable lanes;

// The reads of a region not yet placed in a work item, by length: the longest free read of at
// most n bases (the latest such read when several have that length) in a few word operations.
class FreeReads {
 public:
  void reset(const int64_t* read_off, int64_t r0, int64_t r1) {
    std::fill(std::begin(count_), std::end(count_), 0);
    std::fill(std::begin(bits_), std::end(bits_), 0);
    for (int64_t r = r0; r < r1; ++r) ++count_[read_off[r + 1] - read_off[r]];
    int32_t at = 0;
    for (int n = 0; n <= kMaxReadLen; ++n) {
      start_[n] = at;
      at += count_[n];
      top_[n] = start_[n];
      if (count_[n]) bits_[n >> 6] |= uint64_t(1) << (n & 63);
    }
    reads_.resize(size_t(at));
    for (int64_t r = r0; r < r1; ++r) reads_[size_t(top_[read_off[r + 1] - read_off[r]]++)] = r;
    left_ = r1 - r0;
  }
  bool empty() const { return left_ == 0; }
  // The longest free length of at most n (0 if none).
  int64_t longest_at_most(int64_t n) const {
    if (n <= 0) return 0;
    if (n > kMaxReadLen) n = kMaxReadLen;
    int w = int(n >> 6);
    uint64_t m = bits_[w] & (~uint64_t(0) >> (63 - (n & 63)));
    while (!m) {
      if (--w < 0) return 0;
      m = bits_[w];
    }
    return int64_t(w) * 64 + 63 - __builtin_clzll(m);
  }
  // Removes and returns the latest free read of length n (one must be free).
  int64_t take(int64_t n) {
    const int64_t r = reads_[size_t(--top_[n])];
    if (top_[n] == start_[n]) bits_[n >> 6] &= ~(uint64_t(1) << (n & 63));
    --left_;
    return r;
  }

 private:
  int32_t count_[kMaxReadLen + 1], start_[kMaxReadLen + 1], top_[kMaxReadLen + 1];
  uint64_t bits_[(kMaxReadLen + 64) / 64];
  std::vector<int64_t> reads_;
  int64_t left_ = 0;
};

struct RegionHaps {
  const BatchInput* in;
  const std::vector<uint32_t>* end_state;
  int64_t len(int64_t h) const { return in->hap_off[h + 1] - in->hap_off[h]; }
  const uint8_t* b(int64_t h) const { return in->hap_bases + in->hap_off[h]; }
  const uint8_t* f(int64_t h) const { return in->hap_flags + in->hap_off[h]; }

  // Lexicographic order of kernel.rs:55 (plain) or of (row-end state, interleaved bases and
  // flags) as pd_kernel.rs:67-76 sorts.
  bool less(int64_t a, int64_t c) const {
    if (in->pd) {
      if ((*end_state)[a] != (*end_state)[c]) return (*end_state)[a] < (*end_state)[c];
      const int64_t la = len(a), lc = len(c), n = std::min(la, lc);
      const uint8_t *ba = b(a), *bc = b(c), *fa = f(a), *fc = f(c);
      for (int64_t j = 0; j < n; ++j) {
        if (ba[j] != bc[j]) return ba[j] < bc[j];
        if (fa[j] != fc[j]) return fa[j] < fc[j];
      }
      return la < lc;
    }
    const int64_t la = len(a), lc = len(c);
    const int r = std::memcmp(b(a), b(c), size_t(std::min(la, lc)));
    return r != 0 ? r < 0 : la < lc;
  }

  // Common prefix of two haplotypes (kernel.rs:68-70; zero across PD end states,
  // pd_kernel.rs:115-122).
  int64_t lcp(int64_t a, int64_t c) const {
    if (in->pd && (*end_state)[a] != (*end_state)[c]) return 0;
    const int64_t n = std::min(len(a), len(c));
    const uint8_t *ba = b(a), *bc = b(c);
    int64_t j = 0;
    if (in->pd) {
      const uint8_t *fa = f(a), *fc = f(c);
      while (j < n && ba[j] == bc[j] && fa[j] == fc[j]) ++j;
    } else {
      while (j < n && ba[j] == bc[j]) ++j;
    }
    return j;
  }
};

bool check_offsets(const int64_t* off, int64_t count, const char* what) {
  if (off[0] != 0) {
    set_error(kInvalidArgument, std::string(what) + " offset table must start at zero");
    return false;
  }
  for (int64_t i = 0; i < count; ++i)
    if (off[i + 1] < off[i]) {
      set_error(kInvalidArgument, std::string(what) + " offset table must be non-decreasing");
      return false;
    }
  return true;
}

}  // namespace

int build_plan(const BatchInput& in, HostPlan& p) {
  const int64_t G = in.n_regions;
  if (G < 0) {
    set_error(kInvalidArgument, "negative region count");
    return kInvalidArgument;
  }
  if (!check_offsets(in.region_read, G, "region read") ||
      !check_offsets(in.region_hap, G, "region haplotype"))
    return kInvalidArgument;
  const int64_t NR = in.region_read[G], NH = in.region_hap[G];
  if (!check_offsets(in.read_off, NR, "read") || !check_offsets(in.hap_off, NH, "haplotype"))
    return kInvalidArgument;

  // lib.rs:276-283: every read validated before any haplotype; indices are region-local as the
  // caller's per-region call would report them.
  for (int64_t g = 0; g < G; ++g) {
    for (int64_t r = in.region_read[g]; r < in.region_read[g + 1]; ++r) {
      const int64_t n = in.read_off[r + 1] - in.read_off[r];
      if (n == 0) {
        set_error(kInvalidArgument, "read " + std::to_string(r - in.region_read[g]) +
                                        " has no bases");
        return kInvalidArgument;
      }
      if (n > (int64_t(1) << 30)) {
        set_error(kInvalidArgument, "read too long");
        return kInvalidArgument;
      }
    }
    for (int64_t h = in.region_hap[g]; h < in.region_hap[g + 1]; ++h) {
      if (in.hap_off[h + 1] == in.hap_off[h]) {
        set_error(kInvalidArgument, "haplotype " + std::to_string(h - in.region_hap[g]) +
                                        " has no bases");
        return kInvalidArgument;
      }
      if (in.hap_off[h + 1] - in.hap_off[h] > (int64_t(1) << 30)) {
        set_error(kInvalidArgument, "haplotype too long");
        return kInvalidArgument;
      }
    }
  }

  p = HostPlan{};
  p.n_reads = NR;
  p.n_haps = NH;
  p.read_region.resize(size_t(NR));
  p.read_local.resize(size_t(NR));
  p.read_rows.resize(size_t(NR));
  p.region_out.resize(size_t(G) + 1);
  p.region_hap0.resize(size_t(G));
  p.region_nhap.resize(size_t(G));
  p.plan.resize(size_t(NH));

  std::vector<uint32_t> end_state(in.pd ? size_t(NH) : 0);
  if (in.pd) {
    p.words.resize(size_t(in.hap_off[NH]));
    p.region_classes.assign(size_t(G) * kPdExtraClasses, 0);
    p.region_pd_long.assign(size_t(G), 0);
    const int nt = host_threads(G, 64);
    run_parallel(nt, [&](int t) {
      std::vector<uint32_t> st1, st2;
      for (int64_t g = G * t / nt; g < G * (t + 1) / nt; ++g) {
        // The region's read-base classes beyond A, C, G, T (none when the read bytes are not
        // given: a host-only plan needs no words).
        uint8_t* extra = &p.region_classes[size_t(g) * kPdExtraClasses];
        int n_extra = 0;
        bool is_long = false;
        bool seen[256] = {};
        seen[uint8_t('A')] = seen[uint8_t('C')] = seen[uint8_t('G')] = seen[uint8_t('T')] =
            seen[uint8_t('N')] = true;
        for (int64_t r = in.region_read[g]; r < in.region_read[g + 1]; ++r) {
          is_long |= in.read_off[r + 1] - in.read_off[r] > kMaxReadLen;
          if (!in.bases) continue;
          for (int64_t i = in.read_off[r]; i < in.read_off[r + 1]; ++i) {
            const uint8_t b = in.bases[i];
            if (seen[b]) continue;
            seen[b] = true;
            if (n_extra < kPdExtraClasses)
              extra[n_extra++] = b;
            else
              is_long = true;
          }
        }
        for (int64_t h = in.region_hap[g]; h < in.region_hap[g + 1]; ++h)
          is_long |= in.hap_off[h + 1] - in.hap_off[h] > kMaxStagedHap;
        p.region_pd_long[size_t(g)] = char(is_long);
        for (int64_t h = in.region_hap[g]; h < in.region_hap[g + 1]; ++h) {
          const int64_t o = in.hap_off[h], m = in.hap_off[h + 1] - o;
          st1.resize(size_t(m));
          st2.resize(size_t(m));
          const uint32_t e = column_states(in.hap_flags + o, m, kNormal, st1.data());
          column_states(in.hap_flags + o, m, e, st2.data());
          end_state[size_t(h)] = e;
          for (int64_t j = 0; j < m; ++j) {
            const uint8_t b = in.hap_bases[o + j], f = in.hap_flags[o + j];
            p.words[size_t(o + j)] =
                is_long ? pd_word(b, f, st1[size_t(j)], st2[size_t(j)])
                        : pd_word_classes(pd_class_mask(b, f, extra, n_extra), f, st1[size_t(j)],
                                          st2[size_t(j)]);
          }
        }
      }
    });
  }
  RegionHaps rh{&in, &end_state};

  // Output bases, region by region.
  int64_t out_base = 0;
  for (int64_t g = 0; g < G; ++g) {
    p.region_out[size_t(g)] = out_base;
    out_base += (in.region_read[g + 1] - in.region_read[g]) * (in.region_hap[g + 1] - in.region_hap[g]);
  }
  p.region_out[size_t(G)] = out_base;
  p.n_pairs = out_base;
  if (in.pairs_only && p.n_pairs > kMaxPairList) {
    set_error(kInvalidArgument, "a pairs-only batch holds more than " +
                                    std::to_string(kMaxPairList) + " pairs");
    return kInvalidArgument;
  }

  // Regions plan independently: each range of regions builds its part (snapshot lists, backward
  // sweeps and work items with indices local to the part) and writes its regions' and
  // haplotypes' own entries in place; the parts are then joined in region order, so the plan
  // does not depend on how the regions were split.
  struct Part {
    std::vector<SnapEntry> snaps;
    std::vector<BwdSweep> bsweeps;
    std::vector<int32_t> bwrites, region_bw;
    int32_t max_slots = 0, max_nodes = 0, max_hap_len = 0;
    std::vector<int32_t> long_pairs;
    std::vector<int64_t> region_long_pairs;  // pairs of each region in long_pairs
    std::vector<char> region_is_long;
    int32_t long_max_hap = 0;
    int64_t long_cells = 0;
    int64_t work_nominal = 0, work_prefix = 0, work_forward = 0, work_backward = 0;
    int64_t snap_hbm = 0, snap_lds = 0, snap_restarts = 0;
    std::vector<int32_t> bucket_items[kMaxRowsPerLane];
    int64_t bucket_pairs[kMaxRowsPerLane] = {}, bucket_cells[kMaxRowsPerLane] = {},
            bucket_work[kMaxRowsPerLane] = {};
    std::vector<int64_t> item_costs[kMaxRowsPerLane];
  };

  // Snapshot records of one region's sorted haplotypes hp[0..H) (kernel.rs:236-259, 556-562).
  // A haplotype after the first starts at hp[k].lcp, the column it shares with the previous
  // sorted one (0: from column 0), from the state the last earlier sweep to pass that column left
  // there, rescaled by the two lengths (kernel.rs:641).  Only those states are recorded: sweep j
  // records column c when it is the last sweep to pass c before a haplotype starts at c.  A sweep
  // passes c when it starts before c and computes c, up to the column before its cut.  A sweep
  // that starts before c and stops short of it ends the prefix earlier records of c hold, so a
  // haplotype that would start at c with no record left starts from column 0 (sorted order never
  // leads there: the cut lies past every column shared with a neighbour).
  // A wave holds one record in LDS.  Records take it in order of their last reader, earliest
  // first (the schedule that fits the most intervals in one slot); the others are written to the
  // HBM slot of their column.  A record lives from the sweep that writes it to the start of its
  // last reader, and a sweep restores its start before it records anything.
  auto plan_snapshots = [](HapPlan* hp, int64_t H, const std::map<int64_t, int32_t>& slot_of,
                           int64_t reads, Part& q) {
    struct Rec {
      int32_t writer, col, last;
      bool lds;
    };
    std::vector<Rec> recs;
    std::map<std::pair<int32_t, int32_t>, int32_t> rec_of;  // (writer, column) -> record
    std::vector<int32_t> reads_rec(size_t(H), -1);
    std::map<int64_t, int32_t> pending;    // column -> later haplotypes starting there
    std::map<int64_t, int32_t> last_pass;  // column -> last sweep to pass it, or -1
    for (int64_t k = 1; k < H; ++k)
      if (!hp[k].dup && hp[k].lcp > 0) pending[hp[k].lcp] += 1;
    for (int64_t k = 0; k < H; ++k) {
      if (hp[k].dup) continue;
      if (hp[k].lcp > 0) {
        pending[hp[k].lcp] -= 1;
        const auto it = last_pass.find(hp[k].lcp);
        const int32_t w = it == last_pass.end() ? -1 : it->second;
        if (w < 0) {
          hp[k].lcp = 0;
        } else {
          const auto key = std::make_pair(w, hp[k].lcp);
          auto r = rec_of.find(key);
          if (r == rec_of.end()) {
            r = rec_of.emplace(key, int32_t(recs.size())).first;
            recs.push_back(Rec{w, hp[k].lcp, int32_t(k), false});
          }
          recs[size_t(r->second)].last = int32_t(k);
          reads_rec[size_t(k)] = r->second;
        }
      }
      const int64_t end = hp[k].cut > 0 ? hp[k].cut - 1 : hp[k].len;
      for (auto it = pending.upper_bound(hp[k].lcp); it != pending.end(); ++it)
        if (it->second > 0) last_pass[it->first] = it->first <= end ? int32_t(k) : -1;
    }
    std::vector<int32_t> by_last(recs.size());
    std::iota(by_last.begin(), by_last.end(), 0);
    std::sort(by_last.begin(), by_last.end(), [&](int32_t x, int32_t y) {
      const Rec &a = recs[size_t(x)], &b = recs[size_t(y)];
      return a.last != b.last ? a.last < b.last : a.writer < b.writer;
    });
    int32_t lds_free_from = -1;
    for (int32_t i : by_last) {
      Rec& r = recs[size_t(i)];
      if (r.writer >= lds_free_from) {
        r.lds = true;
        lds_free_from = r.last;
      }
    }
    std::vector<int32_t> by_writer(recs.size());
    std::iota(by_writer.begin(), by_writer.end(), 0);
    std::sort(by_writer.begin(), by_writer.end(), [&](int32_t x, int32_t y) {
      const Rec &a = recs[size_t(x)], &b = recs[size_t(y)];
      return a.writer != b.writer ? a.writer < b.writer : a.col < b.col;
    });
    size_t next = 0;
    int64_t lds_recs = 0, restarts = 0;
    for (int64_t k = 0; k < H; ++k) {
      hp[k].snap0 = int32_t(q.snaps.size());
      hp[k].nsnap = 0;
      for (; next < by_writer.size() && recs[size_t(by_writer[next])].writer == k; ++next) {
        const Rec& r = recs[size_t(by_writer[next])];
        q.snaps.push_back(SnapEntry{r.col, slot_of.at(r.col) | (r.lds ? kSnapLds : 0)});
        hp[k].nsnap += 1;
        lds_recs += r.lds;
      }
      const int32_t ri = reads_rec[size_t(k)];
      if (ri >= 0) {
        const Rec& r = recs[size_t(ri)];
        hp[k].start_slot = slot_of.at(r.col) | (r.lds ? kSnapLds : 0);
        hp[k].start_src_len = hp[r.writer].len;
        restarts += 1;
      }
    }
    q.snap_hbm += reads * (int64_t(recs.size()) - lds_recs);
    q.snap_lds += reads * lds_recs;
    q.snap_restarts += reads * restarts;
  };
  auto plan_range = [&](int64_t g0, int64_t g1, Part& q) {
    std::vector<int64_t> order;
      std::vector<int64_t> lcp;
      std::map<int64_t, int32_t> counts;
      for (int64_t g = g0; g < g1; ++g) {
      const int64_t r0 = in.region_read[g], r1 = in.region_read[g + 1];
      const int64_t h0 = in.region_hap[g], h1 = in.region_hap[g + 1], H = h1 - h0;
      q.region_bw.push_back(int32_t(q.bsweeps.size()));
      p.region_hap0[size_t(g)] = int32_t(h0);
      p.region_nhap[size_t(g)] = int32_t(H);
      // A region the staged sweep cannot hold runs every pair in long_kernel, unshared.
      bool is_long = false;
      for (int64_t r = r0; r < r1; ++r)
        is_long |= in.read_off[r + 1] - in.read_off[r] > kMaxReadLen;
      for (int64_t h = h0; h < h1; ++h) is_long |= rh.len(h) > kMaxStagedHap;
      if (in.pd) is_long |= p.region_pd_long[size_t(g)] != 0;  // also too many read classes
      for (int64_t r = r0; r < r1; ++r) {
        p.read_region[size_t(r)] = int32_t(g);
        p.read_local[size_t(r)] = int32_t(r - r0);
        p.read_rows[size_t(r)] =
            is_long ? 0 : rows_per_lane(in.read_off[r + 1] - in.read_off[r]);
      }
      q.region_long_pairs.push_back(is_long ? (r1 - r0) * H : 0);
      q.region_is_long.push_back(char(is_long));
      if (H == 0) continue;

      // Stable sort, as Rust's sort_by (kernel.rs:55, pd_kernel.rs:72-76).
      order.resize(size_t(H));
      std::iota(order.begin(), order.end(), h0);
      std::stable_sort(order.begin(), order.end(),
                       [&](int64_t a, int64_t c) { return rh.less(a, c); });
      lcp.assign(size_t(H), 0);
      for (int64_t k = 1; k < H; ++k) lcp[size_t(k)] = rh.lcp(order[size_t(k - 1)], order[size_t(k)]);

      // lcp_counts (kernel.rs:556-562) and one scratch slot per distinct start column.
      counts.clear();
      std::vector<char> dup(size_t(H), 0);
      for (int64_t k = 1; k < H; ++k) {
        const int64_t a = order[size_t(k - 1)], c = order[size_t(k)];
        dup[size_t(k)] = rh.len(a) == rh.len(c) && lcp[size_t(k)] == rh.len(c);
        if (!dup[size_t(k)]) counts[lcp[size_t(k)]] += 1;
      }
      std::map<int64_t, int32_t> slot_of;
      for (auto& kv : counts)
        if (kv.first > 0) slot_of.emplace(kv.first, int32_t(slot_of.size()));
      // Beyond kMaxRegionSlots snapshot columns the region shares no prefixes: every haplotype
      // sweeps from column 0 (duplicates still copy).  A long region shares nothing.
      const bool share_prefix = !is_long && slot_of.size() <= size_t(kMaxRegionSlots);
      if (!share_prefix) slot_of.clear();
      q.max_slots = std::max<int32_t>(q.max_slots, int32_t(slot_of.size()));

      for (int64_t k = 0; k < H; ++k) {
        const int64_t h = order[size_t(k)];
        HapPlan& hp = p.plan[size_t(h0 + k)];
        hp.off = in.hap_off[h];
        hp.len = int32_t(rh.len(h));
        hp.lcp = int32_t(lcp[size_t(k)]);
        hp.src = int32_t(h - h0);
        hp.dup = dup[size_t(k)];
        hp.start_slot = -1;
        hp.start_src_len = 0;
        hp.snap0 = int32_t(q.snaps.size());
        hp.nsnap = 0;
        hp.cut = 0;
        hp.cut_node = 0;
        if (is_long) {
          hp.lcp = 0;
          hp.dup = 0;
          q.long_max_hap = std::max(q.long_max_hap, hp.len);
          continue;
        }
        q.max_hap_len = std::max(q.max_hap_len, hp.len);
        if (hp.dup) continue;
        hp.lcp = int32_t(k == 0 || !share_prefix ? 0 : lcp[size_t(k)]);
      }
      if (is_long) {
        int64_t read_bases = 0, hap_bases = 0;
        for (int64_t r = r0; r < r1; ++r) {
          read_bases += in.read_off[r + 1] - in.read_off[r];
          for (int64_t k = 0; k < H; ++k) {
            q.long_pairs.push_back(int32_t(r));
            q.long_pairs.push_back(int32_t(k));
          }
        }
        for (int64_t h = h0; h < h1; ++h) hap_bases += rh.len(h);
        q.long_cells += read_bases * hap_bases;
        q.work_nominal += read_bases * hap_bases;
        q.work_prefix += read_bases * hap_bases;
        q.work_forward += read_bases * hap_bases;
        continue;
      }
      if (!in.pd && H > 1 && !in.pairs_only) {
        // Suffix sharing for plain haplotypes (the reference shares suffixes in its plain kernel
        // only, kernel.rs:53-81); the kernel applies it in the FP32 pass with prefix sharing.
        std::vector<const uint8_t*> hb(static_cast<size_t>(H));
        std::vector<int32_t> hl(static_cast<size_t>(H));
        std::vector<char> dp(static_cast<size_t>(H));
        for (int64_t k = 0; k < H; ++k) {
          const HapPlan& hp = p.plan[size_t(h0 + k)];
          hb[size_t(k)] = in.hap_bases + hp.off;
          hl[size_t(k)] = hp.len;
          dp[size_t(k)] = char(hp.dup);
        }
        const SuffixPlan sp = plan_suffixes(hb, hl, lcp, dp);
        // Beyond kMaxRegionSlots node records the region shares no suffixes.
        if (sp.nodes - 1 <= kMaxRegionSlots) {
        for (int64_t k = 0; k < H; ++k) {
          const int32_t d = sp.depth[size_t(k)];
          if (d > 0) {
            HapPlan& hp = p.plan[size_t(h0 + k)];
            hp.cut = hp.len - d;
            hp.cut_node = sp.node[size_t(k)] - 1;  // node records exclude node 0
          }
        }
        for (const auto& sw : sp.sweeps) {
          const int32_t L = hl[size_t(sw.hap)];
          BwdSweep b;
          b.hap = int32_t(h0 + sw.hap);
          b.right = sw.resume > 0 ? L - sw.resume : L + 1;
          b.left = L - sw.need;
          b.start_node = sw.resume > 0 ? sw.start - 1 : -1;
          b.w0 = int32_t(q.bwrites.size() / 2);
          b.nw = int32_t(sw.node_of.size());
          for (const auto& dn : sw.node_of) {  // ascending depth: descending column
            q.bwrites.push_back(L - dn.first);
            q.bwrites.push_back(dn.second - 1);
          }
          q.bsweeps.push_back(b);
        }
        q.max_nodes = std::max(q.max_nodes, sp.nodes - 1);
        }
      }
      if (share_prefix && H > 1) plan_snapshots(&p.plan[size_t(h0)], H, slot_of, r1 - r0, q);
    }
    q.region_bw.push_back(int32_t(q.bsweeps.size()));
    if (in.pairs_only) return;

    // Work items.  Two reads of a region can share a wave: read A's rows fill lanes [0, L_A),
    // read B's the lanes after, both swept against the same haplotype columns, each on its own
    // anti-diagonal.  A step costs about kStepFixed + kStepPerRow * R VALU instructions
    // (wave-wide moves, the column word, then R cells), and a region's sweeps take
    // sum(W) + H' (L - 1) steps with L = max(L_A, L_B)
    // (H' sweeps, forward over the columns between the shared prefix and the cut, and backward
    // over shared suffixes).  Reads are paired greedily, longest first, each with the free read
    // that lowers the modelled cost most; the candidates are, for every R, the longest free read
    // that fits beside it.
    constexpr int64_t kStepFixed = 20, kStepPerRow = 19;
    FreeReads free_reads;
    for (int64_t g = g0; g < g1; ++g) {
      int64_t cols = 0, swept = 0, hap_bases = 0;
      const int32_t H = p.region_nhap[size_t(g)];
      if (q.region_is_long[size_t(g - g0)]) continue;  // every pair computed in long_kernel
      for (int32_t k = 0; k < H; ++k) {
        const HapPlan& hp = p.plan[size_t(p.region_hap0[size_t(g)] + k)];
        hap_bases += hp.len;
        if (!hp.dup) {
          cols += (hp.cut ? hp.cut - 1 : hp.len) - hp.lcp;  // the FP32 pass stops at its cut
          swept += 1;
        }
      }
      int64_t prefix_cols = 0, fwd_cols = cols, bwd_cols = 0;
      for (int32_t k = 0; k < H; ++k) {
        const HapPlan& hp = p.plan[size_t(p.region_hap0[size_t(g)] + k)];
        if (!hp.dup) prefix_cols += hp.len - hp.lcp;
      }
      for (int32_t w = q.region_bw[size_t(g - g0)]; w < q.region_bw[size_t(g - g0) + 1]; ++w) {
        cols += q.bsweeps[size_t(w)].right - q.bsweeps[size_t(w)].left;
        bwd_cols += q.bsweeps[size_t(w)].right - q.bsweeps[size_t(w)].left;
        swept += 1;
      }
      {
        int64_t read_bases = 0;
        for (int64_t r = in.region_read[g]; r < in.region_read[g + 1]; ++r)
          read_bases += in.read_off[r + 1] - in.read_off[r];
        q.work_nominal += read_bases * hap_bases;
        q.work_prefix += read_bases * prefix_cols;
        q.work_forward += read_bases * fwd_cols;
        q.work_backward += read_bases * bwd_cols;
      }
      auto item_cost = [&](int R, int64_t L) {
        return (cols + swept * (L - 1)) * (kStepFixed + kStepPerRow * R);
      };
      auto single_cost = [&](int64_t n) {
        const int R = rows_per_lane(n);
        return item_cost(R, lanes.of(n, R));
      };
      // Two reads sweep side by side, each on its own anti-diagonal: the ramp is the larger
      // read's lanes.
      auto pair_cost = [&](int64_t na, int64_t nb) {
        for (int R = 1; R <= kMaxRowsPerLane; ++R) {  // rows_per_lane2 through the table
          const int64_t La = lanes.of(na, R), Lb = lanes.of(nb, R);
          if (La + Lb <= 64) return item_cost(R, std::max(La, Lb));
        }
        return INT64_MAX;
      };
      const int64_t r0 = in.region_read[g], r1 = in.region_read[g + 1];
      auto len = [&](int64_t r) { return in.read_off[r + 1] - in.read_off[r]; };
      free_reads.reset(in.read_off, r0, r1);
      while (!free_reads.empty()) {
        const int64_t na = free_reads.longest_at_most(kMaxReadLen);
        const int64_t a = free_reads.take(na);
        const int64_t alone = single_cost(na);
        int64_t best_gain = 0, nb = 0, last = 0;
        for (int R = rows_per_lane(na); R <= kMaxRowsPerLane; ++R) {
          const int64_t spare = 64 - lanes.of(na, R);
          if (spare <= 0) continue;
          const int64_t max_n = spare * R - 1;  // its rows plus one identity row
          const int64_t n = free_reads.longest_at_most(max_n);
          // The gain depends on the partner's length only, so a repeated candidate changes nothing.
          if (n == 0 || n == last) continue;
          last = n;
          const int64_t gain = alone + single_cost(n) - pair_cost(na, n);
          if (gain > best_gain) {
            best_gain = gain;
            nb = n;
          }
        }
        const int64_t b = nb > 0 ? free_reads.take(nb) : -1;
        const int R = b >= 0 ? rows_per_lane2(na, nb) : rows_per_lane(na);
        const int64_t L = std::max(lanes_for(na, R), b >= 0 ? lanes_for(nb, R) : int64_t(0));
        auto& bucket = q.bucket_items[R - 1];
        bucket.push_back(int32_t(a));
        bucket.push_back(int32_t(b));
        q.bucket_pairs[R - 1] += int64_t(H) * (b >= 0 ? 2 : 1);
        q.bucket_cells[R - 1] += (na + nb) * hap_bases;
        q.bucket_work[R - 1] += (na + nb) * cols;
        q.item_costs[R - 1].push_back(item_cost(R, L));
      }
    }
  };
  const int threads = host_threads(G, 256);
  std::vector<Part> parts(static_cast<size_t>(threads));
  run_parallel(threads, [&](int t) {
    plan_range(G * t / threads, G * (t + 1) / threads, parts[size_t(t)]);
  });
  std::vector<int64_t> item_costs[kMaxRowsPerLane];
  p.region_long.push_back(0);  // region g's long pairs: [region_long[g], region_long[g + 1])
  for (int t = 0; t < threads; ++t) {
    Part& q = parts[size_t(t)];
    const int64_t g0 = G * t / threads, g1 = G * (t + 1) / threads;
    const int32_t snap_base = int32_t(p.snaps.size()), sweep_base = int32_t(p.bsweeps.size());
    const int32_t write_base = int32_t(p.bwrites.size() / 2);
    for (int64_t h = in.region_hap[g0]; h < in.region_hap[g1]; ++h) p.plan[size_t(h)].snap0 += snap_base;
    for (BwdSweep& b : q.bsweeps) b.w0 += write_base;
    for (size_t i = 0; i + 1 < q.region_bw.size(); ++i) p.region_bw.push_back(q.region_bw[i] + sweep_base);
    p.snaps.insert(p.snaps.end(), q.snaps.begin(), q.snaps.end());
    p.bsweeps.insert(p.bsweeps.end(), q.bsweeps.begin(), q.bsweeps.end());
    p.bwrites.insert(p.bwrites.end(), q.bwrites.begin(), q.bwrites.end());
    p.max_slots = std::max(p.max_slots, q.max_slots);
    p.max_nodes = std::max(p.max_nodes, q.max_nodes);
    p.max_hap_len = std::max(p.max_hap_len, q.max_hap_len);
    for (int64_t pairs : q.region_long_pairs)
      p.region_long.push_back(p.region_long.back() + pairs);
    p.long_pairs.insert(p.long_pairs.end(), q.long_pairs.begin(), q.long_pairs.end());
    p.long_max_hap = std::max(p.long_max_hap, q.long_max_hap);
    p.long_cells += q.long_cells;
    p.work_nominal += q.work_nominal;
    p.work_prefix += q.work_prefix;
    p.work_forward += q.work_forward;
    p.work_backward += q.work_backward;
    p.snap_hbm += q.snap_hbm;
    p.snap_lds += q.snap_lds;
    p.snap_restarts += q.snap_restarts;
    for (int R = 0; R < kMaxRowsPerLane; ++R) {
      p.bucket_items[R].insert(p.bucket_items[R].end(), q.bucket_items[R].begin(), q.bucket_items[R].end());
      item_costs[R].insert(item_costs[R].end(), q.item_costs[R].begin(), q.item_costs[R].end());
      p.bucket_pairs[R] += q.bucket_pairs[R];
      p.bucket_cells[R] += q.bucket_cells[R];
      p.bucket_work[R] += q.bucket_work[R];
    }
  }
  p.region_bw.push_back(int32_t(p.bsweeps.size()));
  if (in.pairs_only) return kOk;

  // Costliest items first, so the persistent grid's tail is short (a bucket per thread).
  run_parallel(threads > 1 ? kMaxRowsPerLane : 1, [&](int t) {
   for (int R = threads > 1 ? t : 0; R < (threads > 1 ? t + 1 : kMaxRowsPerLane); ++R) {
    auto& items = p.bucket_items[R];
    const auto& c = item_costs[R];
    std::vector<int32_t> idx(c.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int32_t x, int32_t y) { return c[size_t(x)] > c[size_t(y)]; });
    std::vector<int32_t> sorted(items.size());
    for (size_t i = 0; i < idx.size(); ++i) {
      sorted[2 * i] = items[2 * size_t(idx[i])];
      sorted[2 * i + 1] = items[2 * size_t(idx[i]) + 1];
    }
    items.swap(sorted);
   }
  });
  return kOk;
}

}  // namespace fgkl
