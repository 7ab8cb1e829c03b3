// Internal declarations shared by the host planner, the HIP engine and the C-ABI.
#pragma once

#include <algorithm>
#include <cstdint>
#include <thread>
#include <string>
#include <vector>

namespace fgkl {

// Status codes of the C-ABI (include/fgkl.h).
enum Status : int { kOk = 0, kInvalidArgument = 1, kInternal = 2 };

// Thread-local error message for the last failing call on this thread.
void set_error(int code, const std::string& msg);
const char* last_error();

// ---------------------------------------------------------------------------------------------
// Phred tables, built in f64 exactly as model.rs:36-106 builds them.
constexpr int kMaxQual = 254;                                      // model.rs:11
constexpr int kMatchToMatchSize = (kMaxQual + 1) * (kMaxQual + 2) / 2;  // model.rs:39

struct Tables {
  double error_prob[kMaxQual + 1];
  double match_to_match[kMatchToMatchSize];
};
const Tables& tables();

inline int clamp_qual(uint8_t q) { return q > kMaxQual ? kMaxQual : q; }
inline int tri_index(uint8_t a8, uint8_t b8) {  // model.rs:59-64
  int a = clamp_qual(a8), b = clamp_qual(b8);
  int lo = a <= b ? a : b, hi = a <= b ? b : a;
  return ((hi * (hi + 1)) >> 1) + lo;
}

// PD flag bits (pd.rs:15-27) and deletion states (pd_kernel.rs:35-37).
constexpr uint8_t kSnp = 1, kDelStart = 2, kDelEnd = 4;
constexpr uint8_t kAltA = 8, kAltC = 16, kAltG = 32, kAltT = 64;
constexpr uint32_t kNormal = 0, kInsideDel = 1, kAfterDel = 2;

// The per-column word the PD kernel passes between lanes: haplotype byte (bits 0-7), accepted
// alternates (8-11), the deletion state of the column in row 1 (12: inside, 13: after) and in
// rows >= 2 (14: inside, 15: after), its DEL_END flag (16), and bit 18 always set (bit 19 always
// clear), which identity rows read as "inside a deletion, not after one".  A row finds its two
// state bits at its own shift: kRow1Shift, kRowShift or kIdentShift.
constexpr uint32_t kRow1Shift = 12, kRowShift = 14, kIdentShift = 18, kDelEndBit = 16;
inline uint32_t pd_word(uint8_t base, uint8_t flag, uint32_t st1, uint32_t st2) {
  uint32_t alts = (flag & kSnp) ? ((flag >> 3) & 0xFu) : 0u;  // MatchSet::of (pd_kernel.rs:196-214)
  return uint32_t(base) | (alts << 8) | (st1 << kRow1Shift) | (st2 << kRowShift) |
         (uint32_t((flag & kDelEnd) != 0) << kDelEndBit) | (1u << kIdentShift);
}

// A column of a PD region the staged sweep runs carries, in bits 0-11 of its word instead of the
// haplotype byte and alternates, the read-base classes it accepts as a match (MatchSet::accepts,
// pd_kernel.rs:196-214): bit c for class c, where A, C, G, T are classes 0-3 and each other byte
// the region's reads hold (lowercase bases and any other byte) takes one of kPdExtraClasses
// classes from 4 on, in order of first appearance.  A read 'N' takes class 0: its match and
// mismatch priors are equal.  A region whose reads hold more other bytes runs in long_kernel,
// whose words keep the haplotype byte and alternates and whose cells compare bytes.
constexpr int kPdExtraClasses = 8;
inline uint32_t pd_alt_bit(uint8_t b) {  // pd.rs:99-110
  switch (b) {
    case 'A': case 'a': return 1u;
    case 'C': case 'c': return 2u;
    case 'G': case 'g': return 4u;
    case 'T': case 't': return 8u;
    default: return 0u;
  }
}
inline bool pd_accepts(uint8_t read_base, uint8_t hap_base, uint8_t flag) {
  const uint32_t alts = (flag & kSnp) ? ((flag >> 3) & 0xFu) : 0u;
  return read_base == hap_base || read_base == 'N' || hap_base == 'N' ||
         (alts & pd_alt_bit(read_base)) != 0u;
}
inline uint32_t pd_class_mask(uint8_t hap_base, uint8_t flag, const uint8_t* extra, int n_extra) {
  static constexpr uint8_t kStd[4] = {'A', 'C', 'G', 'T'};
  uint32_t m = 0;
  for (int c = 0; c < 4; ++c) m |= uint32_t(pd_accepts(kStd[c], hap_base, flag)) << c;
  for (int c = 0; c < n_extra; ++c) m |= uint32_t(pd_accepts(extra[c], hap_base, flag)) << (4 + c);
  return m;
}
inline uint32_t pd_word_classes(uint32_t classes, uint8_t flag, uint32_t st1, uint32_t st2) {
  return classes | (st1 << kRow1Shift) | (st2 << kRowShift) |
         (uint32_t((flag & kDelEnd) != 0) << kDelEndBit) | (1u << kIdentShift);
}

// ---------------------------------------------------------------------------------------------
// Plan of one batch: haplotypes sorted per region for prefix sharing (kernel.rs:53-81,
// pd_kernel.rs:61-136) and the snapshot columns each sorted haplotype must record
// (kernel.rs:236-259, 556-562), all laid out for the device.

struct HapPlan {        // one per haplotype, region-contiguous, in sorted order
  int64_t off;          // offset of the haplotype in the batch's haplotype byte / word arrays
  int32_t len;          // m
  int32_t lcp;          // common prefix with the previous sorted haplotype (0 for the first)
  int32_t src;          // caller's index of this haplotype within its region
  int32_t dup;          // equal to the previous sorted haplotype: copies its results
  int32_t start_slot;   // snapshot slot holding column `lcp` (kSnapLds: maybe in LDS), or -1
  int32_t start_src_len;  // length of the haplotype that wrote that snapshot
  int32_t snap0;        // first entry in the snapshot list
  int32_t nsnap;        // columns this haplotype must snapshot
  int32_t cut;          // suffix sharing: forward sweep stops at column cut - 1 (0: no cut)
  int32_t cut_node;     // backward node the cut joins (record index)
};
static_assert(sizeof(HapPlan) == 48, "HapPlan layout");

// One backward sweep of a region's suffix plan (plan.rs:14-31): over columns right-1 down to
// left of sorted haplotype `hap` (a global plan index), starting from node `start_node`'s record
// at column `right` or, when -1, from the zero state beyond the last column (then right = len+1
// and the sweep also computes column len, which only trailing insertions leave).  Records node
// bwrites[w0 + i] = (column, node) as it passes those columns, in descending column order.
struct BwdSweep {
  int32_t hap;
  int32_t right;
  int32_t left;
  int32_t start_node;
  int32_t w0;
  int32_t nw;
};

struct SnapEntry {
  int32_t col;   // 1-based haplotype column
  int32_t slot;  // scratch slot, with kSnapLds set when the wave's LDS record may hold it
};

// Set in a snapshot slot (SnapEntry::slot, HapPlan::start_slot) whose record the wave keeps in
// its LDS record when the launch has one; the low bits still name the column's HBM slot, which
// a launch without the LDS record uses.
constexpr int32_t kSnapLds = int32_t(1) << 30;

// A batch as the caller hands it over: regions of reads and haplotypes, concatenated.
struct BatchInput {
  bool pd = false;
  int64_t n_regions = 0;
  const int64_t* region_read = nullptr;  // n_regions + 1, read index ranges
  const int64_t* region_hap = nullptr;   // n_regions + 1, haplotype index ranges
  const int64_t* read_off = nullptr;     // n_reads + 1, byte ranges into the read arrays
  const uint8_t *bases = nullptr, *quals = nullptr, *ins = nullptr, *del = nullptr,
                *gcp = nullptr;
  const int64_t* hap_off = nullptr;  // n_haps + 1, byte ranges into hap_bases / hap_flags
  const uint8_t* hap_bases = nullptr;
  const uint8_t* hap_flags = nullptr;  // PD only
  // Plan only what runs with every pair in a wave of its own need (no read pairing, no suffix
  // sharing); such a batch holds at most kMaxPairList pairs and runs no other way.
  bool pairs_only = false;
};

// Pairs of a batch's one-pair-per-wave list.
constexpr int64_t kMaxPairList = int64_t(1) << 20;

// A wave holds the n read rows plus at least one identity row that reproduces DP row 0, at
// R rows per lane over 64 lanes: reads of up to 64 * 8 - 1 = 511 bases run in the staged sweep.
constexpr int kMaxRowsPerLane = 8;
constexpr int kMaxReadLen = 64 * kMaxRowsPerLane - 1;
// Plain reads whose insertion, deletion and gap-continuation qualities are each the same at every
// base, with a gap-continuation quality of at least this (a continuation probability of at most
// 10^-0.4 = 0.4), keep their insertion values divided by the match-to-insertion probability, which
// saves a product per cell; such reads run with the uniform-gap build of the sweep.
constexpr int kXsMinGcp = 4;
// Haplotypes the staged sweep takes: its LDS holds the longest haplotype of a launch per wave,
// and 2048 columns keep four waves of a 256-thread block within 34 KiB.
constexpr int kMaxStagedHap = 2048;
// A region shares prefixes (suffixes) only while it needs at most this many snapshot slots
// (suffix node records): every wave's scratch holds the batch's largest count.
constexpr int kMaxRegionSlots = 128;

struct HostPlan {
  int64_t n_reads = 0, n_haps = 0, n_pairs = 0;
  std::vector<int32_t> read_region, read_local, read_rows;  // rows per lane R per read
  std::vector<int64_t> region_out;  // output base per region
  std::vector<int32_t> region_hap0, region_nhap;
  std::vector<HapPlan> plan;        // n_haps
  std::vector<SnapEntry> snaps;
  std::vector<uint32_t> words;      // PD column words (same offsets as the haplotype bytes)
  // PD: per region, the bytes of read-base classes 4 on (kPdExtraClasses each, zero-padded), and
  // whether the region runs in long_kernel (too many classes, or a read or haplotype too long).
  std::vector<uint8_t> region_classes;
  std::vector<char> region_pd_long;
  int32_t max_slots = 0;
  int32_t max_hap_len = 0;
  // Suffix sharing (plain batches only): the backward sweeps of region g are
  // bsweeps[region_bw[g] .. region_bw[g + 1]), in the order they must run; their node writes
  // are pairs (column, node) in bwrites; a region uses at most max_nodes node records.
  std::vector<BwdSweep> bsweeps;
  std::vector<int32_t> bwrites;
  std::vector<int32_t> region_bw;
  int32_t max_nodes = 0;
  // DP cells of the batch: nominal (every read base x haplotype base), computed with prefix
  // sharing alone, and with suffix sharing (forward, backward) in the FP32 pass.
  int64_t work_nominal = 0, work_prefix = 0, work_forward = 0, work_backward = 0;
  // Snapshot records one FP32 pass writes (summed over reads): to HBM slots, to the wave's LDS
  // record; and the sweeps that start from a record.
  int64_t snap_hbm = 0, snap_lds = 0, snap_restarts = 0;
  // Regions with a read of more than kMaxReadLen bases or a haplotype of more than
  // kMaxStagedHap run every pair in long_kernel, unshared: pairs (read, sorted haplotype) in
  // region order, region g's from region_long[g] to region_long[g + 1]; their reads have
  // read_rows 0 and appear in no bucket.
  std::vector<int32_t> long_pairs;
  std::vector<int64_t> region_long;
  int32_t long_max_hap = 0;
  int64_t long_cells = 0;
  // Work per rows-per-lane bucket (index R-1), costliest first: items of two int32, a read and
  // a second read of the same region sharing the wave (-1 when the read is alone).
  std::vector<int32_t> bucket_items[kMaxRowsPerLane];
  int64_t bucket_pairs[kMaxRowsPerLane] = {};  // read x haplotype pairs of each bucket
  int64_t bucket_cells[kMaxRowsPerLane] = {};  // their read bases x haplotype bases
  int64_t bucket_work[kMaxRowsPerLane] = {};   // cells the FP32 pass computes (both sharings)
};

// Runs fn(0) .. fn(n - 1) on up to n threads (n <= 1: on this one).
template <typename F>
void run_parallel(int n, F&& fn) {
  if (n <= 1) {
    if (n == 1) fn(0);
    return;
  }
  std::vector<std::thread> pool;
  pool.reserve(size_t(n));
  for (int t = 0; t < n; ++t) pool.emplace_back([&fn, t] { fn(t); });
  for (auto& th : pool) th.join();
}

// Host threads for `work` units, at least `per_thread` units each, at most 16.
inline int host_threads(int64_t work, int64_t per_thread) {
  const int64_t hw = std::max<int64_t>(1, int64_t(std::thread::hardware_concurrency()));
  return int(std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(work / per_thread, hw), 16)));
}

// Validates the batch and builds its plan.  Returns kOk or kInvalidArgument with a message
// set through set_error.
int build_plan(const BatchInput& in, HostPlan& out);

// Rows per lane for a read of n bases (n rows plus one identity row over 64 lanes) and the
// lanes it then occupies.
inline int rows_per_lane(int64_t n) { return int((n + 64) / 64); }
inline int64_t lanes_for(int64_t n, int R) { return (n + R) / R; }
// Rows per lane for two reads sharing a wave (each with its identity row), or 0 if none fits.
inline int rows_per_lane2(int64_t na, int64_t nb) {
  for (int R = 1; R <= kMaxRowsPerLane; ++R)
    if (lanes_for(na, R) + lanes_for(nb, R) <= 64) return R;
  return 0;
}

}  // namespace fgkl
