// The C-ABI of include/fgkl.h: argument validation with the reference's error classes and
// messages, and the calls into the device engine.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fgkl.h"
#include "fgkl_aggregate.hpp"
#include "fgkl_engine.hpp"
#include "fgkl_grad.hpp"
#include "fgkl_internal.hpp"
#include "fgkl_pool.hpp"
#include "fgkl_replay.hpp"
#include "fgkl_sw.hpp"
#include "fgkl_synthetic.hpp"

using namespace fgkl;

struct fgkl_pairhmm {
  int pd;
  int double_precision;
  int double_fallback;
  std::atomic<uint64_t> fallback_pairs{0};
};

struct fgkl_batch {
  std::unique_ptr<DeviceBatch> batch;
};

struct fgkl_synth {
  Batch b;
  std::vector<double> expected;  // recorded likelihoods of a replayed dump
  int64_t stats[3] = {};         // dump lines, skipped lines, dropped regions
};

struct fgkl_grad {
  std::unique_ptr<GradBatch> grad;
};

struct fgkl_sw_batch {
  std::unique_ptr<SwBatch> batch;
};

namespace fgkl {
int dpp_shl_selftest(int32_t* out64);
}

namespace {

// The per-call pools: plain and PD PairHMM batches, Smith-Waterman batches.  Allocated once and
// never destroyed (see fgkl_pool.hpp).
Pool<DeviceBatch>& batch_pool(int pd) {
  static Pool<DeviceBatch>* const pools = new Pool<DeviceBatch>[2];
  return pools[pd ? 1 : 0];
}
Pool<SwBatch>& sw_pool() {
  static Pool<SwBatch>* const pool = new Pool<SwBatch>;
  return *pool;
}

int fail(int code, const std::string& msg) {
  set_error(code, msg);
  return code;
}

// crates/jni/src/lib.rs:343-363
int split(const int32_t* off, int64_t n, int64_t total, std::vector<int64_t>& out) {
  if (n < 1 || !off || off[0] != 0) return fail(kInvalidArgument, "offset table must start at zero");
  if (off[n - 1] < 0 || int64_t(off[n - 1]) != total)
    return fail(kInvalidArgument, "offset table must end at the array length");
  out.resize(size_t(n));
  for (int64_t i = 0; i < n; ++i) {
    if (i > 0 && off[i] < off[i - 1]) return fail(kInvalidArgument, "offset table must be non-decreasing");
    out[size_t(i)] = off[i];
  }
  return kOk;
}

// The checks of PairHmm::compute_log10_likelihoods (lib.rs:54-67, :279-281, jni lib.rs:325-335).
int check_region(int64_t n_reads, const int64_t* read_off, int64_t n_haps, const int64_t* hap_off,
                 int64_t out_len) {
  if (n_reads < 0 || n_haps < 0) return fail(kInvalidArgument, "negative read or haplotype count");
  for (int64_t i = 0; i < n_reads; ++i)
    if (read_off[i + 1] == read_off[i])
      return fail(kInvalidArgument, "read " + std::to_string(i) + " has no bases");
  for (int64_t i = 0; i < n_haps; ++i)
    if (hap_off[i + 1] == hap_off[i])
      return fail(kInvalidArgument, "haplotype " + std::to_string(i) + " has no bases");
  const int64_t expected = n_reads * n_haps;
  if (out_len != expected)
    return fail(kInvalidArgument, "output has " + std::to_string(out_len) +
                                      " elements but reads x haplotypes is " +
                                      std::to_string(expected));
  return kOk;
}

// PairHmm::compute_log10_likelihoods (lib.rs:272-293) / PdPairHmm (pd.rs:69-95) on one region.
int compute_region(int pd, int double_precision, int double_fallback, int64_t n_reads,
                   const int64_t* read_off, const uint8_t* bases, const uint8_t* quals,
                   const uint8_t* ins, const uint8_t* del, const uint8_t* gcp, int64_t n_haps,
                   const int64_t* hap_off, const uint8_t* hap_bases, const uint8_t* hap_flags,
                   double* out, int64_t out_len, uint64_t* fallbacks) {
  if (int rc = check_region(n_reads, read_off, n_haps, hap_off, out_len)) return rc;
  if (n_reads * n_haps == 0) return kOk;
  BatchInput in;
  in.pd = pd != 0;
  in.n_regions = 1;
  const int64_t region_read[2] = {0, n_reads}, region_hap[2] = {0, n_haps};
  in.region_read = region_read;
  in.region_hap = region_hap;
  in.read_off = read_off;
  in.bases = bases; in.quals = quals; in.ins = ins; in.del = del; in.gcp = gcp;
  in.hap_off = hap_off;
  in.hap_bases = hap_bases;
  in.hap_flags = hap_flags;
  in.pairs_only = n_reads * n_haps <= kMaxPairList;
  // A reusable batch of this kind, checked out of the process-wide pool for this call:
  // consecutive calls reuse its device memory, stream and events (the reference keeps its runners
  // per thread for the same reason, kernel.rs:144-191), and no thread owns one when it exits.
  Checkout<DeviceBatch> co(batch_pool(pd));
  std::unique_ptr<DeviceBatch>& b = co.get();
  int rc = b ? b->assign(in) : DeviceBatch::create(in, b);
  const int mode = double_precision ? kModeFp64 : (double_fallback ? kModeFp32Rescue : kModeFp32NoRescue);
  if (rc == kOk) rc = b->run_latency(mode, nullptr);
  if (rc == kOk) {
    if (fallbacks) *fallbacks = uint64_t(b->last_rescued());
    rc = b->results(out, out_len);
  }
  if (rc == kInternal) co.drop();  // a device error: the next call starts from a fresh batch
  return rc;
}

int compute_native(int pd, const uint8_t* bases, int64_t bases_len, const uint8_t* quals,
                   int64_t quals_len, const uint8_t* ins, int64_t ins_len, const uint8_t* del,
                   int64_t del_len, const uint8_t* gcp, int64_t gcp_len, const int32_t* read_offsets,
                   int64_t read_offsets_len, const uint8_t* hap_bases, int64_t hap_bases_len,
                   const uint8_t* hap_flags, int64_t hap_flags_len, const int32_t* hap_offsets,
                   int64_t hap_offsets_len, int double_precision, double* likelihoods,
                   int64_t likelihoods_len) {
  try {
    // ReadBatch::fetch (crates/jni/src/lib.rs:105-133)
    if (quals_len != bases_len || ins_len != bases_len || del_len != bases_len || gcp_len != bases_len)
      return fail(kInvalidArgument, "read bases, qualities and penalties differ in total length");
    std::vector<int64_t> roff, hoff;
    if (int rc = split(read_offsets, read_offsets_len, bases_len, roff)) return rc;
    if (pd && hap_flags_len != hap_bases_len)  // lib.rs:232-237
      return fail(kInvalidArgument, "haplotype bases and PD flags differ in total length");
    if (int rc = split(hap_offsets, hap_offsets_len, hap_bases_len, hoff)) return rc;
    const int64_t n_reads = read_offsets_len - 1, n_haps = hap_offsets_len - 1;
    if (aggregate_enabled()) {
      if (int rc = check_region(n_reads, roff.data(), n_haps, hoff.data(), likelihoods_len)) return rc;
      if (n_reads * n_haps == 0) return kOk;
      RegionCall call;
      call.pd = pd != 0;
      call.double_precision = double_precision != 0;
      call.n_reads = n_reads;
      call.read_off = roff.data();
      call.bases = bases; call.quals = quals; call.ins = ins; call.del = del; call.gcp = gcp;
      call.n_haps = n_haps;
      call.hap_off = hoff.data();
      call.hap_bases = hap_bases;
      call.hap_flags = pd ? hap_flags : nullptr;
      call.out = likelihoods;
      call.out_len = likelihoods_len;
      return aggregate_run(call);
    }
    return compute_region(pd, double_precision, 1, n_reads, roff.data(), bases, quals, ins, del,
                          gcp, n_haps, hoff.data(), hap_bases, pd ? hap_flags : nullptr,
                          likelihoods, likelihoods_len, nullptr);
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

}  // namespace

namespace fgkl {
int compute_region_alone(const RegionCall& c) {
  return compute_region(c.pd, c.double_precision, 1, c.n_reads, c.read_off, c.bases, c.quals,
                        c.ins, c.del, c.gcp, c.n_haps, c.hap_off, c.hap_bases, c.hap_flags, c.out,
                        c.out_len, nullptr);
}
}  // namespace fgkl

extern "C" {

const char* fgkl_last_error(void) { return fgkl::last_error(); }
const char* fgkl_backend_name(void) { return "hip-gfx950"; }
int fgkl_device_count(void) { return device_available(); }

int fgkl_pairhmm_compute_native(const uint8_t* bases, int64_t bases_len, const uint8_t* quals,
                                int64_t quals_len, const uint8_t* ins_gop, int64_t ins_gop_len,
                                const uint8_t* del_gop, int64_t del_gop_len, const uint8_t* gcp,
                                int64_t gcp_len, const int32_t* read_offsets,
                                int64_t read_offsets_len, const uint8_t* hap_bases,
                                int64_t hap_bases_len, const int32_t* hap_offsets,
                                int64_t hap_offsets_len, int double_precision, double* likelihoods,
                                int64_t likelihoods_len) {
  return compute_native(0, bases, bases_len, quals, quals_len, ins_gop, ins_gop_len, del_gop,
                        del_gop_len, gcp, gcp_len, read_offsets, read_offsets_len, hap_bases,
                        hap_bases_len, nullptr, 0, hap_offsets, hap_offsets_len, double_precision,
                        likelihoods, likelihoods_len);
}

int fgkl_pdhmm_compute_native(const uint8_t* bases, int64_t bases_len, const uint8_t* quals,
                              int64_t quals_len, const uint8_t* ins_gop, int64_t ins_gop_len,
                              const uint8_t* del_gop, int64_t del_gop_len, const uint8_t* gcp,
                              int64_t gcp_len, const int32_t* read_offsets,
                              int64_t read_offsets_len, const uint8_t* hap_bases,
                              int64_t hap_bases_len, const uint8_t* hap_flags,
                              int64_t hap_flags_len, const int32_t* hap_offsets,
                              int64_t hap_offsets_len, int double_precision, double* likelihoods,
                              int64_t likelihoods_len) {
  return compute_native(1, bases, bases_len, quals, quals_len, ins_gop, ins_gop_len, del_gop,
                        del_gop_len, gcp, gcp_len, read_offsets, read_offsets_len, hap_bases,
                        hap_bases_len, hap_flags, hap_flags_len, hap_offsets, hap_offsets_len,
                        double_precision, likelihoods, likelihoods_len);
}

fgkl_pairhmm* fgkl_pairhmm_new(int pd, int double_precision, int double_fallback) {
  if (!device_available()) {
    set_error(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    return nullptr;
  }
  auto* h = new fgkl_pairhmm;
  h->pd = pd;
  h->double_precision = double_precision;
  h->double_fallback = double_fallback;
  return h;
}

void fgkl_pairhmm_free(fgkl_pairhmm* hmm) { delete hmm; }

uint64_t fgkl_pairhmm_fallback_pairs(const fgkl_pairhmm* hmm) {
  return hmm ? hmm->fallback_pairs.load(std::memory_order_relaxed) : 0;
}

int fgkl_pairhmm_compute(fgkl_pairhmm* hmm, int64_t n_reads, const int64_t* read_offsets,
                         const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                         const uint8_t* del_gop, const uint8_t* gcp, int64_t n_haps,
                         const int64_t* hap_offsets, const uint8_t* hap_bases,
                         const uint8_t* hap_flags, double* out, int64_t out_len) {
  if (!hmm) return fail(kInvalidArgument, "null PairHMM handle");
  if (hmm->pd && n_haps > 0 && !hap_flags)
    return fail(kInvalidArgument, "a PD PairHMM needs haplotype flags");
  try {
    uint64_t fb = 0;
    const int rc = compute_region(hmm->pd, hmm->double_precision, hmm->double_fallback, n_reads,
                                  read_offsets, bases, quals, ins_gop, del_gop, gcp, n_haps,
                                  hap_offsets, hap_bases, hmm->pd ? hap_flags : nullptr, out,
                                  out_len, &fb);
    if (rc == kOk) hmm->fallback_pairs.fetch_add(fb, std::memory_order_relaxed);
    return rc;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_batch_create(fgkl_batch** out, int pd, int64_t n_regions, const int64_t* region_reads,
                      const int64_t* region_haps, const int64_t* read_offsets,
                      const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                      const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_offsets,
                      const uint8_t* hap_bases, const uint8_t* hap_flags) {
  if (!out) return fail(kInvalidArgument, "null output handle");
  *out = nullptr;
  if (pd && !hap_flags) return fail(kInvalidArgument, "a PD batch needs haplotype flags");
  try {
    BatchInput in;
    in.pd = pd != 0;
    in.n_regions = n_regions;
    in.region_read = region_reads;
    in.region_hap = region_haps;
    in.read_off = read_offsets;
    in.bases = bases; in.quals = quals; in.ins = ins_gop; in.del = del_gop; in.gcp = gcp;
    in.hap_off = hap_offsets;
    in.hap_bases = hap_bases;
    in.hap_flags = pd ? hap_flags : nullptr;
    auto* b = new fgkl_batch;
    if (int rc = DeviceBatch::create(in, b->batch)) {
      delete b;
      return rc;
    }
    *out = b;
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_batch_run_pairs(fgkl_batch* batch, int mode, fgkl_run_stats* stats) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  RunStats st;
  const int rc = batch->batch->run_latency(mode, stats ? &st : nullptr);
  if (rc == kOk && stats) {
    *stats = fgkl_run_stats{};
    stats->total_ms = st.total_ms;
    stats->primary_ms = st.primary_ms;
    stats->rescue_ms = st.rescue_ms;
    stats->rescued_pairs = st.rescued_pairs;
  }
  return rc;
}

int fgkl_batch_run(fgkl_batch* batch, int mode, int share_prefixes, int use_graph,
                   fgkl_run_stats* stats) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  RunStats st;
  const int rc = batch->batch->run(mode, share_prefixes, use_graph, &st);
  if (rc == kOk && stats) {
    stats->total_ms = st.total_ms;
    stats->primary_ms = st.primary_ms;
    stats->rescue_ms = st.rescue_ms;
    stats->rescued_pairs = st.rescued_pairs;
    for (int i = 0; i < 8; ++i) {
      stats->bucket_ms[i] = st.bucket_ms[i];
      stats->bucket_cells[i] = st.bucket_cells[i];
      stats->bucket_work[i] = st.bucket_work[i];
    }
    stats->dense_rescue_ms = st.dense_rescue_ms;
  }
  return rc;
}

int fgkl_batch_results(fgkl_batch* batch, double* likelihoods, int64_t likelihoods_len) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  return batch->batch->results(likelihoods, likelihoods_len);
}

int64_t fgkl_batch_pairs(const fgkl_batch* batch) { return batch ? batch->batch->n_pairs() : 0; }

int fgkl_batch_work(const fgkl_batch* batch, int64_t* cells4) {
  if (!batch || !cells4) return fail(kInvalidArgument, "null argument");
  batch->batch->work(cells4);
  return kOk;
}

int fgkl_batch_set_uniform_gaps(fgkl_batch* batch, int enable) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  if (enable != 0 && enable != 1) return fail(kInvalidArgument, "enable must be 0 or 1");
  batch->batch->set_uniform_gaps(enable != 0);
  return kOk;
}

int fgkl_batch_set_lds_records(fgkl_batch* batch, int enable) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  if (enable != 0 && enable != 1) return fail(kInvalidArgument, "enable must be 0 or 1");
  batch->batch->set_lds_records(enable != 0);
  return kOk;
}

int fgkl_batch_snapshots(const fgkl_batch* batch, int64_t* out4) {
  if (!batch || !out4) return fail(kInvalidArgument, "null argument");
  batch->batch->snapshots(out4);
  return kOk;
}

int fgkl_plan_snapshots(int pd, int64_t n_regions, const int64_t* region_reads,
                        const int64_t* region_haps, const int64_t* read_offsets,
                        const int64_t* hap_offsets, const uint8_t* hap_bases,
                        const uint8_t* hap_flags, int64_t* out3) {
  if (!region_reads || !region_haps || !read_offsets || !hap_offsets || !hap_bases || !out3)
    return fail(kInvalidArgument, "null argument");
  if (pd && !hap_flags) return fail(kInvalidArgument, "a PD batch needs haplotype flags");
  try {
    BatchInput in;
    in.pd = pd != 0;
    in.n_regions = n_regions;
    in.region_read = region_reads;
    in.region_hap = region_haps;
    in.read_off = read_offsets;
    in.hap_off = hap_offsets;
    in.hap_bases = hap_bases;
    in.hap_flags = pd ? hap_flags : nullptr;
    HostPlan plan;
    if (int rc = build_plan(in, plan)) return rc;
    out3[0] = plan.snap_hbm;
    out3[1] = plan.snap_lds;
    out3[2] = plan.snap_restarts;
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, e.what());
  }
}

int fgkl_batch_tune(fgkl_batch* batch, int share_prefixes, fgkl_tune_report* report) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  try {
    TuneReport rep;
    const int rc = batch->batch->tune(share_prefixes, &rep);
    if (rc == kOk && report) {
      *report = fgkl_tune_report{};
      for (int i = 0; i < 8; ++i) {
        const bool used = rep.cfg[i].rows != 0;
        report->rows[i] = used ? rep.cfg[i].rows : 0;
        report->block[i] = used ? rep.cfg[i].block : 0;
        report->boost[i] = used && rep.cfg[i].boost ? 1 : 0;
        report->default_ms[i] = rep.default_ms[i];
        report->tuned_ms[i] = rep.tuned_ms[i];
      }
      report->candidates = rep.candidates;
      report->seconds = rep.seconds;
    }
    return rc;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_batch_set_launch(fgkl_batch* batch, int bucket_rows, int rows, int block, int boost) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  if (boost != 0 && boost != 1) return fail(kInvalidArgument, "boost must be 0 or 1");
  return batch->batch->set_launch(bucket_rows, LaunchCfg{rows, block, boost != 0});
}

int fgkl_batch_get_launch(const fgkl_batch* batch, int bucket_rows, int* rows, int* block,
                          int* boost) {
  if (!batch || !rows || !block || !boost) return fail(kInvalidArgument, "null argument");
  LaunchCfg c;
  if (int rc = batch->batch->get_launch(bucket_rows, &c)) return rc;
  *rows = c.rows;
  *block = c.block;
  *boost = c.boost ? 1 : 0;
  return kOk;
}

void fgkl_batch_free(fgkl_batch* batch) { delete batch; }

fgkl_synth* fgkl_synth_region(uint64_t seed, int64_t num_reads, int64_t read_len,
                              int64_t num_haps, int64_t hap_len, int pd) {
  auto* s = new fgkl_synth;
  if (pd)
    pd_region_generate(s->b, seed, num_reads, read_len, num_haps, hap_len);
  else
    region_generate(s->b, seed, num_reads, read_len, num_haps, hap_len);
  s->b.end_batch();
  return s;
}

fgkl_synth* fgkl_synth_workload(int workload, uint64_t seed, int64_t n_regions) {
  auto* s = new fgkl_synth;
  workload_generate(s->b, workload, seed, n_regions);
  return s;
}

void fgkl_synth_sizes(const fgkl_synth* s, int64_t* n_regions, int64_t* n_reads,
                      int64_t* read_bases, int64_t* n_haps, int64_t* hap_bases) {
  *n_regions = int64_t(s->b.region_reads.size()) - 1;
  *n_reads = int64_t(s->b.read_offsets.size()) - 1;
  *read_bases = int64_t(s->b.bases.size());
  *n_haps = int64_t(s->b.hap_offsets.size()) - 1;
  *hap_bases = int64_t(s->b.hap_bases.size());
}

void fgkl_synth_copy(const fgkl_synth* s, int64_t* region_reads, int64_t* region_haps,
                     int64_t* read_offsets, uint8_t* bases, uint8_t* quals, uint8_t* ins_gop,
                     uint8_t* del_gop, uint8_t* gcp, int64_t* hap_offsets, uint8_t* hap_bases,
                     uint8_t* hap_flags) {
  const Batch& b = s->b;
  std::memcpy(region_reads, b.region_reads.data(), b.region_reads.size() * sizeof(int64_t));
  std::memcpy(region_haps, b.region_haps.data(), b.region_haps.size() * sizeof(int64_t));
  std::memcpy(read_offsets, b.read_offsets.data(), b.read_offsets.size() * sizeof(int64_t));
  std::memcpy(bases, b.bases.data(), b.bases.size());
  std::memcpy(quals, b.quals.data(), b.quals.size());
  std::memcpy(ins_gop, b.ins.data(), b.ins.size());
  std::memcpy(del_gop, b.del.data(), b.del.size());
  std::memcpy(gcp, b.gcp.data(), b.gcp.size());
  std::memcpy(hap_offsets, b.hap_offsets.data(), b.hap_offsets.size() * sizeof(int64_t));
  std::memcpy(hap_bases, b.hap_bases.data(), b.hap_bases.size());
  std::memcpy(hap_flags, b.hap_flags.data(), b.hap_flags.size());
}

void fgkl_synth_free(fgkl_synth* s) { delete s; }

int fgkl_replay_load(const char* path, int pd, int64_t max_regions, fgkl_synth** out) {
  if (!path || !out) return fail(kInvalidArgument, "null argument");
  *out = nullptr;
  try {
    ReplayData d;
    std::string err;
    if (replay_read(path, pd != 0, max_regions, d, err)) return fail(kInvalidArgument, err);
    auto* s = new fgkl_synth;
    s->b = std::move(d.batch);
    s->expected = std::move(d.expected);
    s->stats[0] = d.lines;
    s->stats[1] = d.skipped_lines;
    s->stats[2] = d.dropped_regions;
    *out = s;
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int64_t fgkl_synth_expected(const fgkl_synth* s, double* out, int64_t n) {
  if (!s) return 0;
  const int64_t have = int64_t(s->expected.size());
  if (out) std::memcpy(out, s->expected.data(), size_t(std::min(n, have)) * sizeof(double));
  return have;
}

void fgkl_replay_stats(const fgkl_synth* s, int64_t* lines, int64_t* skipped_lines,
                       int64_t* dropped_regions) {
  *lines = s ? s->stats[0] : 0;
  *skipped_lines = s ? s->stats[1] : 0;
  *dropped_regions = s ? s->stats[2] : 0;
}

int fgkl_replay_write(const char* path, int pd, int64_t n_regions, const int64_t* region_reads,
                      const int64_t* region_haps, const int64_t* read_offsets,
                      const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                      const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_offsets,
                      const uint8_t* hap_bases, const uint8_t* hap_flags, const double* expected,
                      int64_t n_expected) {
  if (!path || n_regions < 0 || (n_regions && (!region_reads || !region_haps)))
    return fail(kInvalidArgument, "null argument");
  try {
    Batch b;
    b.region_reads.assign(region_reads, region_reads + n_regions + 1);
    b.region_haps.assign(region_haps, region_haps + n_regions + 1);
    const int64_t nr = region_reads[n_regions], nh = region_haps[n_regions];
    b.read_offsets.assign(read_offsets, read_offsets + nr + 1);
    b.hap_offsets.assign(hap_offsets, hap_offsets + nh + 1);
    const size_t rb = size_t(read_offsets[nr]), hb = size_t(hap_offsets[nh]);
    b.bases.assign(bases, bases + rb);
    b.quals.assign(quals, quals + rb);
    b.ins.assign(ins_gop, ins_gop + rb);
    b.del.assign(del_gop, del_gop + rb);
    b.gcp.assign(gcp, gcp + rb);
    b.hap_bases.assign(hap_bases, hap_bases + hb);
    if (hap_flags)
      b.hap_flags.assign(hap_flags, hap_flags + hb);
    else
      b.hap_flags.assign(hb, 0);
    std::vector<double> e(expected, expected + (expected ? n_expected : 0));
    std::string err;
    if (replay_write(b, e, path, pd != 0, err)) return fail(kInvalidArgument, err);
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

void fgkl_synth_splitmix(uint64_t seed, int64_t n, uint64_t* out) {
  Rng rng(seed);
  for (int64_t i = 0; i < n; ++i) out[i] = rng.next_u64();
}

int fgkl_selftest_dpp(int32_t* out64) { return dpp_selftest(out64); }

int fgkl_selftest_dpp_shl(int32_t* out64) { return dpp_shl_selftest(out64); }

int fgkl_grad_create(fgkl_grad** out, int64_t n_regions, const int64_t* region_reads,
                     const int64_t* region_haps, const int64_t* read_offsets,
                     const uint8_t* bases, const int64_t* hap_offsets,
                     const uint8_t* hap_bases) {
  if (!out) return fail(kInvalidArgument, "null output handle");
  *out = nullptr;
  try {
    GradInput in;
    in.n_regions = n_regions;
    in.region_read = region_reads;
    in.region_hap = region_haps;
    in.read_off = read_offsets;
    in.bases = bases;
    in.hap_off = hap_offsets;
    in.hap_bases = hap_bases;
    auto* g = new fgkl_grad;
    if (int rc = GradBatch::create(in, g->grad)) {
      delete g;
      return rc;
    }
    *out = g;
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_grad_set_precision(fgkl_grad* g, int fp32) {
  if (!g) return fail(kInvalidArgument, "null gradient handle");
  return g->grad->set_precision(fp32);
}

int64_t fgkl_grad_pairs(const fgkl_grad* g) { return g ? g->grad->n_pairs() : 0; }

int64_t fgkl_grad_read_bases(const fgkl_grad* g) { return g ? g->grad->n_read_bases() : 0; }

int fgkl_grad_forward(fgkl_grad* g, const double* err, const double* ins, const double* del,
                      double gcp, double* log10_likelihoods, double* kernel_ms) {
  if (!g) return fail(kInvalidArgument, "null gradient handle");
  try {
    GradParams p;
    p.err = err; p.ins = ins; p.del = del; p.gcp = gcp;
    return g->grad->forward(p, log10_likelihoods, kernel_ms);
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_grad_backward(fgkl_grad* g, const double* err, const double* ins, const double* del,
                       double gcp, const double* weights, double* log10_likelihoods,
                       double* grad_err, double* grad_ins, double* grad_del, double* grad_gcp,
                       double* kernel_ms) {
  if (!g) return fail(kInvalidArgument, "null gradient handle");
  try {
    GradParams p;
    p.err = err; p.ins = ins; p.del = del; p.gcp = gcp;
    return g->grad->backward(p, weights, log10_likelihoods, grad_err, grad_ins, grad_del,
                             grad_gcp, kernel_ms);
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

void fgkl_grad_free(fgkl_grad* g) { delete g; }

}  // extern "C"

// ---------------------------------------------------------------------------------------------
// Smith-Waterman

namespace {

// Alignment::cigar_string: "<len><op>" per element.
std::string cigar_text(const uint32_t* ops, int64_t n) {
  static const char kOps[] = "MIDNSHP=X";
  std::string s;
  for (int64_t k = 0; k < n; ++k) {
    s += std::to_string(ops[k] >> 4);
    s += kOps[ops[k] & 15u];
  }
  return s;
}

}  // namespace

const char* fgkl_sw_backend_name(void) { return "hip-gfx950"; }

int fgkl_sw_align_native(const uint8_t* reference, int64_t reference_len,
                         const uint8_t* alternate, int64_t alternate_len, int32_t match_value,
                         int32_t mismatch_penalty, int32_t gap_open, int32_t gap_extend,
                         int32_t strategy, int32_t* offset, char* cigar, int64_t cigar_cap) {
  try {
    if (strategy < 0 || strategy > 3)  // crates/jni/src/lib.rs:283-291
      return fail(kInvalidArgument, "unknown overhang strategy code " + std::to_string(strategy));
    if (reference_len < 0 || alternate_len < 0 || (reference_len > 0 && !reference) ||
        (alternate_len > 0 && !alternate))
      return fail(kInvalidArgument, "null or negative-length sequence");
    if (reference_len == 0 || alternate_len == 0)  // Error::EmptySequence
      return fail(kInvalidArgument, "sequences must be non-empty");
    if (!offset || !cigar || cigar_cap < 1) return fail(kInvalidArgument, "no room for the result");
    // A reusable one-pair batch checked out of the process-wide pool for this call (the
    // reference keeps one aligner per thread, crates/jni/src/lib.rs:23-26).
    Checkout<SwBatch> co(sw_pool());
    std::unique_ptr<SwBatch>& sw = co.get();
    if (!sw && !device_available())
      return fail(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    const int64_t ro[2] = {0, reference_len}, ao[2] = {0, alternate_len};
    const int32_t params[4] = {match_value, mismatch_penalty, gap_open, gap_extend};
    std::vector<uint32_t> ops;
    int64_t starts[2] = {0, 0};
    int32_t off = 0;
    int rc = sw ? sw->assign(1, ro, reference, ao, alternate)
                : SwBatch::create(1, ro, reference, ao, alternate, sw);
    if (rc == kOk) rc = sw->run(params, strategy, nullptr);
    if (rc == kOk) {
      ops.resize(static_cast<size_t>(sw->cigar_bound()));
      rc = sw->results(&off, starts, ops.data(), int64_t(ops.size()));
    }
    if (rc == kInternal) co.drop();
    if (rc != kOk) return rc;
    const std::string text = cigar_text(ops.data(), starts[1]);
    if (int64_t(text.size()) + 1 > cigar_cap)
      return fail(kInvalidArgument, "CIGAR text needs " + std::to_string(text.size() + 1) + " bytes");
    std::memcpy(cigar, text.c_str(), text.size() + 1);
    *offset = off;
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_sw_batch_create(fgkl_sw_batch** out, int64_t n_pairs, const int64_t* ref_offsets,
                         const uint8_t* ref, const int64_t* alt_offsets, const uint8_t* alt) {
  try {
    if (!out) return fail(kInvalidArgument, "null output handle");
    *out = nullptr;
    if (!device_available())
      return fail(kInternal, "no HIP device is available: the fgkl GPU engine needs an MI355X (gfx950)");
    std::unique_ptr<fgkl_sw_batch> h(new fgkl_sw_batch);
    if (int rc = SwBatch::create(n_pairs, ref_offsets, ref, alt_offsets, alt, h->batch)) return rc;
    *out = h.release();
    return kOk;
  } catch (const std::exception& e) {
    return fail(kInternal, std::string("fgkl native code failed: ") + e.what());
  }
}

int fgkl_sw_batch_run(fgkl_sw_batch* batch, const int32_t params[4], int strategy,
                      fgkl_sw_stats* stats) {
  if (!batch || !params) return fail(kInvalidArgument, "null batch or weights");
  SwRunStats st;
  if (int rc = batch->batch->run(params, strategy, stats ? &st : nullptr)) return rc;
  if (stats) {
    stats->total_ms = st.total_ms;
    for (int k = 0; k < 8; ++k) {
      stats->bucket_ms[k] = st.bucket_ms[k];
      stats->bucket_pairs[k] = st.bucket_pairs[k];
      stats->bucket_cells[k] = st.bucket_cells[k];
    }
    stats->substring_pairs = st.substring_pairs;
  }
  return kOk;
}

int fgkl_sw_batch_results(fgkl_sw_batch* batch, int32_t* offsets, int64_t* cigar_starts,
                          uint32_t* cigar, int64_t cigar_cap) {
  if (!batch) return fail(kInvalidArgument, "null batch");
  return batch->batch->results(offsets, cigar_starts, cigar, cigar_cap);
}

int64_t fgkl_sw_batch_pairs(const fgkl_sw_batch* batch) { return batch ? batch->batch->n_pairs() : 0; }
int64_t fgkl_sw_batch_cells(const fgkl_sw_batch* batch) { return batch ? batch->batch->cells() : 0; }
int64_t fgkl_sw_batch_cigar_bound(const fgkl_sw_batch* batch) {
  return batch ? batch->batch->cigar_bound() : 0;
}
void fgkl_sw_batch_free(fgkl_sw_batch* batch) { delete batch; }

// ---------------------------------------------------------------------------------------------
// Aggregation of per-call regions

int fgkl_aggregate_config(int max_regions, int max_wait_us) {
  if (max_regions < 0 || max_wait_us < 0)
    return fail(kInvalidArgument, "aggregation needs a non-negative region count and wait");
  aggregate_config(max_regions, max_wait_us);
  return kOk;
}

void fgkl_aggregate_stats(int64_t* batches, int64_t* regions) { aggregate_stats(batches, regions); }

// ---------------------------------------------------------------------------------------------
// Pools of reusable per-call batches

namespace {

// Counts its own destructions, for the pool self-test.
struct PoolProbe {
  explicit PoolProbe(std::atomic<int64_t>* d) : destroyed(d) {}
  ~PoolProbe() { destroyed->fetch_add(1); }
  std::atomic<int64_t>* destroyed;
};

}  // namespace

int64_t fgkl_release_idle(void) {
  try {
    return batch_pool(0).release_idle() + batch_pool(1).release_idle() +
           sw_pool().release_idle() + aggregate_release_idle();
  } catch (const std::exception& e) {
    fail(kInternal, std::string("fgkl native code failed: ") + e.what());
    return -1;
  }
}

int fgkl_pool_stats(int kind, int64_t* idle, int64_t* checked_out, int64_t* reused) {
  if (kind == 0 || kind == 1)
    batch_pool(kind).stats(idle, checked_out, reused);
  else if (kind == 2)
    sw_pool().stats(idle, checked_out, reused);
  else
    return fail(kInvalidArgument, "pool kind must be 0 (PairHMM), 1 (PD-HMM) or 2 (Smith-Waterman)");
  return kOk;
}

int fgkl_selftest_pool(int threads, int rounds, int64_t* out5) {
  if (threads < 1 || rounds < 1 || !out5) return fail(kInvalidArgument, "bad self-test arguments");
  std::atomic<int64_t> created{0}, destroyed{0};
  Pool<PoolProbe> pool;
  {
    std::vector<std::thread> ts;
    for (int t = 0; t < threads; ++t)
      ts.emplace_back([&] {
        for (int r = 0; r < rounds; ++r) {
          Checkout<PoolProbe> co(pool);
          if (!co.get()) {
            co.get().reset(new PoolProbe(&destroyed));
            created.fetch_add(1);
          }
          std::this_thread::yield();
        }
      });
    for (auto& th : ts) th.join();  // every worker has exited here
  }
  int64_t idle = 0, out = 0, reused = 0;
  pool.stats(&idle, &out, &reused);
  out5[0] = created.load();
  out5[1] = destroyed.load();  // destroyed while the workers ran and exited
  out5[2] = idle;
  out5[3] = out;
  out5[4] = reused;
  pool.release_idle();
  if (destroyed.load() != created.load())
    return fail(kInternal, "pool self-test: release_idle left objects alive");
  return kOk;
}
