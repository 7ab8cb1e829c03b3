// Aggregation of per-call regions: concurrent JNI-shaped calls (one region each, as HaplotypeCaller
// threads make them) are gathered for up to a short wait and run as one batch, every pair in a
// wave of its own as a lone call runs it, each caller blocking until its own region's likelihoods
// are written.  A region's results are therefore the same bits whatever it was gathered with.
//
// Leader / follower: a queued caller that finds no batch being gathered for its kind, and one of
// the kind's two batch slots free, becomes the leader.  While the other slot's batch runs it waits
// until `max_regions` calls are queued, that batch finishes or `max_wait_us` has passed (with both
// slots free it does not wait); then it takes the queued calls, runs them in its slot and wakes
// every caller it served, while the next queued caller may lead the next batch in the other slot.  The
// inputs are read in place: their callers block until their results are written.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "fgkl_aggregate.hpp"
#include "fgkl_engine.hpp"

namespace fgkl {
namespace {

struct Job {
  const RegionCall* call;
  int rc = kOk;
  std::string err;
  bool done = false;
};

// Two batches per kind: while one runs, the next leader gathers and uploads into the other.
struct Queue {
  std::vector<Job*> pending;
  bool gathering = false;
  std::unique_ptr<DeviceBatch> batch[2];  // device memory and streams reused from batch to batch
  bool busy[2] = {false, false};
};

struct Aggregator {
  std::mutex mu;
  std::condition_variable cv;
  Queue queues[4];  // [pd][double_precision]
  int max_regions = 0;
  int max_wait_us = 0;
  int64_t batches = 0, regions = 0;
};

// Allocated once and never destroyed: its batches must not be freed by a static destructor at
// process exit, when the HIP runtime may already be gone (fgkl_release_idle frees idle ones).
Aggregator& agg() {
  static Aggregator* const a = new Aggregator;
  return *a;
}

std::once_flag g_env_once;

// FGKL_AGGREGATE=<regions>[:<wait microseconds>] turns aggregation on for programs that cannot
// call fgkl_aggregate_config (a JVM running the unchanged Java bindings).
void read_env() {
  const char* v = std::getenv("FGKL_AGGREGATE");
  if (!v || !*v) return;
  char* end = nullptr;
  const long n = std::strtol(v, &end, 10);
  long w = 1000;
  if (end && *end == ':') w = std::strtol(end + 1, nullptr, 10);
  if (n > 1 && w >= 0) {
    std::lock_guard<std::mutex> lk(agg().mu);
    agg().max_regions = int(n);
    agg().max_wait_us = int(w);
  }
}

template <typename T>
void append(std::vector<T>& dst, const T* src, int64_t n) {
  if (n > 0) dst.insert(dst.end(), src, src + n);
}

// Runs the jobs' regions as one batch; on any failure, each region alone, so an invalid region
// fails only its own call.
void run_jobs(std::unique_ptr<DeviceBatch>& batch, const std::vector<Job*>& jobs) {
  const RegionCall& first = *jobs[0]->call;
  std::vector<int64_t> region_read{0}, region_hap{0}, read_off{0}, hap_off{0};
  std::vector<uint8_t> bases, quals, ins, del, gcp, hap_bases, hap_flags;
  int64_t pairs = 0;
  for (const Job* j : jobs) {
    const RegionCall& c = *j->call;
    const int64_t rb = c.read_off[c.n_reads], hb = c.hap_off[c.n_haps];
    pairs += c.n_reads * c.n_haps;
    for (int64_t r = 1; r <= c.n_reads; ++r) read_off.push_back(read_off.back() - c.read_off[r - 1] + c.read_off[r]);
    for (int64_t h = 1; h <= c.n_haps; ++h) hap_off.push_back(hap_off.back() - c.hap_off[h - 1] + c.hap_off[h]);
    region_read.push_back(region_read.back() + c.n_reads);
    region_hap.push_back(region_hap.back() + c.n_haps);
    append(bases, c.bases, rb);
    append(quals, c.quals, rb);
    append(ins, c.ins, rb);
    append(del, c.del, rb);
    append(gcp, c.gcp, rb);
    append(hap_bases, c.hap_bases, hb);
    if (c.pd) append(hap_flags, c.hap_flags, hb);
  }
  BatchInput in;
  in.pd = first.pd;
  in.n_regions = int64_t(jobs.size());
  in.region_read = region_read.data();
  in.region_hap = region_hap.data();
  in.read_off = read_off.data();
  in.bases = bases.data();
  in.quals = quals.data();
  in.ins = ins.data();
  in.del = del.data();
  in.gcp = gcp.data();
  in.hap_off = hap_off.data();
  in.hap_bases = hap_bases.data();
  in.hap_flags = first.pd ? hap_flags.data() : nullptr;
  // The run a lone call gets (compute_region): every pair in its own wave, or, for a single
  // region beyond the pair list, the batched run with prefix and suffix sharing.  Pair-per-wave
  // results do not depend on the lane layout or the launch, so a region gets the same bits in
  // any gathered batch.
  in.pairs_only = pairs <= kMaxPairList;
  const int mode = first.double_precision ? kModeFp64 : kModeFp32Rescue;
  int rc = batch ? batch->assign(in) : DeviceBatch::create(in, batch);
  std::vector<double> out;
  if (rc == kOk) rc = batch->run_latency(mode, nullptr);
  if (rc == kOk) {
    out.resize(size_t(batch->n_pairs()));
    rc = batch->results(out.data(), int64_t(out.size()));
  }
  if (rc == kOk) {
    size_t o = 0;
    for (Job* j : jobs) {
      std::memcpy(j->call->out, out.data() + o, sizeof(double) * size_t(j->call->out_len));
      o += size_t(j->call->out_len);
      j->rc = kOk;
    }
    return;
  }
  for (Job* j : jobs) {
    j->rc = compute_region_alone(*j->call);
    if (j->rc != kOk) j->err = last_error();
  }
}

}  // namespace

void aggregate_config(int max_regions, int max_wait_us) {
  std::call_once(g_env_once, read_env);
  std::lock_guard<std::mutex> lk(agg().mu);
  agg().max_regions = max_regions > 1 ? max_regions : 0;
  agg().max_wait_us = max_wait_us > 0 ? max_wait_us : 0;
  if (agg().max_regions == 0) {
    // Off: give back the device memory of the idle batches (a running one keeps its own).
    for (Queue& q : agg().queues)
      for (int slot = 0; slot < 2; ++slot)
        if (!q.busy[slot]) q.batch[slot].reset();
  }
}

int64_t aggregate_release_idle() {
  std::vector<std::unique_ptr<DeviceBatch>> drop;
  {
    std::lock_guard<std::mutex> lk(agg().mu);
    for (Queue& q : agg().queues)
      for (int slot = 0; slot < 2; ++slot)
        if (!q.busy[slot] && q.batch[slot]) drop.push_back(std::move(q.batch[slot]));
  }
  return int64_t(drop.size());  // destroyed here, outside the lock
}

bool aggregate_enabled() {
  std::call_once(g_env_once, read_env);
  std::lock_guard<std::mutex> lk(agg().mu);
  return agg().max_regions > 1;
}

void aggregate_stats(int64_t* batches, int64_t* regions) {
  std::lock_guard<std::mutex> lk(agg().mu);
  if (batches) *batches = agg().batches;
  if (regions) *regions = agg().regions;
}

int aggregate_run(const RegionCall& call) {
  Aggregator& a = agg();
  Queue& q = a.queues[(call.pd ? 2 : 0) + (call.double_precision ? 1 : 0)];
  Job job;
  job.call = &call;
  std::unique_lock<std::mutex> lk(a.mu);
  q.pending.push_back(&job);
  a.cv.notify_all();
  while (!job.done) {
    if (!q.gathering && !(q.busy[0] && q.busy[1]) &&
        std::find(q.pending.begin(), q.pending.end(), &job) != q.pending.end()) {
      // Lead: while the kind's other batch runs, gather until this one is full, the other
      // finishes or the wait is over (an idle GPU takes what is queued at once); then take a free
      // slot and let the next leader gather while this batch runs.
      q.gathering = true;
      const auto deadline =
          std::chrono::steady_clock::now() + std::chrono::microseconds(a.max_wait_us);
      while (int(q.pending.size()) < a.max_regions && (q.busy[0] || q.busy[1]) &&
             a.cv.wait_until(lk, deadline) != std::cv_status::timeout) {
      }
      // Calls in arrival order, up to max_regions and the one-pair-per-wave list: every gathered
      // region then runs as a lone call runs it (each pair in its own wave, no sharing), so its
      // bits do not depend on the calls it was gathered with.  A region beyond the list runs on
      // its own, as a lone call does.
      size_t take = 0;
      int64_t pairs = 0;
      while (take < q.pending.size() && take < size_t(std::max(a.max_regions, 1))) {
        const RegionCall& c = *q.pending[take]->call;
        if (take > 0 && pairs + c.n_reads * c.n_haps > kMaxPairList) break;
        pairs += c.n_reads * c.n_haps;
        ++take;
      }
      std::vector<Job*> jobs(q.pending.begin(), q.pending.begin() + long(take));
      q.pending.erase(q.pending.begin(), q.pending.begin() + long(take));
      const int slot = q.busy[0] ? 1 : 0;
      q.busy[slot] = true;
      q.gathering = false;
      ++a.batches;
      a.regions += int64_t(jobs.size());
      a.cv.notify_all();
      lk.unlock();
      try {
        run_jobs(q.batch[slot], jobs);
      } catch (const std::exception& e) {
        for (Job* j : jobs) {
          j->rc = kInternal;
          j->err = std::string("fgkl native code failed: ") + e.what();
        }
      }
      lk.lock();
      for (Job* j : jobs) j->done = true;
      q.busy[slot] = false;
      a.cv.notify_all();
    } else {
      a.cv.wait(lk);
    }
  }
  lk.unlock();
  if (job.rc != kOk) set_error(job.rc, job.err);
  return job.rc;
}

}  // namespace fgkl
