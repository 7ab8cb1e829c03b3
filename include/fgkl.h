/*
 * fgkl-mi355x: the C-ABI of the MI355X (gfx950) PairHMM / PD-HMM engine.
 *
 * Plain pointers and sizes only.  Every function returns FGKL_OK (0), FGKL_INVALID_ARGUMENT (1)
 * or FGKL_INTERNAL_ERROR (2); fgkl_last_error() gives the calling thread's message for the last
 * failure.  The reference maps the same two failure classes to IllegalArgumentException and
 * RuntimeException (crates/jni/src/lib.rs:21-92).
 *
 * Reads are passed as five concatenated byte arrays (bases, base qualities, insertion and
 * deletion gap-open qualities, gap-continuation penalties) delimited by an offset table, and
 * haplotypes (with PD flags) likewise; the log10 likelihood of read r given haplotype h of a
 * region lands at likelihoods[r * H + h] (FgklPairHmm.java:100-105, lib.rs:270-271).
 */
#ifndef FGKL_H
#define FGKL_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FGKL_OK 0
#define FGKL_INVALID_ARGUMENT 1
#define FGKL_INTERNAL_ERROR 2

/* Message of the last failing call on this thread ("" if none). */
const char* fgkl_last_error(void);

/* Replaces FgklPairHmm.backendNative / FgklPdHmm.backendNative
 * (crates/jni/src/lib.rs:149-156, :197-204): the engine's name, "hip-gfx950". */
const char* fgkl_backend_name(void);

/* Number of HIP devices visible to the process (0 on a host without a GPU). */
int fgkl_device_count(void);

/*
 * Replaces Java_com_fulcrumgenomics_fgkl_pairhmm_FgklPairHmm_computeNative
 * (crates/jni/src/lib.rs:159-195; Java declaration FgklPairHmm.java:106).
 * One region: read_offsets has R+1 entries, hap_offsets H+1, likelihoods R*H.
 * Validation follows the reference: the five read arrays must have equal total length
 * (lib.rs:122-131); offset tables start at zero, never decrease and end at the array length
 * (lib.rs:343-363); likelihoods must hold R*H values (lib.rs:325-335); reads and haplotypes
 * must be non-empty (crates/pairhmm/src/lib.rs:54-67, :279-281).  double_precision selects the
 * FP64 engine, otherwise FP32 with FP64 recomputation of pairs below 1e-28
 * (crates/jni/src/lib.rs:319-322).
 */
int fgkl_pairhmm_compute_native(const uint8_t* bases, int64_t bases_len,
                                const uint8_t* quals, int64_t quals_len,
                                const uint8_t* ins_gop, int64_t ins_gop_len,
                                const uint8_t* del_gop, int64_t del_gop_len,
                                const uint8_t* gcp, int64_t gcp_len,
                                const int32_t* read_offsets, int64_t read_offsets_len,
                                const uint8_t* hap_bases, int64_t hap_bases_len,
                                const int32_t* hap_offsets, int64_t hap_offsets_len,
                                int double_precision,
                                double* likelihoods, int64_t likelihoods_len);

/*
 * Replaces Java_com_fulcrumgenomics_fgkl_pdhmm_FgklPdHmm_computeNative
 * (crates/jni/src/lib.rs:207-254; Java declaration FgklPdHmm.java:121).  As above plus the
 * per-base PD flags (pd.rs:15-27), which must be as long as hap_bases (lib.rs:232-237).
 */
int fgkl_pdhmm_compute_native(const uint8_t* bases, int64_t bases_len,
                              const uint8_t* quals, int64_t quals_len,
                              const uint8_t* ins_gop, int64_t ins_gop_len,
                              const uint8_t* del_gop, int64_t del_gop_len,
                              const uint8_t* gcp, int64_t gcp_len,
                              const int32_t* read_offsets, int64_t read_offsets_len,
                              const uint8_t* hap_bases, int64_t hap_bases_len,
                              const uint8_t* hap_flags, int64_t hap_flags_len,
                              const int32_t* hap_offsets, int64_t hap_offsets_len,
                              int double_precision,
                              double* likelihoods, int64_t likelihoods_len);

/*
 * The allocation-free variant of fgkl_pairhmm_compute_native (no counterpart in the reference:
 * it is for callers that cannot allocate device memory, such as serving loops and harnesses that
 * forbid hipMalloc).  Same arguments, validation, messages and results (bit for bit) as
 * fgkl_pairhmm_compute_native; it calls no hipMalloc, hipFree or other allocating routine.
 * Every device byte beyond the caller's arrays is a static array of the library: 32 MiB of
 * scratch per GPU, held from module load to process exit, plus the Phred tables.  The first call
 * on a device (with that device current) looks up their addresses and uploads the tables.
 * Limits: reads of at most 511 bases, haplotypes of at most 2048, at most 1048576 pairs, and a
 * region whose arrays, pair lists and results fit the 32 MiB; beyond them it returns
 * FGKL_INVALID_ARGUMENT with a message naming the limit (fgkl_pairhmm_compute_native takes them).
 * Calls on one device share its scratch: this library serialises them, and work another process
 * or caller enqueues on that memory meanwhile would race.
 */
int fgkl_pairhmm_compute_native_noalloc(const uint8_t* bases, int64_t bases_len,
                                        const uint8_t* quals, int64_t quals_len,
                                        const uint8_t* ins_gop, int64_t ins_gop_len,
                                        const uint8_t* del_gop, int64_t del_gop_len,
                                        const uint8_t* gcp, int64_t gcp_len,
                                        const int32_t* read_offsets, int64_t read_offsets_len,
                                        const uint8_t* hap_bases, int64_t hap_bases_len,
                                        const int32_t* hap_offsets, int64_t hap_offsets_len,
                                        int double_precision, double* likelihoods,
                                        int64_t likelihoods_len);

/*
 * A configured PairHMM or PD-HMM, replacing PairHmm::new / PdPairHmm::new with Config
 * {precision, double_fallback} (crates/pairhmm/src/lib.rs:213-265, pd.rs:41-66).
 * pd selects the PD kernel; double_precision the FP64 engine; double_fallback recomputes
 * underflowing FP32 pairs in FP64 (off: they are left NaN, lib.rs:221-226).
 * Returns NULL (message in fgkl_last_error) if no GPU is available.
 */
typedef struct fgkl_pairhmm fgkl_pairhmm;
fgkl_pairhmm* fgkl_pairhmm_new(int pd, int double_precision, int double_fallback);
void fgkl_pairhmm_free(fgkl_pairhmm* hmm);
/* Replaces PairHmm::fallback_pairs (lib.rs:263-268): FP32 pairs recomputed so far. */
uint64_t fgkl_pairhmm_fallback_pairs(const fgkl_pairhmm* hmm);
/*
 * Replaces PairHmm::compute_log10_likelihoods (lib.rs:272-293) and
 * PdPairHmm::compute_log10_likelihoods (pd.rs:69-95) for one region.  read_offsets has
 * n_reads+1 entries into the five read arrays, hap_offsets n_haps+1 into hap_bases (and
 * hap_flags, NULL unless the handle is PD); out holds n_reads*n_haps values.
 */
int fgkl_pairhmm_compute(fgkl_pairhmm* hmm,
                         int64_t n_reads, const int64_t* read_offsets,
                         const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                         const uint8_t* del_gop, const uint8_t* gcp,
                         int64_t n_haps, const int64_t* hap_offsets,
                         const uint8_t* hap_bases, const uint8_t* hap_flags,
                         double* out, int64_t out_len);

/*
 * Aggregation of per-call regions (no reference counterpart: fgkl serves each per-region call on
 * its own).  With max_regions > 1, concurrent fgkl_pairhmm_compute_native /
 * fgkl_pdhmm_compute_native calls of one kind (plain or PD, float or double) share batches of
 * up to max_regions regions.  A call waits for company only while another batch of its kind is
 * running, and for at most max_wait_us; on an idle GPU it runs at once.  Each call is validated
 * on its own thread and returns when its own likelihoods are written.  A batch that fails reruns
 * its regions one by one, so a bad region fails only its own call.  0 turns aggregation off (the
 * default) and frees the device memory its idle batches hold.  The environment variable
 * FGKL_AGGREGATE=<regions>[:<wait us>] sets it for programs that cannot call this (a JVM with
 * the unchanged Java bindings).  fgkl_aggregate_stats reports the batches run and the regions
 * they held.
 */
int fgkl_aggregate_config(int max_regions, int max_wait_us);
void fgkl_aggregate_stats(int64_t* batches, int64_t* regions);

/*
 * Reusable per-call batches (no reference counterpart; the reference keeps one runner per
 * thread, crates/pairhmm/src/kernel.rs:144-191).  The per-call entry points check a device batch
 * out of a process-wide pool for each call and return it afterwards, so calls reuse device
 * memory, streams and events, and a thread that exits holds none: nothing makes HIP calls at
 * thread exit.  Pooled batches are never freed implicitly (not at thread exit, not from static
 * destructors at process exit); fgkl_release_idle frees the idle ones (per-call pools and idle
 * aggregation batches) and returns how many, or -1 on error.  fgkl_pool_stats reports, for kind
 * 0 (PairHMM), 1 (PD-HMM) or 2 (Smith-Waterman), the idle batches, the checkouts not yet
 * returned and the checkouts served by an idle batch.  fgkl_selftest_pool exercises the pool on
 * the CPU with `threads` threads of `rounds` checkouts each, all of which exit before it reads
 * out5 = {objects created, objects destroyed while the threads ran and exited, idle objects,
 * checkouts outstanding, checkouts served by an idle object}.
 */
int64_t fgkl_release_idle(void);
int fgkl_pool_stats(int kind, int64_t* idle, int64_t* checked_out, int64_t* reused);
int fgkl_selftest_pool(int threads, int rounds, int64_t* out5);

/*
 * Batched engine (no reference counterpart: the reference computes one region per call on the
 * calling thread, crates/pairhmm/src/lib.rs:377-397).  A batch holds many regions resident in
 * device memory; region g owns reads region_reads[g]..region_reads[g+1] and haplotypes
 * region_haps[g]..region_haps[g+1]; results land region after region, each region's block laid
 * out as likelihoods[r * H + h].
 */
typedef struct fgkl_batch fgkl_batch;
#define FGKL_MODE_FP32_RESCUE 0   /* FP32, FP64 recomputation of pairs below 1e-28 */
#define FGKL_MODE_FP64 1          /* FP64 for every pair */
#define FGKL_MODE_FP64_PER_REGION 2 /* FP64, one pair per wave, one synchronous call per region */
#define FGKL_MODE_FP32_NO_RESCUE 3  /* FP32, underflowing pairs left NaN */

typedef struct fgkl_run_stats {
  double total_ms;        /* device time of the run, HIP events on the batch's stream */
  double primary_ms;      /* sweep kernels of the first precision (-1 inside a graph) */
  double rescue_ms;       /* FP64 rescue kernels (-1 inside a graph) */
  int64_t rescued_pairs;  /* FP32 pairs recomputed in FP64 */
  double bucket_ms[8];    /* primary sweep launch per rows-per-lane bucket R=1..8 (eager runs) */
  int64_t bucket_cells[8]; /* DP cells (read bases x haplotype bases) each of those launches covers */
  int64_t bucket_work[8];  /* DP cells each computes with prefix and suffix sharing (FP32 pass) */
  double dense_rescue_ms;  /* part of rescue_ms spent on reads lost against a third or more of
                              their haplotypes, rerun in FP64 with prefix sharing (eager runs) */
} fgkl_run_stats;

int fgkl_batch_create(fgkl_batch** out, int pd, int64_t n_regions,
                      const int64_t* region_reads, const int64_t* region_haps,
                      const int64_t* read_offsets, const uint8_t* bases, const uint8_t* quals,
                      const uint8_t* ins_gop, const uint8_t* del_gop, const uint8_t* gcp,
                      const int64_t* hap_offsets, const uint8_t* hap_bases,
                      const uint8_t* hap_flags);
int fgkl_batch_run(fgkl_batch* batch, int mode, int share_prefixes, int use_graph,
                   fgkl_run_stats* stats);
/*
 * Every (read, haplotype) pair in a wave of its own, without prefix or suffix sharing: more cells,
 * but no wave walks a read through all of its region's haplotypes, so a batch of a few regions
 * that cannot fill the GPU finishes sooner than fgkl_batch_run makes it.  The per-call entry
 * points run this way.  Batches of more than 2^20 pairs run as fgkl_batch_run(mode, 1, 0) does.
 * Stats carry total_ms, primary_ms, rescue_ms and rescued_pairs.
 */
int fgkl_batch_run_pairs(fgkl_batch* batch, int mode, fgkl_run_stats* stats);
int fgkl_batch_results(fgkl_batch* batch, double* likelihoods, int64_t likelihoods_len);
int64_t fgkl_batch_pairs(const fgkl_batch* batch);
/*
 * The work the batch's plan computes, in DP cells (read bases x haplotype columns): [0] nominal
 * (every read base x every haplotype base, as GCUPS counts them), [1] with prefix sharing alone
 * (FP64 mode), [2] forward and [3] backward with suffix sharing too (the FP32 pass with
 * prefix sharing; plain batches only, PD batches report 0 backward and forward = [1]).
 */
int fgkl_batch_work(const fgkl_batch* batch, int64_t* cells4);
/*
 * Plans a batch on the host alone (no device; the same arrays as fgkl_batch_create, of which the
 * read arrays are not read and may be null) and reports its prefix-sharing snapshot records per
 * FP32 pass, summed over reads: out3[0] records written to HBM scratch slots, out3[1] records a
 * wave keeps in its LDS record (launches that cannot hold one write these to HBM too), out3[2]
 * sweeps that resume from a record.  No reference counterpart: the reference keeps its
 * snapshot columns in the CPU workspace (crates/pairhmm/src/kernel.rs:366-396, 655-662).
 */
/*
 * Whether the batch's prefix-sharing launches may keep a snapshot record in each wave's LDS
 * (1, the default: when the plan marks records for it and the record costs the launch no
 * resident workgroups) or write every record to an HBM scratch slot (0).  Results are the same
 * bits either way.  No reference counterpart.
 */
int fgkl_batch_set_lds_records(fgkl_batch* batch, int enable);
/*
 * Whether a plain batch's FP32 main passes run the items whose reads have uniform gap qualities
 * (each read's insertion, deletion and gap-continuation qualities the same at every base, as
 * GATK passes them by default) with the kernel build that keeps one copy of the gap terms per
 * lane (1, the default), or every item with the general build (0).  Results are the same bits
 * either way.  No reference counterpart.
 */
int fgkl_batch_set_uniform_gaps(fgkl_batch* batch, int enable);
/*
 * The batch's snapshot records per FP32 pass as fgkl_plan_snapshots reports them (out4[0..2])
 * and out4[3], the sweep launches so far (eager or captured into a graph) that kept a record in
 * each wave's LDS.
 */
int fgkl_batch_snapshots(const fgkl_batch* batch, int64_t* out4);
int fgkl_plan_snapshots(int pd, int64_t n_regions, const int64_t* region_reads,
                        const int64_t* region_haps, const int64_t* read_offsets,
                        const int64_t* hap_offsets, const uint8_t* hap_bases,
                        const uint8_t* hap_flags, int64_t* out3);

/*
 * Launch autotuner (no reference counterpart: fgkl picks lanes per batch from the CPU's SIMD
 * width, crates/pairhmm/src/lib.rs:408-413, simd.rs:317-328).  For every rows-per-lane bucket
 * R of the FP32 pass it times rows per lane (R or R+1; R only for PD batches, whose last bits
 * may depend on the lane layout under prefix sharing), threads per workgroup (256 or 64) and
 * the natural or occupancy-raised kernel build on this batch, keeps the fastest (a candidate
 * must beat the default by 1%) for this batch and for batches of the same kind (plain / PD)
 * created later in the process on this device, and drops captured graphs.  Every
 * configuration computes the same bits.  set_launch sets one bucket by hand (rows 0 = R;
 * blocks 64, 128 or 256); out-of-range values are rejected with FGKL_INVALID_ARGUMENT.
 */
typedef struct fgkl_tune_report {
  int32_t rows[8];       /* per bucket R=1..8: kernel rows per lane chosen (0: empty bucket) */
  int32_t block[8];      /* threads per workgroup chosen */
  int32_t boost[8];      /* 1 when the occupancy-raised build won */
  double default_ms[8];  /* the bucket's launch at rows R, 256 threads, natural build */
  double tuned_ms[8];    /* with the configuration chosen */
  int32_t candidates;    /* configurations timed (each twice, best kept) */
  double seconds;        /* wall time of the tuning */
} fgkl_tune_report;

int fgkl_batch_tune(fgkl_batch* batch, int share_prefixes, fgkl_tune_report* report);
int fgkl_batch_set_launch(fgkl_batch* batch, int bucket_rows, int rows, int block, int boost);
int fgkl_batch_get_launch(const fgkl_batch* batch, int bucket_rows, int* rows, int* block,
                          int* boost);
void fgkl_batch_free(fgkl_batch* batch);

/*
 * Synthetic data (synthetic.rs:10-334 restated: Region::generate, PdRegion::generate) and the
 * benchmark workloads.  A generated batch is read back with fgkl_synth_sizes/fgkl_synth_copy.
 * workload: 0 one 32x8 region (151 bp reads, 300-350 bp haplotypes), 1 batched uneven regions,
 * 2 underflow-heavy regions, 3 PD regions, 4 adjoint regions.
 */
typedef struct fgkl_synth fgkl_synth;
fgkl_synth* fgkl_synth_region(uint64_t seed, int64_t num_reads, int64_t read_len,
                              int64_t num_haps, int64_t hap_len, int pd);
fgkl_synth* fgkl_synth_workload(int workload, uint64_t seed, int64_t n_regions);
void fgkl_synth_sizes(const fgkl_synth* s, int64_t* n_regions, int64_t* n_reads,
                      int64_t* read_bases, int64_t* n_haps, int64_t* hap_bases);
void fgkl_synth_copy(const fgkl_synth* s, int64_t* region_reads, int64_t* region_haps,
                     int64_t* read_offsets, uint8_t* bases, uint8_t* quals, uint8_t* ins_gop,
                     uint8_t* del_gop, uint8_t* gcp, int64_t* hap_offsets, uint8_t* hap_bases,
                     uint8_t* hap_flags);
void fgkl_synth_free(fgkl_synth* s);

/*
 * GATK results dumps (no C-ABI counterpart: fgkl reads them in its replay tools,
 * crates/pairhmm/src/bin/replay.rs:83-159 and pdreplay.rs:34-115).  fgkl_replay_load reads a
 * --pair-hmm-results-file (pd = 0: space-separated hap, read, Phred+33 quals, ins, del, gcp,
 * log10L) or --pdhmm-results-file (pd = 1: tab-separated, flags as Java's Arrays.toString) dump
 * into a batch handle, regions reconstructed from the repeating haplotype lists, at most
 * max_regions of them (all when <= 0); read it back with fgkl_synth_sizes / fgkl_synth_copy and
 * the recorded likelihoods with fgkl_synth_expected (returns their count).  fgkl_replay_stats
 * reports data lines, skipped malformed lines and dropped incomplete regions.
 * fgkl_replay_write writes a batch and per-pair values in the same format.
 */
int fgkl_replay_load(const char* path, int pd, int64_t max_regions, fgkl_synth** out);
int64_t fgkl_synth_expected(const fgkl_synth* s, double* out, int64_t n);
void fgkl_replay_stats(const fgkl_synth* s, int64_t* lines, int64_t* skipped_lines,
                       int64_t* dropped_regions);
int fgkl_replay_write(const char* path, int pd, int64_t n_regions, const int64_t* region_reads,
                      const int64_t* region_haps, const int64_t* read_offsets,
                      const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                      const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_offsets,
                      const uint8_t* hap_bases, const uint8_t* hap_flags, const double* expected,
                      int64_t n_expected);
/* SplitMix64 (synthetic.rs:19-40): n successive next_u64 draws from seed. */
void fgkl_synth_splitmix(uint64_t seed, int64_t n, uint64_t* out);

/*
 * Measurement (no reference counterpart): the rate at which a full grid of waves issues vector
 * ALU instructions on the current device, compiled as the library's kernels are (packed FP32
 * off).  kind 0: FP32 FMA (v_fma_f32), 1: FP64 FMA (v_fma_f64), 2: 32-bit integer add and xor.
 * Writes lane operations per second (an FMA counts once) and the best launch time in ms.
 */
int fgkl_measure_valu_rate(int kind, double* lane_ops_per_second, double* best_ms);

/* Checks the DPP lane shift the kernels rely on: out[i] receives 1000+i-1, out[0] = -1. */
int fgkl_selftest_dpp(int32_t* out64);
/* Checks the opposite DPP lane shift (the adjoint's): out[i] receives 1000+i+1, out[63] = -1. */
int fgkl_selftest_dpp_shl(int32_t* out64);
/*
 * Gradient engine (no reference counterpart: fgkl computes likelihoods only).  The model is the
 * reference's FP64 forward recurrence (crates/pairhmm/src/reference.rs:20-64) with continuous
 * parameters per read base i: error probability e[i] (priors 1-e match, e/3 mismatch, N matches
 * everything), insertion and deletion gap-open probabilities d_ins[i], d_del[i] (M->M =
 * 1 - d_ins - d_del, M->I = d_ins, M->D = d_del) and one global gap-continuation probability g
 * (I->I = D->D = g, I->M = D->M = 1 - g), the continuous form of the Phred tables
 * (crates/pairhmm/src/model.rs:66-84).  Regions as fgkl_batch_create; pairs land region after
 * region, [r * H + h] inside each.
 *
 * Host arrays: e, d_ins, d_del and grad_err/ins/del hold one double per read base
 * (fgkl_grad_read_bases), weights and log10_likelihoods one per pair (fgkl_grad_pairs); g and
 * *grad_gcp are single values.  Each call copies its inputs to the device, runs, copies the
 * outputs back and returns when they are written; kernel_ms (NULL to skip) receives the device
 * time of the kernels (HIP events).  fgkl_grad_backward writes d/dx of sum_p w_p log10 L_p
 * (w_p = 1 when weights is NULL) and log10 L when log10_likelihoods is not NULL.  A pair with
 * w_p = 0, or whose FP64 likelihood underflows to 0, contributes nothing.  Calls on one handle
 * must not overlap (they share its device buffers).
 *
 * fgkl_grad_set_precision selects the arithmetic of later calls: 0 (the default) runs the
 * forward, the checkpoints and the adjoint in FP64; 1 runs them in FP32 (likelihoods scaled by
 * 2^120), adds each block's gradient terms into FP64 accumulators, and reruns in FP64 every pair
 * whose FP32 likelihood falls below 1e-28 (the threshold the reference's FP32 kernel rescues
 * at, crates/pairhmm/src/kernel.rs:897-914).  Other values return FGKL_INVALID_ARGUMENT.
 */
typedef struct fgkl_grad fgkl_grad;
int fgkl_grad_create(fgkl_grad** out, int64_t n_regions, const int64_t* region_reads,
                     const int64_t* region_haps, const int64_t* read_offsets,
                     const uint8_t* bases, const int64_t* hap_offsets, const uint8_t* hap_bases);
int fgkl_grad_set_precision(fgkl_grad* g, int fp32);
int64_t fgkl_grad_pairs(const fgkl_grad* g);
int64_t fgkl_grad_read_bases(const fgkl_grad* g);
int fgkl_grad_forward(fgkl_grad* g, const double* err, const double* ins, const double* del,
                      double gcp, double* log10_likelihoods, double* kernel_ms);
int fgkl_grad_backward(fgkl_grad* g, const double* err, const double* ins, const double* del,
                       double gcp, const double* weights, double* log10_likelihoods,
                       double* grad_err, double* grad_ins, double* grad_del, double* grad_gcp,
                       double* kernel_ms);
void fgkl_grad_free(fgkl_grad* g);

/*
 * Smith-Waterman: GATK's SmithWatermanJavaAligner (fgkl's crates/smithwaterman, restated from
 * reference.rs:10-203) on the device.  Weights in GATK's SWParameters order {match, mismatch,
 * gap open, gap extend} (lib.rs:41-64); overhang strategies in FgklSmithWaterman's codes
 * (FgklSmithWaterman.java:49-57).  CIGARs are returned as BAM encodes them,
 * (length << 4) | op with op 0 M, 1 I, 2 D, 4 S, or as text ("18M2D20M").
 */
#define FGKL_SW_SOFTCLIP 0
#define FGKL_SW_INDEL 1
#define FGKL_SW_LEADING_INDEL 2
#define FGKL_SW_IGNORE 3

/* Replaces Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_backendNative
 * (crates/jni/src/lib.rs:256-268): "hip-gfx950". */
const char* fgkl_sw_backend_name(void);

/*
 * Replaces Java_com_fulcrumgenomics_fgkl_smithwaterman_FgklSmithWaterman_alignNative
 * (crates/jni/src/lib.rs:271-313; Java declaration FgklSmithWaterman.java:66) and
 * Aligner::align (crates/smithwaterman/src/lib.rs:330-360): one pair.  Writes the reference
 * offset and the CIGAR as NUL-terminated text into cigar[0 .. cigar_cap).  An unknown strategy
 * code or an empty sequence is FGKL_INVALID_ARGUMENT with the reference's messages
 * (lib.rs:283-291, smithwaterman lib.rs:153); so is a CIGAR buffer too small for the text.
 * The pair runs as a batch of one on the calling thread's reusable batch (one per thread, as
 * the reference keeps one aligner per thread, crates/jni/src/lib.rs:23-26).
 */
int fgkl_sw_align_native(const uint8_t* reference, int64_t reference_len,
                         const uint8_t* alternate, int64_t alternate_len, int32_t match_value,
                         int32_t mismatch_penalty, int32_t gap_open, int32_t gap_extend,
                         int32_t strategy, int32_t* offset, char* cigar, int64_t cigar_cap);

/*
 * Batched aligner (no reference counterpart: the reference aligns one pair per call on the
 * calling thread).  Pair p aligns alternate bytes alt[alt_offsets[p] .. alt_offsets[p+1]) to
 * ref[ref_offsets[p] .. ref_offsets[p+1]); both tables have n_pairs + 1 entries starting at 0.
 * The pairs stay resident on the device; fgkl_sw_batch_run aligns all of them with one set of
 * weights and one strategy and may be called again with others (with stats it returns when the
 * alignments are done; with stats NULL it may return earlier, and fgkl_sw_batch_results waits for
 * them).  fgkl_sw_batch_results writes
 * offsets[p] and pair p's CIGAR elements to cigar[cigar_starts[p] .. cigar_starts[p+1])
 * (cigar_starts has n_pairs + 1 entries; cigar_cap >= fgkl_sw_batch_cigar_bound suffices).
 */
typedef struct fgkl_sw_batch fgkl_sw_batch;
typedef struct fgkl_sw_stats {
  double total_ms;           /* device time of the run, HIP events on the batch's stream */
  double bucket_ms[8];       /* launch per rows-per-lane bucket R = 1..8 */
  int64_t bucket_pairs[8];
  int64_t bucket_cells[8];   /* reference x alternate bases of each bucket */
  int64_t substring_pairs;   /* pairs resolved by the exact-substring shortcut */
} fgkl_sw_stats;
int fgkl_sw_batch_create(fgkl_sw_batch** out, int64_t n_pairs, const int64_t* ref_offsets,
                         const uint8_t* ref, const int64_t* alt_offsets, const uint8_t* alt);
int fgkl_sw_batch_run(fgkl_sw_batch* batch, const int32_t params[4], int strategy,
                      fgkl_sw_stats* stats);
int fgkl_sw_batch_results(fgkl_sw_batch* batch, int32_t* offsets, int64_t* cigar_starts,
                          uint32_t* cigar, int64_t cigar_cap);
int64_t fgkl_sw_batch_pairs(const fgkl_sw_batch* batch);
int64_t fgkl_sw_batch_cells(const fgkl_sw_batch* batch);
int64_t fgkl_sw_batch_cigar_bound(const fgkl_sw_batch* batch);
void fgkl_sw_batch_free(fgkl_sw_batch* batch);

#ifdef __cplusplus
}
#endif
#endif /* FGKL_H */
