/*
 * CPU oracle for the PairHMM / PD-HMM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain-C restatement of the
 * reference's scalar double-precision ports and is used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * It is never linked into, loaded by, or called from the product library
 * (fgkl_amd/), which must fail loudly when its HIP code is unavailable.
 *
 * Every function cites the reference file:line it follows.
 */
#ifndef FGKL_ORACLE_H
#define FGKL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* model.rs:38, model.rs:53-55 — 10^(-q/10), q clamped to 254. */
double fgklo_error_prob(uint8_t qual);
/* model.rs:59-64 — quantized match-to-match probability, symmetric. */
double fgklo_match_to_match(uint8_t ins_qual, uint8_t del_qual);
/* model.rs:67-77 — the six transitions [mm, im, mi, ii, md, dd]. */
void fgklo_transitions(uint8_t ins_qual, uint8_t del_qual, uint8_t gcp, double out[6]);
/* model.rs:80-83 — (1 - e, e / 3). */
void fgklo_priors(uint8_t qual, double out[2]);
/* model.rs:87-96 — GATK's approximateLog10SumLog10. */
double fgklo_approx_log10_sum_log10(double a, double b);
/* reference.rs:12 — 2^1020 and reference.rs:15-17 — its log10. */
double fgklo_initial_condition(void);
double fgklo_initial_condition_log10(void);

/* reference.rs:20-64 — scalar FP64 forward algorithm (GATK LoglessPairHMM). */
double fgklo_log10_likelihood(const uint8_t* hap, int64_t m,
                              const uint8_t* bases, const uint8_t* quals,
                              const uint8_t* ins_gop, const uint8_t* del_gop,
                              const uint8_t* gcp, int64_t n);

/* pd.rs:99-110 — read base matches a SNP flag's alternate set. */
int fgklo_base_matches_pd(uint8_t read_base, uint8_t flags);

/* pd_reference.rs:29-118 — scalar FP64 PD forward (GATK LoglessPDPairHMM). */
double fgklo_pd_log10_likelihood(const uint8_t* hap, const uint8_t* flags, int64_t m,
                                 const uint8_t* bases, const uint8_t* quals,
                                 const uint8_t* ins_gop, const uint8_t* del_gop,
                                 const uint8_t* gcp, int64_t n);

/*
 * All pairs of a batch of regions, out laid out region by region as
 * out[region_out_base + r * H + h] (lib.rs:270-271 per region).  Reads of
 * region g are read_region[g]..read_region[g+1] into read_off; haplotypes
 * likewise through hap_region / hap_off.  hap_flags may be NULL (plain) or
 * point at per-base PD flags (PD).  nthreads workers each take whole regions,
 * one independent call per core (README.md:29).
 */
void fgklo_batch(int64_t n_regions,
                 const int64_t* read_region, const int64_t* read_off,
                 const uint8_t* bases, const uint8_t* quals,
                 const uint8_t* ins_gop, const uint8_t* del_gop, const uint8_t* gcp,
                 const int64_t* hap_region, const int64_t* hap_off,
                 const uint8_t* hap_bases, const uint8_t* hap_flags,
                 double* out, int nthreads);

/*
 * Gradient oracle for the calibration adjoint (no reference counterpart).
 * Forward uses the continuous parameterisation mm = 1 - d_ins - d_del of the
 * transitions (model.rs:126-130 states the table is within 1e-6 of it),
 * im = 1 - g, ii = dd = g, mi = d_ins, md = d_del, priors (1 - e, e / 3).
 * Inputs are the per-base probabilities e[n], d_ins[n], d_del[n] and the
 * global gap-continuation probability g.  Returns log10 L and writes
 * d log10 L / d e[i], / d d_ins[i], / d d_del[i] and / d g.
 * Straight reverse-mode differentiation of the full matrices, FP64.
 */
double fgklo_log10_likelihood_grad(const uint8_t* hap, int64_t m,
                                   const uint8_t* bases, int64_t n,
                                   const double* e, const double* d_ins,
                                   const double* d_del, double g,
                                   double* de, double* dins, double* ddel,
                                   double* dg);
/* The forward of fgklo_log10_likelihood_grad alone (for finite differences). */
double fgklo_log10_likelihood_cont(const uint8_t* hap, int64_t m,
                                   const uint8_t* bases, int64_t n,
                                   const double* e, const double* d_ins,
                                   const double* d_del, double g);

/*
 * The CPU baseline (port.c): fgkl's lane-batched FP32 backend restated with AVX-512 (16 lanes)
 * or AVX2 (8 lanes) vectors, one region per worker thread, FTZ/DAZ, prefix sharing and (with
 * share_suffixes, fgkl's default) suffix sharing, lost pairs recomputed in FP64.  Same inputs
 * and output layout as fgklo_batch (no PD).  Returns 0, or -1 when the CPU
 * has neither AVX-512 nor AVX2.  force_avx2 selects the 8-lane build.
 */
int fgklo_port_batch(int64_t n_regions, const int64_t* read_region, const int64_t* read_off,
                     const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                     const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_region,
                     const int64_t* hap_off, const uint8_t* hap_bases, double* out, int nthreads,
                     int force_avx2, int share_suffixes);
/*
 * Smith-Waterman (sw.c): GATK's SmithWatermanJavaAligner (crates/smithwaterman/src/reference.rs).
 * params = {match, mismatch, gap open, gap extend}; strategy 0 SOFTCLIP, 1 INDEL, 2 LEADING_INDEL,
 * 3 IGNORE.  Writes the CIGAR as BAM encodes it, (length << 4 | op) with op 0 M, 1 I, 2 D, 4 S (at
 * most nref + nalt + 2 elements), and the reference offset.  Returns 0, or -1 for an empty
 * sequence.
 */
int fgklo_sw_align(const uint8_t* ref, int64_t nref, const uint8_t* alt, int64_t nalt,
                   const int32_t params[4], int strategy, int32_t* offset, uint32_t* cigar,
                   int32_t* ncigar);

/*
 * Lane-batched SIMD Smith-Waterman (sw_port.c, the CPU baseline): n_pairs pairs as offset tables
 * into ref / alt, one weight set and strategy; pair p's CIGAR goes to cigar[cig_start[p] ..] (room
 * for n + m + 4 elements), its length to ncig[p], its offset to offsets[p].  Returns 0, -1 without
 * AVX2, -2 for an empty sequence.
 */
int fgklo_sw_port_batch(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref,
                        const int64_t* alt_off, const uint8_t* alt, const int32_t params[4],
                        int strategy, int32_t* offsets, const int64_t* cig_start, uint32_t* cigar,
                        int32_t* ncig, int nthreads, int force_avx2);
const char* fgklo_sw_port_backend(void);

/* The vector width fgklo_port_batch runs with on this CPU. */
const char* fgklo_port_backend(void);

#ifdef __cplusplus
}
#endif
#endif
