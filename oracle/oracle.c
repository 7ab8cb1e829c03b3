/*
 * CPU oracle for the PairHMM / PD-HMM hot path: a plain-C restatement of the
 * reference's scalar FP64 ports.
 *
 * TEST INFRASTRUCTURE ONLY — used by tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg as the checker; never by the product library.
 *
 * Built with -O2 -ffp-contract=off so every multiply and add rounds exactly
 * where the reference's Rust/Java ports round (no FMA contraction), and the
 * last-row sum runs left to right as reference.rs:58-62 does.
 *
 * Pinning (see tests/test_oracle.py): the model.rs:108-155 and
 * reference.rs:66-115 / pd_reference.rs:120-184 known-answer tests, an
 * independent exact path-sum enumerator on tiny inputs, closed forms for
 * single-column / single-row cases, PD-without-flags == plain, and the
 * quantisation bound of the Jacobian table against the exact log-sum.
 */
#include "oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>

/* model.rs:10-26 */
#define MAX_QUAL 254
#define TRISTATE_CORRECTION 3.0
#define JACOBIAN_MAX_TOLERANCE 8.0
#define JACOBIAN_TABLE_STEP 0.0001
static const double JACOBIAN_INV_STEP = 1.0 / JACOBIAN_TABLE_STEP;
static const double INV_LN10 = 1.0 / 2.302585092994045684017991454684364208; /* 1 / ln 10 */

/* pd.rs:15-27 */
#define PD_SNP 1
#define PD_DEL_START 2
#define PD_DEL_END 4
#define PD_ALT_A 8
#define PD_ALT_C 16
#define PD_ALT_G 32
#define PD_ALT_T 64

static double g_error_prob[MAX_QUAL + 1];
static double g_match_to_match[((MAX_QUAL + 1) * (MAX_QUAL + 2)) / 2];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;

/* model.rs:104-106 */
static int64_t fast_round(double d) {
    return d > 0.0 ? (int64_t)(d + 0.5) : (int64_t)(d - 0.5);
}

/* model.rs:99-102 */
static double jacobian(double diff) {
    int64_t k = fast_round(diff * JACOBIAN_INV_STEP);
    return log10(1.0 + pow(10.0, -(double)k * JACOBIAN_TABLE_STEP));
}

/* model.rs:87-96 */
double fgklo_approx_log10_sum_log10(double a, double b) {
    if (a > b) return fgklo_approx_log10_sum_log10(b, a);
    if (a == -INFINITY) return b;
    double diff = b - a;
    return diff < JACOBIAN_MAX_TOLERANCE ? b + jacobian(diff) : b;
}

/* model.rs:37-50 */
static void build_tables(void) {
    for (int q = 0; q <= MAX_QUAL; ++q) g_error_prob[q] = pow(10.0, (double)q / -10.0);
    int offset = 0;
    for (int i = 0; i <= MAX_QUAL; ++i) {
        for (int j = 0; j <= i; ++j) {
            double log10_sum = fgklo_approx_log10_sum_log10(-0.1 * (double)i, -0.1 * (double)j);
            double p = pow(10.0, log10_sum);
            if (p > 1.0) p = 1.0;
            double m2m_log10 = log1p(-p) * INV_LN10;
            g_match_to_match[offset + j] = pow(10.0, m2m_log10);
        }
        offset += i + 1;
    }
}

static void tables(void) { pthread_once(&g_once, build_tables); }

/* model.rs:53-55 */
double fgklo_error_prob(uint8_t qual) {
    tables();
    return g_error_prob[qual > MAX_QUAL ? MAX_QUAL : qual];
}

/* model.rs:59-64 */
double fgklo_match_to_match(uint8_t ins_qual, uint8_t del_qual) {
    tables();
    int a = ins_qual > MAX_QUAL ? MAX_QUAL : ins_qual;
    int b = del_qual > MAX_QUAL ? MAX_QUAL : del_qual;
    int lo = a <= b ? a : b, hi = a <= b ? b : a;
    return g_match_to_match[((hi * (hi + 1)) >> 1) + lo];
}

/* model.rs:67-77; index order model.rs:15-21 */
void fgklo_transitions(uint8_t ins_qual, uint8_t del_qual, uint8_t gcp, double t[6]) {
    double gap_continue = fgklo_error_prob(gcp);
    t[0] = fgklo_match_to_match(ins_qual, del_qual); /* MATCH_TO_MATCH */
    t[2] = fgklo_error_prob(ins_qual);               /* MATCH_TO_INSERTION */
    t[4] = fgklo_error_prob(del_qual);               /* MATCH_TO_DELETION */
    t[1] = 1.0 - gap_continue;                       /* INDEL_TO_MATCH */
    t[3] = gap_continue;                             /* INSERTION_TO_INSERTION */
    t[5] = gap_continue;                             /* DELETION_TO_DELETION */
}

/* model.rs:80-83 */
void fgklo_priors(uint8_t qual, double p[2]) {
    double e = fgklo_error_prob(qual);
    p[0] = 1.0 - e;
    p[1] = e / TRISTATE_CORRECTION;
}

/* reference.rs:12 */
double fgklo_initial_condition(void) { return ldexp(1.0, 1020); }
/* reference.rs:15-17 */
double fgklo_initial_condition_log10(void) { return log10(ldexp(1.0, 1020)); }

/* reference.rs:20-64 */
double fgklo_log10_likelihood(const uint8_t* hap, int64_t m,
                              const uint8_t* bases, const uint8_t* quals,
                              const uint8_t* ins_gop, const uint8_t* del_gop,
                              const uint8_t* gcp, int64_t n) {
    const int64_t rows = n + 1, cols = m + 1;
    double* mm = calloc((size_t)(rows * cols), sizeof(double));
    double* xx = calloc((size_t)(rows * cols), sizeof(double));
    double* yy = calloc((size_t)(rows * cols), sizeof(double));
    const double initial = fgklo_initial_condition() / (double)m;
    for (int64_t j = 0; j < cols; ++j) yy[j] = initial;
    for (int64_t i = 1; i < rows; ++i) {
        double t[6], p[2];
        fgklo_transitions(ins_gop[i - 1], del_gop[i - 1], gcp[i - 1], t);
        fgklo_priors(quals[i - 1], p);
        const uint8_t rb = bases[i - 1];
        for (int64_t j = 1; j < cols; ++j) {
            const uint8_t hb = hap[j - 1];
            const double prior = (rb == hb || rb == 'N' || hb == 'N') ? p[0] : p[1];
            const int64_t diag = (i - 1) * cols + (j - 1), up = (i - 1) * cols + j;
            const int64_t left = i * cols + (j - 1), cur = i * cols + j;
            mm[cur] = prior * (mm[diag] * t[0] + xx[diag] * t[1] + yy[diag] * t[1]);
            xx[cur] = mm[up] * t[2] + xx[up] * t[3];
            yy[cur] = mm[left] * t[4] + yy[left] * t[5];
        }
    }
    const int64_t last = (rows - 1) * cols;
    double sum = 0.0;
    for (int64_t j = 1; j < cols; ++j) sum += mm[last + j] + xx[last + j];
    free(mm);
    free(xx);
    free(yy);
    return log10(sum) - fgklo_initial_condition_log10();
}

/* pd.rs:99-110 */
int fgklo_base_matches_pd(uint8_t rb, uint8_t flags) {
    if (!(flags & PD_SNP)) return 0;
    switch (rb) {
        case 'A': case 'a': return (flags & PD_ALT_A) != 0;
        case 'C': case 'c': return (flags & PD_ALT_C) != 0;
        case 'G': case 'g': return (flags & PD_ALT_G) != 0;
        case 'T': case 't': return (flags & PD_ALT_T) != 0;
        default: return 0;
    }
}

static double dmax(double a, double b) { return a >= b ? a : b; }

/* pd_reference.rs:29-118 */
double fgklo_pd_log10_likelihood(const uint8_t* hap, const uint8_t* flags, int64_t m,
                                 const uint8_t* bases, const uint8_t* quals,
                                 const uint8_t* ins_gop, const uint8_t* del_gop,
                                 const uint8_t* gcp, int64_t n) {
    enum { NORMAL, INSIDE_DEL, AFTER_DEL };
    const int64_t rows = n + 1, cols = m + 1;
    const size_t cells = (size_t)(rows * cols);
    double* M = calloc(cells, sizeof(double));
    double* X = calloc(cells, sizeof(double));
    double* Y = calloc(cells, sizeof(double));
    double* BM = calloc(cells, sizeof(double));
    double* BX = calloc(cells, sizeof(double));
    double* BY = calloc(cells, sizeof(double));
    const double initial = fgklo_initial_condition() / (double)m;
    for (int64_t j = 0; j < cols; ++j) Y[j] = initial;
    int state = NORMAL; /* carried from the end of one row into the next (pd_reference.rs:46) */
    for (int64_t i = 1; i < rows; ++i) {
#ifdef FGKLO_PLANT_PD_ROW_RESET
        state = NORMAL; /* a planted mistake, built only by tests/test_oracle.py to show its KAT fails */
#endif
        double t[6], p[2];
        fgklo_transitions(ins_gop[i - 1], del_gop[i - 1], gcp[i - 1], t);
        fgklo_priors(quals[i - 1], p);
        const uint8_t rb = bases[i - 1];
        for (int64_t j = 1; j < cols; ++j) {
            const uint8_t hb = hap[j - 1], fl = flags[j - 1];
            const double prior =
                (rb == hb || rb == 'N' || hb == 'N' || fgklo_base_matches_pd(rb, fl)) ? p[0] : p[1];
            const int64_t diag = (i - 1) * cols + (j - 1), up = (i - 1) * cols + j;
            const int64_t left = i * cols + (j - 1), cur = i * cols + j;
            const int del_end = (fl & PD_DEL_END) != 0;
            if (state == NORMAL) {
                BM[cur] = M[left];
                BY[cur] = Y[left];
                BX[cur] = X[left];
                M[cur] = prior * (M[diag] * t[0] + X[diag] * t[1] + Y[diag] * t[1]);
                Y[cur] = M[left] * t[4] + Y[left] * t[5];
            } else if (state == INSIDE_DEL) {
                BM[cur] = BM[left];
                BY[cur] = BY[left];
                BX[cur] = BX[left];
                M[cur] = prior * (M[diag] * t[0] + X[diag] * t[1] + Y[diag] * t[1]);
                Y[cur] = M[left] * t[4] + Y[left] * t[5];
            } else {
                BM[cur] = dmax(BM[left], M[left]);
                BY[cur] = dmax(BY[left], Y[left]);
                BX[cur] = dmax(BX[left], X[left]);
                M[cur] = prior * (dmax(BM[diag], M[diag]) * t[0] + dmax(BX[diag], X[diag]) * t[1] +
                                  dmax(BY[diag], Y[diag]) * t[1]);
                Y[cur] = dmax(BM[left], M[left]) * t[4] + dmax(BY[left], Y[left]) * t[5];
            }
            X[cur] = del_end ? dmax(BM[up], M[up]) * t[2] + dmax(BX[up], X[up]) * t[3]
                             : M[up] * t[2] + X[up] * t[3];
            if (state == AFTER_DEL) state = NORMAL;
            if (fl & PD_DEL_START) state = INSIDE_DEL;
            if (del_end) state = AFTER_DEL;
        }
    }
    const int64_t last = (rows - 1) * cols;
    double sum = 0.0;
    for (int64_t j = 1; j < cols; ++j) sum += M[last + j] + X[last + j];
    free(M); free(X); free(Y); free(BM); free(BX); free(BY);
    return log10(sum) - fgklo_initial_condition_log10();
}

/* ---- batch driver: whole regions per worker (README.md:29) ---- */
struct batch_job {
    int64_t n_regions;
    const int64_t *read_region, *read_off;
    const uint8_t *bases, *quals, *ins, *del, *gcp;
    const int64_t *hap_region, *hap_off;
    const uint8_t *hap_bases, *hap_flags;
    double* out;
    int64_t* out_base;
    atomic_long next;
};

static void* batch_worker(void* arg) {
    struct batch_job* job = arg;
    for (;;) {
        int64_t g = atomic_fetch_add(&job->next, 1);
        if (g >= job->n_regions) break;
        const int64_t r0 = job->read_region[g], r1 = job->read_region[g + 1];
        const int64_t h0 = job->hap_region[g], h1 = job->hap_region[g + 1];
        const int64_t H = h1 - h0;
        double* out = job->out + job->out_base[g];
        for (int64_t r = r0; r < r1; ++r) {
            const int64_t o = job->read_off[r], n = job->read_off[r + 1] - o;
            for (int64_t h = h0; h < h1; ++h) {
                const int64_t ho = job->hap_off[h], m = job->hap_off[h + 1] - ho;
                double v;
                if (job->hap_flags)
                    v = fgklo_pd_log10_likelihood(job->hap_bases + ho, job->hap_flags + ho, m,
                                                  job->bases + o, job->quals + o, job->ins + o,
                                                  job->del + o, job->gcp + o, n);
                else
                    v = fgklo_log10_likelihood(job->hap_bases + ho, m, job->bases + o,
                                               job->quals + o, job->ins + o, job->del + o,
                                               job->gcp + o, n);
                out[(r - r0) * H + (h - h0)] = v;
            }
        }
    }
    return NULL;
}

void fgklo_batch(int64_t n_regions, const int64_t* read_region, const int64_t* read_off,
                 const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                 const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_region,
                 const int64_t* hap_off, const uint8_t* hap_bases, const uint8_t* hap_flags,
                 double* out, int nthreads) {
    tables();
    struct batch_job job = {n_regions, read_region, read_off, bases, quals, ins_gop, del_gop,
                            gcp, hap_region, hap_off, hap_bases, hap_flags, out, NULL, 0};
    job.out_base = malloc(sizeof(int64_t) * (size_t)(n_regions + 1));
    job.out_base[0] = 0;
    for (int64_t g = 0; g < n_regions; ++g)
        job.out_base[g + 1] = job.out_base[g] + (read_region[g + 1] - read_region[g]) *
                                                    (hap_region[g + 1] - hap_region[g]);
    atomic_init(&job.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, batch_worker, &job);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(job.out_base);
}

/* ---- gradient oracle (continuous parameterisation; no reference counterpart) ---- */
static double cont_forward(const uint8_t* hap, int64_t m, const uint8_t* bases, int64_t n,
                           const double* e, const double* di, const double* dd, double g,
                           double* M, double* X, double* Y) {
    const int64_t cols = m + 1;
    const double initial = fgklo_initial_condition() / (double)m;
    for (int64_t j = 0; j < cols; ++j) { M[j] = 0; X[j] = 0; Y[j] = initial; }
    for (int64_t i = 1; i <= n; ++i) {
        const double mm = 1.0 - di[i - 1] - dd[i - 1], im = 1.0 - g;
        const double pm = 1.0 - e[i - 1], px = e[i - 1] / 3.0;
        const uint8_t rb = bases[i - 1];
        M[i * cols] = X[i * cols] = Y[i * cols] = 0.0;
        for (int64_t j = 1; j < cols; ++j) {
            const uint8_t hb = hap[j - 1];
            const double prior = (rb == hb || rb == 'N' || hb == 'N') ? pm : px;
            const int64_t diag = (i - 1) * cols + (j - 1), up = (i - 1) * cols + j;
            const int64_t left = i * cols + (j - 1), cur = i * cols + j;
            M[cur] = prior * (M[diag] * mm + X[diag] * im + Y[diag] * im);
            X[cur] = M[up] * di[i - 1] + X[up] * g;
            Y[cur] = M[left] * dd[i - 1] + Y[left] * g;
        }
    }
    double sum = 0.0;
    for (int64_t j = 1; j < cols; ++j) sum += M[n * cols + j] + X[n * cols + j];
    return sum;
}

double fgklo_log10_likelihood_cont(const uint8_t* hap, int64_t m, const uint8_t* bases, int64_t n,
                                   const double* e, const double* di, const double* dd,
                                   double g) {
    const size_t cells = (size_t)((n + 1) * (m + 1));
    double* M = malloc(cells * sizeof(double));
    double* X = malloc(cells * sizeof(double));
    double* Y = malloc(cells * sizeof(double));
    double s = cont_forward(hap, m, bases, n, e, di, dd, g, M, X, Y);
    free(M); free(X); free(Y);
    return log10(s) - fgklo_initial_condition_log10();
}

double fgklo_log10_likelihood_grad(const uint8_t* hap, int64_t m, const uint8_t* bases,
                                   int64_t n, const double* e, const double* di,
                                   const double* dd, double g, double* de, double* ddi,
                                   double* ddd, double* dg) {
    const int64_t cols = m + 1;
    const size_t cells = (size_t)((n + 1) * cols);
    double* M = malloc(cells * sizeof(double));
    double* X = malloc(cells * sizeof(double));
    double* Y = malloc(cells * sizeof(double));
    double* AM = calloc(cells, sizeof(double));
    double* AX = calloc(cells, sizeof(double));
    double* AY = calloc(cells, sizeof(double));
    const double s = cont_forward(hap, m, bases, n, e, di, dd, g, M, X, Y);
    /* Seed with 1/S so every adjoint is d ln S / d cell (no overflow at 2^1020 scale). */
    const double seed = 1.0 / s;
    for (int64_t j = 1; j < cols; ++j) { AM[n * cols + j] = seed; AX[n * cols + j] = seed; }
    double g_acc = 0.0, im_acc = 0.0;
    for (int64_t i = n; i >= 1; --i) {
        const double mm = 1.0 - di[i - 1] - dd[i - 1], im = 1.0 - g;
        const double pm = 1.0 - e[i - 1], px = e[i - 1] / 3.0;
        const uint8_t rb = bases[i - 1];
        double e_acc = 0.0, mm_acc = 0.0, mi_acc = 0.0, md_acc = 0.0;
        for (int64_t j = m; j >= 1; --j) {
            const uint8_t hb = hap[j - 1];
            const int match = (rb == hb || rb == 'N' || hb == 'N');
            const double prior = match ? pm : px;
            const int64_t diag = (i - 1) * cols + (j - 1), up = (i - 1) * cols + j;
            const int64_t left = i * cols + (j - 1), cur = i * cols + j;
            const double ay = AY[cur], ax = AX[cur], am = AM[cur];
            /* Y = M_left * md + Y_left * g */
            AM[left] += ay * dd[i - 1];
            AY[left] += ay * g;
            md_acc += ay * M[left];
            g_acc += ay * Y[left];
            /* X = M_up * mi + X_up * g */
            AM[up] += ax * di[i - 1];
            AX[up] += ax * g;
            mi_acc += ax * M[up];
            g_acc += ax * X[up];
            /* M = prior * (M_d * mm + X_d * im + Y_d * im) */
            const double tsum = M[diag] * mm + X[diag] * im + Y[diag] * im;
            const double dprior = am * tsum;
            e_acc += match ? -dprior : dprior / 3.0;
            const double ap = am * prior;
            AM[diag] += ap * mm;
            AX[diag] += ap * im;
            AY[diag] += ap * im;
            mm_acc += ap * M[diag];
            im_acc += ap * (X[diag] + Y[diag]);
        }
        /* mm = 1 - d_ins - d_del, mi = d_ins, md = d_del */
        de[i - 1] = e_acc * INV_LN10;
        ddi[i - 1] = (mi_acc - mm_acc) * INV_LN10;
        ddd[i - 1] = (md_acc - mm_acc) * INV_LN10;
    }
    /* im = 1 - g */
    *dg = (g_acc - im_acc) * INV_LN10;
    free(M); free(X); free(Y); free(AM); free(AX); free(AY);
    return log10(s) - fgklo_initial_condition_log10();
}
