/*
 * Lane-batched SIMD restatement of fgkl's CPU PairHMM backend: the CPU baseline that bench.py
 * times on the host cores beside the GPU.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY.  Like the scalar oracle beside it, this is never
 * linked into, loaded by or called from the product library (fgkl_amd/).
 *
 * fgkl's fastest x86 backend runs 16 FP32 lanes (AVX-512; AVX2: 8), one read per lane, and
 * sweeps every haplotype of a region row by row with lexicographic prefix sharing, flushing
 * denormals to zero; pairs whose FP32 sum underflows are recomputed in FP64 (lib.rs:377-549,
 * kernel.rs, simd.rs:203-234, 317-328).  fgkl is Rust and cannot be built here, so this is a C
 * restatement of that design with GCC vector types, compiled for AVX-512 and for AVX2 and
 * picked at run time.  Simplification: lost pairs are recomputed one by one with the scalar
 * FP64 oracle instead of fgkl's dense/sparse FP64 lane batches (lib.rs:500-549); on the
 * benchmark workload they are a few dozen of 7.4 million pairs.
 */
#define _GNU_SOURCE
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <xmmintrin.h>

#include "oracle.h"

struct port_job {
    int64_t n_regions;
    const int64_t *read_region, *read_off;
    const uint8_t *bases, *quals, *ins, *del, *gcp;
    const int64_t *hap_region, *hap_off;
    const uint8_t* hap_bases;
    double* out;
    const int64_t* out_base;
    int share_suffixes;
    atomic_long next;
};

/* One backward sweep of the suffix plan (plan.rs:14-31): along sorted haplotype `hap`, depths
 * resume+1..need (depth d is column len - d), starting from node `start` (depth resume) and
 * storing node node_of[d] (>= 0) at each depth a cut or a later sweep reads. */
struct port_sweep {
    int hap, resume, need, start;
    int* node_of; /* [need + 1] */
};

/* One region's haplotypes in sorted order with their LCPs, duplicates and byte codes, its reads
 * by length, longest first (kernel.rs:53-107, lib.rs:377-384), and its suffix plan. */
struct port_plan {
    int64_t H, nr, r0;
    const uint8_t** hap;
    int32_t *hap_len, *lcp, *hap_caller, *lcp_count;
    uint8_t* dup;
    int64_t* read_order;
    int max_len, ncodes;
    uint8_t code_of[256], code_byte[256];
    /* suffix sharing: per sorted haplotype the cut column (0: none) and node it joins */
    int32_t *cut_col, *cut_node;
    struct port_sweep* sweeps;
    int nsweeps, nodes;
};

/* Per-thread buffers, grown on demand. */
struct port_ws {
    float *trans, *prior, *endm, *row_a, *row_b, *snap, *snap_acc, *nodes;
    uint8_t* end_row;
    int32_t *cnt, *slot_of, *snap_cols, *snap_len;
    double* res;
    size_t cap[13];
};

static void* grow(void* p, size_t* cap, size_t want) {
    if (want <= *cap) return p;
    free(p);
    void* q = NULL;
    if (posix_memalign(&q, 64, want ? want : 64) != 0) abort();
    *cap = want;
    return q;
}

static void port_ws_reserve(struct port_ws* w, int rows, int max_len, int ncodes, int H, int L) {
    const size_t f = sizeof(float);
    w->trans = grow(w->trans, &w->cap[0], f * (size_t)rows * 6 * L);
    w->prior = grow(w->prior, &w->cap[1], f * (size_t)rows * (size_t)ncodes * L);
    w->endm = grow(w->endm, &w->cap[2], f * (size_t)rows * L);
    w->row_a = grow(w->row_a, &w->cap[3], f * (size_t)(max_len + 1) * 3 * L);
    w->row_b = grow(w->row_b, &w->cap[4], f * (size_t)(max_len + 1) * 3 * L);
    /* at most one snapshot slot per haplotype */
    w->snap = grow(w->snap, &w->cap[5], f * (size_t)H * 3 * (size_t)(rows + 1) * L);
    w->snap_acc = grow(w->snap_acc, &w->cap[6], f * (size_t)H * L);
    w->end_row = grow(w->end_row, &w->cap[7], (size_t)rows);
    w->cnt = grow(w->cnt, &w->cap[8], sizeof(int32_t) * (size_t)(max_len + 1));
    w->slot_of = grow(w->slot_of, &w->cap[9], sizeof(int32_t) * (size_t)(max_len + 1));
    w->snap_cols = grow(w->snap_cols, &w->cap[10], sizeof(int32_t) * (size_t)(max_len + 1) * 2);
    w->snap_len = w->snap_cols + (max_len + 1);
    w->res = grow(w->res, &w->cap[11], sizeof(double) * (size_t)L * (size_t)H);
}

/* Backward nodes: per node B_M and B_Y of rows 0..rows, then the row-0 start sum S. */
static void port_ws_reserve_nodes(struct port_ws* w, int rows, int nodes, int L) {
    w->nodes = grow(w->nodes, &w->cap[12],
                    sizeof(float) * (size_t)nodes * (size_t)(2 * (rows + 1) + 1) * L);
}

static void port_ws_free(struct port_ws* w) {
    free(w->trans); free(w->prior); free(w->endm); free(w->row_a); free(w->row_b);
    free(w->snap); free(w->snap_acc); free(w->end_row); free(w->cnt); free(w->slot_of);
    free(w->snap_cols); free(w->res); free(w->nodes);
}

#define PORT_LANES 16
#define PORT_TARGET __attribute__((target("avx512f,fma")))
#define PORT_FN port_region_avx512
#include "port_body.h"
#undef PORT_LANES
#undef PORT_TARGET
#undef PORT_FN

#define PORT_LANES 8
#define PORT_TARGET __attribute__((target("avx2,fma")))
#define PORT_FN port_region_avx2
#include "port_body.h"
#undef PORT_LANES
#undef PORT_TARGET
#undef PORT_FN

static const struct port_job* g_sort_job;

static int hap_cmp(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    const struct port_job* J = g_sort_job;
    const int64_t lx = J->hap_off[x + 1] - J->hap_off[x], ly = J->hap_off[y + 1] - J->hap_off[y];
    const int c = memcmp(J->hap_bases + J->hap_off[x], J->hap_bases + J->hap_off[y],
                         (size_t)(lx < ly ? lx : ly));
    if (c) return c;
    if (lx != ly) return lx < ly ? -1 : 1;
    return x < y ? -1 : (x > y);
}

static const struct port_job* g_len_job;

static int read_cmp(const void* a, const void* b) {  /* longest first, then index */
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    const struct port_job* J = g_len_job;
    const int64_t lx = J->read_off[x + 1] - J->read_off[x], ly = J->read_off[y + 1] - J->read_off[y];
    if (lx != ly) return lx > ly ? -1 : 1;
    return x < y ? -1 : (x > y);
}

static pthread_mutex_t g_sort_mutex = PTHREAD_MUTEX_INITIALIZER;

static void plan_region(const struct port_job* J, int64_t g, struct port_plan* P) {
    const int64_t h0 = J->hap_region[g], h1 = J->hap_region[g + 1];
    const int64_t r0 = J->read_region[g], r1 = J->read_region[g + 1];
    const int64_t H = h1 - h0, nr = r1 - r0;
    P->H = H;
    P->nr = nr;
    P->r0 = r0;
    int64_t* hidx = malloc(sizeof(int64_t) * (size_t)(H ? H : 1));
    for (int64_t k = 0; k < H; ++k) hidx[k] = h0 + k;
    P->read_order = malloc(sizeof(int64_t) * (size_t)(nr ? nr : 1));
    for (int64_t r = 0; r < nr; ++r) P->read_order[r] = r0 + r;
    pthread_mutex_lock(&g_sort_mutex);  /* qsort has no context argument in C11 */
    g_sort_job = J;
    qsort(hidx, (size_t)H, sizeof(int64_t), hap_cmp);
    g_len_job = J;
    qsort(P->read_order, (size_t)nr, sizeof(int64_t), read_cmp);
    pthread_mutex_unlock(&g_sort_mutex);
    P->hap = malloc(sizeof(uint8_t*) * (size_t)(H ? H : 1));
    P->hap_len = malloc(sizeof(int32_t) * (size_t)(H ? H : 1) * 3);
    P->lcp = P->hap_len + H;
    P->hap_caller = P->lcp + H;
    P->dup = malloc((size_t)(H ? H : 1));
    P->max_len = 0;
    memset(P->code_of, 0, sizeof(P->code_of));
    int seen[256] = {0};
    P->ncodes = 0;
    for (int64_t k = 0; k < H; ++k) {
        const int64_t h = hidx[k];
        P->hap[k] = J->hap_bases + J->hap_off[h];
        P->hap_len[k] = (int32_t)(J->hap_off[h + 1] - J->hap_off[h]);
        P->hap_caller[k] = (int32_t)(h - h0);
        if (P->hap_len[k] > P->max_len) P->max_len = P->hap_len[k];
        int32_t l = 0;
        if (k) {
            const int32_t n = P->hap_len[k] < P->hap_len[k - 1] ? P->hap_len[k] : P->hap_len[k - 1];
            while (l < n && P->hap[k][l] == P->hap[k - 1][l]) ++l;
        }
        P->lcp[k] = l;
        P->dup[k] = k && l == P->hap_len[k] && l == P->hap_len[k - 1];
        for (int32_t j = 0; j < P->hap_len[k]; ++j) {
            const uint8_t b = P->hap[k][j];
            if (!seen[b]) {
                seen[b] = 1;
                P->code_of[b] = (uint8_t)P->ncodes;
                P->code_byte[P->ncodes++] = b;
            }
        }
    }
    P->lcp_count = calloc((size_t)P->max_len + 1, sizeof(int32_t));
    for (int64_t k = 1; k < H; ++k)
        if (!P->dup[k]) P->lcp_count[P->lcp[k]] += 1;
    free(hidx);
    P->sweeps = NULL;
    P->nsweeps = 0;
    P->nodes = 1;
    P->cut_col = NULL;
}

static void plan_free(struct port_plan* P) {
    free(P->hap); free(P->hap_len); free(P->dup); free(P->read_order); free(P->lcp_count);
    free(P->cut_col);
    for (int s = 0; s < P->nsweeps; ++s) free(P->sweeps[s].node_of);
    free(P->sweeps);
}

static int common_suffix(const uint8_t* a, int na, const uint8_t* b, int nb) {
    int n = 0;
    while (n < na && n < nb && a[na - 1 - n] == b[nb - 1 - n]) ++n;
    return n;
}

static const struct port_plan* g_rev_plan;

static int rev_cmp(const void* x, const void* y) {  /* reversed bases, then sorted index */
    const int a = *(const int*)x, b = *(const int*)y;
    const struct port_plan* P = g_rev_plan;
    const uint8_t *ha = P->hap[a], *hb = P->hap[b];
    const int na = P->hap_len[a], nb = P->hap_len[b];
    for (int i = 0; i < na && i < nb; ++i) {
        const uint8_t ca = ha[na - 1 - i], cb = hb[nb - 1 - i];
        if (ca != cb) return ca < cb ? -1 : 1;
    }
    if (na != nb) return na < nb ? -1 : 1;
    return a < b ? -1 : (a > b);
}

/* The node holding depth `depth` of the suffix at position r of the reversed order: the one the
 * latest sweep at or before r wrote for a suffix r shares that deep (plan.rs:205-230). */
static int node_for(struct port_plan* P, const int* sweep_at, const int* lcs, int r, int depth) {
    if (depth == 0) return 0;
    int shared = 1 << 30;
    for (int p = r; p >= 0; --p) {
        const int s = sweep_at[p];
        if (s >= 0 && P->sweeps[s].resume < depth && depth <= P->sweeps[s].need) {
            int* id = &P->sweeps[s].node_of[depth];
            if (*id < 0) *id = P->nodes++;
            return *id;
        }
        if (lcs[p] < shared) shared = lcs[p];
        if (p == 0 || shared < depth) break;
    }
    abort();  /* the plan never reads a depth no sweep computes */
}

/* Suffix sharing (plan.rs:55-104, 146-203): cut each distinct haplotype at depth
 * min(longest shared suffix, len - longest shared prefix - 1) when that saves at least 4
 * columns, and plan the backward sweeps in reversed-bases order, each resuming from the deepest
 * depth already computed along its suffix. */
static void plan_suffixes(struct port_plan* P) {
    const int H = (int)P->H;
    P->cut_col = calloc((size_t)(2 * H + 1), sizeof(int32_t));
    P->cut_node = P->cut_col + H;
    P->sweeps = calloc((size_t)(H + 1), sizeof(struct port_sweep));
    P->nsweeps = 0;
    P->nodes = 1;
    int* distinct = malloc(sizeof(int) * (size_t)(H + 1) * 6);
    int* order = distinct + H + 1;
    int* lcs = order + H + 1;
    int* pos = lcs + H + 1;
    int* depth = pos + H + 1;
    int* sweep_at = depth + H + 1;
    int nd = 0;
    for (int k = 0; k < H; ++k)
        if (!P->dup[k]) distinct[nd++] = k;
    memcpy(order, distinct, sizeof(int) * (size_t)nd);
    pthread_mutex_lock(&g_sort_mutex);
    g_rev_plan = P;
    qsort(order, (size_t)nd, sizeof(int), rev_cmp);
    pthread_mutex_unlock(&g_sort_mutex);
    for (int r = 0; r < nd; ++r) {
        lcs[r] = r ? common_suffix(P->hap[order[r - 1]], P->hap_len[order[r - 1]],
                                   P->hap[order[r]], P->hap_len[order[r]]) : 0;
        pos[order[r]] = r;
        sweep_at[r] = -1;
    }
    for (int t = 0; t < nd; ++t) {
        const int k = distinct[t];
        depth[k] = 0;
        const int next_lcp = t + 1 < nd ? P->lcp[distinct[t + 1]] : 0;
        const int shared_prefix = P->lcp[k] > next_lcp ? P->lcp[k] : next_lcp;
        const int len = P->hap_len[k];
        if (shared_prefix < len) {
            const int r = pos[k];
            int s = lcs[r];
            if (r + 1 < nd && lcs[r + 1] > s) s = lcs[r + 1];
            const int d = s < len - shared_prefix - 1 ? s : len - shared_prefix - 1;
            if (d >= 4) depth[k] = d;
        }
    }
    int reach = 0;
    for (int r = 0; r < nd; ++r) {
        const int k = order[r];
        const int resume = lcs[r] < reach ? lcs[r] : reach;
        if (depth[k] > resume) {
            struct port_sweep* sw = &P->sweeps[P->nsweeps];
            sw->hap = k;
            sw->resume = resume;
            sw->need = depth[k];
            sw->node_of = malloc(sizeof(int) * (size_t)(depth[k] + 1));
            for (int d = 0; d <= depth[k]; ++d) sw->node_of[d] = -1;
            sweep_at[r] = P->nsweeps++;
        }
        reach = resume > depth[k] ? resume : depth[k];
    }
    for (int s = 0; s < P->nsweeps; ++s)
        P->sweeps[s].start = node_for(P, sweep_at, lcs, pos[P->sweeps[s].hap], P->sweeps[s].resume);
    for (int t = 0; t < nd; ++t) {
        const int k = distinct[t];
        if (depth[k] > 0) {
            P->cut_col[k] = P->hap_len[k] - depth[k];
            P->cut_node[k] = node_for(P, sweep_at, lcs, pos[k], depth[k]);
        }
    }
    free(distinct);
}

static int g_use_avx512 = -1;

static void* port_worker(void* arg) {
    struct port_job* J = arg;
    _mm_setcsr(_mm_getcsr() | 0x8040);  /* FTZ + DAZ, as fgkl on x86 (simd.rs:203-234) */
    struct port_ws ws;
    memset(&ws, 0, sizeof(ws));
    for (;;) {
        const int64_t g = atomic_fetch_add(&J->next, 1);
        if (g >= J->n_regions) break;
        struct port_plan P;
        plan_region(J, g, &P);
        if (J->share_suffixes) plan_suffixes(&P);
        if (P.H > 0 && P.nr > 0) {
            if (g_use_avx512)
                port_region_avx512(J, &P, &ws, J->out + J->out_base[g]);
            else
                port_region_avx2(J, &P, &ws, J->out + J->out_base[g]);
        }
        plan_free(&P);
    }
    port_ws_free(&ws);
    return NULL;
}

const char* fgklo_port_backend(void) {
    if (g_use_avx512 < 0) {
        __builtin_cpu_init();
        g_use_avx512 = __builtin_cpu_supports("avx512f") ? 1 : 0;
    }
    return g_use_avx512 ? "avx512 (16 f32 lanes)" : "avx2 (8 f32 lanes)";
}

int fgklo_port_batch(int64_t n_regions, const int64_t* read_region, const int64_t* read_off,
                     const uint8_t* bases, const uint8_t* quals, const uint8_t* ins_gop,
                     const uint8_t* del_gop, const uint8_t* gcp, const int64_t* hap_region,
                     const int64_t* hap_off, const uint8_t* hap_bases, double* out, int nthreads,
                     int force_avx2, int share_suffixes) {
    fgklo_port_backend();
    if (force_avx2) g_use_avx512 = 0;
    if (!g_use_avx512 && !__builtin_cpu_supports("avx2")) return -1;
    int64_t* base = malloc(sizeof(int64_t) * (size_t)(n_regions + 1));
    base[0] = 0;
    for (int64_t g = 0; g < n_regions; ++g)
        base[g + 1] = base[g] + (read_region[g + 1] - read_region[g]) *
                                    (hap_region[g + 1] - hap_region[g]);
    struct port_job J = {n_regions, read_region, read_off, bases, quals, ins_gop, del_gop, gcp,
                         hap_region, hap_off, hap_bases, out, base, share_suffixes != 0, 0};
    atomic_init(&J.next, 0);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, port_worker, &J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(base);
    if (force_avx2) g_use_avx512 = -1;
    return 0;
}
