/*
 * Lane-batched SIMD restatement of fgkl's CPU PairHMM backend: the CPU baseline that bench.py
 * times on the GPU box's host cores.
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
    atomic_long next;
};

/* One region's haplotypes in sorted order with their LCPs, duplicates and byte codes, and its
 * reads by length, longest first (kernel.rs:53-107, lib.rs:377-384). */
struct port_plan {
    int64_t H, nr, r0;
    const uint8_t** hap;
    int32_t *hap_len, *lcp, *hap_caller, *lcp_count;
    uint8_t* dup;
    int64_t* read_order;
    int max_len, ncodes;
    uint8_t code_of[256], code_byte[256];
};

/* Per-thread buffers, grown on demand. */
struct port_ws {
    float *trans, *prior, *endm, *row_a, *row_b, *snap, *snap_acc;
    uint8_t* end_row;
    int32_t *cnt, *slot_of, *snap_cols, *snap_len;
    double* res;
    size_t cap[12];
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

static void port_ws_free(struct port_ws* w) {
    free(w->trans); free(w->prior); free(w->endm); free(w->row_a); free(w->row_b);
    free(w->snap); free(w->snap_acc); free(w->end_row); free(w->cnt); free(w->slot_of);
    free(w->snap_cols); free(w->res);
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
}

static void plan_free(struct port_plan* P) {
    free(P->hap); free(P->hap_len); free(P->dup); free(P->read_order); free(P->lcp_count);
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
                     int force_avx2) {
    fgklo_port_backend();
    if (force_avx2) g_use_avx512 = 0;
    if (!g_use_avx512 && !__builtin_cpu_supports("avx2")) return -1;
    int64_t* base = malloc(sizeof(int64_t) * (size_t)(n_regions + 1));
    base[0] = 0;
    for (int64_t g = 0; g < n_regions; ++g)
        base[g + 1] = base[g] + (read_region[g + 1] - read_region[g]) *
                                    (hap_region[g + 1] - hap_region[g]);
    struct port_job J = {n_regions, read_region, read_off, bases, quals, ins_gop, del_gop, gcp,
                         hap_region, hap_off, hap_bases, out, base, 0};
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
