/*
 * Lane-batched SIMD Smith-Waterman: the CPU baseline bench.py times for the aligner on the host
 * cores beside the GPU.
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY.  Never linked into, loaded by or called from the product
 * library (fgkl_amd/).
 *
 * fgkl's fastest CPU aligner fills GATK's matrix anti-diagonal by anti-diagonal in 16-bit SIMD
 * lanes with a 32-bit redo of saturating pairs (crates/smithwaterman/src/diag.rs, simd.rs); it is
 * Rust and cannot be built here.  This baseline vectorises across pairs instead (one pair per
 * 32-bit lane, 16 lanes on AVX-512, 8 on AVX2, picked at run time; rows outer, columns inner, one
 * trace byte per cell and lane) and traces each pair back on its own, threads pulling groups of
 * pairs of similar size.  Its results equal the scalar oracle's (oracle/sw.c) and the reference's.
 */
#define _GNU_SOURCE
#include <limits.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <immintrin.h>

#include "oracle.h"

#define SWP_CUTOFF (-100000000)
#define SWP_LOW (INT_MIN / 2)
#define SWP_CAT_(a, b) a##b
#define SWP_CAT(a, b) SWP_CAT_(a, b)

enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4 };

struct swp_job {
    int64_t n_pairs;
    const int64_t *ref_off, *alt_off;
    const uint8_t *ref, *alt;
    const int32_t* params;
    int strategy;
    int32_t* offsets;
    const int64_t* cig_start; /* pair p's elements from cigar[cig_start[p]] (room for n + m + 4) */
    uint32_t* cigar;
    int32_t* ncig;
    const int64_t* order;     /* pairs by size, largest first */
    int lanes;
    atomic_long next;
};

struct swp_ws {
    uint8_t* trace;
    void* rows;
    int32_t* bases;
    size_t cap_trace, cap_rows, cap_bases;
};

static void* swp_grow(void* p, size_t* cap, size_t need) {
    if (need <= *cap) return p;
    free(p);
    *cap = need;
    return aligned_alloc(64, (need + 63) & ~(size_t)63);
}

static void swp_reserve(struct swp_ws* w, size_t trace, size_t rows, size_t bases) {
    w->trace = swp_grow(w->trace, &w->cap_trace, trace);
    w->rows = swp_grow(w->rows, &w->cap_rows, rows);
    w->bases = swp_grow(w->bases, &w->cap_bases, bases);
}

static int64_t swp_last_index_of(const uint8_t* r, int64_t n, const uint8_t* a, int64_t m) {
    if (m > n) return -1;
    for (int64_t s = n - m; s >= 0; --s)
        if (memcmp(r + s, a, (size_t)m) == 0) return s;
    return -1;
}

/* calculateCigar (reference.rs:102-203) over one lane's trace bytes: bit 0 not diagonal, bit 1
 * insertion over deletion, bits 2 / 3 the vertical / horizontal gap extends the one before. */
static void swp_traceback(const struct swp_job* J, int64_t p, const uint8_t* tr, int L,
                          size_t cols, int n, int m, int cb, int ci, int lb, int ld, int lj) {
    uint32_t* out = J->cigar + J->cig_start[p];
    int k = 0;
    if (J->strategy == 0 || J->strategy == 3) {
        const int64_t s = swp_last_index_of(J->ref + J->ref_off[p], n, J->alt + J->alt_off[p], m);
        if (s >= 0) {
            out[0] = ((uint32_t)m << 4) | OP_M;
            J->ncig[p] = 1;
            J->offsets[p] = (int32_t)s;
            return;
        }
    }
#define BITS(i, j) (tr[((size_t)(i) * cols + (size_t)(j)) * (size_t)L])
    int p1 = n, p2 = m, seg = 0;
    if (J->strategy != 1) {
        p1 = ci;
        if (J->strategy != 2 && (lb > cb || (lb == cb && ld < abs(p1 - p2)))) {
            p1 = n;
            p2 = lj;
            seg = m - lj;
        }
    }
    if (seg > 0 && J->strategy == 0) {
        out[k++] = ((uint32_t)seg << 4) | OP_S;
        seg = 0;
    }
    int state = OP_M;
    do {
        const uint8_t b = BITS(p1, p2);
        int next, step = 1;
        if (!(b & 1)) {
            next = OP_M;
            --p1;
            --p2;
        } else if (b & 2) {
            for (int c = p2; c > 1 && (BITS(p1, c) & 8); --c) ++step;
            next = OP_I;
            p2 -= step;
        } else {
            for (int c = p1; c > 1 && (BITS(c, p2) & 4); --c) ++step;
            next = OP_D;
            p1 -= step;
        }
        if (next == state) {
            seg += step;
        } else {
            if (seg > 0) out[k++] = ((uint32_t)seg << 4) | (uint32_t)state;
            seg = step;
            state = next;
        }
    } while (p1 > 0 && p2 > 0);
#undef BITS
    int off = 0;
    if (J->strategy == 0) {
        out[k++] = ((uint32_t)seg << 4) | (uint32_t)state;
        if (p2 > 0) out[k++] = ((uint32_t)p2 << 4) | OP_S;
        off = p1;
    } else if (J->strategy == 3) {
        out[k++] = ((uint32_t)(seg + p2) << 4) | (uint32_t)state;
        off = p1 - p2;
    } else {
        out[k++] = ((uint32_t)seg << 4) | (uint32_t)state;
        if (p1 > 0)
            out[k++] = ((uint32_t)p1 << 4) | OP_D;
        else if (p2 > 0)
            out[k++] = ((uint32_t)p2 << 4) | OP_I;
    }
    for (int a = 0, z = k - 1; a < z; ++a, --z) {
        const uint32_t t = out[a];
        out[a] = out[z];
        out[z] = t;
    }
    J->ncig[p] = k;
    J->offsets[p] = off;
}

/* Per width: lane-wise signed maximum (vpmaxsd), which GCC does not form from a compare and a
 * blend of its own vector types. */
#define SWP_LANES 16
#define SWP_TARGET __attribute__((target("avx512f,avx512bw")))
#define SWP_FN swp_group_avx512
#define SWP_MAX(a, b) ((VI)_mm512_max_epi32((__m512i)(a), (__m512i)(b)))
#include "sw_port_body.h"
#undef SWP_LANES
#undef SWP_TARGET
#undef SWP_FN
#undef SWP_MAX

#define SWP_LANES 8
#define SWP_TARGET __attribute__((target("avx2")))
#define SWP_FN swp_group_avx2
#define SWP_MAX(a, b) ((VI)_mm256_max_epi32((__m256i)(a), (__m256i)(b)))
#include "sw_port_body.h"
#undef SWP_LANES
#undef SWP_TARGET
#undef SWP_FN
#undef SWP_MAX

static int g_sw_avx512 = -1;

const char* fgklo_sw_port_backend(void) {
    if (g_sw_avx512 < 0) {
        __builtin_cpu_init();
        g_sw_avx512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
    }
    return g_sw_avx512 ? "avx512 (16 i32 lanes)" : "avx2 (8 i32 lanes)";
}

static void* swp_worker(void* arg) {
    struct swp_job* J = arg;
    struct swp_ws w = {0};
    const int L = J->lanes;
    for (;;) {
        const int64_t g = atomic_fetch_add(&J->next, 1);
        const int64_t first = g * L;
        if (first >= J->n_pairs) break;
        const int count = (int)(J->n_pairs - first < L ? J->n_pairs - first : L);
        if (L == 16)
            swp_group_avx512(J, J->order + first, count, &w);
        else
            swp_group_avx2(J, J->order + first, count, &w);
    }
    free(w.trace);
    free(w.rows);
    free(w.bases);
    return NULL;
}

static const struct swp_job* g_swp_sort;
static pthread_mutex_t g_swp_mutex = PTHREAD_MUTEX_INITIALIZER;

static int swp_cmp(const void* a, const void* b) {
    const int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    const struct swp_job* J = g_swp_sort;
    const int64_t nx = J->ref_off[x + 1] - J->ref_off[x], ny = J->ref_off[y + 1] - J->ref_off[y];
    const int64_t mx = J->alt_off[x + 1] - J->alt_off[x], my = J->alt_off[y + 1] - J->alt_off[y];
    if (nx != ny) return nx > ny ? -1 : 1;
    if (mx != my) return mx > my ? -1 : 1;
    return x < y ? -1 : x > y;
}

int fgklo_sw_port_batch(int64_t n_pairs, const int64_t* ref_off, const uint8_t* ref,
                        const int64_t* alt_off, const uint8_t* alt, const int32_t params[4],
                        int strategy, int32_t* offsets, const int64_t* cig_start, uint32_t* cigar,
                        int32_t* ncig, int nthreads, int force_avx2) {
    fgklo_sw_port_backend();
    const int use512 = g_sw_avx512 && !force_avx2;
    if (!use512 && !__builtin_cpu_supports("avx2")) return -1;
    for (int64_t p = 0; p < n_pairs; ++p)
        if (ref_off[p + 1] == ref_off[p] || alt_off[p + 1] == alt_off[p]) return -2;
    int64_t* order = malloc(sizeof(int64_t) * (size_t)(n_pairs > 0 ? n_pairs : 1));
    for (int64_t p = 0; p < n_pairs; ++p) order[p] = p;
    struct swp_job J = {n_pairs, ref_off, alt_off, ref, alt, params, strategy, offsets,
                        cig_start, cigar, ncig, order, use512 ? 16 : 8, 0};
    atomic_init(&J.next, 0);
    pthread_mutex_lock(&g_swp_mutex);
    g_swp_sort = &J;
    qsort(order, (size_t)n_pairs, sizeof(int64_t), swp_cmp);
    pthread_mutex_unlock(&g_swp_mutex);
    if (nthreads < 1) nthreads = 1;
    pthread_t* th = malloc(sizeof(pthread_t) * (size_t)nthreads);
    for (int i = 0; i < nthreads; ++i) pthread_create(&th[i], NULL, swp_worker, &J);
    for (int i = 0; i < nthreads; ++i) pthread_join(th[i], NULL);
    free(th);
    free(order);
    return 0;
}
