/*
 * CPU oracle for Smith-Waterman: GATK's SmithWatermanJavaAligner as fgkl's scalar port states it
 * (crates/smithwaterman/src/reference.rs:10-203, lib.rs:20-30 for the constants).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ and bench.py as the checker, never by the product
 * library (fgkl_amd/).
 */
#include <limits.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define MATRIX_MIN_CUTOFF (-100000000)  /* lib.rs:20 */
#define LOW_INIT_VALUE (INT_MIN / 2)    /* lib.rs:22 */

enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4 };  /* BAM's operation codes ("MIDNSHP=X") */
enum { SOFTCLIP = 0, INDEL = 1, LEADING_INDEL = 2, IGNORE = 3 };

/* The last start of `alt` inside `ref`, or -1 (lib.rs last_index_of). */
static int64_t last_index_of(const uint8_t* ref, int64_t n, const uint8_t* alt, int64_t m) {
    if (m > n) return -1;
    for (int64_t s = n - m; s >= 0; --s)
        if (memcmp(ref + s, alt, (size_t)m) == 0) return s;
    return -1;
}

/* reference.rs:34-100 */
static void calculate_matrix(const uint8_t* ref, int64_t nref, const uint8_t* alt, int64_t nalt,
                             int32_t* sw, int32_t* bt, int strategy, const int32_t p[4]) {
    const int64_t ncol = nalt + 1, nrow = nref + 1;
    int32_t* best_gap_v = malloc(sizeof(int32_t) * (size_t)(ncol + 1));
    int32_t* gap_size_v = calloc((size_t)(ncol + 1), sizeof(int32_t));
    int32_t* best_gap_h = malloc(sizeof(int32_t) * (size_t)(nrow + 1));
    int32_t* gap_size_h = calloc((size_t)(nrow + 1), sizeof(int32_t));
    for (int64_t j = 0; j <= ncol; ++j) best_gap_v[j] = LOW_INIT_VALUE;
    for (int64_t i = 0; i <= nrow; ++i) best_gap_h[i] = LOW_INIT_VALUE;
    const int32_t w_match = p[0], w_mismatch = p[1], w_open = p[2], w_extend = p[3];
    if (strategy == INDEL || strategy == LEADING_INDEL) {
        sw[1] = w_open;
        int32_t cur = w_open;
        for (int64_t j = 2; j < ncol; ++j) {
            cur += w_extend;
            sw[j] = cur;
        }
        sw[ncol] = w_open;
        cur = w_open;
        for (int64_t i = 2; i < nrow; ++i) {
            cur += w_extend;
            sw[i * ncol] = cur;
        }
    }
    for (int64_t i = 1; i < nrow; ++i) {
        const uint8_t a = ref[i - 1];
        for (int64_t j = 1; j < ncol; ++j) {
            const uint8_t b = alt[j - 1];
            const int32_t step_diag = sw[(i - 1) * ncol + j - 1] + (a == b ? w_match : w_mismatch);
            int32_t prev_gap = sw[(i - 1) * ncol + j] + w_open;
            best_gap_v[j] += w_extend;
            if (prev_gap > best_gap_v[j]) {
                best_gap_v[j] = prev_gap;
                gap_size_v[j] = 1;
            } else {
                gap_size_v[j] += 1;
            }
            const int32_t step_down = best_gap_v[j], kd = gap_size_v[j];
            prev_gap = sw[i * ncol + j - 1] + w_open;
            best_gap_h[i] += w_extend;
            if (prev_gap > best_gap_h[i]) {
                best_gap_h[i] = prev_gap;
                gap_size_h[i] = 1;
            } else {
                gap_size_h[i] += 1;
            }
            const int32_t step_right = best_gap_h[i], ki = gap_size_h[i];
            const int64_t cell = i * ncol + j;
            if (step_diag >= step_down && step_diag >= step_right) {
                sw[cell] = step_diag > MATRIX_MIN_CUTOFF ? step_diag : MATRIX_MIN_CUTOFF;
                bt[cell] = 0;
            } else if (step_right >= step_down) {
                sw[cell] = step_right > MATRIX_MIN_CUTOFF ? step_right : MATRIX_MIN_CUTOFF;
                bt[cell] = -ki;
            } else {
                sw[cell] = step_down > MATRIX_MIN_CUTOFF ? step_down : MATRIX_MIN_CUTOFF;
                bt[cell] = kd;
            }
        }
    }
    free(best_gap_v); free(gap_size_v); free(best_gap_h); free(gap_size_h);
}

static void push(uint32_t* out, int* n, uint32_t len, int op) { out[(*n)++] = (len << 4) | (uint32_t)op; }

/* reference.rs:102-203 */
static int32_t calculate_cigar(const int32_t* sw, const int32_t* bt, int64_t n, int64_t m,
                               int strategy, uint32_t* cig, int* ncig) {
    const int64_t ref_length = n - 1, alt_length = m - 1;
    int64_t p1 = 0, p2 = alt_length, segment = 0;
    int32_t max_score = INT_MIN;
    if (strategy == INDEL) {
        p1 = ref_length;
    } else {
        for (int64_t i = 1; i < n; ++i) {
            const int32_t cur = sw[i * m + alt_length];
            if (cur >= max_score) {
                p1 = i;
                max_score = cur;
            }
        }
        if (strategy != LEADING_INDEL) {
            for (int64_t j = 1; j < m; ++j) {
                const int32_t cur = sw[ref_length * m + j];
                const int64_t dj = ref_length - j, dp = p1 - p2;
                if (cur > max_score || (cur == max_score && llabs(dj) < llabs(dp))) {
                    p1 = ref_length;
                    p2 = j;
                    max_score = cur;
                    segment = alt_length - j;
                }
            }
        }
    }
    *ncig = 0;
    if (segment > 0 && strategy == SOFTCLIP) {
        push(cig, ncig, (uint32_t)segment, OP_S);
        segment = 0;
    }
    int state = OP_M;
    for (;;) {
        const int32_t btr = bt[p1 * m + p2];
        int new_state;
        int64_t step;
        if (btr > 0) {
            new_state = OP_D;
            step = btr;
        } else if (btr < 0) {
            new_state = OP_I;
            step = -btr;
        } else {
            new_state = OP_M;
            step = 1;
        }
        if (new_state == OP_M) {
            p1 -= 1;
            p2 -= 1;
        } else if (new_state == OP_I) {
            p2 -= step;
        } else {
            p1 -= step;
        }
        if (new_state == state) {
            segment += step;
        } else {
            if (segment > 0) push(cig, ncig, (uint32_t)segment, state);
            segment = step;
            state = new_state;
        }
        if (!(p1 > 0 && p2 > 0)) break;
    }
    int32_t offset;
    if (strategy == SOFTCLIP) {
        push(cig, ncig, (uint32_t)segment, state);
        if (p2 > 0) push(cig, ncig, (uint32_t)p2, OP_S);
        offset = (int32_t)p1;
    } else if (strategy == IGNORE) {
        push(cig, ncig, (uint32_t)(segment + p2), state);
        offset = (int32_t)(p1 - p2);
    } else {
        push(cig, ncig, (uint32_t)segment, state);
        if (p1 > 0)
            push(cig, ncig, (uint32_t)p1, OP_D);
        else if (p2 > 0)
            push(cig, ncig, (uint32_t)p2, OP_I);
        offset = 0;
    }
    for (int a = 0, b = *ncig - 1; a < b; ++a, --b) {  /* alternate order */
        const uint32_t t = cig[a];
        cig[a] = cig[b];
        cig[b] = t;
    }
    return offset;
}

int fgklo_sw_align(const uint8_t* ref, int64_t nref, const uint8_t* alt, int64_t nalt,
                   const int32_t params[4], int strategy, int32_t* offset, uint32_t* cigar,
                   int32_t* ncigar) {
    if (nref <= 0 || nalt <= 0) return -1;
    if (strategy == SOFTCLIP || strategy == IGNORE) {
        const int64_t s = last_index_of(ref, nref, alt, nalt);
        if (s >= 0) {
            cigar[0] = ((uint32_t)nalt << 4) | OP_M;
            *ncigar = 1;
            *offset = (int32_t)s;
            return 0;
        }
    }
    const int64_t n = nref + 1, m = nalt + 1;
    int32_t* sw = calloc((size_t)(n * m), sizeof(int32_t));
    int32_t* bt = calloc((size_t)(n * m), sizeof(int32_t));
    calculate_matrix(ref, nref, alt, nalt, sw, bt, strategy, params);
    int nc = 0;
    *offset = calculate_cigar(sw, bt, n, m, strategy, cigar, &nc);
    *ncigar = nc;
    free(sw);
    free(bt);
    return 0;
}
