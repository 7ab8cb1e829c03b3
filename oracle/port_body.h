/*
 * One region through the lane-batched FP32 sweep, compiled once per vector width by port.c
 * (PORT_LANES floats per vector, PORT_TARGET the instruction set, PORT_FN the function name).
 *
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY (the CPU baseline); see port.c.
 *
 * Restates the structure of fgkl's CPU kernel: reads sorted by length, PORT_LANES reads per
 * batch, one lane each (lib.rs:399-460); lane-interleaved per-row tables with zeros past a
 * read's end (kernel.rs:485-529); haplotypes in sorted order, each swept row by row over its
 * columns after the shared prefix, starting from the column snapshot an earlier haplotype left,
 * rescaled by the length ratio (kernel.rs:220-278, 610-759); the cell recurrence in the FMA
 * order of kernel.rs:1011-1013; results through finish_lane's FP32 rule (kernel.rs:897-914).
 * With a suffix plan, backward sweeps first compute, once per shared suffix, the scaled mass
 * from each cell to the end of the read (kernel.rs:579-602, 761-891, 1029-1126), and each cut
 * haplotype stops its forward sweep one column before its cut and joins there in FP64
 * (kernel.rs:744-791).
 */

PORT_TARGET static void PORT_FN(const struct port_job* J, const struct port_plan* P,
                                struct port_ws* ws, double* out) {
    typedef float V __attribute__((vector_size(PORT_LANES * sizeof(float)), may_alias));
    enum { L = PORT_LANES };
    const int64_t H = P->H, nr = P->nr;

    for (int64_t b0 = 0; b0 < nr; b0 += L) {
        const int nl = (int)(nr - b0 < L ? nr - b0 : L);
        const int64_t* rd = P->read_order + b0;
        const int rows = (int)(J->read_off[rd[0] + 1] - J->read_off[rd[0]]);
        port_ws_reserve(ws, rows, P->max_len, P->ncodes, (int)H, L);

        /* Tables (kernel.rs:485-529): zeros everywhere a lane has no read row. */
        float* trans = ws->trans;  /* [row][6][L] */
        float* prior = ws->prior;  /* [row][code][L] */
        float* endm = ws->endm;    /* [row][L] */
        memset(trans, 0, sizeof(float) * (size_t)rows * 6 * L);
        memset(prior, 0, sizeof(float) * (size_t)rows * (size_t)P->ncodes * L);
        memset(endm, 0, sizeof(float) * (size_t)rows * L);
        memset(ws->end_row, 0, (size_t)rows);
        for (int l = 0; l < nl; ++l) {
            const int64_t o = J->read_off[rd[l]];
            const int n = (int)(J->read_off[rd[l] + 1] - o);
            for (int r = 0; r < n; ++r) {
                double t[6], pr[2];
                fgklo_transitions(J->ins[o + r], J->del[o + r], J->gcp[o + r], t);
                fgklo_priors(J->quals[o + r], pr);
                for (int s = 0; s < 6; ++s) trans[((size_t)r * 6 + s) * L + l] = (float)t[s];
                const uint8_t b = J->bases[o + r];
                for (int c = 0; c < P->ncodes; ++c) {
                    const uint8_t hb = P->code_byte[c];
                    const int match = b == hb || b == 'N' || hb == 'N';
                    prior[((size_t)r * P->ncodes + c) * L + l] = (float)(match ? pr[0] : pr[1]);
                }
            }
            endm[(size_t)(n - 1) * L + l] = 1.0f;
            ws->end_row[n - 1] = 1;
        }

        /* Snapshot bookkeeping (kernel.rs:236-259): a column keeps a snapshot while a later
         * haplotype still starts there. */
        int32_t* cnt = ws->cnt;
        memcpy(cnt, P->lcp_count, sizeof(int32_t) * (size_t)(P->max_len + 1));
        for (int c = 0; c <= P->max_len; ++c) ws->slot_of[c] = -1;
        int nslots = 0;
        const size_t col_stride = (size_t)(rows + 1) * L;  /* one state of one snapshot */
        double* res = ws->res;  /* [lane][sorted hap] */

        /* Backward nodes: node n holds, at its column, B_M and B_Y of rows 0..rows and the sum S
         * of the weights of paths starting in row 0 at or after that column, all scaled by 2^120
         * (the forward initial value's scale, simd.rs:109). */
        const size_t nstride = (size_t)(2 * (rows + 1) + 1) * L;
        const V bscale = (V){} + 0x1p120f;
        if (P->nsweeps > 0 || P->nodes > 1 || (P->cut_col && H > 0)) {
            port_ws_reserve_nodes(ws, rows, P->nodes, L);
            /* Node 0, the last column: only trailing insertions lead on (kernel.rs:579-602). */
            float* n0 = ws->nodes;
            memset(n0, 0, sizeof(float) * nstride);
            V x_below = (V){};
            for (int i = rows; i >= 1; --i) {
                const V sink = bscale * *(const V*)(endm + (size_t)(i - 1) * L);
                V mi = (V){}, ii = (V){};
                if (i < rows) {
                    mi = *(const V*)(trans + ((size_t)i * 6 + 2) * L);
                    ii = *(const V*)(trans + ((size_t)i * 6 + 3) * L);
                }
                *(V*)(n0 + (size_t)i * L) = x_below * mi + sink;
                x_below = x_below * ii + sink;
            }
            /* The sweeps, rows bottom-up, each row right to left over depths resume+1..need. */
            float* bprev = ws->row_a;  /* [column][B_M, B_X][L] of row i + 1 */
            float* bcur = ws->row_b;
            for (int w = 0; w < P->nsweeps; ++w) {
                const struct port_sweep* sw = &P->sweeps[w];
                const uint8_t* hap = P->hap[sw->hap];
                const int len = P->hap_len[sw->hap];
                const int right = len - sw->resume, left = len - sw->need;
                const float* st = ws->nodes + (size_t)sw->start * nstride;
                for (int j = left; j < right; ++j) {
                    *(V*)(bprev + (size_t)j * 2 * L) = (V){};
                    *(V*)(bprev + (size_t)j * 2 * L + L) = (V){};
                }
                V starts = *(const V*)(st + (size_t)2 * (rows + 1) * L);
                for (int i = rows; i >= 0; --i) {
                    V mm = (V){}, im = (V){}, mi = (V){}, ii = (V){}, md = (V){}, dd = (V){};
                    if (i < rows) {
                        const float* t = trans + (size_t)i * 6 * L;
                        mm = *(const V*)(t); im = *(const V*)(t + L);
                        mi = *(const V*)(t + 2 * L); ii = *(const V*)(t + 3 * L);
                    }
                    if (i >= 1) {
                        const float* t = trans + (size_t)(i - 1) * 6 * L;
                        md = *(const V*)(t + 4 * L); dd = *(const V*)(t + 5 * L);
                    }
                    const float* prow = prior + (size_t)(i < rows ? i : rows - 1) * P->ncodes * L;
                    const int has_sink = i >= 1 && ws->end_row[i - 1];
                    V sink = (V){};
                    if (has_sink) sink = bscale * *(const V*)(endm + (size_t)(i - 1) * L);
                    V bm_diag = i < rows ? *(const V*)(st + (size_t)(i + 1) * L) : (V){};
                    V by_right = *(const V*)(st + (size_t)(rows + 1 + i) * L);
                    V bm_last = (V){};
                    int hi = right;
                    for (int d = sw->resume + 1;; ++d) {
                        const int write = d <= sw->need && sw->node_of[d] >= 0;
                        if (d <= sw->need && !write) continue;
                        const int lo = d <= sw->need ? len - d : left;
                        for (int j = hi - 1; j >= lo; --j) {
                            const V pri = *(const V*)(prow + (size_t)P->code_of[hap[j]] * L);
                            if (i == 0) {  /* row 0: paths start here and enter M of row 1 */
                                const V by = pri * bm_diag * im;
                                starts += by;
                                bm_diag = *(const V*)(bprev + (size_t)j * 2 * L);
                                by_right = by;
                            } else {
                                const V bm_below = *(const V*)(bprev + (size_t)j * 2 * L);
                                const V bx_below = *(const V*)(bprev + (size_t)j * 2 * L + L);
                                const V into = pri * bm_diag;
                                const V ind = into * im;
                                V bx = bx_below * ii + ind;
                                const V by = by_right * dd + ind;
                                V bm = by_right * md + (bx_below * mi + into * mm);
                                bm += sink;
                                bx += sink;
                                *(V*)(bcur + (size_t)j * 2 * L) = bm;
                                *(V*)(bcur + (size_t)j * 2 * L + L) = bx;
                                bm_diag = bm_below;
                                by_right = by;
                                bm_last = bm;
                            }
                        }
                        hi = lo;
                        if (!write) break;
                        float* nd = ws->nodes + (size_t)sw->node_of[d] * nstride;
                        *(V*)(nd + (size_t)(rows + 1 + i) * L) = by_right;
                        if (i == 0)
                            *(V*)(nd + (size_t)2 * (rows + 1) * L) = starts;
                        else
                            *(V*)(nd + (size_t)i * L) = bm_last;
                    }
                    if (i > 0) {
                        float* tmp = bprev; bprev = bcur; bcur = tmp;
                    }
                }
            }
        }

        for (int64_t k = 0; k < H; ++k) {
            if (P->dup[k]) {
                for (int l = 0; l < nl; ++l) res[l * H + k] = res[l * H + k - 1];
                continue;
            }
            const int start = k ? P->lcp[k] : 0;
            if (k) cnt[start] -= 1;
            const uint8_t* hap = P->hap[k];
            const int len = P->hap_len[k];
            const int cut = P->cut_col ? P->cut_col[k] : 0;  /* 0: sweep to the end */
            const int stop = cut ? cut - 1 : len;
            const float* node = cut ? ws->nodes + (size_t)P->cut_node[k] * nstride : NULL;
            double cross[L];
            for (int l = 0; l < L; ++l) cross[l] = 0.0;
            int nq = 0;
            for (int q = start + 1; q <= stop; ++q) {
                if (cnt[q] > 0) {
                    if (ws->slot_of[q] < 0) ws->slot_of[q] = nslots++;
                    ws->snap_cols[nq++] = q;
                }
            }
            const V init = (V){} + (float)(0x1p120 / (double)len);  /* kernel.rs:640 */
            const V zero = (V){};
            /* The start column: the origin (column 0) or an earlier haplotype's snapshot. */
            const float* sm = NULL;
            const float* sx = NULL;
            const float* sy = NULL;
            V scale = (V){} + 1.0f, acc;
            if (start > 0) {
                const int s = ws->slot_of[start];
                sm = ws->snap + (size_t)s * 3 * col_stride;
                sx = sm + col_stride;
                sy = sx + col_stride;
                scale = (V){} + (float)((double)ws->snap_len[s] / (double)len);  /* kernel.rs:641 */
                acc = *(const V*)(ws->snap_acc + (size_t)s * L) * scale;
            } else {
                acc = zero;
            }
            float* prev = ws->row_a;
            float* cur = ws->row_b;
            for (int j = start; j <= stop; ++j) {  /* DP row 0: (0, 0, C/m) */
                *(V*)(prev + (size_t)j * 3 * L) = zero;
                *(V*)(prev + (size_t)j * 3 * L + L) = zero;
                *(V*)(prev + (size_t)j * 3 * L + 2 * L) = init;
            }
            for (int a = 0; a < nq; ++a) {
                const int s = ws->slot_of[ws->snap_cols[a]];
                float* d = ws->snap + (size_t)s * 3 * col_stride;
                *(V*)(d) = zero;
                *(V*)(d + col_stride) = zero;
                *(V*)(d + 2 * col_stride) = init;
                *(V*)(ws->snap_acc + (size_t)s * L) = acc;
                ws->snap_len[s] = len;
            }
            for (int i = 1; i <= rows; ++i) {
                const int r = i - 1;
                const float* t = trans + (size_t)r * 6 * L;
                const V mm = *(const V*)(t), im = *(const V*)(t + L), mi = *(const V*)(t + 2 * L);
                const V ii = *(const V*)(t + 3 * L), md = *(const V*)(t + 4 * L);
                const V dd = *(const V*)(t + 5 * L);
                const float* prow = prior + (size_t)r * P->ncodes * L;
                V dm, dx, dy, lm, ly;
                if (sm) {
                    dm = *(const V*)(sm + (size_t)(i - 1) * L) * scale;
                    dx = *(const V*)(sx + (size_t)(i - 1) * L) * scale;
                    dy = *(const V*)(sy + (size_t)(i - 1) * L) * scale;
                    lm = *(const V*)(sm + (size_t)i * L) * scale;
                    ly = *(const V*)(sy + (size_t)i * L) * scale;
                } else {
                    dm = zero; dx = zero; dy = i == 1 ? init : zero;
                    lm = zero; ly = zero;
                }
                const int end = ws->end_row[r];
                const V e = *(const V*)(endm + (size_t)r * L);
                V rowsum = zero;
                int a = 0;
                for (int j = start + 1; j <= stop; ++j) {
                    const float* up = prev + (size_t)j * 3 * L;
                    const V um = *(const V*)(up), ux = *(const V*)(up + L);
                    const V uy = *(const V*)(up + 2 * L);
                    const V pri = *(const V*)(prow + (size_t)P->code_of[hap[j - 1]] * L);
                    const V m = pri * (dy * im + (dx * im + dm * mm));
                    const V x = ux * ii + um * mi;
                    const V y = ly * dd + lm * md;
                    float* c = cur + (size_t)j * 3 * L;
                    *(V*)(c) = m;
                    *(V*)(c + L) = x;
                    *(V*)(c + 2 * L) = y;
                    if (end) {
                        rowsum += m + x;
                        if (a < nq && ws->snap_cols[a] == j) {
                            float* sa = ws->snap_acc + (size_t)ws->slot_of[j] * L;
                            *(V*)sa = e * rowsum + *(V*)sa;
                            ++a;
                        }
                    }
                    dm = um; dx = ux; dy = uy;
                    lm = m; ly = y;
                }
                if (end) acc = e * rowsum + acc;
                if (node) {
                    /* The values entering the cut column (kernel.rs:744-755): match from the
                     * diagonal, deletion from the left; insertions stay within a column. */
                    const V pri = *(const V*)(prow + (size_t)P->code_of[hap[cut - 1]] * L);
                    const V jm = pri * (dy * im + (dx * im + dm * mm));
                    const V jy = ly * dd + lm * md;
                    const V nm = *(const V*)(node + (size_t)i * L);
                    const V ny = *(const V*)(node + (size_t)(rows + 1 + i) * L);
                    for (int l = 0; l < L; ++l)
                        cross[l] += (double)jm[l] * (double)nm[l] + (double)jy[l] * (double)ny[l];
                }
                for (int q = 0; q < nq; ++q) {
                    const int col = ws->snap_cols[q];
                    float* d = ws->snap + (size_t)ws->slot_of[col] * 3 * col_stride;
                    const float* c = cur + (size_t)col * 3 * L;
                    *(V*)(d + (size_t)i * L) = *(const V*)(c);
                    *(V*)(d + col_stride + (size_t)i * L) = *(const V*)(c + L);
                    *(V*)(d + 2 * col_stride + (size_t)i * L) = *(const V*)(c + 2 * L);
                }
                float* tmp = prev; prev = cur; cur = tmp;
            }
            if (node) {  /* the join (kernel.rs:767-791), in FP64 */
                const double init_d = (double)(float)(0x1p120 / (double)len);
                const float* ns = node + (size_t)2 * (rows + 1) * L;
                for (int l = 0; l < nl; ++l)
                    res[l * H + k] = (double)acc[l] + (cross[l] + init_d * (double)ns[l]) / 0x1p120;
            } else {
                for (int l = 0; l < nl; ++l) res[l * H + k] = (double)acc[l];
            }
        }
        /* finish_lane (kernel.rs:897-914) and the FP64 recomputation of lost pairs. */
        for (int l = 0; l < nl; ++l) {
            const int64_t rl = rd[l] - P->r0;
            const int64_t o = J->read_off[rd[l]], n = J->read_off[rd[l] + 1] - o;
            for (int64_t k = 0; k < H; ++k) {
                const double raw = res[l * H + k];
                double v;
                if (!(raw >= 1e-28)) {
                    v = fgklo_log10_likelihood(P->hap[k], P->hap_len[k], J->bases + o,
                                               J->quals + o, J->ins + o, J->del + o,
                                               J->gcp + o, n);
                } else {
                    v = log10(raw) - 120.0 * 0.30102999566398119521;
                }
                out[rl * H + P->hap_caller[k]] = v;
            }
        }
    }
}
