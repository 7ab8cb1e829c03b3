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

        for (int64_t k = 0; k < H; ++k) {
            if (P->dup[k]) {
                for (int l = 0; l < nl; ++l) res[l * H + k] = res[l * H + k - 1];
                continue;
            }
            const int start = k ? P->lcp[k] : 0;
            if (k) cnt[start] -= 1;
            const uint8_t* hap = P->hap[k];
            const int len = P->hap_len[k];
            int nq = 0;
            for (int q = start + 1; q <= len; ++q) {
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
            for (int j = start; j <= len; ++j) {  /* DP row 0: (0, 0, C/m) */
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
                for (int j = start + 1; j <= len; ++j) {
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
            for (int l = 0; l < nl; ++l) res[l * H + k] = (double)acc[l];
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
