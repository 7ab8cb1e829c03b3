/*
 * The lane-batched Smith-Waterman group kernel, included once per vector width by sw_port.c with
 * SWP_LANES (int32 lanes), SWP_TARGET (target attribute) and SWP_FN (function name) defined.
 * TEST / MEASUREMENT INFRASTRUCTURE ONLY.
 */
typedef int32_t SWP_CAT(swp_vi, SWP_LANES) __attribute__((vector_size(4 * SWP_LANES)));
typedef int8_t SWP_CAT(swp_vb, SWP_LANES) __attribute__((vector_size(SWP_LANES)));
#define VI SWP_CAT(swp_vi, SWP_LANES)
#define VB SWP_CAT(swp_vb, SWP_LANES)
/* Lane-wise select and maximum (comparisons give all-ones or zero per lane). */
#define VSEL(mask, a, b) (((a) & (mask)) | ((b) & ~(mask)))
#define VMAX(a, b) SWP_MAX(a, b)

/* Fills the score recurrences of up to SWP_LANES pairs at once (one pair per lane, row by row over
 * the group's largest matrix) and then traces each pair back on its own. */
SWP_TARGET static void SWP_FN(const struct swp_job* J, const int64_t* pairs, int count,
                              struct swp_ws* w) {
    enum { L = SWP_LANES };
    int n[L], m[L], N = 0, M = 0;
    for (int l = 0; l < L; ++l) {
        const int64_t p = pairs[l < count ? l : 0];
        n[l] = (int)(J->ref_off[p + 1] - J->ref_off[p]);
        m[l] = (int)(J->alt_off[p + 1] - J->alt_off[p]);
        if (n[l] > N) N = n[l];
        if (m[l] > M) M = m[l];
    }
    const size_t cols = (size_t)M + 1;
    swp_reserve(w, (size_t)(N + 1) * cols * L, (size_t)(M + 1) * L * sizeof(int32_t) * 3,
                (size_t)(N + M + 2) * L * sizeof(int32_t));
    VI* hp = (VI*)w->rows;       /* H of the previous row, columns 0..M */
    VI* hc = hp + cols;          /* H of this row */
    VI* vg = hc + cols;          /* best vertical gap per column */
    int32_t* altb = w->bases;    /* [j][l] */
    int32_t* refb = w->bases + (size_t)(M + 1) * L;  /* [i][l] */
    uint8_t* tr = w->trace;      /* [(i * cols + j) * L + l] */
    for (int l = 0; l < L; ++l) {
        const int64_t p = pairs[l < count ? l : 0];
        const uint8_t* a = J->alt + J->alt_off[p];
        const uint8_t* r = J->ref + J->ref_off[p];
        for (int j = 0; j < M; ++j) altb[(size_t)(j + 1) * L + l] = j < m[l] ? a[j] : 0x100;
        for (int i = 0; i < N; ++i) refb[(size_t)(i + 1) * L + l] = i < n[l] ? r[i] : 0x200;
    }
    const int32_t wm = J->params[0], wx = J->params[1], wo = J->params[2], we = J->params[3];
    const int gaps = J->strategy == 1 || J->strategy == 2;
    const VI vwm = (VI){} + wm, vwx = (VI){} + wx, vwo = (VI){} + wo, vwe = (VI){} + we;
    const VI vcut = (VI){} + SWP_CUTOFF, vlow = (VI){} + SWP_LOW;
    const VI one = (VI){} + 1, two = (VI){} + 2, four = (VI){} + 4, eight = (VI){} + 8;
    for (int j = 0; j <= M; ++j) {
        hp[j] = (VI){} + (gaps && j > 0 ? wo + (j - 1) * we : 0);
        vg[j] = vlow;
    }
    int cb[L], ci[L], lb[L], ld[L], lj[L];
    for (int l = 0; l < L; ++l) {
        cb[l] = INT32_MIN;
        ci[l] = 0;
        lb[l] = INT32_MIN;
        ld[l] = INT32_MAX;
        lj[l] = 0;
    }
    for (int i = 1; i <= N; ++i) {
        const VI ra = *(const VI*)(refb + (size_t)i * L);
        hc[0] = (VI){} + (gaps ? wo + (i - 1) * we : 0);
        VI z = vlow;
        uint8_t* trow = tr + (size_t)i * cols * L;
        for (int j = 1; j <= M; ++j) {
            const VI ab = *(const VI*)(altb + (size_t)j * L);
            const VI s = VSEL(ra == ab, vwm, vwx);
            const VI d = hp[j - 1] + s;
            const VI to = hp[j] + vwo, te = vg[j] + vwe;
            const VI v = VMAX(to, te);
            vg[j] = v;
            const VI ho = hc[j - 1] + vwo, he = z + vwe;
            z = VMAX(ho, he);
            const VI mx = VMAX(v, z);
            VI h = VMAX(d, mx);
            h = VMAX(h, vcut);
            hc[j] = h;
            const VI bits = ((d < mx) & one) | ((z >= v) & two) | ((to <= te) & four) |
                            ((ho <= he) & eight);
            *(VB*)(trow + (size_t)j * L) = __builtin_convertvector(bits, VB);
        }
        for (int l = 0; l < L; ++l) {
            if (i > n[l]) continue;
            const int32_t c = ((const int32_t*)&hc[m[l]])[l];
            if (c >= cb[l]) {
                cb[l] = c;
                ci[l] = i;
            }
            if (i == n[l]) {
                for (int j = 1; j <= m[l]; ++j) {
                    const int32_t c2 = ((const int32_t*)&hc[j])[l];
                    const int dist = abs(n[l] - j);
                    if (c2 > lb[l] || (c2 == lb[l] && dist < ld[l])) {
                        lb[l] = c2;
                        ld[l] = dist;
                        lj[l] = j;
                    }
                }
            }
        }
        VI* t = hp;
        hp = hc;
        hc = t;
    }
    for (int l = 0; l < count; ++l)
        swp_traceback(J, pairs[l], tr + l, L, cols, n[l], m[l], cb[l], ci[l], lb[l], ld[l], lj[l]);
}

#undef VI
#undef VB
#undef VSEL
#undef VMAX
