/* Row loops of the reference's SIMD batch kernels, written once over a small
 * set of lane operations and included once per instruction set by
 * refsimd_baseline.c (no include guard on purpose).  TEST / BENCH
 * INFRASTRUCTURE: the CPU timing baseline only.
 *
 * Before inclusion define:
 *   RS_SFX         function-name suffix
 *   RS_TARGET      __attribute__((target(...))) of the instantiation
 *   V, M8, M16     vector type, 8-bit lane mask type, 16-bit lane mask type
 *   RS_W8, RS_W16  lanes per vector at 8 and 16 bits
 *   LD(p), ST(p,v), ZERO(), OR(a,b)
 *   S8(x), ADDS8, SUBS8, MAX8, MINU8, EQ8, GT8, SEL8(m,a,b), MSB8(v), MAND8
 *   S16(x), ADDS16, SUBS16, MAX16, MIN16, EQ16, GT16, SEL16, MAND16, MOR16
 *   MFROM8(p), MFROM16(p)   mask from a lane array of 0 / -1
 * SEL(m, a, b) is a where m is set, else b.
 *
 * i8 kernel: banded_swa/kernel.rs:29-284 (sw_kernel_with_ws).  i16 kernel:
 * banded_swa/kernel_i16.rs:44-272 (sw_kernel_i16).  Per chunk of lanes:
 * full qmax x tmax sweep with band masking, per-lane scalar z-drop after each
 * row, 0-based ends, gscore / gtle / max_off 0 (the reference's semantics,
 * not ksw_extend2's). */

#define RS_CAT2(a, b) a##b
#define RS_CAT(a, b) RS_CAT2(a, b)

/* qs/ts: SoA codes [pos * RS_W8 + lane] (0xFF padding); H, E: qmax * RS_W8 bytes */
RS_TARGET static void RS_CAT(rs_kern8_, RS_SFX)(const int8_t *qs, const int8_t *ts, int lanes, int qmax,
                                                int tmax, const int8_t *ql, const int8_t *tl,
                                                const int8_t *h0, const int8_t *wv, int o_del,
                                                int e_del, int o_ins, int e_ins, int zdrop,
                                                int8_t match, int8_t mis, int8_t *H, int8_t *E,
                                                int32_t *out) {
    const int W = RS_W8;
    const V zero = ZERO();
    const V vmatch = S8(match), vmis = S8(mis), one = S8(1);
    const V oed = S8((int8_t)(o_del + e_del)), oei = S8((int8_t)(o_ins + e_ins));
    const V ed = S8((int8_t)e_del), ei = S8((int8_t)e_ins);
    int8_t lane_v[64] = {0}, beg[64] = {0}, end[64] = {0}, cb[64], ce[64], tm[64] = {0}, msv[64];
    int term[64] = {0}, nterm = 0;
    /* first row (:103-118): signed saturating steps, no clamp at 0 */
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? h0[k] : 0;
    V hv = LD(lane_v);
    if (qmax > 0) ST(H, hv);
    V hp = SUBS8(hv, oei);
    if (qmax > 1) ST(H + W, hp);
    for (int j = 2; j < qmax; ++j) {
        hp = SUBS8(hp, ei);
        ST(H + (size_t)j * W, hp);
    }
    for (int j = 0; j < qmax; ++j) ST(E + (size_t)j * W, zero);
    V maxs = hv, maxi = zero, maxj = zero;
    for (int k = 0; k < W; ++k) end[k] = k < lanes ? ql[k] : 0;
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? ql[k] : 0;
    const V vql = LD(lane_v);
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? wv[k] : 0;
    const V vw = LD(lane_v);
    for (int i = 0; i < tmax; ++i) {
        if (nterm > lanes / 2) break; /* :144-147 */
        V f = zero, hdiag = LD(H);
        const V s1 = LD(ts + (size_t)i * W);
        const V vi = S8((int8_t)i);
        const V vb = MAX8(LD(beg), SUBS8(vi, vw));
        V ve = MINU8(LD(end), ADDS8(ADDS8(vi, vw), one));
        ve = MINU8(ve, vql);
        ST(cb, vb);
        ST(ce, ve);
        for (int k = 0; k < lanes; ++k) tm[k] = (!term[k] && i < tl[k]) ? -1 : 0;
        const M8 tmask = MFROM8(tm);
        const V bm1 = SUBS8(vb, one);
        for (int j = 0; j < qmax; ++j) {
            const V htop = LD(H + (size_t)j * W), ep = LD(E + (size_t)j * W);
            const V hd = hdiag;
            hdiag = htop;
            const V s2 = LD(qs + (size_t)j * W);
            const V sc = SEL8(EQ8(s1, s2), vmatch, vmis);
            V m = ADDS8(hd, sc);
            m = SEL8(MSB8(OR(s1, s2)), zero, m); /* 0xFF padding on either side */
            m = MAX8(m, zero);
            const V e = MAX8(SUBS8(m, oed), SUBS8(ep, ed));
            f = MAX8(SUBS8(m, oei), SUBS8(f, ei));
            V h = MAX8(MAX8(m, e), f);
            const V vj = S8((int8_t)j);
            const M8 in = MAND8(MAND8(GT8(vj, bm1), GT8(ve, vj)), tmask);
            h = SEL8(in, h, zero);
            ST(H + (size_t)j * W, h);
            ST(E + (size_t)j * W, SEL8(in, e, zero));
            const M8 gt = GT8(h, maxs);
            maxs = MAX8(h, maxs);
            maxi = SEL8(gt, vi, maxi);
            maxj = SEL8(gt, vj, maxj);
        }
        if (zdrop > 0) { /* :228-259, per lane after the row */
            ST(msv, maxs);
            for (int k = 0; k < lanes; ++k) {
                if (term[k] || i == 0 || i >= tl[k]) continue;
                int rm = 0;
                for (int j = cb[k] < 0 ? qmax : cb[k]; j < ce[k]; ++j)
                    rm = rm > H[(size_t)j * W + k] ? rm : H[(size_t)j * W + k];
                if (rm == 0 || msv[k] - rm > zdrop) {
                    term[k] = 1;
                    ++nterm;
                }
            }
        }
    }
    int8_t a[64], b[64], c[64];
    ST(a, maxs);
    ST(b, maxi);
    ST(c, maxj);
    for (int k = 0; k < lanes; ++k) {
        int32_t *o = out + 6 * k;
        o[0] = a[k];
        o[1] = b[k];
        o[2] = 0;
        o[3] = c[k];
        o[4] = 0;
        o[5] = 0;
    }
}

/* qs/ts: i16 SoA codes (the byte SoA widened, 0xFF -> 255); H, E: qmax * RS_W16 int16 */
RS_TARGET static void RS_CAT(rs_kern16_, RS_SFX)(const int16_t *qs, const int16_t *ts, int lanes, int qmax,
                                                 int tmax, const int8_t *ql, const int8_t *tl,
                                                 const int16_t *h0, const int8_t *wv, int o_del,
                                                 int e_del, int o_ins, int e_ins, int zdrop,
                                                 int16_t match, int16_t mis, int16_t amb, int16_t *H,
                                                 int16_t *E, int32_t *out) {
    const int W = RS_W16;
    const V zero = ZERO();
    const V vmatch = S16(match), vmis = S16(mis), vamb = S16(amb), three = S16(3), one = S16(1);
    const V oed = S16((int16_t)(o_del + e_del)), oei = S16((int16_t)(o_ins + e_ins));
    const V ed = S16((int16_t)e_del), ei = S16((int16_t)e_ins);
    int16_t lane_v[32] = {0}, beg[32] = {0}, end[32] = {0}, tm[32] = {0}, msv[32], rmv[32];
    int term[32] = {0}, nterm = 0;
    /* first row (:94-108): saturating steps clamped at 0 */
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? h0[k] : 0;
    const V hv = LD(lane_v);
    if (qmax > 0) ST(H, hv);
    V hp = MAX16(SUBS16(hv, oei), zero);
    if (qmax > 1) ST(H + W, hp);
    for (int j = 2; j < qmax; ++j) {
        hp = MAX16(SUBS16(hp, ei), zero);
        ST(H + (size_t)j * W, hp);
    }
    for (int j = 0; j < qmax; ++j) ST(E + (size_t)j * W, zero);
    V maxs = hv, maxi = S16(-1), maxj = S16(-1);
    for (int k = 0; k < lanes; ++k) end[k] = ql[k];
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? ql[k] : 0;
    const V vql = LD(lane_v);
    for (int k = 0; k < W; ++k) lane_v[k] = k < lanes ? wv[k] : 0;
    const V vw = LD(lane_v);
    for (int i = 0; i < tmax; ++i) {
        if (nterm == lanes) break;
        V f = zero, hdiag = LD(H);
        const V s1 = LD(ts + (size_t)i * W);
        const V vi = S16((int16_t)i);
        const V vb = MAX16(LD(beg), SUBS16(vi, vw));
        V ve = MIN16(LD(end), ADDS16(ADDS16(vi, vw), one));
        ve = MIN16(ve, vql);
        ST(beg, vb); /* the band is written back (:150-160) */
        ST(end, ve);
        for (int k = 0; k < lanes; ++k) tm[k] = (!term[k] && i < tl[k]) ? -1 : 0;
        const M16 tmask = MFROM16(tm);
        const V bm1 = SUBS16(vb, one);
        V rowmax = zero;
        for (int j = 0; j < qmax; ++j) {
            const V htop = LD(H + (size_t)j * W), ep = LD(E + (size_t)j * W);
            const V hd = hdiag;
            hdiag = htop;
            const V s2 = LD(qs + (size_t)j * W);
            V sc = SEL16(EQ16(s1, s2), vmatch, vmis);
            sc = SEL16(MOR16(GT16(s2, three), GT16(s1, three)), vamb, sc);
            const V m = MAX16(ADDS16(hd, sc), zero);
            const V e = MAX16(MAX16(SUBS16(m, oed), zero), SUBS16(ep, ed));
            f = MAX16(MAX16(SUBS16(m, oei), zero), SUBS16(f, ei));
            V h = MAX16(MAX16(MAX16(m, e), f), zero);
            const V vj = S16((int16_t)j);
            const M16 in = MAND16(MAND16(GT16(vj, bm1), GT16(ve, vj)), tmask);
            h = SEL16(in, h, zero);
            ST(H + (size_t)j * W, h);
            ST(E + (size_t)j * W, SEL16(in, e, zero));
            const M16 gt = GT16(h, maxs);
            maxs = MAX16(h, maxs);
            maxi = SEL16(gt, vi, maxi);
            maxj = SEL16(gt, vj, maxj);
            rowmax = MAX16(rowmax, h);
        }
        if (zdrop > 0) { /* :233-247 */
            ST(msv, maxs);
            ST(rmv, rowmax);
            for (int k = 0; k < lanes; ++k)
                if (!term[k] && i > 0 && i < tl[k] && msv[k] - rmv[k] > zdrop) {
                    term[k] = 1;
                    ++nterm;
                }
        }
    }
    int16_t a[32], b[32], c[32];
    ST(a, maxs);
    ST(b, maxi);
    ST(c, maxj);
    for (int k = 0; k < lanes; ++k) {
        int32_t *o = out + 6 * k;
        o[0] = a[k];
        o[1] = b[k];
        o[2] = 0;
        o[3] = c[k];
        o[4] = 0;
        o[5] = 0;
    }
}
