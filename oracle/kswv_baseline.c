/* CPU timing baseline: a C restatement of the reference's horizontal SIMD
 * mate-rescue path (kswv), run over every 32-job chunk of a batch on all
 * host threads.  TEST / BENCH INFRASTRUCTURE ONLY (bench.py's cpu_baseline
 * leg and tests/test_kswv_baseline.py); the library never calls it.
 *
 *   caller        paired/mate_rescue.rs:704-783 (execute_mate_rescue_batch_simd:
 *                 chunks of the engine's batch size in input order, h0 = 0,
 *                 ambig 0, match 1, mismatch -4, gaps 6/1/6/1)
 *   dispatch      batch_extension/dispatch.rs:270-395 (dispatch_kswv_soa)
 *   SoA transpose core/alignment/workspace.rs:768-840 (0xFF padding, lane
 *                 lengths min(len, 127), h0 as the lane's xtra)
 *   entry         core/alignment/kswv/shared.rs:8-89 (32 SeqPairs, inactive
 *                 lanes zero, nrow / ncol = the chunk's longest target / query)
 *   kernel        core/alignment/kswv_avx2.rs:93-635 (batch_ksw_align_avx2)
 *
 * The reference's default x86 build runs this AVX2 kernel (32 x u8 lanes);
 * its AVX-512 kernel needs the non-default `avx512` cargo feature.
 * Semantics kept on purpose: full nrow x ncol sweep per chunk, scores biased
 * by shift with unsigned saturation, the row-maximum log zeroed for lanes
 * still running, second-best windows taken from every other lane and mixed
 * 16-bit / 8-bit blend masks, tb = te - qe and qb = 0 as the approximate
 * start.  One departure: a chunk whose targets are all empty returns zero
 * records (the reference indexes row -1 there).
 *
 * kv_chunk_scalar restates the same chunk semantics lane by lane in plain C
 * and pins the vector code (tests/test_kswv_baseline.py). */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define KW 32 /* u8 lanes of one AVX2 register = jobs per chunk */
#define KV_XSUBO 0x40000
#define KV_XSTOP 0x20000

typedef struct {
    int32_t score, te, qe, score2, te2, tb, qb;
} kv_res; /* KswResult, kswv_batch.rs:58-69 */

typedef struct {
    int8_t match, mis, ambig;
    int32_t o_del, e_del, o_ins, e_ins;
} kv_score;

/* scoring table of the shuffle (kswv_avx2.rs:126-143): xor 0 match, 1-3
 * mismatch, 4-12 ambig, + shift (wrapping) in the first 16 bytes */
static void kv_table(const kv_score *sc, int8_t temp[KW], uint8_t *shift, int *qmax) {
    int mdiff = sc->match, sv = sc->match;
    if (sc->mis > mdiff) mdiff = sc->mis;
    if (sc->ambig > mdiff) mdiff = sc->ambig;
    if (sc->mis < sv) sv = sc->mis;
    if (sc->ambig < sv) sv = sc->ambig;
    *shift = (uint8_t)(256 - sv);
    *qmax = mdiff;
    memset(temp, 0, KW);
    temp[0] = sc->match;
    temp[1] = temp[2] = temp[3] = sc->mis;
    for (int i = 4; i <= 12; ++i) temp[i] = sc->ambig;
    for (int i = 0; i < 16; ++i) temp[i] = (int8_t)(temp[i] + (int8_t)*shift);
    for (int i = 16; i < KW; ++i) temp[i] = temp[i - 16];
}

/* minsc / endsc per lane from the lane's xtra (:151-178) */
static void kv_thresholds(const int32_t *xtra, uint8_t minsc[KW], uint8_t endsc[KW]) {
    for (int i = 0; i < KW; ++i) {
        minsc[i] = 0;
        endsc[i] = 255;
        const int32_t x = xtra[i];
        const uint32_t v = (x & KV_XSUBO) ? (uint32_t)(x & 0xffff) : 0x10000u;
        if (v <= 255) minsc[i] = (uint8_t)v;
        if (x & KV_XSTOP) {
            const uint32_t e = (uint32_t)(x & 0xffff);
            if (e <= 255) endsc[i] = (uint8_t)e;
        }
    }
}

/* ---- AVX2 kernel --------------------------------------------------------- */

#define KV_AVX2 __attribute__((target("avx2")))

KV_AVX2 static inline __m256i kv_gt_u8(__m256i a, __m256i b) {
    const __m256i s = _mm256_set1_epi8((char)0x80);
    return _mm256_cmpgt_epi8(_mm256_xor_si256(a, s), _mm256_xor_si256(b, s));
}
KV_AVX2 static inline __m256i kv_ge_u8(__m256i a, __m256i b) {
    return _mm256_cmpeq_epi8(_mm256_max_epu8(a, b), a);
}
#define LDV(p) _mm256_loadu_si256((const __m256i *)(p))
#define STV(p, v) _mm256_storeu_si256((__m256i *)(p), (v))

/* s1: target SoA [nrow][KW], s2: query SoA [ncol][KW]; rlen / xtra per lane;
 * h0, h1, f: (ncol + 1) * KW bytes, rm: (nrow + 1) * KW bytes */
KV_AVX2 static void kv_chunk_avx2(const uint8_t *s1soa, const uint8_t *s2soa, int nrow, int ncol,
                                  const int32_t *rlen, const int32_t *xtra, const kv_score *sc,
                                  uint8_t *h0b, uint8_t *h1b, uint8_t *fb, uint8_t *rmb,
                                  kv_res *res) {
    int8_t temp[KW];
    uint8_t shift;
    int qmax;
    kv_table(sc, temp, &shift, &qmax);
    const __m256i zero = _mm256_setzero_si256(), one = _mm256_set1_epi8(1), ones = _mm256_set1_epi8(-1);
    const __m256i perm = LDV(temp), sft = _mm256_set1_epi8((char)shift);
    const __m256i cmax = _mm256_set1_epi8((char)255), five = _mm256_set1_epi8(5);
    uint8_t minsc[KW], endsc[KW];
    kv_thresholds(xtra, minsc, endsc);
    const __m256i minsc256 = LDV(minsc), endsc256 = LDV(endsc);
    const __m256i e_del = _mm256_set1_epi8((char)sc->e_del);
    const __m256i oe_del = _mm256_set1_epi8((char)(sc->o_del + sc->e_del));
    const __m256i e_ins = _mm256_set1_epi8((char)sc->e_ins);
    const __m256i oe_ins = _mm256_set1_epi8((char)(sc->o_ins + sc->e_ins));
    __m256i gmax = zero, te_lo = _mm256_set1_epi16(-1), te_hi = _mm256_set1_epi16(-1);
    memset(h0b, 0, (size_t)(ncol + 1) * KW);
    memset(h1b, 0, (size_t)(ncol + 1) * KW);
    memset(fb, 0, (size_t)(ncol + 1) * KW);
    memset(rmb, 0, (size_t)(nrow + 1) * KW);
    __m256i pimax = zero, mask = zero, minsc_msk = zero, qe = zero;
    __m256i exit0 = ones;
    int limit = nrow;
    uint8_t *H0 = h0b, *H1 = h1b;
    for (int i = 0; i < nrow; ++i) {
        __m256i e11 = zero, imax = zero, iqe = ones, l256 = zero;
        const __m256i i256 = _mm256_set1_epi16((short)i);
        const __m256i s1 = LDV(s1soa + (size_t)i * KW);
        for (int j = 0; j < ncol; ++j) {
            const __m256i h00 = LDV(H0 + (size_t)j * KW), s2 = LDV(s2soa + (size_t)j * KW);
            const __m256i f11 = LDV(fb + (size_t)(j + 1) * KW);
            __m256i sbt = _mm256_shuffle_epi8(perm, _mm256_xor_si256(s1, s2));
            sbt = _mm256_blendv_epi8(sbt, sft, _mm256_cmpeq_epi8(s2, five));
            __m256i m11 = _mm256_adds_epu8(h00, sbt);
            m11 = _mm256_blendv_epi8(m11, zero, _mm256_or_si256(s1, s2));
            m11 = _mm256_subs_epu8(m11, sft);
            __m256i h11 = _mm256_max_epu8(_mm256_max_epu8(m11, e11), f11);
            const __m256i c0 = kv_gt_u8(h11, imax);
            imax = _mm256_max_epu8(imax, h11);
            iqe = _mm256_blendv_epi8(iqe, l256, c0);
            e11 = _mm256_max_epu8(_mm256_subs_epu8(h11, oe_ins), _mm256_subs_epu8(e11, e_ins));
            const __m256i f21 = _mm256_max_epu8(_mm256_subs_epu8(h11, oe_del), _mm256_subs_epu8(f11, e_del));
            STV(H1 + (size_t)(j + 1) * KW, h11);
            STV(fb + (size_t)(j + 1) * KW, f21);
            l256 = _mm256_add_epi8(l256, one);
        }
        if (i > 0) { /* :337-352 */
            __m256i msk = _mm256_or_si256(kv_gt_u8(imax, pimax), mask);
            __m256i t = _mm256_blendv_epi8(pimax, zero, msk);
            t = _mm256_blendv_epi8(t, zero, exit0);
            STV(rmb + (size_t)(i - 1) * KW, t);
            mask = _mm256_andnot_si256(msk, ones);
        }
        pimax = imax;
        minsc_msk = _mm256_or_si256(minsc_msk, kv_ge_u8(imax, minsc256));
        __m256i c0 = _mm256_and_si256(kv_gt_u8(imax, gmax), exit0);
        gmax = _mm256_blendv_epi8(gmax, imax, c0);
        qe = _mm256_blendv_epi8(qe, iqe, c0);
        const __m128i lo = _mm256_extracti128_si256(c0, 0), hi = _mm256_extracti128_si256(c0, 1);
        const __m256i mlo = _mm256_set_m128i(_mm_unpackhi_epi8(lo, lo), _mm_unpacklo_epi8(lo, lo));
        const __m256i mhi = _mm256_set_m128i(_mm_unpackhi_epi8(hi, hi), _mm_unpacklo_epi8(hi, hi));
        te_lo = _mm256_blendv_epi8(te_lo, i256, mlo);
        te_hi = _mm256_blendv_epi8(te_hi, i256, mhi);
        const __m256i c1 = kv_ge_u8(gmax, endsc256);
        const __m256i c2 = kv_ge_u8(_mm256_adds_epu8(gmax, sft), cmax);
        exit0 = _mm256_andnot_si256(_mm256_or_si256(c1, c2), exit0);
        if (_mm256_movemask_epi8(exit0) == 0) {
            limit = i + 1;
            break;
        }
        uint8_t *tmp = H0;
        H0 = H1;
        H1 = tmp;
    }
    (void)minsc_msk;
    STV(rmb + (size_t)(limit - 1) * KW, _mm256_blendv_epi8(pimax, zero, mask));
    uint8_t score_a[KW], qe_a[KW];
    int16_t te_a[KW];
    STV(score_a, gmax);
    STV(te_a, te_lo);
    STV(te_a + 16, te_hi);
    STV(qe_a, qe);
    memset(res, 0, sizeof(kv_res) * KW);
    int live = 0;
    for (int l = 0; l < KW; ++l) {
        const int s = score_a[l];
        res[l].score = s + shift < 255 ? s : 255;
        res[l].te = te_a[l];
        res[l].qe = qe_a[l];
        if (res[l].score != 255) ++live;
    }
    if (live == 0) return;
    /* second best (:521-610): windows from lanes 0, 2, 4, ... */
    int16_t low[16], high[16], rl[16];
    int maxl = 0, minh = nrow;
    for (int i = 0; i < 16; ++i) {
        const int seq = 2 * i;
        const int val = (score_a[seq] + qmax - 1) / qmax;
        const int16_t tv = te_a[seq];
        int16_t lo16 = (int16_t)(tv - (int16_t)val), hi16 = (int16_t)(tv + (int16_t)val);
        low[i] = lo16 > 0 ? lo16 : 0;
        high[i] = hi16 < nrow - 1 ? hi16 : (int16_t)(nrow - 1);
        rl[i] = (int16_t)rlen[seq];
        if (qe_a[seq] != 0) {
            if (low[i] > maxl) maxl = low[i];
            if (high[i] < minh) minh = high[i];
        }
    }
    __m256i mx = zero, te2 = _mm256_set1_epi16(-1);
    const __m256i low256 = LDV(low), high256 = LDV(high), rlen256 = LDV(rl);
    for (int i = 0; i < maxl; ++i) {
        const __m256i i256 = _mm256_set1_epi16((short)i), rmax = LDV(rmb + (size_t)i * KW);
        const __m256i cm = _mm256_and_si256(_mm256_cmpgt_epi16(low256, i256), kv_gt_u8(rmax, mx));
        mx = _mm256_blendv_epi8(mx, rmax, cm);
        te2 = _mm256_blendv_epi8(te2, i256, cm);
    }
    for (int i = minh + 1 > 0 ? minh + 1 : 0; i < limit; ++i) {
        const __m256i i256 = _mm256_set1_epi16((short)i), rmax = LDV(rmb + (size_t)i * KW);
        const __m256i m1 = _mm256_and_si256(_mm256_cmpgt_epi16(i256, high256), _mm256_cmpgt_epi16(rlen256, i256));
        const __m256i cm = _mm256_and_si256(m1, kv_gt_u8(rmax, mx));
        mx = _mm256_blendv_epi8(mx, rmax, cm);
        te2 = _mm256_blendv_epi8(te2, i256, cm);
    }
    uint8_t s2a[KW];
    int16_t t2a[16];
    STV(s2a, mx);
    STV(t2a, te2);
    for (int i = 0; i < KW; ++i) {
        if (qe_a[i] != 0) {
            res[i].score2 = s2a[i] == 0 ? -1 : s2a[i];
            res[i].te2 = t2a[i / 2];
        } else {
            res[i].score2 = res[i].te2 = -1;
        }
        if (res[i].score > 0 && res[i].te >= 0 && res[i].qe >= 0) {
            res[i].qb = 0;
            res[i].tb = res[i].te - res[i].qe > 0 ? res[i].te - res[i].qe : 0;
        } else {
            res[i].tb = res[i].qb = -1;
        }
    }
}

/* ---- the same chunk semantics, lane by lane in plain C ------------------- */

static inline uint8_t u_adds(uint8_t a, uint8_t b) { return (uint8_t)(a + b > 255 ? 255 : a + b); }
static inline uint8_t u_subs(uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : 0); }
static inline uint8_t u_max(uint8_t a, uint8_t b) { return a > b ? a : b; }

void kv_chunk_scalar(const uint8_t *s1soa, const uint8_t *s2soa, int nrow, int ncol,
                     const int32_t *rlen, const int32_t *xtra, const kv_score *sc, uint8_t *h0b,
                     uint8_t *h1b, uint8_t *fb, uint8_t *rmb, kv_res *res) {
    int8_t temp[KW];
    uint8_t shift;
    int qmax;
    kv_table(sc, temp, &shift, &qmax);
    uint8_t minsc[KW], endsc[KW];
    kv_thresholds(xtra, minsc, endsc);
    const uint8_t ed = (uint8_t)sc->e_del, oed = (uint8_t)(sc->o_del + sc->e_del);
    const uint8_t ei = (uint8_t)sc->e_ins, oei = (uint8_t)(sc->o_ins + sc->e_ins);
    memset(h0b, 0, (size_t)(ncol + 1) * KW);
    memset(h1b, 0, (size_t)(ncol + 1) * KW);
    memset(fb, 0, (size_t)(ncol + 1) * KW);
    memset(rmb, 0, (size_t)(nrow + 1) * KW);
    uint8_t gmax[KW] = {0}, pimax[KW] = {0}, mask[KW] = {0}, qe[KW] = {0}, exit0[KW];
    int16_t te[KW];
    for (int l = 0; l < KW; ++l) {
        exit0[l] = 0xFF;
        te[l] = -1;
    }
    int limit = nrow;
    uint8_t *H0 = h0b, *H1 = h1b;
    for (int i = 0; i < nrow; ++i) {
        uint8_t imax[KW], iqe[KW];
        for (int l = 0; l < KW; ++l) {
            uint8_t e = 0, im = 0, iq = 0xFF;
            const uint8_t s1 = s1soa[(size_t)i * KW + l];
            for (int j = 0; j < ncol; ++j) {
                const uint8_t s2 = s2soa[(size_t)j * KW + l];
                const uint8_t x = s1 ^ s2;
                uint8_t sbt = (x & 0x80) ? 0 : (uint8_t)temp[x & 15]; /* pshufb */
                if (s2 == 5) sbt = shift;
                uint8_t m = u_adds(H0[(size_t)j * KW + l], sbt);
                if ((s1 | s2) & 0x80) m = 0;
                m = u_subs(m, shift);
                const uint8_t f = fb[(size_t)(j + 1) * KW + l];
                const uint8_t h = u_max(u_max(m, e), f);
                if (h > im) iq = (uint8_t)j;
                im = u_max(im, h);
                e = u_max(u_subs(h, oei), u_subs(e, ei));
                H1[(size_t)(j + 1) * KW + l] = h;
                fb[(size_t)(j + 1) * KW + l] = u_max(u_subs(h, oed), u_subs(f, ed));
            }
            imax[l] = im;
            iqe[l] = iq;
        }
        int any = 0;
        for (int l = 0; l < KW; ++l) {
            if (i > 0) {
                const int msk = imax[l] > pimax[l] || mask[l];
                uint8_t t = msk ? 0 : pimax[l];
                if (exit0[l]) t = 0;
                rmb[(size_t)(i - 1) * KW + l] = t;
                mask[l] = msk ? 0 : 0xFF;
            }
            pimax[l] = imax[l];
            if (imax[l] > gmax[l] && exit0[l]) {
                gmax[l] = imax[l];
                qe[l] = iqe[l];
                te[l] = (int16_t)i;
            }
            if (gmax[l] >= endsc[l] || u_adds(gmax[l], shift) >= 255) exit0[l] = 0;
            any |= exit0[l];
        }
        if (!any) {
            limit = i + 1;
            break;
        }
        uint8_t *tmp = H0;
        H0 = H1;
        H1 = tmp;
    }
    for (int l = 0; l < KW; ++l) rmb[(size_t)(limit - 1) * KW + l] = mask[l] ? 0 : pimax[l];
    memset(res, 0, sizeof(kv_res) * KW);
    int live = 0;
    for (int l = 0; l < KW; ++l) {
        res[l].score = gmax[l] + shift < 255 ? gmax[l] : 255;
        res[l].te = te[l];
        res[l].qe = qe[l];
        if (res[l].score != 255) ++live;
    }
    if (live == 0) return;
    int16_t low[16], high[16], rl[16];
    int maxl = 0, minh = nrow;
    for (int i = 0; i < 16; ++i) {
        const int seq = 2 * i;
        const int val = (gmax[seq] + qmax - 1) / qmax;
        const int16_t lo16 = (int16_t)(te[seq] - (int16_t)val), hi16 = (int16_t)(te[seq] + (int16_t)val);
        low[i] = lo16 > 0 ? lo16 : 0;
        high[i] = hi16 < nrow - 1 ? hi16 : (int16_t)(nrow - 1);
        rl[i] = (int16_t)rlen[seq];
        if (qe[seq] != 0) {
            if (low[i] > maxl) maxl = low[i];
            if (high[i] < minh) minh = high[i];
        }
    }
    /* byte-wise blends: byte b of the 16-bit window mask belongs to lane b / 2,
     * byte b of the score mask to sequence b; te2 is 16 little-endian 16-bit lanes */
    uint8_t mx[KW] = {0}, t2b[KW];
    memset(t2b, 0xFF, sizeof t2b);
    for (int pass = 0; pass < 2; ++pass) {
        const int i0 = pass ? (minh + 1 > 0 ? minh + 1 : 0) : 0, i1 = pass ? limit : maxl;
        for (int i = i0; i < i1; ++i) {
            const uint8_t ib[2] = {(uint8_t)(i & 0xFF), (uint8_t)((i >> 8) & 0xFF)};
            for (int b = 0; b < KW; ++b) {
                const int k = b / 2;
                const int win = pass ? (i > high[k] && rl[k] > i) : (low[k] > i);
                const uint8_t r = rmb[(size_t)i * KW + b];
                if (win && r > mx[b]) {
                    mx[b] = r;
                    t2b[b] = ib[b & 1];
                }
            }
        }
    }
    for (int i = 0; i < KW; ++i) {
        if (qe[i] != 0) {
            res[i].score2 = mx[i] == 0 ? -1 : mx[i];
            const int k = i / 2;
            res[i].te2 = (int16_t)(t2b[2 * k] | (t2b[2 * k + 1] << 8));
        } else {
            res[i].score2 = res[i].te2 = -1;
        }
        if (res[i].score > 0 && res[i].te >= 0 && res[i].qe >= 0) {
            res[i].qb = 0;
            res[i].tb = res[i].te - res[i].qe > 0 ? res[i].te - res[i].qe : 0;
        } else {
            res[i].tb = res[i].qb = -1;
        }
    }
}

/* ---- batch driver -------------------------------------------------------- */

typedef struct {
    int64_t c0, c1; /* chunk range */
    int64_t n;
    const uint8_t *q, *t;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen, *xtra;
    kv_score sc;
    int simd;
    kv_res *out;
} kv_job;

static void *kv_worker(void *arg) {
    kv_job *jb = (kv_job *)arg;
    size_t hcap = 0, rcap = 0, qcap = 0, tcap = 0;
    uint8_t *h0 = NULL, *h1 = NULL, *f = NULL, *rm = NULL, *qs = NULL, *ts = NULL;
    for (int64_t c = jb->c0; c < jb->c1; ++c) {
        const int64_t k0 = c * KW;
        const int cnt = (int)(jb->n - k0 < KW ? jb->n - k0 : KW);
        int mq = 0, mt = 0;
        int32_t rlen[KW] = {0}, xtra[KW] = {0};
        for (int l = 0; l < cnt; ++l) {
            const int ql = jb->qlen[k0 + l], tl = jb->tlen[k0 + l];
            if (ql > mq) mq = ql;
            if (tl > mt) mt = tl;
            rlen[l] = tl < 127 ? tl : 127; /* ksw_tlen as i8 */
            xtra[l] = jb->xtra ? (jb->xtra[k0 + l] < 127 ? jb->xtra[k0 + l] : 127) : 0;
        }
        /* workspace.rs:805-824: SoA with 0xFF padding */
        const size_t qn = (size_t)KW * (mq > 0 ? mq : 1), tn = (size_t)KW * (mt > 0 ? mt : 1);
        if (qn > qcap) { free(qs); qs = (uint8_t *)malloc(qn); qcap = qn; }
        if (tn > tcap) { free(ts); ts = (uint8_t *)malloc(tn); tcap = tn; }
        if (!qs || !ts) break; /* out of memory: the remaining records stay as they are */
        memset(qs, 0xFF, qn);
        memset(ts, 0xFF, tn);
        for (int l = 0; l < cnt; ++l) {
            const uint8_t *qq = jb->q + jb->qoff[k0 + l], *tt = jb->t + jb->toff[k0 + l];
            for (int p = 0; p < jb->qlen[k0 + l]; ++p) qs[(size_t)p * KW + l] = qq[p];
            for (int p = 0; p < jb->tlen[k0 + l]; ++p) ts[(size_t)p * KW + l] = tt[p];
        }
        kv_res res[KW];
        if (mt == 0) {
            memset(res, 0, sizeof res);
        } else {
            const size_t hn = (size_t)(mq + 1) * KW, rn = (size_t)(mt + 1) * KW;
            if (hn > hcap) {
                free(h0); free(h1); free(f);
                h0 = (uint8_t *)malloc(hn);
                h1 = (uint8_t *)malloc(hn);
                f = (uint8_t *)malloc(hn);
                hcap = hn;
            }
            if (rn > rcap) { free(rm); rm = (uint8_t *)malloc(rn); rcap = rn; }
            if (!h0 || !h1 || !f || !rm) break;
            if (jb->simd)
                kv_chunk_avx2(ts, qs, mt, mq, rlen, xtra, &jb->sc, h0, h1, f, rm, res);
            else
                kv_chunk_scalar(ts, qs, mt, mq, rlen, xtra, &jb->sc, h0, h1, f, rm, res);
        }
        memcpy(jb->out + k0, res, sizeof(kv_res) * (size_t)cnt);
    }
    free(h0); free(h1); free(f); free(rm); free(qs); free(ts);
    return NULL;
}

/* Every job of a batch through the kswv path: chunks of 32 jobs in input
 * order, split over nthreads (<= 0: all cores) in contiguous runs of chunks.
 * xtra: per-job h0 (the lane's xtra), or NULL for 0 as mate rescue passes.
 * simd 1: the AVX2 kernel (the caller checks the CPU), 0: the plain-C twin.
 * Returns 0, or -1 when AVX2 was asked for and is absent, or on bad input. */
int kv_batch(int64_t n, const uint8_t *q, const int64_t *qoff, const int32_t *qlen,
             const uint8_t *t, const int64_t *toff, const int32_t *tlen, const int32_t *xtra,
             int32_t match, int32_t mis, int32_t ambig, int32_t o_del, int32_t e_del,
             int32_t o_ins, int32_t e_ins, int simd, kv_res *out, int nthreads) {
    if (n <= 0) return 0;
    if (simd) {
        __builtin_cpu_init();
        if (!__builtin_cpu_supports("avx2")) return -1;
    }
    int mdiff = match;
    if (mis > mdiff) mdiff = mis;
    if (ambig > mdiff) mdiff = ambig;
    if (mdiff <= 0) return -1; /* the reference divides by it */
    for (int64_t k = 0; k < n; ++k)
        if (qlen[k] < 0 || qlen[k] > 255 || tlen[k] < 0 || tlen[k] > 32767) return -1;
    const int64_t chunks = (n + KW - 1) / KW;
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads > 256) nthreads = 256;
    if (nthreads > chunks) nthreads = (int)chunks;
    if (nthreads < 1) nthreads = 1;
    kv_job jobs[256];
    pthread_t th[256];
    const int64_t per = (chunks + nthreads - 1) / nthreads;
    const kv_score sc = {(int8_t)match, (int8_t)mis, (int8_t)ambig, o_del, e_del, o_ins, e_ins};
    for (int x = 0; x < nthreads; ++x) {
        const int64_t c0 = per * x, c1 = per * (x + 1) < chunks ? per * (x + 1) : chunks;
        kv_job jb = {c0, c1 > c0 ? c1 : c0, n, q, t, qoff, toff, qlen, tlen, xtra, sc, simd, out};
        jobs[x] = jb;
    }
    for (int x = 1; x < nthreads; ++x) pthread_create(&th[x], NULL, kv_worker, &jobs[x]);
    kv_worker(&jobs[0]);
    for (int x = 1; x < nthreads; ++x) pthread_join(th[x], NULL);
    return 0;
}
