/* CPU timing baseline: a C restatement of the reference's SIMD batch path
 * for seed extension, run over every chunk of a batch on all host threads.
 * TEST / BENCH INFRASTRUCTURE ONLY (bench.py's cpu_baseline leg and
 * tests/test_refsimd_baseline.py); the library never calls it.
 *
 *   batch policy   batch_extension/dispatch.rs:95-268 (i8 kernel when every
 *                  job has max(qlen, tlen) <= 128, else i16; lane lengths
 *                  min(len, 127) as i8, h0 as i8 / i16, w as i8)
 *   SoA chunks     batch_extension/soa.rs:11-71 (W jobs per chunk in input
 *                  order, rows capped at 128, 0xFF padding)
 *   i8 kernel      banded_swa/kernel.rs:29-284
 *   i16 kernel     banded_swa/kernel_i16.rs:44-272
 *
 * Two departures, both so that every job gets a result: every chunk is
 * scored (the reference dispatch scores only the first W jobs,
 * dispatch.rs:126-146), and a chunk sweeps its own capped qmax x tmax (the
 * reference passes the whole batch's uncapped maxima with the capped SoA).
 * The E rows start at zero (the reference's i8 kernel reuses a workspace).
 *
 * The SIMD kernels are refsimd_kernel.h instantiated for AVX-512BW (64 x i8,
 * 32 x i16 lanes) and AVX2 (32 x i8, 16 x i16); rs_chunk8_scalar /
 * rs_chunk16_scalar restate the same chunk semantics lane by lane in plain C
 * and pin the vector code (tests/test_refsimd_baseline.py). */
#include <immintrin.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- AVX-512BW instantiation ---- */
#define RS_SFX avx512
#define RS_TARGET __attribute__((target("avx512f,avx512bw")))
#define V __m512i
#define M8 __mmask64
#define M16 __mmask32
#define RS_W8 64
#define RS_W16 32
#define LD(p) _mm512_loadu_si512((const void *)(p))
#define ST(p, v) _mm512_storeu_si512((void *)(p), (v))
#define ZERO() _mm512_setzero_si512()
#define OR(a, b) _mm512_or_si512((a), (b))
#define S8(x) _mm512_set1_epi8((char)(x))
#define ADDS8(a, b) _mm512_adds_epi8((a), (b))
#define SUBS8(a, b) _mm512_subs_epi8((a), (b))
#define MAX8(a, b) _mm512_max_epi8((a), (b))
#define MINU8(a, b) _mm512_min_epu8((a), (b))
#define EQ8(a, b) _mm512_cmpeq_epi8_mask((a), (b))
#define GT8(a, b) _mm512_cmpgt_epi8_mask((a), (b))
#define SEL8(m, a, b) _mm512_mask_blend_epi8((m), (b), (a))
#define MSB8(v) _mm512_movepi8_mask(v)
#define MAND8(a, b) ((a) & (b))
#define MFROM8(p) _mm512_movepi8_mask(LD(p))
#define S16(x) _mm512_set1_epi16((short)(x))
#define ADDS16(a, b) _mm512_adds_epi16((a), (b))
#define SUBS16(a, b) _mm512_subs_epi16((a), (b))
#define MAX16(a, b) _mm512_max_epi16((a), (b))
#define MIN16(a, b) _mm512_min_epi16((a), (b))
#define EQ16(a, b) _mm512_cmpeq_epi16_mask((a), (b))
#define GT16(a, b) _mm512_cmpgt_epi16_mask((a), (b))
#define SEL16(m, a, b) _mm512_mask_blend_epi16((m), (b), (a))
#define MAND16(a, b) ((a) & (b))
#define MOR16(a, b) ((a) | (b))
#define MFROM16(p) _mm512_movepi16_mask(LD(p))
#include "refsimd_kernel.h"
#undef RS_SFX
#undef RS_TARGET
#undef V
#undef M8
#undef M16
#undef RS_W8
#undef RS_W16
#undef LD
#undef ST
#undef ZERO
#undef OR
#undef S8
#undef ADDS8
#undef SUBS8
#undef MAX8
#undef MINU8
#undef EQ8
#undef GT8
#undef SEL8
#undef MSB8
#undef MAND8
#undef MFROM8
#undef S16
#undef ADDS16
#undef SUBS16
#undef MAX16
#undef MIN16
#undef EQ16
#undef GT16
#undef SEL16
#undef MAND16
#undef MOR16
#undef MFROM16

/* ---- AVX2 instantiation ---- */
#define RS_SFX avx2
#define RS_TARGET __attribute__((target("avx2")))
#define V __m256i
#define M8 __m256i
#define M16 __m256i
#define RS_W8 32
#define RS_W16 16
#define LD(p) _mm256_loadu_si256((const __m256i *)(const void *)(p))
#define ST(p, v) _mm256_storeu_si256((__m256i *)(void *)(p), (v))
#define ZERO() _mm256_setzero_si256()
#define OR(a, b) _mm256_or_si256((a), (b))
#define S8(x) _mm256_set1_epi8((char)(x))
#define ADDS8(a, b) _mm256_adds_epi8((a), (b))
#define SUBS8(a, b) _mm256_subs_epi8((a), (b))
#define MAX8(a, b) _mm256_max_epi8((a), (b))
#define MINU8(a, b) _mm256_min_epu8((a), (b))
#define EQ8(a, b) _mm256_cmpeq_epi8((a), (b))
#define GT8(a, b) _mm256_cmpgt_epi8((a), (b))
#define SEL8(m, a, b) _mm256_blendv_epi8((b), (a), (m))
#define MSB8(v) (v)
#define MAND8(a, b) _mm256_and_si256((a), (b))
#define MFROM8(p) LD(p)
#define S16(x) _mm256_set1_epi16((short)(x))
#define ADDS16(a, b) _mm256_adds_epi16((a), (b))
#define SUBS16(a, b) _mm256_subs_epi16((a), (b))
#define MAX16(a, b) _mm256_max_epi16((a), (b))
#define MIN16(a, b) _mm256_min_epi16((a), (b))
#define EQ16(a, b) _mm256_cmpeq_epi16((a), (b))
#define GT16(a, b) _mm256_cmpgt_epi16((a), (b))
#define SEL16(m, a, b) _mm256_blendv_epi8((b), (a), (m))
#define MAND16(a, b) _mm256_and_si256((a), (b))
#define MOR16(a, b) _mm256_or_si256((a), (b))
#define MFROM16(p) LD(p)
#include "refsimd_kernel.h"

/* ---- lane-by-lane restatement of the same chunk semantics ---- */
static int8_t sat8(int x) { return (int8_t)(x > 127 ? 127 : (x < -128 ? -128 : x)); }
static int16_t sat16(int x) { return (int16_t)(x > 32767 ? 32767 : (x < -32768 ? -32768 : x)); }
static int mx(int a, int b) { return a > b ? a : b; }

static void rs_chunk8_scalar(const int8_t *qs, const int8_t *ts, int W, int lanes, int qmax, int tmax,
                             const int8_t *ql, const int8_t *tl, const int8_t *h0, const int8_t *wv,
                             int o_del, int e_del, int o_ins, int e_ins, int zdrop, int8_t match,
                             int8_t mis, int8_t *H, int8_t *E, int32_t *out) {
    const int8_t oed = (int8_t)(o_del + e_del), oei = (int8_t)(o_ins + e_ins);
    const int8_t ed = (int8_t)e_del, ei = (int8_t)e_ins;
    int8_t maxs[64], maxi[64], maxj[64];
    int term[64] = {0}, nterm = 0;
    for (int k = 0; k < W; ++k) {
        const int8_t h = k < lanes ? h0[k] : 0;
        if (qmax > 0) H[k] = h;
        int8_t p = sat8(h - oei);
        if (qmax > 1) H[W + k] = p;
        for (int j = 2; j < qmax; ++j) H[(size_t)j * W + k] = p = sat8(p - ei);
        for (int j = 0; j < qmax; ++j) E[(size_t)j * W + k] = 0;
        maxs[k] = h;
        maxi[k] = maxj[k] = 0;
    }
    for (int i = 0; i < tmax; ++i) {
        if (nterm > lanes / 2) break;
        int cb[64], ce[64];
        for (int k = 0; k < W; ++k) {
            const int8_t w = k < lanes ? wv[k] : 0, q = k < lanes ? ql[k] : 0;
            const int8_t b = (int8_t)mx(0, sat8(i - w));
            const uint8_t e1 = (uint8_t)sat8(sat8(i + w) + 1);
            uint8_t e = (uint8_t)q < e1 ? (uint8_t)q : e1; /* unsigned minima with end = qlen */
            e = (uint8_t)q < e ? (uint8_t)q : e;
            cb[k] = b;
            ce[k] = (int8_t)e;
            const int on = k < lanes && !term[k] && i < tl[k];
            int8_t f = 0, hdiag = H[k];
            const int8_t s1 = ts[(size_t)i * W + k];
            for (int j = 0; j < qmax; ++j) {
                const int8_t htop = H[(size_t)j * W + k], ep = E[(size_t)j * W + k];
                const int8_t hd = hdiag;
                hdiag = htop;
                const int8_t s2 = qs[(size_t)j * W + k];
                int8_t m = sat8(hd + (s1 == s2 ? match : mis));
                if ((s1 | s2) < 0) m = 0;
                m = (int8_t)mx(m, 0);
                const int8_t ev = (int8_t)mx(sat8(m - oed), sat8(ep - ed));
                f = (int8_t)mx(sat8(m - oei), sat8(f - ei));
                int8_t h = (int8_t)mx(mx(m, ev), f);
                const int in = on && j > sat8(b - 1) && (int8_t)e > j;
                h = in ? h : 0;
                H[(size_t)j * W + k] = h;
                E[(size_t)j * W + k] = in ? ev : 0;
                if (h > maxs[k]) {
                    maxs[k] = h;
                    maxi[k] = (int8_t)i;
                    maxj[k] = (int8_t)j;
                }
            }
        }
        if (zdrop > 0)
            for (int k = 0; k < lanes; ++k) {
                if (term[k] || i == 0 || i >= tl[k]) continue;
                int rm = 0;
                for (int j = cb[k] < 0 ? qmax : cb[k]; j < ce[k]; ++j) rm = mx(rm, H[(size_t)j * W + k]);
                if (rm == 0 || maxs[k] - rm > zdrop) {
                    term[k] = 1;
                    ++nterm;
                }
            }
    }
    for (int k = 0; k < lanes; ++k) {
        int32_t *o = out + 6 * k;
        o[0] = maxs[k];
        o[1] = maxi[k];
        o[2] = 0;
        o[3] = maxj[k];
        o[4] = 0;
        o[5] = 0;
    }
}

static void rs_chunk16_scalar(const int16_t *qs, const int16_t *ts, int W, int lanes, int qmax, int tmax,
                              const int8_t *ql, const int8_t *tl, const int16_t *h0, const int8_t *wv,
                              int o_del, int e_del, int o_ins, int e_ins, int zdrop, int16_t match,
                              int16_t mis, int16_t amb, int16_t *H, int16_t *E, int32_t *out) {
    const int16_t oed = (int16_t)(o_del + e_del), oei = (int16_t)(o_ins + e_ins);
    const int16_t ed = (int16_t)e_del, ei = (int16_t)e_ins;
    int16_t maxs[32], maxi[32], maxj[32], beg[32] = {0}, end[32] = {0};
    int term[32] = {0}, nterm = 0;
    for (int k = 0; k < W; ++k) {
        const int16_t h = k < lanes ? h0[k] : 0;
        if (qmax > 0) H[k] = h;
        int16_t p = (int16_t)mx(sat16(h - oei), 0);
        if (qmax > 1) H[W + k] = p;
        for (int j = 2; j < qmax; ++j) H[(size_t)j * W + k] = p = (int16_t)mx(sat16(p - ei), 0);
        for (int j = 0; j < qmax; ++j) E[(size_t)j * W + k] = 0;
        maxs[k] = h;
        maxi[k] = maxj[k] = -1;
        end[k] = k < lanes ? ql[k] : 0;
    }
    for (int i = 0; i < tmax; ++i) {
        if (nterm == lanes) break;
        int16_t rowmax[32];
        for (int k = 0; k < W; ++k) {
            const int16_t w = k < lanes ? wv[k] : 0, q = k < lanes ? ql[k] : 0;
            const int16_t b = (int16_t)mx(beg[k], sat16(i - w));
            int16_t e = sat16(sat16(i + w) + 1);
            e = end[k] < e ? end[k] : e;
            e = q < e ? q : e;
            beg[k] = b;
            end[k] = e;
            const int on = k < lanes && !term[k] && i < tl[k];
            int16_t f = 0, hdiag = H[k], rm = 0;
            const int16_t s1 = ts[(size_t)i * W + k];
            for (int j = 0; j < qmax; ++j) {
                const int16_t htop = H[(size_t)j * W + k], ep = E[(size_t)j * W + k];
                const int16_t hd = hdiag;
                hdiag = htop;
                const int16_t s2 = qs[(size_t)j * W + k];
                int16_t sc = s1 == s2 ? match : mis;
                if (s2 > 3 || s1 > 3) sc = amb;
                const int16_t m = (int16_t)mx(sat16(hd + sc), 0);
                const int16_t ev = (int16_t)mx(mx(sat16(m - oed), 0), sat16(ep - ed));
                f = (int16_t)mx(mx(sat16(m - oei), 0), sat16(f - ei));
                int16_t h = (int16_t)mx(mx(mx(m, ev), f), 0);
                const int in = on && j > sat16(b - 1) && e > j;
                h = in ? h : 0;
                H[(size_t)j * W + k] = h;
                E[(size_t)j * W + k] = in ? ev : 0;
                if (h > maxs[k]) {
                    maxs[k] = h;
                    maxi[k] = (int16_t)i;
                    maxj[k] = (int16_t)j;
                }
                rm = (int16_t)mx(rm, h);
            }
            rowmax[k] = rm;
        }
        if (zdrop > 0)
            for (int k = 0; k < lanes; ++k)
                if (!term[k] && i > 0 && i < tl[k] && maxs[k] - rowmax[k] > zdrop) {
                    term[k] = 1;
                    ++nterm;
                }
    }
    for (int k = 0; k < lanes; ++k) {
        int32_t *o = out + 6 * k;
        o[0] = maxs[k];
        o[1] = maxi[k];
        o[2] = 0;
        o[3] = maxj[k];
        o[4] = 0;
        o[5] = 0;
    }
}

/* ---- batch driver ---- */
enum { RS_AUTO = 0, RS_AVX2 = 1, RS_AVX512 = 2, RS_SCALAR512 = 3, RS_SCALAR2 = 4 };

typedef struct {
    int64_t n;
    const uint8_t *qseq, *tseq;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen, *h0, *w;
    int o_del, e_del, o_ins, e_ins, zdrop;
    int8_t match, mis, amb;
    int isa, wide, W;
    int64_t nchunks;
    int64_t next; /* chunk counter (atomic) */
    int32_t *out;
} RsBatch;

static void rs_run_chunk(const RsBatch *b, int64_t c, uint8_t *qs, uint8_t *ts, int16_t *qs16, int16_t *ts16,
                         void *Hbuf, void *Ebuf) {
    const int W = b->W;
    const int64_t j0 = c * W;
    const int lanes = (int)(b->n - j0 < W ? b->n - j0 : W);
    int qmax = 0, tmax = 0;
    int8_t ql[64], tl[64], wv[64], h8[64];
    int16_t h16[64];
    for (int k = 0; k < lanes; ++k) {
        const int64_t j = j0 + k;
        qmax = mx(qmax, b->qlen[j]);
        tmax = mx(tmax, b->tlen[j]);
        ql[k] = (int8_t)(b->qlen[j] < 127 ? b->qlen[j] : 127);
        tl[k] = (int8_t)(b->tlen[j] < 127 ? b->tlen[j] : 127);
        wv[k] = (int8_t)b->w[j];
        h8[k] = (int8_t)b->h0[j];
        h16[k] = (int16_t)b->h0[j];
    }
    qmax = qmax < 128 ? qmax : 128;
    tmax = tmax < 128 ? tmax : 128;
    /* soa.rs:26-66: interleave, 0xFF padding */
    memset(qs, 0xFF, (size_t)qmax * W);
    memset(ts, 0xFF, (size_t)tmax * W);
    for (int k = 0; k < lanes; ++k) {
        const int64_t j = j0 + k;
        const int qn = b->qlen[j] < qmax ? b->qlen[j] : qmax, tn = b->tlen[j] < tmax ? b->tlen[j] : tmax;
        for (int p = 0; p < qn; ++p) qs[(size_t)p * W + k] = b->qseq[b->qoff[j] + p];
        for (int p = 0; p < tn; ++p) ts[(size_t)p * W + k] = b->tseq[b->toff[j] + p];
    }
    int32_t *out = b->out + 6 * j0;
    if (qmax == 0 || tmax == 0) { /* the reference kernels return no record */
        memset(out, 0, sizeof(int32_t) * 6 * (size_t)lanes);
        return;
    }
    if (!b->wide) {
        const int8_t *q = (const int8_t *)qs, *t = (const int8_t *)ts;
        int8_t *H = (int8_t *)Hbuf, *E = (int8_t *)Ebuf;
        if (b->isa == RS_AVX512)
            rs_kern8_avx512(q, t, lanes, qmax, tmax, ql, tl, h8, wv, b->o_del, b->e_del, b->o_ins, b->e_ins,
                            b->zdrop, b->match, b->mis, H, E, out);
        else if (b->isa == RS_AVX2)
            rs_kern8_avx2(q, t, lanes, qmax, tmax, ql, tl, h8, wv, b->o_del, b->e_del, b->o_ins, b->e_ins,
                          b->zdrop, b->match, b->mis, H, E, out);
        else
            rs_chunk8_scalar(q, t, W, lanes, qmax, tmax, ql, tl, h8, wv, b->o_del, b->e_del, b->o_ins,
                             b->e_ins, b->zdrop, b->match, b->mis, H, E, out);
    } else {
        /* dispatch.rs:187-188: the byte SoA widened */
        for (size_t p = 0; p < (size_t)qmax * W; ++p) qs16[p] = qs[p];
        for (size_t p = 0; p < (size_t)tmax * W; ++p) ts16[p] = ts[p];
        int16_t *H = (int16_t *)Hbuf, *E = (int16_t *)Ebuf;
        if (b->isa == RS_AVX512)
            rs_kern16_avx512(qs16, ts16, lanes, qmax, tmax, ql, tl, h16, wv, b->o_del, b->e_del, b->o_ins,
                             b->e_ins, b->zdrop, b->match, b->mis, b->amb, H, E, out);
        else if (b->isa == RS_AVX2)
            rs_kern16_avx2(qs16, ts16, lanes, qmax, tmax, ql, tl, h16, wv, b->o_del, b->e_del, b->o_ins,
                           b->e_ins, b->zdrop, b->match, b->mis, b->amb, H, E, out);
        else
            rs_chunk16_scalar(qs16, ts16, W, lanes, qmax, tmax, ql, tl, h16, wv, b->o_del, b->e_del,
                              b->o_ins, b->e_ins, b->zdrop, b->match, b->mis, b->amb, H, E, out);
    }
}

static void *rs_worker(void *arg) {
    RsBatch *b = (RsBatch *)arg;
    const size_t cap = (size_t)128 * 64;
    uint8_t *qs = (uint8_t *)malloc(cap), *ts = (uint8_t *)malloc(cap);
    int16_t *qs16 = (int16_t *)malloc(cap * 2), *ts16 = (int16_t *)malloc(cap * 2);
    void *H = malloc(cap * 2), *E = malloc(cap * 2);
    if (qs && ts && qs16 && ts16 && H && E)
        for (;;) {
            const int64_t c = __atomic_fetch_add(&b->next, 1, __ATOMIC_RELAXED);
            if (c >= b->nchunks) break;
            rs_run_chunk(b, c, qs, ts, qs16, ts16, H, E);
        }
    free(qs);
    free(ts);
    free(qs16);
    free(ts16);
    free(H);
    free(E);
    return NULL;
}

/* isa: 0 auto (AVX-512BW, else AVX2), 1 AVX2, 2 AVX-512BW, 3 / 4 the lane-by-lane
 * restatement with the AVX-512 / AVX2 chunk widths.  info[0] <- the isa run,
 * info[1] <- 1 when the i16 kernel ran.  Returns -1 when the isa is missing. */
int ora_refsimd_batch(int64_t n, const uint8_t *qseq, const int64_t *qoff, const int32_t *qlen,
                      const uint8_t *tseq, const int64_t *toff, const int32_t *tlen, const int32_t *h0,
                      const int32_t *w, int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                      int32_t zdrop, const int8_t *mat, int32_t isa, int32_t nthreads, int32_t *out,
                      int32_t *info) {
    __builtin_cpu_init();
    const int has512 = __builtin_cpu_supports("avx512bw") && __builtin_cpu_supports("avx512f");
    const int has2 = __builtin_cpu_supports("avx2");
    if (isa == RS_AUTO) isa = has512 ? RS_AVX512 : (has2 ? RS_AVX2 : RS_SCALAR2);
    if ((isa == RS_AVX512 && !has512) || (isa == RS_AVX2 && !has2)) return -1;
    RsBatch b;
    memset(&b, 0, sizeof b);
    b.n = n;
    b.qseq = qseq;
    b.tseq = tseq;
    b.qoff = qoff;
    b.toff = toff;
    b.qlen = qlen;
    b.tlen = tlen;
    b.h0 = h0;
    b.w = w;
    b.o_del = o_del;
    b.e_del = e_del;
    b.o_ins = o_ins;
    b.e_ins = e_ins;
    b.zdrop = zdrop;
    b.match = mat[0];  /* kernel.rs:84-85, kernel_i16.rs:65-67 */
    b.mis = mat[1];
    b.amb = mat[24];
    b.isa = isa;
    int maxlen = 0;
    for (int64_t j = 0; j < n; ++j) maxlen = mx(maxlen, mx(qlen[j], tlen[j]));
    b.wide = maxlen > 128; /* dispatch.rs:100-107 */
    const int w512 = isa == RS_AVX512 || isa == RS_SCALAR512;
    b.W = b.wide ? (w512 ? 32 : 16) : (w512 ? 64 : 32);
    b.nchunks = (n + b.W - 1) / b.W;
    b.out = out;
    if (info) {
        info[0] = isa;
        info[1] = b.wide;
    }
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    int started = 0;
    for (int t = 1; t < nthreads; ++t)
        if (pthread_create(&th[started], NULL, rs_worker, &b) == 0) ++started;
    rs_worker(&b);
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
    return 0;
}
