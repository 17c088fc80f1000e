/*
 * CPU oracle: restatement of FerrousAlign's ksw_align2 in its 16-bit and
 * 8-bit modes (src/core/alignment/ksw_affine_gap.rs:78-151 ksw_qinit,
 * :166-203 ksw_qalloc, :236-259 max_16_u8 / max_8_u16, :267-494 ksw_u8_impl,
 * :503-735 ksw_i16_impl, :773-861 ksw_align2), as executed by SimdEngine128
 * (src/core/compute/simd_abstraction/engine128.rs): P lanes per 128-bit
 * register (8 x i16, or 16 x u8 with KSW_XBYTE), query position k lives in
 * lane k / slen, segment k % slen with slen = ceil(qlen / P).
 * TEST INFRASTRUCTURE ONLY.
 *
 * Semantics kept on purpose (they change results):
 *   - 16-bit: adds/subs are SIGNED saturating (_mm_adds_epi16/_mm_subs_epi16),
 *     H is never clamped at 0;
 *   - 8-bit: UNSIGNED saturating bytes with the profile biased by
 *     shift = (u8)(256 - smallest matrix entry): H = subs(adds(H, s + shift),
 *     shift), E and F clamp at 0; a best score with gmax + shift >= 255 stops
 *     the pass and reports 255 (no qe, score2, te2);
 *   - the row maximum is taken before the lazy-F pass, Hmax after it;
 *   - the lazy-F pass only raises H by a decaying F and never updates E, and
 *     stops after the first pass (of at most P) that changed no H;
 *   - ties in qe go to the smallest query position;
 *   - the reverse pass accepts rr.score >= r.score.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define KSW_XBYTE 0x10000u
#define KSW_XSTOP 0x20000u
#define KSW_XSUBO 0x40000u
#define KSW_XSTART 0x80000u

typedef struct {
    int32_t score, te, qe, score2, te2, tb, qb;
} ora_kswr; /* Kswr, ksw_affine_gap.rs:51-61 */

#define P 8 /* i16 lanes of one 128-bit vector */

static inline int16_t sat16(int32_t x) {
    return (int16_t)(x > 32767 ? 32767 : (x < -32768 ? -32768 : x));
}
static inline int16_t max16(int16_t a, int16_t b) { return a > b ? a : b; }

typedef struct {
    int qlen, slen;
    int16_t *qp; /* [m][slen][P] */
    int qmax;    /* q.max = largest matrix entry (at least 0), as u8 */
} prof_t;

/* ksw_qinit for size 2 (:78-151): qp[a][i][p] = mat[a*m + query[i + p*slen]], 0 past qlen */
static int qinit16(prof_t *q, int qlen, const uint8_t *query, int m, const int8_t *mat) {
    q->qlen = qlen;
    q->slen = (qlen + P - 1) / P;
    int mx = 0;
    for (int k = 0; k < m * m; ++k)
        if (mat[k] > mx) mx = mat[k];
    q->qmax = (uint8_t)mx;
    q->qp = (int16_t *)malloc(sizeof(int16_t) * (size_t)m * q->slen * P + 2);
    if (!q->qp) return -1;
    int16_t *t = q->qp;
    for (int a = 0; a < m; ++a)
        for (int i = 0; i < q->slen; ++i)
            for (int p = 0; p < P; ++p) {
                int k = i + p * q->slen;
                *t++ = (int16_t)(k < qlen ? mat[a * m + query[k]] : 0);
            }
    return 0;
}

/* ksw_i16_impl (:503-735).  *rows receives the number of target rows run. */
static ora_kswr ksw_i16(const prof_t *q, int tlen, const uint8_t *target, int o_del, int e_del,
                        int o_ins, int e_ins, uint32_t xtra, int *rows) {
    ora_kswr r = {0, -1, -1, -1, -1, -1, -1}; /* KSW_DEFR */
    const uint32_t minsc = (xtra & KSW_XSUBO) ? (xtra & 0xffff) : 0x10000;
    const uint32_t endsc = (xtra & KSW_XSTOP) ? (xtra & 0xffff) : 0x10000;
    const int slen = q->slen;
    const int n = slen * P;
    int16_t *H0 = (int16_t *)calloc((size_t)n, 2), *H1 = (int16_t *)calloc((size_t)n, 2);
    int16_t *E = (int16_t *)calloc((size_t)n, 2), *Hmax = (int16_t *)calloc((size_t)n, 2);
    uint64_t *b = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(tlen > 0 ? tlen : 1));
    int nb = 0;
    int gmax = 0, te = -1;
    const int16_t oe_del = (int16_t)(o_del + e_del), oe_ins = (int16_t)(o_ins + e_ins);
    int nrow = 0;
    for (int i = 0; i < tlen; ++i) {
        ++nrow;
        int16_t f[P], mx[P], h[P];
        for (int p = 0; p < P; ++p) {
            f[p] = 0;
            mx[p] = 0;
            h[p] = p == 0 ? 0 : H0[(slen - 1) * P + p - 1]; /* slli by one lane */
        }
        const int16_t *s = q->qp + (size_t)target[i] * slen * P;
        for (int j = 0; j < slen; ++j) {
            for (int p = 0; p < P; ++p) {
                int16_t hv = sat16((int32_t)h[p] + s[j * P + p]);
                int16_t e = E[j * P + p];
                hv = max16(hv, e);
                hv = max16(hv, f[p]);
                mx[p] = max16(mx[p], hv);
                H1[j * P + p] = hv;
                E[j * P + p] = max16(sat16((int32_t)e - e_del), sat16((int32_t)hv - oe_del));
                f[p] = max16(sat16((int32_t)f[p] - e_ins), sat16((int32_t)hv - oe_ins));
                h[p] = H0[j * P + p];
            }
        }
        /* lazy F (:607-633): up to 8 passes, stop after a pass that changed nothing */
        for (int k = 0; k < P; ++k) {
            for (int p = P - 1; p > 0; --p) f[p] = f[p - 1];
            f[0] = 0;
            int changed = 0;
            for (int j = 0; j < slen; ++j) {
                for (int p = 0; p < P; ++p) {
                    int16_t hb = H1[j * P + p];
                    int16_t hn = max16(hb, f[p]);
                    H1[j * P + p] = hn;
                    f[p] = sat16((int32_t)f[p] - e_ins);
                    if (hb != hn) changed = 1;
                }
            }
            if (!changed) break;
        }
        /* max_8_u16: unsigned horizontal max (lanes are >= 0 here) */
        uint16_t um = 0;
        for (int p = 0; p < P; ++p)
            if ((uint16_t)mx[p] > um) um = (uint16_t)mx[p];
        const int imax = um;
        if ((uint32_t)imax >= minsc) { /* :639-655 */
            if (nb == 0) {
                b[nb++] = ((uint64_t)imax << 32) | (uint32_t)i;
            } else if ((int32_t)(b[nb - 1] & 0xFFFFFFFFu) + 1 != i) {
                b[nb++] = ((uint64_t)imax << 32) | (uint32_t)i;
            } else if ((int32_t)(b[nb - 1] >> 32) < imax) {
                b[nb - 1] = ((uint64_t)imax << 32) | (uint32_t)i;
            }
        }
        if (imax > gmax) { /* :657-668 */
            gmax = imax;
            te = i;
            memcpy(Hmax, H1, (size_t)n * 2);
            if ((uint32_t)gmax >= 0x7FFF || (uint32_t)gmax >= endsc) break;
        }
        int16_t *tmp = H0;
        H0 = H1;
        H1 = tmp;
    }
    r.score = gmax;
    r.te = te;
    if (r.score != 0x7FFF) { /* :674-732 */
        int max_qe = -1, max_val = -1;
        for (int k = 0; k < n; ++k) {
            int seg = k / P, lane = k % P;
            int pos = seg + lane * slen;
            if (pos >= q->qlen) continue;
            int v = Hmax[seg * P + lane];
            if (v > max_val) {
                max_val = v;
                max_qe = pos;
            } else if (v == max_val && pos < max_qe) {
                max_qe = pos;
            }
        }
        r.qe = max_qe;
        if (nb > 0) {
            int iv = (r.score + q->qmax - 1) / q->qmax;
            int low = te - iv, high = te + iv;
            for (int x = 0; x < nb; ++x) {
                int e = (int32_t)(b[x] & 0xFFFFFFFFu), sc = (int32_t)(b[x] >> 32);
                if ((e < low || e > high) && sc > r.score2) {
                    r.score2 = sc;
                    r.te2 = e;
                }
            }
        }
    }
    free(H0);
    free(H1);
    free(E);
    free(Hmax);
    free(b);
    if (rows) *rows = nrow;
    return r;
}

/* ---- 8-bit mode -------------------------------------------------------- */

#define P8 16 /* u8 lanes of one 128-bit vector */

static inline uint8_t addsu8(uint8_t a, uint8_t b) {
    unsigned x = (unsigned)a + b;
    return (uint8_t)(x > 255 ? 255 : x);
}
static inline uint8_t subsu8(uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : 0); }
static inline uint8_t maxu8(uint8_t a, uint8_t b) { return a > b ? a : b; }

typedef struct {
    int qlen, slen;
    uint8_t *qp; /* [m][slen][P8], biased by shift */
    int qmax;    /* q.max, as u8 */
    uint8_t shift;
} prof8_t;

/* ksw_qinit for size 1 (:78-151): min over the matrix starts at 127, max at 0;
 * shift = (u8)(256 - min); qp[a][i][p] = (u8)(mat[a*m + query[i + p*slen]] +
 * shift), shift itself past qlen (the i8 sum wraps, as Rust's release build) */
static int qinit8(prof8_t *q, int qlen, const uint8_t *query, int m, const int8_t *mat) {
    q->qlen = qlen;
    q->slen = (qlen + P8 - 1) / P8;
    int mn = 127, mx = 0;
    for (int k = 0; k < m * m; ++k) {
        if (mat[k] < mn) mn = mat[k];
        if (mat[k] > mx) mx = mat[k];
    }
    q->shift = (uint8_t)(256 - mn);
    q->qmax = (uint8_t)mx;
    q->qp = (uint8_t *)malloc((size_t)m * q->slen * P8 + 2);
    if (!q->qp) return -1;
    uint8_t *t = q->qp;
    for (int a = 0; a < m; ++a)
        for (int i = 0; i < q->slen; ++i)
            for (int p = 0; p < P8; ++p) {
                int k = i + p * q->slen;
                *t++ = (uint8_t)((k < qlen ? mat[a * m + query[k]] : 0) + q->shift);
            }
    return 0;
}

/* ksw_u8_impl (:267-494).  *rows receives the number of target rows run. */
static ora_kswr ksw_u8(const prof8_t *q, int tlen, const uint8_t *target, int o_del, int e_del,
                       int o_ins, int e_ins, uint32_t xtra, int *rows) {
    ora_kswr r = {0, -1, -1, -1, -1, -1, -1}; /* KSW_DEFR */
    const uint32_t minsc = (xtra & KSW_XSUBO) ? (xtra & 0xffff) : 0x10000;
    const uint32_t endsc = (xtra & KSW_XSTOP) ? (xtra & 0xffff) : 0x10000;
    const int slen = q->slen;
    const int n = slen * P8;
    uint8_t *H0 = (uint8_t *)calloc((size_t)n, 1), *H1 = (uint8_t *)calloc((size_t)n, 1);
    uint8_t *E = (uint8_t *)calloc((size_t)n, 1), *Hmax = (uint8_t *)calloc((size_t)n, 1);
    uint64_t *b = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(tlen > 0 ? tlen : 1));
    int nb = 0;
    int gmax = 0, te = -1;
    /* set1_epi8 of the penalties, read back as unsigned bytes by subs_epu8 */
    const uint8_t oe_del = (uint8_t)(o_del + e_del), ed = (uint8_t)e_del;
    const uint8_t oe_ins = (uint8_t)(o_ins + e_ins), ei = (uint8_t)e_ins;
    const uint8_t shift = q->shift;
    int nrow = 0;
    for (int i = 0; i < tlen; ++i) {
        ++nrow;
        uint8_t f[P8], mx[P8], h[P8];
        for (int p = 0; p < P8; ++p) {
            f[p] = 0;
            mx[p] = 0;
            h[p] = p == 0 ? 0 : H0[(slen - 1) * P8 + p - 1]; /* slli by one byte */
        }
        const uint8_t *s = q->qp + (size_t)target[i] * slen * P8;
        for (int j = 0; j < slen; ++j) {
            for (int p = 0; p < P8; ++p) {
                uint8_t hv = subsu8(addsu8(h[p], s[j * P8 + p]), shift);
                const uint8_t e = E[j * P8 + p];
                hv = maxu8(hv, e);
                hv = maxu8(hv, f[p]);
                mx[p] = maxu8(mx[p], hv);
                H1[j * P8 + p] = hv;
                E[j * P8 + p] = maxu8(subsu8(e, ed), subsu8(hv, oe_del));
                f[p] = maxu8(subsu8(f[p], ei), subsu8(hv, oe_ins));
                h[p] = H0[j * P8 + p];
            }
        }
        /* lazy F (:359-387): up to 16 passes, stop after a pass that changed nothing */
        for (int k = 0; k < P8; ++k) {
            for (int p = P8 - 1; p > 0; --p) f[p] = f[p - 1];
            f[0] = 0;
            int changed = 0;
            for (int j = 0; j < slen; ++j) {
                for (int p = 0; p < P8; ++p) {
                    const uint8_t hb = H1[j * P8 + p];
                    const uint8_t hn = maxu8(hb, f[p]);
                    H1[j * P8 + p] = hn;
                    f[p] = subsu8(f[p], ei);
                    if (hb != hn) changed = 1;
                }
            }
            if (!changed) break;
        }
        uint8_t um = 0; /* max_16_u8 */
        for (int p = 0; p < P8; ++p) um = maxu8(um, mx[p]);
        const int imax = um;
        if ((uint32_t)imax >= minsc) { /* :392-405 */
            if (nb == 0) {
                b[nb++] = ((uint64_t)imax << 32) | (uint32_t)i;
            } else if ((int32_t)(b[nb - 1] & 0xFFFFFFFFu) + 1 != i) {
                b[nb++] = ((uint64_t)imax << 32) | (uint32_t)i;
            } else if ((int32_t)(b[nb - 1] >> 32) < imax) {
                b[nb - 1] = ((uint64_t)imax << 32) | (uint32_t)i;
            }
        }
        if (imax > gmax) { /* :408-422 */
            gmax = imax;
            te = i;
            memcpy(Hmax, H1, (size_t)n);
            if (gmax + (int)shift >= 255 || (uint32_t)gmax >= endsc) break;
        }
        uint8_t *tmp = H0;
        H0 = H1;
        H1 = tmp;
    }
    r.score = gmax + (int)shift < 255 ? gmax : 255;
    r.te = te;
    if (r.score != 255) { /* :434-491 */
        int max_qe = -1, max_val = -1;
        for (int k = 0; k < n; ++k) {
            int seg = k / P8, lane = k % P8;
            int pos = seg + lane * slen;
            if (pos >= q->qlen) continue;
            int v = Hmax[seg * P8 + lane];
            if (v > max_val) {
                max_val = v;
                max_qe = pos;
            } else if (v == max_val && pos < max_qe) {
                max_qe = pos;
            }
        }
        r.qe = max_qe;
        if (nb > 0) {
            int iv = (r.score + q->qmax - 1) / q->qmax;
            int low = te - iv, high = te + iv;
            for (int x = 0; x < nb; ++x) {
                int e = (int32_t)(b[x] & 0xFFFFFFFFu), sc = (int32_t)(b[x] >> 32);
                if ((e < low || e > high) && sc > r.score2) {
                    r.score2 = sc;
                    r.te2 = e;
                }
            }
        }
    }
    free(H0);
    free(H1);
    free(E);
    free(Hmax);
    free(b);
    if (rows) *rows = nrow;
    return r;
}

/* one pass of the mode xtra selects (KSW_XBYTE: 8-bit) over query[0, qlen) */
static int run_pass(int byte_mode, int qlen, const uint8_t *query, int tlen, const uint8_t *target,
                    int m, const int8_t *mat, int o_del, int e_del, int o_ins, int e_ins,
                    uint32_t xtra, ora_kswr *r, int *rows) {
    if (byte_mode) {
        prof8_t q;
        if (qinit8(&q, qlen, query, m, mat)) return -1;
        *r = ksw_u8(&q, tlen, target, o_del, e_del, o_ins, e_ins, xtra, rows);
        free(q.qp);
    } else {
        prof_t q;
        if (qinit16(&q, qlen, query, m, mat)) return -1;
        *r = ksw_i16(&q, tlen, target, o_del, e_del, o_ins, e_ins, xtra, rows);
        free(q.qp);
    }
    return 0;
}

/* ksw_align2 (:773-861) in the mode xtra selects (KSW_XBYTE: 8-bit scores,
 * else 16-bit).  Returns 0, or -1 on bad input (qlen <= 0, tlen < 0,
 * code >= m, no positive matrix entry).  cells: target rows run times query
 * length, forward plus reverse pass. */
int ora_ksw_align2(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                   int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                   int32_t e_ins, uint32_t xtra, ora_kswr *out, uint64_t *cells) {
    if (qlen <= 0 || tlen < 0) return -1;
    for (int k = 0; k < qlen; ++k)
        if (query[k] >= (uint32_t)m) return -1;
    for (int k = 0; k < tlen; ++k)
        if (target[k] >= (uint32_t)m) return -1;
    int mx = 0;
    for (int k = 0; k < m * m; ++k)
        if (mat[k] > mx) mx = mat[k];
    if (mx == 0) return -1;
    const int byte_mode = (xtra & KSW_XBYTE) != 0;
    int rows = 0, rows2 = 0, qlen2 = 0;
    ora_kswr r;
    if (run_pass(byte_mode, qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, xtra, &r,
                 &rows))
        return -1;
    const int find_start =
        (xtra & KSW_XSTART) && !((xtra & KSW_XSUBO) && r.score < (int32_t)(xtra & 0xffff));
    if (find_start && r.qe >= 0 && r.te >= 0) {
        const int ql = r.qe + 1, tl = r.te + 1;
        uint8_t *qr = (uint8_t *)malloc((size_t)ql);
        uint8_t *tr = (uint8_t *)malloc((size_t)tlen + 1);
        for (int k = 0; k < ql; ++k) qr[k] = query[ql - 1 - k];
        for (int k = 0; k < tlen; ++k) tr[k] = k < tl ? target[tl - 1 - k] : target[k];
        ora_kswr rr;
        if (run_pass(byte_mode, ql, qr, tlen, tr, m, mat, o_del, e_del, o_ins, e_ins,
                     KSW_XSTOP | (uint32_t)r.score, &rr, &rows2) == 0) {
            qlen2 = ql;
            if (rr.score >= r.score) {
                r.tb = r.te - rr.te;
                r.qb = r.qe - rr.qe;
            }
        }
        free(qr);
        free(tr);
    }
    *out = r;
    if (cells) *cells = (uint64_t)rows * (uint64_t)qlen + (uint64_t)rows2 * (uint64_t)qlen2;
    return 0;
}

/* ksw_align2 (:773-861) in its 16-bit mode only: -1 when KSW_XBYTE is set. */
int ora_ksw_align2_i16(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                       int32_t m, const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                       int32_t e_ins, uint32_t xtra, ora_kswr *out, uint64_t *cells) {
    if (xtra & KSW_XBYTE) return -1;
    return ora_ksw_align2(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, xtra, out,
                          cells);
}

/* Batch over fixed-stride buffers with POSIX threads (<= 0: all cores). */
#include <pthread.h>
#include <unistd.h>
typedef struct {
    int64_t lo, hi;
    const uint8_t *q, *t;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen;
    int32_t m;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins;
    uint32_t xtra;
    ora_kswr *out;
    uint64_t *cells;
    int32_t *rc;
} align_job_t;

static void *align_worker(void *arg) {
    align_job_t *j = (align_job_t *)arg;
    for (int64_t k = j->lo; k < j->hi; ++k) {
        ora_kswr r = {0, 0, 0, 0, 0, 0, 0};
        uint64_t c = 0;
        int code = ora_ksw_align2(j->qlen[k], j->q + j->qoff[k], j->tlen[k], j->t + j->toff[k],
                                  j->m, j->mat, j->o_del, j->e_del, j->o_ins, j->e_ins, j->xtra,
                                  &r, &c);
        j->out[k] = r;
        if (j->cells) j->cells[k] = c;
        if (j->rc) j->rc[k] = code;
    }
    return NULL;
}

void ora_ksw_align2_batch(int64_t n, const uint8_t *q, const int64_t *qoff, const int32_t *qlen,
                          const uint8_t *t, const int64_t *toff, const int32_t *tlen, int32_t m,
                          const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                          int32_t e_ins, uint32_t xtra, ora_kswr *out, uint64_t *cells,
                          int32_t *rc, int nthreads) {
    if (n <= 0) return;
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads > 256) nthreads = 256;
    if (nthreads > n) nthreads = (int)n;
    if (nthreads < 1) nthreads = 1;
    align_job_t jobs[256];
    pthread_t th[256];
    int64_t per = (n + nthreads - 1) / nthreads;
    for (int x = 0; x < nthreads; ++x) {
        align_job_t jb = {per * x, per * (x + 1) < n ? per * (x + 1) : n, q, t, qoff, toff, qlen,
                          tlen, m, mat, o_del, e_del, o_ins, e_ins, xtra, out, cells, rc};
        jobs[x] = jb;
    }
    for (int x = 1; x < nthreads; ++x) pthread_create(&th[x], NULL, align_worker, &jobs[x]);
    align_worker(&jobs[0]);
    for (int x = 1; x < nthreads; ++x) pthread_join(th[x], NULL);
}
