/*
 * CPU oracle: restatement of FerrousAlign's `ksw_extend2`.
 * TEST INFRASTRUCTURE ONLY (see ksw_oracle.h).
 *
 * Follows src/core/alignment/ksw_affine_gap.rs:952-1169 statement by
 * statement; the line numbers in the comments below point there.
 */
#include "ksw_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

typedef struct {
    int32_t h, e;
} eh_t; /* EhT, :919-923 */

/* Rust `f64 as i32`: truncation toward zero, saturating, NaN -> 0. */
static int32_t f64_as_i32(double x) {
    if (isnan(x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (int32_t)(-2147483647 - 1);
    return (int32_t)x; /* C truncates toward zero, like Rust */
}

/* rows (optional): target rows the row loop ran, the break row included. */
static int extend2_core(int32_t qlen, const uint8_t *query, int32_t tlen,
                        const uint8_t *target, int32_t m, const int8_t *mat,
                        int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                        int32_t w, int32_t end_bonus, int32_t zdrop, int32_t h0,
                        ora_ext_result *out, uint64_t *cells, int32_t *rows) {
    if (h0 <= 0) return ORA_ERR_H0;   /* :968 assert!(h0 > 0) */
    if (qlen <= 0) return ORA_ERR_QLEN; /* :993 eh[1] needs qlen >= 1 */
    if (tlen < 0) return ORA_ERR_TLEN;
    for (int32_t j = 0; j < qlen; ++j)
        if (query[j] >= (uint32_t)m) return ORA_ERR_CODE;
    for (int32_t i = 0; i < tlen; ++i)
        if (target[i] >= (uint32_t)m) return ORA_ERR_CODE;

    const int32_t oe_del = o_del + e_del; /* :973 */
    const int32_t oe_ins = o_ins + e_ins; /* :974 */
    const size_t qn = (size_t)qlen;

    /* query profile qp[k*qlen + j] = mat[k*m + query[j]]  (:976-989) */
    int8_t *qp = (int8_t *)malloc(qn * (size_t)m);
    eh_t *eh = (eh_t *)calloc(qn + 1, sizeof(eh_t));
    if (!qp || !eh) {
        free(qp);
        free(eh);
        return -1;
    }
    {
        size_t idx = 0;
        for (int32_t k = 0; k < m; ++k) {
            const int8_t *p = mat + (size_t)k * m;
            for (size_t j = 0; j < qn; ++j) qp[idx++] = p[query[j]];
        }
    }

    /* first row (:992-998) */
    eh[0].h = h0;
    eh[1].h = h0 > oe_ins ? h0 - oe_ins : 0;
    for (int32_t j = 2; j <= qlen && eh[j - 1].h > e_ins; ++j)
        eh[j].h = eh[j - 1].h - e_ins;

    /* w capping (:1001-1018) */
    int32_t max_score = 0;
    for (int32_t i = 0; i < m * m; ++i)
        if (mat[i] > (int8_t)max_score) max_score = mat[i];
    int32_t max_ins = f64_as_i32(((double)qlen * (double)max_score +
                                  (double)end_bonus - (double)o_ins) /
                                     (double)e_ins + 1.0);
    if (max_ins < 1) max_ins = 1;
    if (w > max_ins) w = max_ins;
    int32_t max_del = f64_as_i32(((double)qlen * (double)max_score +
                                  (double)end_bonus - (double)o_del) /
                                     (double)e_del + 1.0);
    if (max_del < 1) max_del = 1;
    if (w > max_del) w = max_del;

    /* DP loop (:1021-1159) */
    int32_t max = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1,
            max_off = 0;
    int32_t beg = 0, end = qlen;
    uint64_t ncell = 0;

    int32_t nrows = 0;
    for (int32_t i = 0; i < tlen; ++i) {
        int32_t f = 0, m_row = 0, mj = -1;
        nrows = i + 1;
        const int8_t *q = qp + (size_t)target[i] * qn;

        if (beg < i - w) beg = i - w;         /* :1038-1040 */
        if (end > i + w + 1) end = i + w + 1; /* :1041-1043 */
        if (end > qlen) end = qlen;           /* :1044-1046 */

        int32_t h1; /* :1049-1054 */
        if (beg == 0) {
            int32_t v = h0 - (o_del + e_del * (i + 1));
            h1 = v < 0 ? 0 : v;
        } else {
            h1 = 0;
        }
        if (end > beg) ncell += (uint64_t)(end - beg);

        for (int32_t j = beg; j < end; ++j) { /* :1056-1104 */
            eh_t *p = &eh[j];
            int32_t M = p->h;
            int32_t e = p->e;
            p->h = h1;
            M = M != 0 ? M + q[j] : 0;
            int32_t h = M > e ? M : e;
            h = h > f ? h : f;
            h1 = h;
            if (m_row < h) {
                mj = j;
                m_row = h;
            }
            int32_t t = M - oe_del;
            if (t < 0) t = 0;
            int32_t en = e - e_del;
            if (en < t) en = t;
            p->e = en;
            t = M - oe_ins;
            if (t < 0) t = 0;
            f -= e_ins;
            if (f < t) f = t;
        }
        eh[end].h = h1; /* :1106-1107 */
        eh[end].e = 0;

        /* :1110-1120 ; (beg..end).contains(qlen-1) implies end == qlen */
        if ((beg <= qlen - 1 && qlen - 1 < end) || end == qlen) {
            int32_t h_end = end == qlen ? h1 : eh[qlen - 1].h;
            if (gscore < h_end) {
                gscore = h_end;
                max_ie = i;
            }
        }

        if (m_row == 0) break; /* :1122-1124 */

        if (m_row > max) { /* :1126-1133 */
            max = m_row;
            max_i = i;
            max_j = mj;
            int32_t off = mj - i;
            if (off < 0) off = -off;
            if (max_off < off) max_off = off;
        } else if (zdrop > 0) { /* :1134-1145 */
            int32_t di = i - max_i, dj = mj - max_j;
            if (di > dj) {
                if (max - m_row - (di - dj) * e_del > zdrop) break;
            } else if (max - m_row - (dj - di) * e_ins > zdrop) {
                break;
            }
        }

        /* :1148-1158 */
        int32_t nb = beg;
        while (nb < end && eh[nb].h == 0 && eh[nb].e == 0) ++nb;
        beg = nb;
        int32_t ne = end;
        while (ne > beg && eh[ne].h == 0 && eh[ne].e == 0) --ne;
        end = ne + 2 < qlen ? ne + 2 : qlen;
    }

    free(qp);
    free(eh);
    out->score = max;
    out->qle = max_j + 1;
    out->tle = max_i + 1;
    out->gtle = max_ie + 1;
    out->gscore = gscore;
    out->max_off = max_off;
    if (cells) *cells = ncell;
    if (rows) *rows = nrows;
    return ORA_OK;
}

int ora_ksw_extend2(int32_t qlen, const uint8_t *query, int32_t tlen,
                    const uint8_t *target, int32_t m, const int8_t *mat,
                    int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                    int32_t w, int32_t end_bonus, int32_t zdrop, int32_t h0,
                    ora_ext_result *out, uint64_t *cells) {
    return extend2_core(qlen, query, tlen, target, m, mat, o_del, e_del, o_ins, e_ins, w,
                        end_bonus, zdrop, h0, out, cells, NULL);
}

/* ------------------------------------------------------------------ */
/* batch driver                                                        */

typedef struct {
    int64_t lo, hi;
    const uint8_t *qseq, *tseq;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen, *h0, *w;
    int32_t m;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins, end_bonus, zdrop;
    ora_ext_result *out;
    uint64_t *cells;
    int32_t *rc;
    int32_t *rows;
    int64_t nbad;
} batch_job_t;

static void *batch_worker(void *arg) {
    batch_job_t *b = (batch_job_t *)arg;
    for (int64_t k = b->lo; k < b->hi; ++k) {
        ora_ext_result r = {0, 0, 0, 0, 0, 0};
        uint64_t c = 0;
        int32_t nr = 0;
        int code = extend2_core(b->qlen[k], b->qseq + b->qoff[k], b->tlen[k],
                                b->tseq + b->toff[k], b->m, b->mat, b->o_del,
                                b->e_del, b->o_ins, b->e_ins, b->w[k],
                                b->end_bonus, b->zdrop, b->h0[k], &r, &c, &nr);
        if (code != ORA_OK) {
            ++b->nbad;
            memset(&r, 0, sizeof r);
            c = 0;
        }
        b->out[k] = r;
        if (b->cells) b->cells[k] = c;
        if (b->rc) b->rc[k] = code;
        if (b->rows) b->rows[k] = code == ORA_OK ? nr : 0;
    }
    return NULL;
}

int64_t ora_extend_batch(int64_t n, const uint8_t *qseq, const int64_t *qoff,
                         const int32_t *qlen, const uint8_t *tseq,
                         const int64_t *toff, const int32_t *tlen,
                         const int32_t *h0, const int32_t *w, int32_t m,
                         const int8_t *mat, int32_t o_del, int32_t e_del,
                         int32_t o_ins, int32_t e_ins, int32_t end_bonus,
                         int32_t zdrop, ora_ext_result *out, uint64_t *cells,
                         int32_t *rc, int nthreads) {
    return ora_extend_batch_rows(n, qseq, qoff, qlen, tseq, toff, tlen, h0, w, m, mat, o_del,
                                 e_del, o_ins, e_ins, end_bonus, zdrop, out, cells, rc, NULL,
                                 nthreads);
}

int64_t ora_extend_batch_rows(int64_t n, const uint8_t *qseq, const int64_t *qoff,
                              const int32_t *qlen, const uint8_t *tseq,
                              const int64_t *toff, const int32_t *tlen,
                              const int32_t *h0, const int32_t *w, int32_t m,
                              const int8_t *mat, int32_t o_del, int32_t e_del,
                              int32_t o_ins, int32_t e_ins, int32_t end_bonus,
                              int32_t zdrop, ora_ext_result *out, uint64_t *cells,
                              int32_t *rc, int32_t *rows, int nthreads) {
    if (n <= 0) return 0;
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    if ((int64_t)nthreads > n) nthreads = (int)n;
    batch_job_t jobs[256];
    pthread_t th[256];
    int64_t per = (n + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        batch_job_t *b = &jobs[t];
        b->lo = per * t;
        b->hi = b->lo + per < n ? b->lo + per : n;
        b->qseq = qseq;
        b->tseq = tseq;
        b->qoff = qoff;
        b->toff = toff;
        b->qlen = qlen;
        b->tlen = tlen;
        b->h0 = h0;
        b->w = w;
        b->m = m;
        b->mat = mat;
        b->o_del = o_del;
        b->e_del = e_del;
        b->o_ins = o_ins;
        b->e_ins = e_ins;
        b->end_bonus = end_bonus;
        b->zdrop = zdrop;
        b->out = out;
        b->cells = cells;
        b->rc = rc;
        b->rows = rows;
        b->nbad = 0;
    }
    if (nthreads == 1) {
        batch_worker(&jobs[0]);
    } else {
        for (int t = 0; t < nthreads; ++t)
            pthread_create(&th[t], NULL, batch_worker, &jobs[t]);
        for (int t = 0; t < nthreads; ++t) pthread_join(th[t], NULL);
    }
    int64_t nbad = 0;
    for (int t = 0; t < nthreads; ++t) nbad += jobs[t].nbad;
    return nbad;
}
