/*
 * CPU oracle: restatement of FerrousAlign's scalar_banded_swa (banded local
 * Smith-Waterman with traceback), src/core/alignment/banded_swa/scalar/
 * implementation.rs:5-440.  TEST INFRASTRUCTURE ONLY: only tests/, the smoke
 * check and bench.py's CPU baseline may call it; the library never does.
 *
 * Traceback codes (banded_swa/types.rs:101-103): MATCH 0, DEL 1, INS 2.
 * CIGAR ops are written as (len << 4) | op with op M=0, I=1, D=2, S=4 (the
 * BAM encoding); the reference's (u8 op char, i32 len) pairs map one to one.
 */
#include <stdlib.h>
#include <string.h>

#include "ksw_oracle.h"

enum { TB_MATCH = 0, TB_DEL = 1, TB_INS = 2 };
enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4 };

static void push_op(uint32_t *cigar, int32_t *n, int op, int32_t len) {
    cigar[(*n)++] = ((uint32_t)len << 4) | (uint32_t)op;
}

int ora_scalar_banded_swa(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                          const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                          int32_t e_ins, int32_t zdrop, int32_t pen_clip3, int32_t w, int32_t h0,
                          ora_ext_result *out, uint32_t *cigar, int32_t *n_cigar,
                          uint8_t *ref_aligned, int32_t *n_ref, uint8_t *query_aligned,
                          int32_t *n_query, uint64_t *cells) {
    *n_cigar = *n_ref = *n_query = 0;
    if (cells) *cells = 0;
    memset(out, 0, sizeof(*out));
    /* :17-30 empty sequences: zero score, empty CIGAR */
    if (tlen <= 0 || qlen <= 0) return ORA_OK;
    const int32_t oe_del = o_del + e_del, oe_ins = o_ins + e_ins; /* :50-51 */
    const size_t qn = (size_t)qlen, W = qn + 1;
    int8_t *qp = (int8_t *)malloc(5 * qn);
    int32_t *eh_h = (int32_t *)calloc(W, sizeof(int32_t));
    int32_t *eh_e = (int32_t *)calloc(W, sizeof(int32_t));
    uint8_t *tb = (uint8_t *)calloc((size_t)(tlen + 1) * W, 1); /* :57 (tlen+1) x (qlen+1), MATCH */
    if (!qp || !eh_h || !eh_e || !tb) {
        free(qp), free(eh_h), free(eh_e), free(tb);
        return -1;
    }
    /* query profile qp[k*qlen + j] = mat[k*5 + min(query[j], 4)] (:60-70) */
    for (int k = 0; k < 5; ++k)
        for (size_t j = 0; j < qn; ++j) qp[k * qn + j] = mat[k * 5 + (query[j] > 4 ? 4 : query[j])];
    /* first row (:73-81) */
    eh_h[0] = h0;
    eh_h[1] = h0 > oe_ins ? h0 - oe_ins : 0;
    for (size_t j = 2; j <= qn; ++j) eh_h[j] = eh_h[j - 1] > e_ins ? eh_h[j - 1] - e_ins : 0;
    /* traceback edges (:84-91) */
    for (size_t j = 1; j <= qn; ++j) tb[j] = TB_INS;
    for (int32_t i = 1; i <= tlen; ++i) tb[(size_t)i * W] = TB_DEL;

    int32_t max_score = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
    int32_t beg = 0, end = qlen;
    uint64_t ncell = 0;
    for (int32_t i = 0; i < tlen; ++i) { /* :102-225 */
        int32_t f = 0, h1, m_val = 0, mj = -1;
        const int32_t tbase = target[i] > 4 ? 4 : target[i];
        const int8_t *q = qp + (size_t)tbase * qn;
        int32_t cb = beg, ce = end;
        if (cb < i - w) cb = i - w;
        if (ce > i + w + 1) ce = i + w + 1;
        if (ce > qlen) ce = qlen;
        if (cb == 0) {
            h1 = h0 - (o_del + e_del * (i + 1));
            if (h1 < 0) h1 = 0;
        } else {
            h1 = 0;
        }
        if (ce > cb) ncell += (uint64_t)(ce - cb);
        for (int32_t j = cb; j < ce; ++j) {
            int32_t m = eh_h[j], e = eh_e[j];
            eh_h[j] = h1;
            m += q[j];
            if (m < 0) m = 0; /* local: clamp to 0 */
            int32_t h;
            uint8_t code;
            if (m >= e && m >= f) {
                h = m;
                code = TB_MATCH;
            } else if (e >= f) {
                h = e;
                code = TB_DEL;
            } else {
                h = f;
                code = TB_INS;
            }
            h1 = h;
            tb[(size_t)(i + 1) * W + (size_t)(j + 1)] = code;
            if (m_val < h) {
                m_val = h;
                mj = j;
            }
            int32_t t = m - oe_del;
            if (t < 0) t = 0;
            e -= e_del;
            if (e < t) e = t;
            eh_e[j] = e;
            t = m - oe_ins;
            if (t < 0) t = 0;
            f -= e_ins;
            if (f < t) f = t;
        }
        eh_h[ce] = h1;
        eh_e[ce] = 0;
        if (ce == qlen && gscore < h1) {
            gscore = h1;
            max_ie = i;
        }
        if (m_val == 0) break;
        if (m_val > max_score) {
            max_score = m_val;
            max_i = i;
            max_j = mj;
            const int32_t off = mj > i ? mj - i : i - mj;
            if (max_off < off) max_off = off;
        } else if (zdrop > 0) {
            const int32_t di = i - max_i, dj = mj - max_j;
            if (di > dj) {
                if (max_score - m_val - (di - dj) * e_del > zdrop) break;
            } else if (max_score - m_val - (dj - di) * e_ins > zdrop) {
                break;
            }
        }
        int32_t nb = cb;
        while (nb < ce && eh_h[nb] == 0 && eh_e[nb] == 0) ++nb;
        beg = nb;
        int32_t ne = ce;
        while (ne >= beg && eh_h[ne] == 0 && eh_e[ne] == 0) --ne; /* >= (:221) */
        end = ne + 2 < qlen ? ne + 2 : qlen;
    }

    /* traceback from (max_i + 1, max_j + 1) (:227-330), runs collected backwards */
    uint32_t *rc = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)qlen + (size_t)tlen + 4));
    uint8_t *ra = (uint8_t *)malloc((size_t)qlen + (size_t)tlen + 2);
    uint8_t *qa = (uint8_t *)malloc((size_t)qlen + (size_t)tlen + 2);
    int32_t nrc = 0, nra = 0, nqa = 0;
    int32_t ci = max_i + 1, cj = max_j + 1, iter = 0;
    while (ci > 0 || cj > 0) {
        if (iter >= 10000) break; /* MAX_TRACEBACK_ITERATIONS */
        ++iter;
        const int32_t pi = ci, pj = cj;
        const uint8_t code = tb[(size_t)ci * W + (size_t)cj];
        if (code == TB_MATCH) {
            int32_t cnt = 0;
            while (ci > 0 && cj > 0 && tb[(size_t)ci * W + (size_t)cj] == TB_MATCH) {
                ++cnt;
                ra[nra++] = target[ci - 1];
                qa[nqa++] = query[cj - 1];
                --ci;
                --cj;
            }
            if (cnt > 0) push_op(rc, &nrc, OP_M, cnt);
        } else if (code == TB_DEL) {
            int32_t cnt = 0;
            while (ci > 0 && tb[(size_t)ci * W + (size_t)cj] == TB_DEL) {
                ++cnt;
                ra[nra++] = target[ci - 1];
                --ci;
            }
            if (cnt > 0) push_op(rc, &nrc, OP_D, cnt);
        } else if (code == TB_INS) {
            int32_t cnt = 0;
            while (cj > 0 && tb[(size_t)ci * W + (size_t)cj] == TB_INS) {
                ++cnt;
                qa[nqa++] = query[cj - 1];
                --cj;
            }
            if (cnt > 0) push_op(rc, &nrc, OP_I, cnt);
        } else {
            break;
        }
        if (ci == pi && cj == pj) break; /* no progress */
    }
    /* reverse the runs and the aligned bases (:333-337) */
    const int32_t query_start = cj, query_end = max_j + 1;
    int32_t nc = 0;
    if (query_start > 0) push_op(cigar, &nc, OP_S, query_start); /* 5' soft clip (:357-359) */
    for (int32_t k = nrc - 1; k >= 0; --k) cigar[nc++] = rc[k];
    if (query_end < qlen) { /* 3' end (:367-414) */
        const int clip3 = gscore <= 0 || gscore <= max_score - pen_clip3;
        push_op(cigar, &nc, clip3 ? OP_S : OP_M, qlen - query_end);
    }
    *n_cigar = nc;
    for (int32_t k = 0; k < nra; ++k) ref_aligned[k] = ra[nra - 1 - k];
    for (int32_t k = 0; k < nqa; ++k) query_aligned[k] = qa[nqa - 1 - k];
    *n_ref = nra;
    *n_query = nqa;
    out->score = max_score; /* :431-438 */
    out->tle = max_i + 1;
    out->qle = max_j + 1;
    out->gtle = max_ie + 1;
    out->gscore = gscore;
    out->max_off = max_off;
    if (cells) *cells = ncell;
    free(rc), free(ra), free(qa);
    free(qp), free(eh_h), free(eh_e), free(tb);
    return ORA_OK;
}

/* ------------------------------------------------------------------ */
/* batch driver (threads over pairs), for the CPU baseline            */
/* ------------------------------------------------------------------ */
#include <pthread.h>
#include <unistd.h>

typedef struct {
    int64_t lo, hi;
    const uint8_t *q, *t;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen, *w, *h0;
    const int8_t *mat;
    int32_t o_del, e_del, o_ins, e_ins, zdrop, pen_clip3;
    ora_ext_result *out;
    uint64_t *cells;
    int32_t *n_cigar;
} swtb_job;

static void *swtb_worker(void *arg) {
    swtb_job *j = (swtb_job *)arg;
    uint32_t *cig = (uint32_t *)malloc(sizeof(uint32_t) * (256 + 2048 + 8));
    uint8_t *ra = (uint8_t *)malloc(256 + 2048 + 8), *qa = (uint8_t *)malloc(256 + 2048 + 8);
    for (int64_t k = j->lo; k < j->hi; ++k) {
        int32_t nc, nr, nq;
        uint64_t c = 0;
        ora_scalar_banded_swa(j->qlen[k], j->q + j->qoff[k], j->tlen[k], j->t + j->toff[k], j->mat,
                              j->o_del, j->e_del, j->o_ins, j->e_ins, j->zdrop, j->pen_clip3, j->w[k],
                              j->h0[k], &j->out[k], cig, &nc, ra, &nr, qa, &nq, &c);
        if (j->cells) j->cells[k] = c;
        if (j->n_cigar) j->n_cigar[k] = nc;
    }
    free(cig), free(ra), free(qa);
    return NULL;
}

int64_t ora_scalar_banded_swa_batch(int64_t n, const uint8_t *q, const int64_t *qoff,
                                    const int32_t *qlen, const uint8_t *t, const int64_t *toff,
                                    const int32_t *tlen, const int32_t *w, const int32_t *h0,
                                    const int8_t *mat, int32_t o_del, int32_t e_del,
                                    int32_t o_ins, int32_t e_ins, int32_t zdrop,
                                    int32_t pen_clip3, ora_ext_result *out, uint64_t *cells,
                                    int32_t *n_cigar, int nthreads) {
    if (nthreads <= 0) nthreads = (int)sysconf(_SC_NPROCESSORS_ONLN);
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    swtb_job jobs[64];
    const int64_t per = (n + nthreads - 1) / nthreads;
    int nt = 0;
    for (int i = 0; i < nthreads; ++i) {
        const int64_t lo = per * i, hi = lo + per < n ? lo + per : n;
        if (lo >= hi) break;
        swtb_job x = {lo, hi, q, t, qoff, toff, qlen, tlen, w, h0, mat, o_del, e_del, o_ins, e_ins,
                      zdrop, pen_clip3, out, cells, n_cigar};
        jobs[nt] = x;
        pthread_create(&th[nt], NULL, swtb_worker, &jobs[nt]);
        ++nt;
    }
    for (int i = 0; i < nt; ++i) pthread_join(th[i], NULL);
    return 0;
}
