/*
 * CPU oracle for the banded seed-extension path.  TEST INFRASTRUCTURE ONLY:
 * this code is the checker that the GPU kernels are compared against.  Only
 * the tests, the smoke check and the CPU-baseline leg of the benchmark may
 * load it; the product library never links or calls it.
 *
 * Every function restates the reference Rust source it cites (paths are
 * relative to the FerrousAlign repository root).
 */
#ifndef FA_KSW_ORACLE_H
#define FA_KSW_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* KswExtendResult, src/core/alignment/ksw_affine_gap.rs:902-916 */
typedef struct {
    int32_t score, qle, tle, gtle, gscore, max_off;
} ora_ext_result;

/* Return codes (the reference panics or reads out of bounds instead). */
#define ORA_OK 0
#define ORA_ERR_H0 (-2)      /* h0 <= 0: assert at ksw_affine_gap.rs:968 */
#define ORA_ERR_QLEN (-3)    /* qlen <= 0: eh[1] out of bounds at :993 */
#define ORA_ERR_CODE (-4)    /* base code >= m: qp index out of bounds at :985 */
#define ORA_ERR_TLEN (-5)    /* tlen < 0 */

/* ksw_extend2, src/core/alignment/ksw_affine_gap.rs:952-1169.
 * `cells` (may be NULL) receives the number of DP cells the row loop visits:
 * the sum over executed rows of (end - beg) for the band [beg, end) of that
 * row, counting the row on which the loop breaks. */
int ora_ksw_extend2(int32_t qlen, const uint8_t *query, int32_t tlen,
                    const uint8_t *target, int32_t m, const int8_t *mat,
                    int32_t o_del, int32_t e_del, int32_t o_ins, int32_t e_ins,
                    int32_t w, int32_t end_bonus, int32_t zdrop, int32_t h0,
                    ora_ext_result *out, uint64_t *cells);

/* Batch driver over contiguous u8 code buffers with per-pair offsets
 * (the layout of ExtensionJobBatch, batch_extension/types.rs:67-98).
 * Runs pairs [0, n) with `nthreads` POSIX threads (<=0: one per core).
 * Per-pair return codes go to `rc` (may be NULL).  Returns the number of
 * pairs whose code was not ORA_OK. */
int64_t ora_extend_batch(int64_t n, const uint8_t *qseq, const int64_t *qoff,
                         const int32_t *qlen, const uint8_t *tseq,
                         const int64_t *toff, const int32_t *tlen,
                         const int32_t *h0, const int32_t *w, int32_t m,
                         const int8_t *mat, int32_t o_del, int32_t e_del,
                         int32_t o_ins, int32_t e_ins, int32_t end_bonus,
                         int32_t zdrop, ora_ext_result *out, uint64_t *cells,
                         int32_t *rc, int nthreads);
/* As ora_extend_batch, plus rows[k] (optional): target rows job k's row loop
 * ran, the row it broke on included (0 for an invalid job). */
int64_t ora_extend_batch_rows(int64_t n, const uint8_t *qseq, const int64_t *qoff,
                              const int32_t *qlen, const uint8_t *tseq,
                              const int64_t *toff, const int32_t *tlen,
                              const int32_t *h0, const int32_t *w, int32_t m,
                              const int8_t *mat, int32_t o_del, int32_t e_del,
                              int32_t o_ins, int32_t e_ins, int32_t end_bonus,
                              int32_t zdrop, ora_ext_result *out, uint64_t *cells,
                              int32_t *rc, int32_t *rows, int nthreads);

/* Target segment of a 2-bit .pac reference (pac_oracle.c): `len` bases from
 * `start` in the bidirectional coordinates [0, 2 l_pac) (>= l_pac: reverse
 * complement), written backwards when `reverse`.  Returns the number of bases,
 * or -1 for a segment outside [0, 2 l_pac) or across l_pac. */
int64_t ora_pac_segment(const uint8_t *pac, int64_t l_pac, int64_t start, int64_t len,
                        int reverse, uint8_t *out);

/* scalar_banded_swa (swtb_oracle.c): banded local SW with traceback, score
 * record in `out`, CIGAR as (len << 4) | op (M 0, I 1, D 2, S 4) into
 * cigar[0 .. *n_cigar), aligned reference and query bases (capacity
 * qlen + tlen + 2 each, CIGAR capacity qlen + tlen + 4).  Returns ORA_OK. */
int ora_scalar_banded_swa(int32_t qlen, const uint8_t *query, int32_t tlen, const uint8_t *target,
                          const int8_t *mat, int32_t o_del, int32_t e_del, int32_t o_ins,
                          int32_t e_ins, int32_t zdrop, int32_t pen_clip3, int32_t w, int32_t h0,
                          ora_ext_result *out, uint32_t *cigar, int32_t *n_cigar,
                          uint8_t *ref_aligned, int32_t *n_ref, uint8_t *query_aligned,
                          int32_t *n_query, uint64_t *cells);
/* The same over a batch on `nthreads` host threads (0 = all); records and
 * cell counts only (CPU baseline). */
int64_t ora_scalar_banded_swa_batch(int64_t n, const uint8_t *q, const int64_t *qoff,
                                    const int32_t *qlen, const uint8_t *t, const int64_t *toff,
                                    const int32_t *tlen, const int32_t *w, const int32_t *h0,
                                    const int8_t *mat, int32_t o_del, int32_t e_del,
                                    int32_t o_ins, int32_t e_ins, int32_t zdrop,
                                    int32_t pen_clip3, ora_ext_result *out, uint64_t *cells,
                                    int32_t *n_cigar, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
