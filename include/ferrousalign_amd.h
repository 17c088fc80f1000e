/*
 * ferrousalign_amd — C ABI of the MI355X (gfx950) banded seed-extension path.
 *
 * Every entry point computes, for each extension pair, the result of the
 * reference's `ksw_extend2` (src/core/alignment/ksw_affine_gap.rs:952-1169):
 * banded affine-gap extension from score h0 with adaptive band, z-drop and
 * end-bonus width capping, returning score, qle, tle, gtle, gscore, max_off.
 * Plain pointers and sizes only; no C++ or framework types cross this ABI.
 *
 * Threading: every call is synchronous; a context may be used by one host
 * thread at a time (give each thread its own context, as each rayon worker
 * holds its own workspace in the reference, workspace.rs:155-160).
 * Errors: every call returns FA_OK (0) or a negative FA_ERR_* code and never
 * aborts the process.
 */
#ifndef FERROUSALIGN_AMD_H
#define FERROUSALIGN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------- */
#define FA_OK 0
#define FA_ERR_ARG (-1)      /* null pointer, negative count, bad flag */
#define FA_ERR_H0 (-2)       /* h0 <= 0 (the reference asserts h0 > 0, ksw_affine_gap.rs:968) */
#define FA_ERR_QLEN (-3)     /* qlen <= 0 (reference panics, :993) or qlen > FA_MAX_QLEN */
#define FA_ERR_CODE (-4)     /* base code > 4 (reference indexes out of bounds, :985) */
#define FA_ERR_TLEN (-5)     /* tlen < 0 or tlen > FA_MAX_TLEN */
#define FA_ERR_BAND (-6)     /* w < 0 (reference indexes eh[] out of bounds, :1041-1107) */
#define FA_ERR_SCORING (-7)  /* m != 5, negative penalty, N row of mat not uniform, score range */
#define FA_ERR_DEVICE (-8)   /* HIP runtime error */
#define FA_ERR_NODEVICE (-9) /* no usable gfx950 device */
#define FA_ERR_OVERFLOW (-10) /* 8-bit path only: a pair exceeded 8-bit range (FA_SCORE_U8 without rerun) */
#define FA_ERR_SEGMENT (-11)  /* reference target segment outside [0, 2 l_pac) or across l_pac */
#define FA_ERR_SCRATCH (-12)  /* batch exceeds the fixed scratch of fa_execute_batch_device_noalloc */

/* Lengths.  The extension entry points (fa_execute_batch_simd_scoring,
 * fa_batch_upload and the SoA entry points) take queries up to FA_MAX_QLEN
 * and targets up to FA_MAX_TLEN bases: pairs up to the _PACKED limits run in
 * the lane-group kernels, longer ones in a kernel of one wavefront per pair
 * (the reference's ksw_extend2 has no length limit; these keep a pair's
 * cell count within 32 bits).  The chain path, the allocation-free entry
 * point, CIGAR regeneration and ksw_align2 take the _PACKED limits. */
#define FA_MAX_QLEN 32767
#define FA_MAX_TLEN 65536
#define FA_MAX_QLEN_PACKED 255
#define FA_MAX_TLEN_PACKED 2048

/* ---- flags -------------------------------------------------------------- */
/* score path.  FA_SCORE_I16 (default): pairs whose scores provably fit run
 * the packed 16-bit kernel (two pairs per lane group); the others fall back
 * to the 32-bit kernel.  CDNA4 has no packed 8-bit integer ALU, so the 8-bit
 * entry points (FA_SCORE_U8, the reference's i8 kernels) take the same
 * 16-bit/32-bit route; every path returns the same exact records. */
#define FA_SCORE_I16 0x0u
#define FA_SCORE_U8 0x1u
#define FA_SCORE_AUTO 0x2u
#define FA_SCORE_I32 0x4u       /* force the 32-bit kernel for every pair */
/* packing of pairs into lane groups (output order never changes) */
#define FA_PACK_LENGTH 0x00u    /* sort by (qlen, tlen): the plain kernel */
#define FA_PACK_NONE 0x10u      /* input order */
#define FA_PACK_COST_ORDER 0x30u /* caller-supplied per-pair cost (fa_batch_set_cost), e.g. predicted
                                   DP cells: inside each kernel whose lane groups hold 64 or more
                                   columns, pairs in descending cost, so that a wavefront's pairs
                                   end together; length order for the narrower kernels */
#define FA_PACK_COST 0x20u      /* group by a caller-supplied per-pair cost (fa_batch_set_cost): cost
                                   classes e^(1/2) apart, length order inside a class */
/* launch strategy */
#define FA_LAUNCH_PLAIN 0x000u  /* one launch per query-length bucket */
#define FA_LAUNCH_PERSISTENT 0x100u /* persistent waves pulling chunks from a device queue */
#define FA_LAUNCH_GRAPH 0x200u  /* replay the launch sequence as a captured hipGraph */
#define FA_RUN_ALIGN2 0x1000u  /* run/time the batch's mate-rescue local SW (fa_batch_set_align2) */
#define FA_LAUNCH_SINGLE 0x400u /* one launch for the whole batch at the slot count of its
                                   longest query (no query-length buckets) */
#define FA_PLAN_FRESH 0x2000u  /* plan the packing again even when the batch's plan for these
                                  flags is current (per-batch planning cost in a timed step) */
#define FA_ALIGN2_FUSED 0x4000u /* with FA_RUN_ALIGN2: run each pair's XSTART reverse pass in the
                                    forward kernel (default: a second launch over the pairs that
                                    need it, bucketed by the stripe length of qe + 1; same results) */

/* ---- records ------------------------------------------------------------ */
/* BandedPairWiseSW, src/core/alignment/banded_swa/types.rs:106-199 */
typedef struct {
    int32_t m;          /* alphabet size; must be 5 (A,C,G,T,N) */
    int32_t end_bonus;
    int32_t zdrop;
    int32_t pen_clip5;  /* carried for the caller; not used by extension */
    int32_t pen_clip3;
    int32_t o_del;
    int32_t o_ins;
    int32_t e_del;
    int32_t e_ins;
    int8_t mat[25];     /* mat[target_base * 5 + query_base] */
} fa_banded_pairwise_sw;

/* OutScore, src/core/alignment/banded_swa/types.rs:77-84 (same field order).
 * target_end_pos = tle, gtarget_end_pos = gtle, query_end_pos = qle,
 * global_score = gscore, max_offset = max_off of KswExtendResult
 * (ksw_affine_gap.rs:902-916). */
typedef struct {
    int32_t score;
    int32_t target_end_pos;
    int32_t gtarget_end_pos;
    int32_t query_end_pos;
    int32_t global_score;
    int32_t max_offset;
} fa_out_score;

typedef struct fa_ctx fa_ctx;
typedef struct fa_batch fa_batch;
typedef struct fa_ref fa_ref;

/* ---- context ------------------------------------------------------------ */
/* Opens `device` (HIP ordinal).  Replaces the CPU-SIMD engine selection of
 * ComputeBackend::Gpu (src/core/compute/mod.rs:67-128, :141-167). */
int fa_ctx_create(int device, fa_ctx **out);
int fa_ctx_destroy(fa_ctx *ctx);
const char *fa_strerror(int code);
/* Library and device description (null-terminated, <= cap bytes). */
int fa_ctx_describe(fa_ctx *ctx, char *buf, size_t cap);

/* ---- pipeline-level entry point ----------------------------------------- */
/* Replaces execute_batch_simd_scoring (src/pipelines/linear/batch_extension/
 * dispatch.rs:31-86) over an ExtensionJobBatch (batch_extension/types.rs:
 * 20-197): job k's query is query_seqs[query_offset[k] .. + query_len[k]]
 * and its target ref_seqs[ref_offset[k] .. + ref_len[k]], u8 codes
 * A=0 C=1 G=2 T=3 N=4 (src/core/compute/encoding.rs:208-216).  Every job is
 * scored (no first-W truncation, no 127 clamp).  out[k] receives job k's
 * OutScore; cells[k] (optional) the DP cells its row loop visited.
 * per_job_rc (optional) receives each job's validation code; invalid jobs get
 * a zeroed record and the call returns the first error code.  The call's
 * device arrays stay with the context for the next call (grow-only: about
 * 180 bytes per job plus 6 bytes per 16 bases, 1.3 GB for 4M pairs of 100 +
 * 200 bases; freed by fa_ctx_destroy). */
int fa_execute_batch_simd_scoring(fa_ctx *ctx, const fa_banded_pairwise_sw *sw,
                                  int64_t njobs, const uint8_t *query_seqs,
                                  const int64_t *query_offset, const int32_t *query_len,
                                  const uint8_t *ref_seqs, const int64_t *ref_offset,
                                  const int32_t *ref_len, const int32_t *h0,
                                  const int32_t *band_width, uint32_t flags,
                                  fa_out_score *out, uint32_t *cells, int32_t *per_job_rc);

/* ---- concurrent callers: coalescing aggregator ---------------------------- */
/* The reference calls execute_batch_simd_scoring from every rayon worker at
 * once, each call synchronous (src/pipelines/linear/single_end.rs:92-110),
 * and a GPU pays off only on large batches (src/core/compute/mod.rs:277-287).
 * An aggregator owns one context on `device` and one worker thread: calls of
 * fa_agg_execute from any number of host threads are coalesced, the worker
 * taking every job submitted so far once min_jobs jobs are pending or
 * max_wait_us after the oldest pending call arrived, and running them as one
 * batch per (scoring, flags).  Each call returns when its own jobs are done,
 * with exactly what fa_execute_batch_simd_scoring would return for them (same
 * arguments and outputs; the return code is the call's own first invalid
 * job's, or a device error of the run).  Destroy only when no call is in
 * flight; calls still pending then are run first. */
typedef struct fa_agg fa_agg;
int fa_agg_create(int device, int32_t min_jobs, int32_t max_wait_us, fa_agg **out);
int fa_agg_destroy(fa_agg *agg);
int fa_agg_execute(fa_agg *agg, const fa_banded_pairwise_sw *sw, int64_t njobs,
                   const uint8_t *query_seqs, const int64_t *query_offset,
                   const int32_t *query_len, const uint8_t *ref_seqs, const int64_t *ref_offset,
                   const int32_t *ref_len, const int32_t *h0, const int32_t *band_width,
                   uint32_t flags, fa_out_score *out, uint32_t *cells, int32_t *per_job_rc);
/* Calls served, batches run and jobs scored since fa_agg_create. */
int fa_agg_stats(fa_agg *agg, int64_t *calls, int64_t *runs, int64_t *jobs);

/* ---- allocation-free form (no counterpart in the reference) -------------- */
/* For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * computation and job layout as fa_execute_batch_simd_scoring, but every
 * pointer is device memory (d_out[k] = job k's OutScore, d_cells and d_job_rc
 * optional) and the call allocates nothing: validation, 2-bit packing,
 * packing order and launch metadata live in one fixed per-device scratch
 * array (64 MiB, set up by fa_ctx_create).  A job is limited to
 * ref_len <= max_ref_len (<= FA_MAX_TLEN_PACKED; longer jobs get FA_ERR_TLEN,
 * and so do queries over FA_MAX_QLEN_PACKED); a
 * batch needs about njobs * (6 * (16 + 2 * ceil(max_ref_len / 32)) + 52)
 * bytes of it, and a larger batch is rejected with FA_ERR_SCRATCH before
 * anything runs.  Synchronous on the context's stream.  Calls on one device
 * share the scratch: calls that overlap on one device race. */
int fa_execute_batch_device_noalloc(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                                    const uint8_t *d_query_seqs, const int64_t *d_query_offset,
                                    const int32_t *d_query_len, const uint8_t *d_ref_seqs,
                                    const int64_t *d_ref_offset, const int32_t *d_ref_len,
                                    const int32_t *d_h0, const int32_t *d_band_width,
                                    int32_t max_ref_len, fa_out_score *d_out, uint32_t *d_cells,
                                    int32_t *d_job_rc);

/* ---- kernel-level entry points (the banded_swa batch surface) ----------- */
/* Replace simd_banded_swa_batch{16,32,64}_soa (8-bit; isa_sse_neon.rs:10-16,
 * isa_avx2.rs:64-70, isa_avx512_int8.rs:7-13) and
 * simd_banded_swa_batch{8,16,32}_int16_soa (16-bit; isa_sse_neon.rs:26-32,
 * isa_avx2.rs:72-78, isa_avx512_int16.rs:9-15), generated by
 * banded_swa/shared.rs:161-389.  Inputs use the SoAInputs layout
 * (shared.rs:12-35): query_soa[pos * lanes + lane], target_soa likewise,
 * padded with 0xFF past each lane's length; per-lane lengths, h0 and w are
 * int32 here (the reference's i8 lanes clamp at 127).  end_bonus is passed
 * explicitly (the reference carries it only in BandedPairWiseSW).
 * num_jobs <= lanes; out[0..num_jobs). */
int fa_simd_banded_swa_batch_soa(fa_ctx *ctx, int32_t lanes, const uint8_t *query_soa,
                                 const uint8_t *target_soa, const int32_t *qlen,
                                 const int32_t *tlen, const int32_t *h0, const int32_t *w,
                                 int32_t num_jobs, int32_t o_del, int32_t e_del, int32_t o_ins,
                                 int32_t e_ins, int32_t zdrop, int32_t end_bonus,
                                 const int8_t *mat, int32_t m, fa_out_score *out);
int fa_simd_banded_swa_batch_int16_soa(fa_ctx *ctx, int32_t lanes, const int16_t *query_soa,
                                       const int16_t *target_soa, const int32_t *qlen,
                                       const int32_t *tlen, const int32_t *h0,
                                       const int32_t *w, int32_t num_jobs, int32_t o_del,
                                       int32_t e_del, int32_t o_ins, int32_t e_ins,
                                       int32_t zdrop, int32_t end_bonus, const int8_t *mat,
                                       int32_t m, fa_out_score *out);

/* ---- sequence layout in HBM --------------------------------------------- */
/* 2-bit packing used on the device: base p of a sequence is
 * (w2[p/16] >> 2*(p%16)) & 3, or N (code 4) when (wn[p/32] >> p%32) & 1.
 * w2 must hold 2*ceil(len/32) words and wn ceil(len/32) words.  The 2-bit
 * order inside a word differs from the reference's .pac (MSB-first, 4 bases
 * per byte, index/bntseq.rs:563-575); N has no 2-bit code there.
 * FA_ERR_CODE when a code exceeds 4 (w2 and wn are then undefined). */
int fa_pack_2bit(const uint8_t *codes, int32_t len, uint32_t *w2, uint32_t *wn);
int fa_unpack_2bit(const uint32_t *w2, const uint32_t *wn, int32_t len, uint8_t *codes);

/* ---- device-resident batches --------------------------------------------- */
/* Validate, 2-bit pack and upload a batch once; run it many times with the
 * sequences resident in HBM; fetch results.  Same job layout as
 * fa_execute_batch_simd_scoring.  The batch keeps its results on the device
 * until fa_batch_download. */
int fa_batch_upload(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                    const uint8_t *query_seqs, const int64_t *query_offset,
                    const int32_t *query_len, const uint8_t *ref_seqs,
                    const int64_t *ref_offset, const int32_t *ref_len, const int32_t *h0,
                    const int32_t *band_width, fa_batch **out);
/* ---- CIGAR regeneration: banded local SW with traceback --------------------
 * Batched scalar_banded_swa (src/core/alignment/banded_swa/scalar/
 * implementation.rs:5-440), the aligner that region.rs:1058 calls once per
 * alignment region to rebuild its CIGAR.  Job j aligns query_len[j] codes at
 * query_seqs + query_offset[j] against ref_len[j] codes at ref_seqs +
 * ref_offset[j] with band band_width[j] (used as given) and start score
 * h0[j] >= 0.  Outputs per job: the OutScore record, the CIGAR as
 * (len << 4) | op (M 0, I 1, D 2, S 4; the reference's (op char, len) pairs)
 * at cigar + cigar_offset[j] with room for query_len + ref_len + 4 entries,
 * its length in n_cigar[j], and the aligned reference and query bases at
 * ref_aligned / query_aligned + aligned_offset[j] with room for
 * query_len + ref_len + 2 codes each (lengths n_ref_aligned[j],
 * n_query_aligned[j]).  An empty query or reference gives the reference's
 * zero record and an empty CIGAR.  cells may be NULL.  Limits as for the
 * lane-group kernels (query <= FA_MAX_QLEN_PACKED, reference <=
 * FA_MAX_TLEN_PACKED, uniform N row).  The call's arrays stay with the
 * context for the next call (grow-only, freed with it): about 140 bytes of
 * pinned host memory per job, the packed sequences, and on the device the
 * traceback codes (about 2.5 KB per 150-base job) and the outputs.  Calls on
 * one context do not overlap. */
int fa_scalar_banded_swa_batch(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                               const uint8_t *query_seqs, const int64_t *query_offset,
                               const int32_t *query_len, const uint8_t *ref_seqs,
                               const int64_t *ref_offset, const int32_t *ref_len,
                               const int32_t *band_width, const int32_t *h0, fa_out_score *out,
                               uint32_t *cells, uint32_t *cigar, const int64_t *cigar_offset,
                               int32_t *n_cigar, uint8_t *ref_aligned, uint8_t *query_aligned,
                               const int64_t *aligned_offset, int32_t *n_ref_aligned,
                               int32_t *n_query_aligned);

/* Same computation as fa_scalar_banded_swa_batch, with each job's CIGAR and
 * aligned bases written back to back in job order instead of into caller
 * slots: the reference returns them as one Vec per alignment
 * (scalar/implementation.rs:5-13), and a packed buffer plus offsets is their
 * flat form.  cigar_offset, ref_aligned_offset and query_aligned_offset get
 * njobs + 1 entries: job k's entries are [off[k], off[k+1]).  Capacities in
 * entries (CIGAR words; bytes for each aligned buffer); FA_ERR_ARG when a
 * batch needs more (sum of qlen + tlen + 4 and qlen + tlen + 2 is always
 * enough).  Only the filled part of each buffer is written. */
int fa_scalar_banded_swa_batch_dense(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                                     const uint8_t *query_seqs, const int64_t *query_offset,
                                     const int32_t *query_len, const uint8_t *ref_seqs,
                                     const int64_t *ref_offset, const int32_t *ref_len,
                                     const int32_t *band_width, const int32_t *h0,
                                     fa_out_score *out, uint32_t *cells, uint32_t *cigar,
                                     int64_t cigar_cap, int64_t *cigar_offset,
                                     uint8_t *ref_aligned, int64_t *ref_aligned_offset,
                                     uint8_t *query_aligned, int64_t *query_aligned_offset,
                                     int64_t aligned_cap);

/* Kernel time (HIP events on the context's stream) of the last
 * fa_scalar_banded_swa_batch{,_dense} call: the DP launches and the traceback launch,
 * with the inputs already in HBM (measurement aid). */
int fa_scalar_banded_swa_last_ms(fa_ctx *ctx, double *dp_ms, double *traceback_ms);

/* ---- targets from a device-resident reference ------------------------------
 * The reference's 2-bit .pac (src/pipelines/linear/index/bntseq.rs:559-575:
 * base p in byte p/4, first base in the two most significant bits, codes
 * A=0 C=1 G=2 T=3), l_pac forward bases, kept in HBM; replaces the host-side
 * segment fetch get_reference_segment (:451-522) on the extension path. */
int fa_ref_upload(fa_ctx *ctx, const uint8_t *pac, int64_t l_pac, fa_ref **out);
int fa_ref_free(fa_ctx *ctx, fa_ref *ref);
/* Number of forward bases (l_pac) of an uploaded reference. */
int fa_ref_length(const fa_ref *ref, int64_t *l_pac);
/* As fa_batch_upload, but target j is read on the device from the reference:
 * the segment [seg_beg[j], seg_beg[j] + seg_len[j]) of the bidirectional
 * coordinate space [0, 2 l_pac) (coordinates >= l_pac are the reverse
 * complement, as get_reference_segment), read backwards when seg_rev[j] != 0
 * (a left-extension job, collect_soa.rs:156-176).  A segment outside
 * [0, 2 l_pac) or across l_pac fails with FA_ERR_SEGMENT.  `ref` must stay
 * alive only during this call. */
int fa_batch_upload_ref(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                        const uint8_t *query_seqs, const int64_t *query_offset,
                        const int32_t *query_len, const fa_ref *ref, const int64_t *seg_beg,
                        const int32_t *seg_len, const uint8_t *seg_rev, const int32_t *h0,
                        const int32_t *band_width, fa_batch **out);
/* The batch's targets as base codes, target j at codes + offset[j] (checking aid). */
int fa_batch_download_targets(fa_ctx *ctx, fa_batch *b, uint8_t *codes, const int64_t *offset);
/* Per-pair packing cost (e.g. predicted DP cells) used by FA_PACK_COST. */
int fa_batch_set_cost(fa_ctx *ctx, fa_batch *b, const float *cost);
/* Launch configuration: lanes per pair (4, 8 or 16, widened per query length
 * where that group is too narrow; 0 = the narrowest group that holds each
 * query length, the default) and the persistent-queue chunk size in pairs
 * (0 = default). */
int fa_batch_set_launch(fa_ctx *ctx, fa_batch *b, int32_t lanes_per_pair, int32_t chunk_pairs,
                        int32_t waves_per_cu);
/* Number of streams (1..4, default 4) the query-length buckets of one run
 * are spread over. */
int fa_batch_set_streams(fa_ctx *ctx, fa_batch *b, int32_t streams);
int fa_batch_run(fa_ctx *ctx, fa_batch *b, uint32_t flags);
/* Runs the batch `reps` times back to back and reports the mean time of one
 * run measured with HIP events on the batch's stream (kernel-only, ms) and the
 * mean duration of the dominant kernel launch (ms). */
int fa_batch_time(fa_ctx *ctx, fa_batch *b, uint32_t flags, int32_t reps, double *ms_per_run,
                  double *ms_dominant_launch, int32_t *launches_per_run);
/* Runs each kernel launch of one run once, alone on the batch's stream and
 * timed with HIP events: launch k covers packed positions [first[k],
 * first[k] + count[k]) of the pair order with slot count slots[k] and kernel
 * kernel[k] = (kind << 8) | lanes per pair, kind 0 = 32-bit extension,
 * 1 = packed 16-bit extension, 2 = mate-rescue local SW.  Fills at most cap
 * entries (any array may be NULL); *nlaunch receives the number of
 * launches. */
int fa_batch_profile_launches(fa_ctx *ctx, fa_batch *b, uint32_t flags, int32_t cap,
                              int32_t *nlaunch, int32_t *slots, int32_t *kernel, int64_t *first,
                              int64_t *count, double *ms);
/* VALU throughput probe (no counterpart in the reference; measurement only):
 * the integer add/max rate this device sustains in packed 16-bit form
 * (packed16 != 0, two 16-bit ops per lane per instruction) or 32-bit form,
 * in 10^12 lane-ops/s, the measured ceiling beside the published VALU peak
 * that the extension kernels' roofline is quoted against. */
int fa_valu_peak(fa_ctx *ctx, int32_t packed16, double *tera_ops);
/* HBM copy probe (no counterpart in the reference; measurement only): copies
 * `bytes` (a multiple of 16) between two device buffers with one vector-copy
 * kernel launch, twice, and reports the second launch's read+write rate in
 * GB/s.  Its per-dispatch FETCH_SIZE/WRITE_SIZE counters are the reference
 * point that scales the HBM byte counters of the other kernels. */
int fa_hbm_copy(fa_ctx *ctx, int64_t bytes, double *gbytes_per_s);
/* The pair order (packed position -> job index) used under `flags`. */
int fa_batch_get_order(fa_ctx *ctx, fa_batch *b, uint32_t flags, uint32_t *order);
int fa_batch_download(fa_ctx *ctx, fa_batch *b, fa_out_score *out, uint32_t *cells);
/* Overflow fallbacks of the last extension run: the number of pairs whose
 * scores could leave the packed 16-bit kernel's range (h0 + qlen * max(mat)
 * >= 4096 and the other bounds of the 16-bit route) and that therefore ran
 * the 32-bit kernel, the reference's i16-after-i8 rerun moved one level up
 * (dispatch.rs:100-107).  The 8-bit entry points share the 16-bit route
 * (gfx950 has no packed 8-bit add/max), so no 8-bit rerun is counted; 0
 * before the first run and after an FA_RUN_ALIGN2 run. */
int fa_batch_overflow_count(fa_ctx *ctx, fa_batch *b, int64_t *count);
int fa_batch_free(fa_ctx *ctx, fa_batch *b);

/* ---- mate-rescue local Smith-Waterman (ksw_align2, 16-bit and 8-bit) ---- */
/* Kswr, src/core/alignment/ksw_affine_gap.rs:51-61. */
typedef struct {
    int32_t score, te, qe, score2, te2, tb, qb;
} fa_kswr;
/* Replaces the per-job ksw_align2 call of execute_mate_rescue_batch_scalar
 * (src/pipelines/linear/paired/mate_rescue.rs:640-685; ksw_affine_gap.rs:
 * 773-861) for a whole batch: out[k] is exactly what ksw_align2 with the
 * reference's SSE engine returns for job k, including score2/te2 and the
 * XSTART reverse pass, in the mode xtra selects: 16-bit (8 lanes,
 * ksw_i16_impl :503-735) or, with KSW_XBYTE, 8-bit (16 lanes, ksw_u8_impl
 * :267-494; a best score reaching 255 - shift is reported as 255 with
 * qe = -1, as the reference does; mate_rescue.rs:314-321 sets KSW_XBYTE when
 * qlen * match < 250).  The N row of mat must score 0 (true of the
 * mate-rescue matrix, mate_rescue.rs:116-122); 1 <= qlen <= FA_MAX_QLEN_PACKED,
 * tlen <= 2032.  cells[k] (optional): target rows run times query length
 * over both passes. */
int fa_ksw_align2_batch(fa_ctx *ctx, int64_t n, const uint8_t *query_seqs,
                        const int64_t *query_offset, const int32_t *query_len,
                        const uint8_t *target_seqs, const int64_t *target_offset,
                        const int32_t *target_len, int32_t m, const int8_t *mat, int32_t o_del,
                        int32_t e_del, int32_t o_ins, int32_t e_ins, uint32_t xtra, fa_kswr *out,
                        uint32_t *cells);
/* Device-resident form: attach mate-rescue parameters to an uploaded batch,
 * then fa_batch_run / fa_batch_time with FA_RUN_ALIGN2. */
int fa_batch_set_align2(fa_ctx *ctx, fa_batch *b, const int8_t *mat, int32_t m, int32_t o_del,
                        int32_t e_del, int32_t o_ins, int32_t e_ins, uint32_t xtra);
int fa_batch_download_align2(fa_ctx *ctx, fa_batch *b, fa_kswr *out, uint32_t *cells);

/* ---- learned work predictor (no counterpart in the reference) ------------ */
/* A 6-32-32-1 MLP (FP16 storage, FP32 accumulation) predicting log(1 + DP
 * cells) of each pair from: qlen, tlen, h0, w (after end-bonus capping),
 * mismatches among the first min(32, qlen, tlen) positions of the ungapped
 * diagonal, and query GC fraction.  Used only to order pairs for packing
 * (FA_PACK_COST); it never changes an alignment result.  Parameters: 1,313
 * floats W1[32][6], b1[32], W2[32][32], b2[32], W3[32], b3; norm: 6 feature
 * means then 6 inverse standard deviations. */
typedef struct fa_mlp fa_mlp;
int fa_mlp_create(fa_ctx *ctx, fa_mlp **out);
int fa_mlp_destroy(fa_ctx *ctx, fa_mlp *m);
int fa_mlp_set_params(fa_ctx *ctx, fa_mlp *m, const float *params, const float *norm);
int fa_mlp_get_params(fa_ctx *ctx, fa_mlp *m, float *params);
/* Raw features of every pair of a resident batch, computed on the device
 * from its 2-bit sequences; copies them to feat[n][6] when feat != NULL. */
int fa_batch_features(fa_ctx *ctx, fa_batch *b, float *feat);
/* Predicted cells per pair (to cells_pred when non-NULL).  set_cost bit 0:
 * the predictions become the batch's packing cost (they stay in device
 * memory, where the packing plan reads them); bit 1: compute the batch's
 * features again, in the forward kernel (what a new batch costs; features
 * that fa_batch_features or training stored are used otherwise). */
int fa_mlp_predict(fa_ctx *ctx, fa_mlp *m, fa_batch *b, float *cells_pred, int32_t set_cost);
/* Mean loss 0.5*(y - log1p(cells))^2 over pairs[0..count) of the batch and
 * its gradient (labels: the cells recorded by the batch's last run). */
int fa_mlp_gradient(fa_ctx *ctx, fa_mlp *m, fa_batch *b, const uint32_t *pairs, int64_t count,
                    float *grad, float *mean_loss);
/* One pass of Adam over order[0..count) in minibatches (labels: the cells
 * recorded by the batch's last run); *mean_loss = mean training loss. */
int fa_mlp_train(fa_ctx *ctx, fa_mlp *m, fa_batch *b, const uint32_t *order, int64_t count,
                 int32_t minibatch, float lr, float *mean_loss);

/* ---- extension jobs from chains, and regions from their scores ----------
 * The callers either side of the extension batch: the job producer
 * collect_extension_jobs_batch_soa (src/pipelines/linear/batch_extension/
 * collect_soa.rs:28-228) and the consumer merge_scores_to_regions
 * (src/pipelines/linear/region.rs:737-890).  Layout (SoA, like the
 * reference's SoAChainBatch / SoASeedBatch): read r owns chains
 * [read_chain_start[r], +read_chain_count[r]); chain c owns the seed indices
 * chain_seeds[chain_seed_start[c] .. +chain_seed_count[c]) into seeds[]. */
typedef struct fa_mem_opt {   /* the MemOpt fields the path reads (mem_opt.rs) */
    int32_t a, b, o_del, e_del, o_ins, e_ins, w, zdrop, pen_clip5, pen_clip3;
} fa_mem_opt;
typedef struct fa_seed {      /* Seed: query_pos, ref_pos (FM coordinates), len */
    int64_t ref_pos;
    int32_t query_pos, len;
} fa_seed;
typedef struct fa_ext_job {   /* one extension job (batch_extension/types.rs:20-98) */
    int32_t read_idx, chain_idx;  /* chain index local to its read, as the reference */
    int64_t seed_idx;             /* into seeds[] */
    int32_t qbeg, qlen;           /* query segment of the read; a left job reads it backwards */
    int32_t h0, w;
    int64_t seg_beg;              /* target: .pac segment [seg_beg, +seg_len) in [0, 2 l_pac), */
    int32_t seg_len;              /* read backwards for a left job (seg_rev = 1) */
    int32_t seg_rev;
} fa_ext_job;
typedef struct fa_region {    /* AlignmentRegion boundaries and scores (region.rs:47-120) */
    int64_t rb, re;               /* FM coordinates, [rb, re) */
    int32_t qb, qe;
    int32_t score, truesc, seedcov, seedlen0;
    int32_t chain_idx;            /* local to its read */
    int32_t w;
    int64_t seed_idx;             /* best seed of the chain; -1: the chain gave no region */
    float frac_rep;
    int32_t pad;
} fa_region;

/* Host-side producer (collect_soa.rs:28-228): for every kept chain the
 * reference window [rmax0, rmax1) (cal_max_gap margins, chaining.rs:1108-1116,
 * kept on one strand), then per seed in reverse order a left job (query
 * prefix reversed, window prefix backwards, h0 = seed_len * a) and a right
 * job (query suffix, window suffix, h0 = 0 as the reference: fa_extend_chains
 * replaces it, see there).  Capacity: the total seed count of all chains
 * (jobs per direction and seed mappings).  Per chain c: mappings
 * [map_start[c], +map_count[c]) give the seed and its left/right job index
 * (-1 = none); rmax[2c], rmax[2c+1] the window (-1 when the chain made no
 * jobs).  FA_ERR_ARG for a seed outside its read (query_pos < 0, len < 0 or
 * query_pos + len > query_len).  Runs on the host only. */
int fa_collect_extension_jobs(const fa_mem_opt *opt, int64_t l_pac, int64_t nreads,
                              const int32_t *query_len, const int64_t *read_chain_start,
                              const int32_t *read_chain_count, const uint8_t *chain_kept,
                              const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                              const int64_t *chain_seeds, const fa_seed *seeds, int64_t capacity,
                              fa_ext_job *left, int64_t *n_left, fa_ext_job *right,
                              int64_t *n_right, int64_t *map_start, int32_t *map_count,
                              int64_t *map_seed, int64_t *map_left, int64_t *map_right,
                              int64_t *rmax);
/* Host-side consumer (region.rs:737-890): per chain with a window, one region
 * per seed mapping from the seed (qb, qe, rb, re, seed score) extended by its
 * left score (pen_clip5 decision: local qle/tle or global to the query start
 * with gtle) and right score (pen_clip3 decision), seedcov over the chain's
 * seeds inside the region, and the chain keeps the first highest-scoring
 * region.  right_from_left = 1: the right scores started from the left score
 * (fa_extend_chains), so the region score is the right score; 0: they started
 * from 0 and are added to it (the reference).  regions[c].seed_idx = -1 for
 * a chain without a region.  FM-to-chromosome coordinates are not filled. */
int fa_merge_scores_to_regions(const fa_mem_opt *opt, int64_t nreads, const int32_t *query_len,
                               const int64_t *read_chain_start, const int32_t *read_chain_count,
                               const float *chain_frac_rep, const int64_t *chain_seed_start,
                               const int32_t *chain_seed_count, const int64_t *chain_seeds,
                               const fa_seed *seeds, const int64_t *map_start,
                               const int32_t *map_count, const int64_t *map_seed,
                               const int64_t *map_left, const int64_t *map_right,
                               const int64_t *rmax, const fa_out_score *left_scores,
                               const fa_out_score *right_scores, int32_t right_from_left,
                               fa_region *regions);
/* ---- chain extension on the device --------------------------------------
 * The producer (as fa_collect_extension_jobs), both extension batches and the
 * merge (as fa_merge_scores_to_regions) on the GPU, with targets cut from the
 * resident .pac (fa_ref_upload).  The right jobs start from the seed's
 * left-extended score (or seed_len * a without a left job), as BWA-MEM2's
 * mem_chain2aln chains them: the reference's h0 = 0 right jobs are outside
 * ksw_extend2's h0 > 0 domain (ksw_affine_gap.rs:968), and region score and
 * clipping then use the right score alone (right_from_left = 1).  sw supplies
 * the scoring and end_bonus of both batches.  Replaces the producer / batch /
 * consumer sequence of collect_soa.rs:28-228, dispatch.rs:31-268 and
 * region.rs:737-890 for one chain batch.
 *
 * Every chain owned by a read gets one region (seed_idx -1: none); chains in
 * no read's range get the empty region.  Errors: FA_ERR_ARG for a chain in
 * two reads' ranges, a range or index outside its array, or a seed outside
 * its read; FA_ERR_SCORING for opt or (when there is a job) sw; otherwise the
 * error of the first invalid job in the reference's job order (left batch
 * first), as fa_batch_upload_ref would report it. */
typedef struct fa_chain_batch fa_chain_batch;
/* Device copy of reads (query_bytes codes; read r = [query_offset[r],
 * +query_len[r])), chains and seeds (n_chain_seeds indices into nseeds
 * seeds); chain_frac_rep may be NULL (frac_rep 0). */
int fa_chains_upload(fa_ctx *ctx, int64_t nreads, const uint8_t *query_seqs, int64_t query_bytes,
                     const int64_t *query_offset, const int32_t *query_len,
                     const int64_t *read_chain_start, const int32_t *read_chain_count,
                     int64_t nchains, const uint8_t *chain_kept, const float *chain_frac_rep,
                     const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                     int64_t n_chain_seeds, const int64_t *chain_seeds, int64_t nseeds,
                     const fa_seed *seeds, fa_chain_batch **out);
/* Runs the path on an uploaded batch; the regions stay on the device until
 * fa_chains_download.  stage_ms: NULL, or FA_CHAIN_STAGES entries of which
 * [1..6] receive the device time in ms of producer, left preparation, left
 * batch, right preparation, right batch, merge. */
#define FA_CHAIN_STAGES 8
int fa_chains_extend(fa_ctx *ctx, fa_chain_batch *cb, const fa_ref *ref,
                     const fa_banded_pairwise_sw *sw, const fa_mem_opt *opt, int64_t *n_left,
                     int64_t *n_right, double *stage_ms);
/* One fa_region per chain. */
int fa_chains_download(fa_ctx *ctx, fa_chain_batch *cb, fa_region *regions);
int fa_chains_free(fa_ctx *ctx, fa_chain_batch *cb);
/* Upload, extend, download over host buffers; stage_ms[0] and [7] receive the
 * host wall time of the upload and the download. */
int fa_extend_chains(fa_ctx *ctx, const fa_ref *ref, const fa_banded_pairwise_sw *sw,
                     const fa_mem_opt *opt, int64_t nreads, const uint8_t *query_seqs,
                     int64_t query_bytes, const int64_t *query_offset, const int32_t *query_len,
                     const int64_t *read_chain_start, const int32_t *read_chain_count,
                     int64_t nchains, const uint8_t *chain_kept, const float *chain_frac_rep,
                     const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                     int64_t n_chain_seeds, const int64_t *chain_seeds, int64_t nseeds,
                     const fa_seed *seeds, fa_region *regions, int64_t *n_left, int64_t *n_right,
                     double *stage_ms);

#ifdef __cplusplus
}
#endif
#endif /* FERROUSALIGN_AMD_H */
