// Kernel-side argument block and launch table of the banded seed-extension
// kernels (internal to the library; not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

// Device layout of a batch (all arrays in HBM):
//   seq2  : 2-bit packed bases, 16 per 32-bit word, base p of a sequence
//           starting at word `off` is (seq2[off + p/16] >> 2*(p%16)) & 3;
//           every sequence starts on an even word (32-base boundary)
//   seqn  : N mask, 1 bit per base, the mask of that sequence starts at word
//           off/2: base p is N when (seqn[off/2 + p/32] >> p%32) & 1
//   qo/to : per-pair word offsets of query and target in seq2
//   qlen, tlen, h0 : per pair; w : per pair, already capped by end_bonus
//   order : permutation of pair ids giving the packing into lane groups
//   out   : 6 int32 per pair in OutScore field order; cells : DP cells;
//           or, with orec, one 32-byte record per pair: the six OutScore
//           fields, the DP cells and a zero word (cells unused)
struct ExtArgs {
    const uint32_t *seq2;
    const uint32_t *seqn;
    const uint32_t *qo;
    const uint32_t *to;
    const int32_t *qlen;
    const int32_t *tlen;
    const int32_t *h0;
    const int32_t *w;
    const uint32_t *order;
    uint32_t *meta;      // 8 words per packed position: pid, qo, to, qlen, tlen, h0, w, nfree
                         // (nfree bit 0: the query has no N base, bit 1: the target has
                         // none; their N-mask words are then not read)
    const uint32_t *params;  // optional: the same 8 words per pair in job order (one 32-byte
                             // read per pair for meta_gather instead of six scattered ones)
    int32_t *out;
    uint32_t *cells;
    uint32_t *queue;     // persistent launch: work counter (one per launch)
    uint32_t *ovf;       // 8-bit path: per-pair overflow flags
    uint32_t profw[5];   // per query code q: bytes mat[b*5+q]+128 for target b = 0..3
    uint32_t nrow;       // mat[20]+128 (the N target row is uniform)
    int32_t o_del, e_del, o_ins, e_ins, zdrop;
    int32_t mguard;      // 16-bit kernel: K = max(2, 1 + max(mat)), M = min(H + s, H * K)
    int32_t tpad;        // LDS bytes of staged target codes (one per row, multiple of 16, > tlen):
                         // per pair, or per lane group of two pairs in the packed kernel
    int32_t chunk;       // persistent launch: pairs per queue pull (multiple of 64/L)
    int32_t orec;        // out holds 32-byte records (see above)
};

#ifdef __HIPCC__
// Result of pair pid.  A 32-byte record fills one aligned 32-byte sector of
// HBM with two 16-byte stores; the 24-byte OutScore plus a separate cell
// word touch two or three sectors when pairs finish in packed order.
__device__ __forceinline__ void store_result(const ExtArgs &a, uint32_t pid, int score, int tle,
                                             int gtle, int qle, int gscore, int max_off,
                                             uint32_t cells) {
    if (a.orec) {
        int4 *o = reinterpret_cast<int4 *>(a.out) + 2 * (size_t)pid;
        o[0] = make_int4(score, tle, gtle, qle);
        o[1] = make_int4(gscore, max_off, (int)cells, 0);
        return;
    }
    int32_t *o = a.out + (size_t)pid * 6;
    o[0] = score;
    o[1] = tle;
    o[2] = gtle;
    o[3] = qle;
    o[4] = gscore;
    o[5] = max_off;
    if (a.cells) a.cells[pid] = cells;
}
#endif

// Smallest instantiated slot count S with L*S >= qlen+1, or -1.
int extend_slot_count(int L, int qlen);
hipError_t extend_launch_plain(int L, int S, const ExtArgs &a, uint32_t first, uint32_t count,
                               hipStream_t st);
// Persistent launch over packed positions [first, first+count) using work
// counter a.queue[qidx] (must be zero); grid = resident blocks per CU (capped
// by blocks_per_cu when > 0) times `cus`.
hipError_t extend_launch_persistent(int L, int S, const ExtArgs &a, uint32_t first,
                                    uint32_t count, uint32_t qidx, int cus, int blocks_per_cu,
                                    hipStream_t st);
hipError_t queue_reset(uint32_t *q, uint32_t n, hipStream_t st);

// Fills a.meta for packed positions [0, n) from a.order and the per-pair
// arrays (or a.params when set), so that a wavefront reads its pairs'
// parameters as one contiguous block.
hipError_t meta_gather(const ExtArgs &a, uint32_t n, hipStream_t st);
// Writes params (8 words per pair, job order) from the per-pair arrays of a,
// with each pair's nfree bits from its N-mask words.
hipError_t params_pack(const ExtArgs &a, uint32_t n, uint32_t *params, hipStream_t st);
// 32-byte records of pairs [0, n) to out (6 int32 per pair) and cells, both
// in job order (coalesced on both sides).
hipError_t records_split(const int32_t *rec, uint32_t n, int32_t *out, uint32_t *cells,
                         hipStream_t st);

// 16-bit packed kernel (two pairs per lane group): slot count for L lanes,
// or -1 (wide: the pair may take the slot counts above 16, which need
// Hmax < 2048), and its launcher (plain or persistent).
int extend16_slot_count(int L, int qlen, bool wide);
hipError_t extend16_launch(int L, int S, const ExtArgs &a, uint32_t first, uint32_t count,
                           bool persistent, uint32_t qidx, int cus, int blocks_per_cu,
                           hipStream_t st);

// Long pairs (query > 255 or target > 2048 bases): one wavefront per pair,
// eh[] in `scratch` (extend_long_scratch bytes for `count` pairs whose
// longest query is qmax), the query's codes in LDS (qmax <= 65519).
size_t extend_long_scratch(int qmax, uint32_t count);
hipError_t extend_long_launch(const ExtArgs &a, uint32_t first, uint32_t count, int qmax,
                              void *scratch, size_t scratch_bytes, hipStream_t st);

}  // namespace fa
