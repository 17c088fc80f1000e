// Mate-rescue local SW kernel (ksw_align2, 16-bit and 8-bit modes): argument
// block and launcher (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

constexpr uint32_t kKswXByte = 0x10000u;
constexpr uint32_t kKswXStop = 0x20000u;
constexpr uint32_t kKswXSubo = 0x40000u;
constexpr uint32_t kKswXStart = 0x80000u;

struct MateArgs {
    const uint32_t *seq2;
    const uint32_t *seqn;
    const uint32_t *qo;
    const uint32_t *to;
    const int32_t *qlen;
    const int32_t *tlen;
    const uint32_t *order;
    int32_t *out;         // 7 int32 per pair: score, te, qe, score2, te2, tb, qb
    uint32_t *cells;
    uint16_t *imax_log;   // per launch slot: log_stride row maxima
    uint32_t log_stride;
    uint32_t profw[5];    // per query code q: bytes mat[b*5+q]+128 for target b = 0..3
    int32_t o_del, e_del, o_ins, e_ins;
    uint32_t xtra;
    int32_t qmax;         // largest matrix entry (>= 1)
    int32_t byte_mode;    // 1: the 8-bit mode (KSW_XBYTE)
    int32_t ovf;          // 8-bit mode: 255 - shift, the best score that saturates
    int32_t split_rev;    // 1: the forward kernel leaves the XSTART pass to mate16_rev_launch
    const uint32_t *rorder;  // reverse launches: pair ids in launch order
};

// Longest stripe (segments per lane) a launch takes: qlen <= 256.
constexpr int kMateMaxSlen = 32;   // 16-bit mode, 8 lanes
constexpr int kMate8MaxSlen = 16;  // 8-bit mode, 16 lanes
// Launch over packed positions [first, first+count), two pairs per 8-lane
// group, 16 pairs per 64-thread block: every pair of [first, first+count) must
// have slen = ceil(qlen / 8) = `slen` (1..32), in the 8-bit mode ceil(qlen / 16)
// (1..16, the reference's 16 byte lanes, two per GPU lane); imax_log must hold
// count * log_stride entries.
hipError_t mate16_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                         hipStream_t st);
// The XSTART reverse pass alone, for the pairs rorder[first, first+count),
// whose forward records (score, te, qe) are in out: every pair of the range
// must need the pass and have a reverse stripe length ceil((qe + 1) / lanes)
// equal to `slen`.  Writes tb and qb and adds the pass's cells.
hipError_t mate16_rev_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st);
// Reverse-pass sort keys of pairs [0, n) from their forward records: bin
// (kMateMaxSlen - reverse stripe length, 63 for a pair without the pass) in
// bits 15-20, 32767 - forward score below; ids[k] = k; hist[bin] counts the
// pairs of each bin (64 entries, zero before the call).
constexpr int kMateRevKeyBits = 21;
hipError_t mate_rev_keys(const MateArgs &a, uint32_t n, uint32_t *keys, uint32_t *ids,
                         uint32_t *hist, hipStream_t st);

}  // namespace fa
