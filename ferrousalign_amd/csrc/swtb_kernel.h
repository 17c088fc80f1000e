// Banded local Smith-Waterman with traceback (scalar_banded_swa semantics):
// argument block and launchers (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

// Device layout: sequences as in ExtArgs (2-bit words + N mask).  Per pair p
// (job order): qo/to word offsets, qlen, tlen, w, h0, and the traceback
// block: tlen rows of L * ((S + 15) / 16) words from tbo[p], lane l's S
// columns (2 bits each, column l*S + s at bits 2*(s % 16) of word
// l * ((S+15)/16) + s/16) of DP row i at word i * L * ((S+15)/16).
// lsl[p] = (L << 8) | S of the kernel that ran the pair, | kSwtbBand for the
// band-window kernel, whose rows store the L*S columns of their window
// (starting at column (i / S) * S - w[p]) instead of columns 0 .. L*S-1.
struct SwtbArgs {
    const uint32_t *seq2;
    const uint32_t *seqn;
    const uint32_t *qo;
    const uint32_t *to;
    const int32_t *qlen;
    const int32_t *tlen;
    const int32_t *w;
    const int32_t *h0;
    const uint32_t *order;   // packed position -> pair id
    const uint64_t *tbo;     // traceback block offset (words) per pair
    const uint32_t *lsl;     // (L << 8) | S per pair
    uint32_t *tb;            // 128-byte aligned, readable 128 bytes past the last block (whole lines)
    int32_t *out;            // 6 int32 per pair, OutScore order
    uint32_t *cells;
    // traceback outputs, per pair at the given offsets
    const uint64_t *cig_off;   // capacity qlen + tlen + 4 ops
    uint32_t *cigar;           // (len << 4) | op, op M 0, I 1, D 2, S 4
    int32_t *n_cigar;
    const uint64_t *aln_off;   // capacity qlen + tlen + 2 bases
    uint8_t *ref_aln;
    uint8_t *qry_aln;
    int32_t *n_ref;
    int32_t *n_qry;
    uint32_t profw[5];       // per query code q: bytes mat[b*5+q]+128 for target b = 0..3
    uint32_t nrow;           // mat[20]+128 (uniform N target row)
    int32_t o_del, e_del, o_ins, e_ins, zdrop, pen_clip3;
    int32_t tpad;            // LDS bytes of staged target per pair
    int32_t qpad;            // LDS bytes of staged query per pair (band-window kernel)
};

constexpr uint32_t kSwtbBand = 1u << 16;

// Slot count S for L lanes with L*S >= qlen+1, or -1 (S <= 32).
int swtb_slot_count(int L, int qlen);
// Band-window configuration (L lanes x S slots) for a pair whose band
// [i - w, i + w + 1] fits 2w + S + 1 <= L*S columns and L*S < full_cols (the
// full-width layout's columns); false when the full-width kernel is the better fit.
bool swtb_band_config(int qlen, int w, int full_cols, int *L, int *S);
hipError_t swtb_dp_launch(int L, int S, bool band, const SwtbArgs &a, uint32_t first,
                          uint32_t count, hipStream_t st);
hipError_t swtb_traceback_launch(const SwtbArgs &a, uint32_t n, hipStream_t st);
// Gather each pair's CIGAR, aligned reference and aligned query into dense
// buffers at dense[3p], dense[3p+1], dense[3p+2] (prefix sums of the counts).
hipError_t swtb_compact_launch(const SwtbArgs &a, uint32_t n, const uint64_t *dense,
                               uint32_t *cigar, uint8_t *ref_aln, uint8_t *qry_aln,
                               hipStream_t st);

}  // namespace fa
