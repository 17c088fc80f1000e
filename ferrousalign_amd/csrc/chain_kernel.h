// Device side of the chain-extension path (internal to the library; reached
// through fa_chains_extend / fa_extend_chains): the extension-job producer,
// the packing of job queries and targets into the extension kernels' 2-bit
// layout, per-job sort keys for the length packing, the right jobs' h0 from
// the left scores, and the per-chain region merge.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ferrousalign_amd.h"

namespace fa {

constexpr int kChainRanks = 64;  // distinct kernel choices a sort key can name (6 bits)

// Device inputs: reads, chains, seeds (the fa_extend_chains arrays) and the
// exclusive prefix of chain_seed_count over chains (slot_base, nchains + 1
// entries): chain c's seed k owns job slot slot_base[c] + k on each side.
struct ChainIn {
    const uint8_t *qseq;  // readable from qseq - 64 to qseq + qbytes + 64 (padding)
    int64_t qbytes;
    const int64_t *qoff;
    const int32_t *qlen;
    int64_t nreads;
    const int64_t *rcs;
    const int32_t *rcc;
    int64_t nchains;
    const uint8_t *kept;
    const float *frac;  // may be null (frac_rep 0)
    const int64_t *cstart;
    const int32_t *ccount;
    const int64_t *slot_base;
    const int64_t *cseeds;
    int64_t ncseeds;
    const fa_seed *seeds;
    int64_t nseeds;
};

// MemOpt fields of the producer and the merge, and the extension scoring
// fields the batch needs (band capping by end_bonus, 16-bit eligibility).
struct ChainOpt {
    int32_t a, o_del, e_del, o_ins, e_ins, w, pen_clip5, pen_clip3;
    int64_t l_pac;
    int32_t max_score, end_bonus, sw_o_del, sw_e_del, sw_o_ins, sw_e_ins;
    int32_t k16;      // max(2, 1 + max(mat)): the 16-bit kernel's guard factor
    int32_t sw16ok;   // gap costs allow the 16-bit kernel
};

// Kernel rank of a job: rank[fits16 level][qlen] (levels as fa_host::fits16;
// ranks order the kernels widest first).
struct ChainRanks {
    uint8_t r[3][256];
};

// One side (left or right) of the job slots.  qlen[s] == 0: no job in slot s.
struct ChainSide {
    int32_t *qlen, *tlen, *h0, *w;
    int64_t *qsrc;        // offset in qseq of the job's query (a left job reads it backwards)
    int64_t *seg;
    uint64_t *okey;       // (read << 33) | (position of the job in the read's job order << 4)
    uint8_t *err;         // right side: first validation failure known before h0 (negated code)
    uint64_t *words;      // seq2 words of the job (0: none)
    uint64_t *wend;       // inclusive prefix of words
    uint32_t *qo, *to;
    uint32_t *seq2, *seqn;
    uint32_t *key, *slot;  // sort input
    int32_t *out;          // 6 per slot, OutScore order
};

// Status block; keys: all ones = no error.
struct ChainStatus {
    unsigned long long left_key, right_key;  // okey | -code of the first invalid job
    uint32_t arg_err;
    uint32_t tmax[2];
    uint32_t any_job;  // some chain gave a job (either side, valid or not)
    unsigned long long count[2][kChainRanks];
    unsigned long long work[2][kChainRanks];
};

hipError_t chain_init(ChainStatus *st, int32_t *chain_read, int64_t nchains, hipStream_t s);
hipError_t chain_owner(const ChainIn &in, int32_t *chain_read, ChainStatus *st, hipStream_t s);
hipError_t chain_jobs(const ChainIn &in, const ChainOpt &o, const int32_t *chain_read,
                      uint8_t *chain_state, ChainSide L, ChainSide R, ChainStatus *st,
                      hipStream_t s);
// Queries and targets of both sides (left: query and target read backwards;
// right: forwards) into seq2/seqn at the scanned offsets.  pac must be
// readable 8 bytes past its last byte (fa_ref_upload pads it).
hipError_t chain_fill(const ChainIn &in, const uint8_t *pac, int64_t l_pac, ChainSide L, ChainSide R,
                      int64_t nslots, ChainStatus *st, hipStream_t s);
// Sort keys and rank counts of side `sd`; for the right side (sd == 1) first
// sets h0 from the left scores (or seed_len * a) and finishes validation.
hipError_t chain_keys(const ChainOpt &o, const ChainRanks &rk, ChainSide S, const ChainSide &Lft,
                      int sd, int64_t nslots, ChainStatus *st, hipStream_t s);
hipError_t chain_merge(const ChainIn &in, const ChainOpt &o, const int32_t *chain_read,
                       const uint8_t *chain_state, const ChainSide &L, const ChainSide &R,
                       fa_region *regions, hipStream_t s);

}  // namespace fa
