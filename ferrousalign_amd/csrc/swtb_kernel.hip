// Banded local Smith-Waterman with traceback for gfx950, exact to the
// reference's scalar_banded_swa (src/core/alignment/banded_swa/scalar/
// implementation.rs:5-440), the aligner behind CIGAR regeneration
// (src/pipelines/linear/region.rs:1058).
//
// DP kernel: the row-parallel schedule of the 32-bit extension kernel (one
// lane group of L lanes per pair, S contiguous columns per lane, F as an
// exclusive max-plus prefix with one DPP scan, DPP all-reduces for the row
// maximum and the next band), with this aligner's recurrences:
//   * M = max(eh.h + score, 0) (no M = 0 rule for eh.h = 0);
//   * H = max(M, E, F) and a traceback code per cell: MATCH when M >= E and
//     M >= F, else DEL when E >= F, else INS (:147-158);
//   * the band w is used as given (no end_bonus cap).
// Every executed row writes the 2-bit codes of the lane's S columns to the
// pair's traceback block in HBM, 0 (MATCH, the reference's initial value)
// outside the row's band.  F at the band start is the prefix identity
// instead of the reference's f = 0; both lose to M >= 0 and E >= 0, so H and
// the codes agree.
//
// Band-window variant (BAND): the band of row i lies inside [i - w, i + w + 1]
// (:108-113), so when 2w + S + 1 <= L*S the group holds only a window of
// L*S columns, starting at (i / S) * S - w for the rows of block i / S; every
// S rows each lane takes over the registers of the next lane (one DPP per
// array and slot) and the last lane loads the S columns entering the window.
// A column beyond the window has never been inside a band, so it enters with
// its first-row values (:73-81) and E = 0, exactly what the reference's eh[]
// holds there; a column leaving on the left is never read again (the band
// start never decreases).  Traceback codes are stored per window slot.
//
// Traceback kernel: one lane per pair walks the codes from (max_i+1, max_j+1)
// to (0, 0) exactly as the reference (runs of equal codes, row 0 = INS,
// column 0 = DEL, 10,000-step cap, stop without progress), then adds the 5'
// soft clip and the 3' clip-or-extend decision (:355-414).
//
// Aligned bases: each M or D step of the walk emits target base ci-1 and
// decrements ci, each M or I step emits query base cj-1 and decrements cj, so
// a walk from (max_i+1, max_j+1) that ends at (ci, cj) emits target
// max_i .. ci and query max_j .. cj, and after the reversal (:333-337) the
// aligned reference is target[ci .. max_i] and the aligned query
// query[cj .. max_j].  The walk only counts them; the bases kernel writes them
// with one wavefront per pair, coalesced, instead of one byte per step.  (The
// walk always ends at (0, 0): every step makes progress, the edges are DEL and
// INS, and qlen + tlen < 10,000.)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "swtb_kernel.h"

namespace fa {

namespace {

constexpr int kNeg = -(1 << 28);
constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140;
constexpr int kRowShr = 0x110;
constexpr int kRowShl1 = 0x101;  // lane i reads lane i+1 of its 16-lane row
enum { TB_MATCH = 0, TB_DEL = 1, TB_INS = 2 };
enum { OP_M = 0, OP_I = 1, OP_D = 2, OP_S = 4 };

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

template <int CTRL>
__device__ __forceinline__ int dpp(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xF, 0xF, false);
}
template <int L>
__device__ __forceinline__ int group_allmax(int x) {
    x = imax(x, dpp<kQuadXor1>(x, x));
    if constexpr (L >= 4) x = imax(x, dpp<kQuadXor2>(x, x));
    if constexpr (L >= 8) x = imax(x, dpp<kRowHalfMirror>(x, x));
    if constexpr (L >= 16) x = imax(x, dpp<kRowMirror>(x, x));
    return x;
}
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {  // one v_pk_max_u16
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_max(__builtin_bit_cast(u2, a),
                                                                  __builtin_bit_cast(u2, b)));
}
template <int L>
__device__ __forceinline__ uint32_t group_allmax_pk(uint32_t x) {
    x = pk_max_u16(x, (uint32_t)dpp<kQuadXor1>((int)x, (int)x));
    if constexpr (L >= 4) x = pk_max_u16(x, (uint32_t)dpp<kQuadXor2>((int)x, (int)x));
    if constexpr (L >= 8) x = pk_max_u16(x, (uint32_t)dpp<kRowHalfMirror>((int)x, (int)x));
    if constexpr (L >= 16) x = pk_max_u16(x, (uint32_t)dpp<kRowMirror>((int)x, (int)x));
    return x;
}
template <int L, int K>
__device__ __forceinline__ int group_shr(int x, int l, int ident) {
    int y = dpp<kRowShr + K>(ident, x);
    if constexpr (L < 16) y = (l >= K) ? y : ident;
    return y;
}
template <int L>
__device__ __forceinline__ int group_excl_scan_max(int x, int l) {
    int y = group_shr<L, 1>(x, l, kNeg);
    if constexpr (L > 1) y = imax(y, group_shr<L, 1>(y, l, kNeg));
    if constexpr (L > 2) y = imax(y, group_shr<L, 2>(y, l, kNeg));
    if constexpr (L > 4) y = imax(y, group_shr<L, 4>(y, l, kNeg));
    if constexpr (L > 8) y = imax(y, group_shr<L, 8>(y, l, kNeg));
    return y;
}
__device__ __forceinline__ int seq_code(const uint32_t *seq2, const uint32_t *seqn, uint32_t off,
                                        int pos) {
    const int code = (int)((seq2[off + (pos >> 4)] >> ((pos & 15) * 2)) & 3u);
    return ((seqn[(off >> 1) + (pos >> 5)] >> (pos & 31)) & 1u) ? 4 : code;
}

}  // namespace

// The 16-slot band windows at 3 waves per SIMD (168 VGPRs, a 12-byte spill)
// rather than 2 (204 VGPRs): measured 0.2 ms faster on 1M region-like pairs.
template <int L, int S, bool BAND>
__global__ __launch_bounds__(64, (BAND && S == 16) ? 3 : 1) void swtb_dp_kernel(SwtbArgs a, uint32_t first,
                                                                                uint32_t count) {
    extern __shared__ uint8_t lds[];
    constexpr int G = 64 / L;
    constexpr int NWL = (S + 15) / 16;
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / L;
    const int l = lane % L;
    const uint32_t slot = blockIdx.x * (uint32_t)G + (uint32_t)g;
    const bool valid = slot < count;
    const uint32_t pid = valid ? a.order[first + slot] : 0u;
    const int qlen = valid ? a.qlen[pid] : 1;
    const int tlen = valid ? a.tlen[pid] : 0;
    const int h0 = valid ? a.h0[pid] : 0;
    const int w = valid ? a.w[pid] : 0;
    const uint32_t qo = valid ? a.qo[pid] : 0u, to = valid ? a.to[pid] : 0u;
    uint32_t *tbp = a.tb + (valid ? a.tbo[pid] : 0ull) + (uint32_t)(l * NWL);
    const int o_del = a.o_del, e_del = a.e_del, o_ins = a.o_ins, e_ins = a.e_ins;
    const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;

    uint8_t *tcodes = lds + g * (a.tpad + (BAND ? a.qpad : 0));
    uint8_t *qcodes = tcodes + a.tpad;  // BAND: the query's codes
    for (int k = l; k < tlen; k += L) tcodes[k] = (uint8_t)seq_code(a.seq2, a.seqn, to, k);
    if constexpr (BAND) {
        for (int k = l; k < qlen; k += L) qcodes[k] = (uint8_t)seq_code(a.seq2, a.seqn, qo, k);
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    }

    // column of this lane's slot 0 (BAND: the window of rows [0, S) starts at -w)
    int cb = (BAND ? -w : 0) + l * S;
    const int v1 = h0 > oe_ins ? h0 - oe_ins : 0;
    auto col_prof = [&](int c) -> uint32_t {
        int code = 4;
        if (c >= 0 && c < qlen) code = BAND ? (int)qcodes[c] : seq_code(a.seq2, a.seqn, qo, c);
        return a.profw[code];
    };
    auto col_init = [&](int c) -> int {  // first-row eh.h (:73-81); 0 outside [0, qlen]
        if (c < 0 || c > qlen) return 0;
        return c == 0 ? h0 : imax(v1 - (c - 1) * e_ins, 0);
    };
    uint32_t prof[S];
    int Hh[S], Ee[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        prof[s] = col_prof(cb + s);
        Hh[s] = col_init(cb + s);
        Ee[s] = 0;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();

    int beg = 0, end = qlen;
    int best = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
    uint32_t ncell = 0;
    bool alive = valid && tlen > 0;
    for (int i = 0;; ++i) {
        if (__builtin_amdgcn_ballot_w64(alive) == 0) break;
        if constexpr (BAND) {
            if (i > 0 && i % S == 0) {  // the window moves S columns right
                cb += S;
                const bool last = l == L - 1;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const uint32_t p2 = (uint32_t)dpp<kRowShl1>(0, (int)prof[s]);
                    const int h2 = dpp<kRowShl1>(0, Hh[s]);
                    const int e2 = dpp<kRowShl1>(0, Ee[s]);
                    prof[s] = last ? col_prof(cb + s) : p2;
                    Hh[s] = last ? col_init(cb + s) : h2;
                    Ee[s] = last ? 0 : e2;
                }
            }
        }
        const int tbase = alive ? tcodes[i] : 0;
        int rb = imax(beg, i - w);
        int re = imin(imin(end, i + w + 1), qlen);
        if (!alive) {
            rb = 1 << 20;
            re = -4;
        }
        const int h1i = rb == 0 ? imax(h0 - (o_del + e_del * (i + 1)), 0) : 0;
        const uint32_t sel = 0x0C0C0C00u | (uint32_t)tbase;

        // pass 1: M and the lane-local exclusive prefix of G(c) = max(M - oe_ins,
        // 0) + c * e_ins, kept relative to the lane's cb * e_ins (G - cb * e_ins
        // = max(M + s * e_ins - oe_ins, s * e_ins): wave-uniform constants per
        // slot); band tests against the lane-relative bound rb - cb
        const int cbe = cb * e_ins, trb = rb - cb;
        int M[S], pre[S];
        int run = kNeg;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int sb = (int)__builtin_amdgcn_perm(a.nrow, prof[s], sel);
            M[s] = imax(Hh[s] + sb - 128, 0);  // :144-146
            const int gk = imax(M[s] + (s * e_ins - oe_ins), s * e_ins);
            pre[s] = run;
            run = imax(run, s >= trb ? gk : kNeg);
        }
        // the scan on absolute values; kNeg stays far below any G either way
        const int carry = group_excl_scan_max<L>(run + cbe, l) - cbe;

        // pass 2: H, the code, E, the row maximum; then, with H of the column
        // before (the previous slot's, or for slot 0 the previous lane's last
        // one, after the loop), the write-back eh[c].h <- H(i, c-1) and the
        // non-zero columns of the next band, so each slot's band tests serve
        // all three
        int kmax = 0, hend = 0, Hprev = 0;
        uint32_t fl = 0;
        uint32_t codes[NWL];
#pragma unroll
        for (int k = 0; k < NWL; ++k) codes[k] = 0u;
        // band tests of column c = cb + s as tests of s against rb - cb and
        // re - cb; column-valued fields from per-lane constants
        const int tre = re - cb, k1023 = 1023 - cb, k1 = cb + 1;
        auto write_back = [&](int s, int hp) {
            const bool ge_rb = s >= trb, le_re = s <= tre;
            const int hv = s <= trb ? h1i : hp;
            Hh[s] = (ge_rb && le_re) ? hv : Hh[s];
            hend = s == tre ? Hh[s] : hend;
            const bool nz = (Hh[s] | Ee[s]) != 0;
            const uint32_t f = (nz && ge_rb && s < tre) ? (uint32_t)(k1023 - s) : 0u;  // 1023 - c
            const uint32_t la = (nz && ge_rb && le_re) ? (uint32_t)(k1 + s) : 0u;      // c + 1
            fl = pk_max_u16(fl, f | (la << 16));
        };
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int F = imax(carry, pre[s]) + (e_ins - s * e_ins);  // - (c - 1) * e_ins, relative
            const int m = M[s], e = Ee[s];
            const int h = imax(imax(m, e), F);
            const uint32_t code = (m >= e && m >= F) ? TB_MATCH : (e >= F ? TB_DEL : TB_INS);
            const int en = imax(imax(e - e_del, m - oe_del), 0);
            const bool inb = s >= trb && s < tre;
            // row-argmax key (h << 10) | (1023 - c), here with 1023 - s (cb
            // subtracted after the slots: every key of the lane shares it)
            kmax = imax(kmax, inb ? ((h << 10) | (1023 - s)) : 0);
            Ee[s] = inb ? en : (s == tre ? 0 : e);
            codes[s / 16] |= (inb ? code : 0u) << (2 * (s % 16));
            if (s > 0) write_back(s, Hprev);
            Hprev = h;
        }
        kmax = kmax ? kmax - cb : 0;  // in-band columns lie in [0, qlen], so 1023 - c >= 0
        write_back(0, dpp<kRowShr + 1>(0, Hprev));
        if (alive) {
#pragma unroll
            for (int k = 0; k < NWL; ++k) tbp[(size_t)i * (L * NWL) + k] = codes[k];
        }
        kmax = group_allmax<L>(kmax);
        hend = group_allmax<L>(hend);
        fl = group_allmax_pk<L>(fl);

        if (alive) {
            ncell += (uint32_t)imax(re - rb, 0);
            if (re == qlen && gscore < hend) {  // :191-194
                gscore = hend;
                max_ie = i;
            }
            const int m_row = kmax >> 10;
            const int mj = 1023 - (kmax & 1023);
            bool brk = m_row == 0;  // :196-198
            if (!brk) {
                if (m_row > best) {
                    best = m_row;
                    max_i = i;
                    max_j = mj;
                    max_off = imax(max_off, mj > i ? mj - i : i - mj);
                } else if (a.zdrop > 0) {
                    const int di = i - max_i, dj = mj - max_j;
                    if (di > dj) {
                        if (best - m_row - (di - dj) * e_del > a.zdrop) brk = true;
                    } else if (best - m_row - (dj - di) * e_ins > a.zdrop) {
                        brk = true;
                    }
                }
            }
            if (!brk) {  // :215-224
                const int fenc = (int)(fl & 0xFFFFu);
                const int nb = fenc ? 1023 - fenc : re;
                const int ne = imax(nb, (int)(fl >> 16) - 1);
                beg = nb;
                end = imin(ne + 2, qlen);
            }
            if (brk || i + 1 >= tlen) alive = false;
        }
    }
    if (valid && l == 0) {
        int32_t *o = a.out + (size_t)pid * 6;
        o[0] = best;
        o[1] = max_i + 1;
        o[2] = max_ie + 1;
        o[3] = max_j + 1;
        o[4] = gscore;
        o[5] = max_off;
        if (a.cells) a.cells[pid] = ncell;
    }
}

namespace {

// Codes of one pair, read through a copy of the 128-byte line that holds the
// current word: the walk steps back about one row at a time and a row of a
// band-window pair is 16 or 32 bytes, so one line serves several steps.  The
// copy sits in LDS, word k of the line at line[k * kTbLanes] (a column per
// lane: the lanes' reads of the same k fall in different banks).
constexpr int kTbLanes = 256;  // threads per block of the traceback kernel
struct TbView {
    const uint32_t *tb0;  // the traceback buffer (256-byte aligned, lines whole)
    uint64_t base;        // word offset of the pair's block in it
    uint32_t *line;       // this lane's column of the LDS copy
    int L, S, nwl, band_w;
    bool band;
    uint32_t smul;             // (n * smul) >> 20 == n / S for 0 <= n < 4096 (S <= 32)
    uint64_t held = ~0ull;     // line in the copy
    uint64_t last = ~0ull;     // word last read (runs of I reuse it)
    uint32_t word = 0;
    __device__ int code(int ci, int cj) {
        if (ci == 0) return cj > 0 ? TB_INS : TB_MATCH;
        if (cj == 0) return TB_DEL;
        const int i = ci - 1;
        int j = cj - 1;
        // row i < 2048 and column j < L * S <= 512: divisions by S as products
        if (band) {  // slot in the window of row i; outside it: never in a band
            j -= (int)(((uint32_t)i * smul) >> 20) * S - band_w;
            if (j < 0 || j >= L * S) return TB_MATCH;
        }
        const int lane = (int)(((uint32_t)j * smul) >> 20), s = j - lane * S;
        const uint64_t w = base + (uint64_t)i * (uint64_t)(L * nwl) + (uint64_t)(lane * nwl + s / 16);
        if (w != last) {
            if ((w >> 5) != held) {
                held = w >> 5;
                const uint4 *src = reinterpret_cast<const uint4 *>(tb0) + held * 8;
                uint4 v[8];
#pragma unroll
                for (int k = 0; k < 8; ++k) v[k] = src[k];
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    line[(4 * k + 0) * kTbLanes] = v[k].x;
                    line[(4 * k + 1) * kTbLanes] = v[k].y;
                    line[(4 * k + 2) * kTbLanes] = v[k].z;
                    line[(4 * k + 3) * kTbLanes] = v[k].w;
                }
            }
            word = line[(w & 31u) * kTbLanes];
            last = w;
        }
        return (int)((word >> (2 * (s % 16))) & 3u);
    }
};

}  // namespace

__global__ __launch_bounds__(kTbLanes) void swtb_traceback_kernel(SwtbArgs a, uint32_t n) {
    __shared__ uint32_t lines[32 * kTbLanes];
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int qlen = a.qlen[p], tlen = a.tlen[p];
    uint32_t *cig = a.cigar + a.cig_off[p];
    if (qlen <= 0 || tlen <= 0) {  // :17-30
        a.n_cigar[p] = 0;
        a.n_ref[p] = 0;
        a.n_qry[p] = 0;
        return;
    }
    const int32_t *o = a.out + (size_t)p * 6;
    const int score = o[0], max_i = o[1] - 1, max_j = o[3] - 1, gscore = o[4];
    const uint32_t lsl = a.lsl[p];
    TbView v{a.tb, a.tbo[p], lines + threadIdx.x, (int)((lsl >> 8) & 0xFFu), (int)(lsl & 0xFFu),
             (int)((lsl & 0xFFu) + 15) / 16, a.w[p], (lsl & kSwtbBand) != 0,
             (1u << 20) / (lsl & 0xFFu) + 1u};
    // runs are produced end to start: keep them from index 1 (cig[0] is kept
    // for the 5' clip), reverse afterwards
    // The reference's walk (:250-326): an outer iteration reads the code at
    // (ci, cj) and follows the run of that code (M: while ci, cj > 0; D: while
    // ci > 0; I: while cj > 0), emitting it when not empty; it stops at (0, 0),
    // at 10,000 iterations, on a code other than M, D, I, or when a run made no
    // progress.  Here one loop step is one cell, so the lanes of a wavefront,
    // whose runs differ, run the same instructions instead of each run kind's
    // loop in turn.
    int nrc = 0;
    int ci = max_i + 1, cj = max_j + 1, iter = 0;
    int op = -1, cnt = 0, pi = 0, pj = 0;  // op < 0: no run open
    for (;;) {
        if (op < 0) {
            if (ci == 0 && cj == 0) break;
            if (iter >= 10000) break;
            ++iter;
            pi = ci;
            pj = cj;
            op = v.code(ci, cj);
            if (op > TB_INS) break;
            cnt = 0;
        }
        const int c = v.code(ci, cj);  // at a run's start the word just read
        const bool ext = c == op && (op == TB_MATCH ? (ci > 0 && cj > 0) : (op == TB_DEL ? ci > 0 : cj > 0));
        if (ext) {
            ++cnt;
            ci -= op != TB_INS;
            cj -= op != TB_DEL;
            continue;
        }
        if (cnt > 0) cig[1 + nrc++] = ((uint32_t)cnt << 4) | (op == TB_MATCH ? OP_M : (op == TB_DEL ? OP_D : OP_I));
        op = -1;
        if (ci == pi && cj == pj) break;
    }
    for (int x = 0, y = nrc - 1; x < y; ++x, --y) {
        const uint32_t t = cig[1 + x];
        cig[1 + x] = cig[1 + y];
        cig[1 + y] = t;
    }
    int nc;
    const int query_start = cj, query_end = max_j + 1;
    if (query_start > 0) {  // 5' soft clip
        cig[0] = ((uint32_t)query_start << 4) | OP_S;
        nc = 1 + nrc;
    } else {
        for (int x = 0; x < nrc; ++x) cig[x] = cig[x + 1];
        nc = nrc;
    }
    if (query_end < qlen) {  // 3' clip or extend (:367-414)
        const bool clip3 = gscore <= 0 || gscore <= score - a.pen_clip3;
        cig[nc++] = ((uint32_t)(qlen - query_end) << 4) | (clip3 ? OP_S : OP_M);
    }
    a.n_cigar[p] = nc;
    a.n_ref[p] = max_i + 1 - ci;  // target [ci, max_i]
    a.n_qry[p] = max_j + 1 - cj;  // query [cj, max_j]
}

// Aligned bases (see the traceback kernel, which ran before on the same
// stream): one wavefront per pair writes target[ci .. max_i] and
// query[cj .. max_j] as codes 0..4.
__global__ __launch_bounds__(256) void swtb_bases_kernel(SwtbArgs a, uint32_t n) {
    const uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (p >= n) return;
    const int qlen = a.qlen[p], tlen = a.tlen[p];
    if (qlen <= 0 || tlen <= 0) return;
    const int32_t *o = a.out + (size_t)p * 6;
    const int nr = a.n_ref[p], nq = a.n_qry[p];
    const int r0 = o[1] - nr, q0 = o[3] - nq;  // ci, cj where the walk ended
    uint8_t *ra = a.ref_aln + a.aln_off[p];
    uint8_t *qa = a.qry_aln + a.aln_off[p];
    const uint32_t to = a.to[p], qo = a.qo[p];
    for (int k = lane; k < nr; k += 64) ra[k] = (uint8_t)seq_code(a.seq2, a.seqn, to, r0 + k);
    for (int k = lane; k < nq; k += 64) qa[k] = (uint8_t)seq_code(a.seq2, a.seqn, qo, q0 + k);
}

// One wavefront per pair copies its CIGAR and aligned bases to the dense
// buffers.
__global__ __launch_bounds__(256) void swtb_compact_kernel(SwtbArgs a, uint32_t n,
                                                           const uint64_t *dense, uint32_t *cigar,
                                                           uint8_t *ref_aln, uint8_t *qry_aln) {
    const uint32_t p = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    if (p >= n) return;
    const int nc = a.n_cigar[p], nr = a.n_ref[p], nq = a.n_qry[p];
    const uint32_t *cs = a.cigar + a.cig_off[p];
    const uint8_t *rs = a.ref_aln + a.aln_off[p], *qs = a.qry_aln + a.aln_off[p];
    uint32_t *cd = cigar + dense[3 * (size_t)p];
    uint8_t *rd = ref_aln + dense[3 * (size_t)p + 1], *qd = qry_aln + dense[3 * (size_t)p + 2];
    for (int k = lane; k < nc; k += 64) cd[k] = cs[k];
    for (int k = lane; k < nr; k += 64) rd[k] = rs[k];
    for (int k = lane; k < nq; k += 64) qd[k] = qs[k];
}

template <int L, int S, bool BAND>
static hipError_t launch_dp(const SwtbArgs &a, uint32_t first, uint32_t count, hipStream_t st) {
    constexpr int G = 64 / L;
    if (count == 0) return hipSuccess;
    const size_t lds = (size_t)G * (size_t)(a.tpad + (BAND ? a.qpad : 0));
    hipLaunchKernelGGL((swtb_dp_kernel<L, S, BAND>), dim3((count + G - 1) / G), dim3(64), lds, st, a,
                       first, count);
    return hipGetLastError();
}

int swtb_slot_count(int L, int qlen) {
    static const int sl[] = {1, 2, 4, 8, 12, 16, 24, 32};
    if (L != 4 && L != 8 && L != 16) return -1;
    for (int s : sl)
        if (L * s >= qlen + 1) return s;
    return -1;
}

bool swtb_band_config(int qlen, int w, int full_cols, int *L, int *S) {
    static const int cfg[][2] = {{4, 8}, {4, 16}, {8, 16}};
    if (w < 0 || w > 1024) return false;
    for (const auto &c : cfg)
        if (2 * w + c[1] + 1 <= c[0] * c[1] && c[0] * c[1] < full_cols && qlen > 0) {
            *L = c[0];
            *S = c[1];
            return true;
        }
    return false;
}

hipError_t swtb_dp_launch(int L, int S, bool band, const SwtbArgs &a, uint32_t first,
                          uint32_t count, hipStream_t st) {
    if (band) {
        if (L == 4 && S == 8) return launch_dp<4, 8, true>(a, first, count, st);
        if (L == 4 && S == 16) return launch_dp<4, 16, true>(a, first, count, st);
        if (L == 8 && S == 16) return launch_dp<8, 16, true>(a, first, count, st);
        return hipErrorInvalidValue;
    }
#define FA_SW_CASE(LL, SS) \
    if (L == LL && S == SS) return launch_dp<LL, SS, false>(a, first, count, st);
#define FA_SW_LIST(LL)                                                                  \
    FA_SW_CASE(LL, 1) FA_SW_CASE(LL, 2) FA_SW_CASE(LL, 4) FA_SW_CASE(LL, 8) FA_SW_CASE(LL, 12) \
    FA_SW_CASE(LL, 16) FA_SW_CASE(LL, 24) FA_SW_CASE(LL, 32)
    FA_SW_LIST(4)
    FA_SW_LIST(8)
    FA_SW_LIST(16)
#undef FA_SW_LIST
#undef FA_SW_CASE
    return hipErrorInvalidValue;
}

hipError_t swtb_compact_launch(const SwtbArgs &a, uint32_t n, const uint64_t *dense,
                               uint32_t *cigar, uint8_t *ref_aln, uint8_t *qry_aln,
                               hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(swtb_compact_kernel, dim3((n + 3) / 4), dim3(256), 0, st, a, n, dense, cigar,
                       ref_aln, qry_aln);
    return hipGetLastError();
}

hipError_t swtb_traceback_launch(const SwtbArgs &a, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(swtb_traceback_kernel, dim3((n + kTbLanes - 1) / kTbLanes), dim3(kTbLanes), 0, st, a, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(swtb_bases_kernel, dim3((n + 3) / 4), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

}  // namespace fa
