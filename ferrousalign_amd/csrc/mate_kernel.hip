// Mate-rescue local Smith-Waterman for gfx950, exact to the reference's
// ksw_align2 in 16-bit mode (src/core/alignment/ksw_affine_gap.rs:503-735,
// :773-861) as executed by its 8-lane SSE engine (SimdEngine128).
//
// The reference's result depends on its stripe geometry (query position k in
// SSE lane k / slen, segment k % slen, slen = ceil(qlen / 8)), so the kernel
// keeps that geometry: a group of 8 GPU lanes per pair, GPU lane p playing
// SSE lane p.  Lane p holds its slen segments of H, E and Hmax in VGPRs and
// walks them in order for every target row, exactly as the SSE loop walks
// its vectors; the cross-lane steps of the SSE code (the one-lane shift of
// H(i-1, -1) and of F in the lazy-F passes, the horizontal maxima) are DPP
// moves inside the 8-lane group.  Arithmetic is signed 16-bit saturating
// (each add/sub clamped to [-32768, 32767]), H is not clamped at 0, the row
// maximum is taken before the lazy-F pass and Hmax after it, as in the
// reference.  Each pair's per-row maxima go to a log in HBM; after the
// forward pass one lane replays the reference's suboptimal-hit list from
// that log to get score2/te2.  The XSTART reverse pass reads the reversed
// query and target prefixes through index arithmetic.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mate_kernel.h"

namespace fa {

namespace {

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int sat16(int x) { return imin(imax(x, -32768), 32767); }

constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowShr1 = 0x111;

template <int CTRL>
__device__ __forceinline__ int dpp(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xF, 0xF, false);
}
// max over the 8 lanes of a group, every lane gets it
__device__ __forceinline__ int grp8_max(int x) {
    x = imax(x, dpp<kQuadXor1>(x, x));
    x = imax(x, dpp<kQuadXor2>(x, x));
    x = imax(x, dpp<kRowHalfMirror>(x, x));
    return x;
}
// value of lane p-1 inside the group (lane 0 gets 0): the SSE byte shift by one lane
__device__ __forceinline__ int grp8_shift_up(int x, int p) {
    int y = dpp<kRowShr1>(0, x);
    return p == 0 ? 0 : y;
}

__device__ __forceinline__ int seq_code(const uint32_t *seq2, const uint32_t *seqn, uint32_t off,
                                        int pos) {
    const uint32_t c = (seq2[off + (pos >> 4)] >> ((pos & 15) * 2)) & 3u;
    return ((seqn[(off >> 1) + (pos >> 5)] >> (pos & 31)) & 1u) ? 4 : (int)c;
}

struct PassResult {
    int score, te, qe, rows;
};

// One ksw_i16 pass over `tlen` rows for the group's pair.
//   qlen  : query length of this pass (profile positions k < qlen are real)
//   qrev  : reverse pass: query position k is original position qlen-1-k,
//           target row i is original i <= trev ? trev - i : i
//   endsc : XSTOP threshold (0x10000 = none); log: per-row maxima or null
// Returns score, te, qe (qe of the pass's Hmax, or -1 if score == 0x7FFF).
template <int SLEN>
__device__ __forceinline__ PassResult ksw_pass(const MateArgs &a, const uint8_t *tcodes, int p,
                                               bool active, uint32_t qo, int qlen, bool qrev,
                                               int tlen, int trev, uint32_t endsc,
                                               uint16_t *log) {
    const int slen = (qlen + 7) >> 3;
    uint32_t prof[SLEN];
    int H[SLEN], E[SLEN], Hm[SLEN];
#pragma unroll
    for (int j = 0; j < SLEN; ++j) {
        const int k = j + p * slen;
        uint32_t pw = 0x80808080u;  // position past the query: score 0 for every target base
        if (active && j < slen && k < qlen) {
            const int orig = qrev ? qlen - 1 - k : k;
            pw = a.profw[seq_code(a.seq2, a.seqn, qo, orig)];
        }
        prof[j] = pw;
        H[j] = 0;
        E[j] = 0;
        Hm[j] = 0;
    }
    const int e_del = a.e_del, e_ins = a.e_ins;
    const int oe_del = a.o_del + a.e_del, oe_ins = a.o_ins + a.e_ins;
    int gmax = 0, te = -1, rows = 0;
    bool alive = active && tlen > 0;
    // H(i-1, slen-1) of lane p-1, carried from the previous row
    int hlast = 0;
    for (int i = 0;; ++i) {
        if (__builtin_amdgcn_ballot_w64(alive) == 0) break;
        int tb = 0;
        if (alive) {
            const int ti = (qrev && i <= trev) ? trev - i : i;
            tb = tcodes[ti];
        }
        const uint32_t sel = 0x0C0C0C00u | (uint32_t)tb;
        int h = grp8_shift_up(hlast, p);
        int f = 0, mx = 0;
#pragma unroll
        for (int j = 0; j < SLEN; ++j) {
            if (j < slen) {
                const int sb = (int)__builtin_amdgcn_perm(0x80u, prof[j], sel);
                int hv = sat16(h + sb - 128);
                const int e = E[j];
                hv = imax(imax(hv, e), f);
                mx = imax(mx, hv);
                h = H[j];  // H(i-1, j): next segment's diagonal
                H[j] = hv;
                E[j] = imax(sat16(e - e_del), sat16(hv - oe_del));
                f = imax(sat16(f - e_ins), sat16(hv - oe_ins));
            }
        }
        // lazy F: shift F one lane up, raise H, decay F; stop after a pass without change
        bool lazy = alive;
        for (int pass = 0; pass < 8; ++pass) {
            if (__builtin_amdgcn_ballot_w64(lazy) == 0) break;
            f = grp8_shift_up(f, p);
            int changed = 0;
#pragma unroll
            for (int j = 0; j < SLEN; ++j) {
                if (j < slen) {
                    const int hn = imax(H[j], f);
                    changed |= hn != H[j];
                    if (lazy) H[j] = hn;
                    f = sat16(f - e_ins);
                }
            }
            if (lazy && grp8_max(changed) == 0) lazy = false;
        }
        // carry H(i, slen-1) for the next row's shift
#pragma unroll
        for (int j = 0; j < SLEN; ++j)
            if (j == slen - 1) hlast = H[j];
        const int rmax = grp8_max(mx);  // lanes are >= 0: unsigned max == signed max
        if (alive) {
            ++rows;
            if (log && p == 0) log[i] = (uint16_t)rmax;
            if (rmax > gmax) {
                gmax = rmax;
                te = i;
#pragma unroll
                for (int j = 0; j < SLEN; ++j) Hm[j] = H[j];
                if ((uint32_t)gmax >= 0x7FFFu || (uint32_t)gmax >= endsc) alive = false;
            }
            if (i + 1 >= tlen) alive = false;
        }
    }
    PassResult r{gmax, te, -1, rows};
    if (gmax != 0x7FFF) {
        // qe: largest Hmax over real positions, smallest position on ties
        int key = 0;
#pragma unroll
        for (int j = 0; j < SLEN; ++j) {
            const int k = j + p * slen;
            if (j < slen && k < qlen) key = imax(key, ((Hm[j] + 32769) << 12) | (4095 - k));
        }
        key = grp8_max(key);
        r.qe = key ? 4095 - (key & 4095) : -1;
    }
    return r;
}

}  // namespace

template <int SLEN>
__global__ __launch_bounds__(256) void mate_kernel(MateArgs a, uint32_t first, uint32_t count) {
    extern __shared__ uint8_t lds[];
    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)(threadIdx.x >> 6);
    const int g = lane >> 3, p = lane & 7;
    const uint32_t slot = (blockIdx.x * 4u + (uint32_t)wave) * 8u + (uint32_t)g;
    const bool valid = slot < count;
    const uint32_t pid = valid ? a.order[first + slot] : 0u;
    const int qlen = valid ? a.qlen[pid] : 1;
    const int tlen = valid ? a.tlen[pid] : 0;
    const uint32_t qo = valid ? a.qo[pid] : 0u, to = valid ? a.to[pid] : 0u;

    uint8_t *tcodes = lds + (wave * 8 + g) * a.tpad;
    {
        const int nw = (tlen + 15) >> 4;
        for (int k = p; k < nw; k += 8) {
            const uint32_t w2 = a.seq2[to + k];
            const uint32_t wn = (a.seqn[(to >> 1) + (k >> 1)] >> ((k & 1) * 16)) & 0xFFFFu;
            uint32_t v[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                uint32_t x = 0;
#pragma unroll
                for (int y = 0; y < 4; ++y) {
                    const int q = b * 4 + y;
                    uint32_t c = (w2 >> (q * 2)) & 3u;
                    c = ((wn >> q) & 1u) ? 4u : c;
                    x |= c << (y * 8);
                }
                v[b] = x;
            }
            uint32_t *dst = reinterpret_cast<uint32_t *>(tcodes + k * 16);
            dst[0] = v[0];
            dst[1] = v[1];
            dst[2] = v[2];
            dst[3] = v[3];
        }
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();

    uint16_t *log = valid ? a.imax_log + (size_t)slot * (size_t)a.log_stride : nullptr;
    const uint32_t xtra = a.xtra;
    const uint32_t minsc = (xtra & kKswXSubo) ? (xtra & 0xffffu) : 0x10000u;
    const uint32_t endsc = (xtra & kKswXStop) ? (xtra & 0xffffu) : 0x10000u;

    PassResult r = ksw_pass<SLEN>(a, tcodes, p, valid, qo, qlen, false, tlen, 0, endsc, log);
    int score2 = -1, te2 = -1, tb = -1, qb = -1;
    uint32_t cells = (uint32_t)r.rows * (uint32_t)qlen;

    // suboptimal hits (:639-655, :717-731), replayed from the row log
    if (valid && p == 0 && r.score != 0x7FFF && minsc <= 0xFFFFu) {
        const int iv = (r.score + a.qmax - 1) / a.qmax;
        const int low = r.te - iv, high = r.te + iv;
        int cs = -1, ce = -1;  // pending (last) entry of the list
        for (int i = 0; i < r.rows; ++i) {
            const int im = log[i];
            if ((uint32_t)im < minsc) continue;
            if (cs < 0 || ce + 1 != i) {
                if (cs >= 0 && (ce < low || ce > high) && cs > score2) {
                    score2 = cs;
                    te2 = ce;
                }
                cs = im;
                ce = i;
            } else if (cs < im) {
                cs = im;
                ce = i;
            }
        }
        if (cs >= 0 && (ce < low || ce > high) && cs > score2) {
            score2 = cs;
            te2 = ce;
        }
    }

    // XSTART reverse pass (:804-855)
    const bool find_start =
        (xtra & kKswXStart) && !((xtra & kKswXSubo) && r.score < (int)(xtra & 0xffffu));
    const bool rev = valid && find_start && r.qe >= 0 && r.te >= 0;
    const int qlen2 = rev ? r.qe + 1 : 1;
    // xtra2 = KSW_XSTOP | score: the pass stops once its maximum reaches `score`
    PassResult rr = ksw_pass<SLEN>(a, tcodes, p, rev, qo, qlen2, true, tlen, r.te,
                                   (uint32_t)r.score & 0xffffu, nullptr);
    if (rev) {
        cells += (uint32_t)rr.rows * (uint32_t)qlen2;
        if (rr.score >= r.score) {
            tb = r.te - rr.te;
            qb = r.qe - rr.qe;
        }
    }
    if (valid && p == 0) {
        int32_t *o = a.out + (size_t)pid * 7;
        o[0] = r.score;
        o[1] = r.te;
        o[2] = r.qe;
        o[3] = score2;
        o[4] = te2;
        o[5] = tb;
        o[6] = qb;
        if (a.cells) a.cells[pid] = cells;
    }
}

template <int SLEN>
static hipError_t launch(const MateArgs &a, uint32_t first, uint32_t count, hipStream_t st) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + 31) / 32;
    const size_t lds = (size_t)32 * a.tpad;
    hipLaunchKernelGGL((mate_kernel<SLEN>), dim3(blocks), dim3(256), lds, st, a, first, count);
    return hipGetLastError();
}

int mate_slen_bucket(int qlen) {
    const int slen = (qlen + 7) / 8;
    static const int b[] = {4, 8, 12, 16, 19, 24, 32};
    for (int s : b)
        if (slen <= s) return s;
    return -1;
}

hipError_t mate_launch(int SLEN, const MateArgs &a, uint32_t first, uint32_t count,
                       hipStream_t st) {
    switch (SLEN) {
        case 4: return launch<4>(a, first, count, st);
        case 8: return launch<8>(a, first, count, st);
        case 12: return launch<12>(a, first, count, st);
        case 16: return launch<16>(a, first, count, st);
        case 19: return launch<19>(a, first, count, st);
        case 24: return launch<24>(a, first, count, st);
        case 32: return launch<32>(a, first, count, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace fa
