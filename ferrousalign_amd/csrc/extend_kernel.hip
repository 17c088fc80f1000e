// Banded seed-extension kernels for gfx950 (CDNA4), exact to the reference's
// ksw_extend2 (src/core/alignment/ksw_affine_gap.rs:952-1169).
//
// Schedule: one lane group of L lanes (L = 4, 8 or 16, inside one 16-lane DPP
// row) per extension pair, 64/L pairs per wavefront.  Query columns are
// owned contiguously: lane l of a group holds columns c = l*S + s,
// s = 0..S-1, of the reference's eh[] array (0..qlen), with H, E and the
// per-column score profile in VGPRs for the whole pair.  Rows are processed
// in order (the adaptive band of row i+1 depends on all of row i):
//   * M and E of a row depend only on the previous row, so every lane computes
//     its columns independently;
//   * F is rebuilt as an exclusive max-plus prefix: with
//     G_k = max(M_k - oe_ins, 0) + k*e_ins, F(j) = max_{beg<=k<j} G_k -
//     (j-1)*e_ins for j > beg (F(beg) = 0).  Each lane scans its own slots
//     and a DPP row_shr scan combines the lanes;
//   * the row maximum and its first column come from one DPP all-reduce of
//     the key (H << 10) | (1023 - j);
//   * the next band [beg, end) comes from one packed 16-bit all-reduce of the
//     first and last non-zero column.
// No MFMA: the dynamic program is max-plus, not a dense contraction.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "extend_kernel.h"

namespace fa {

namespace {

constexpr int kNeg = -(1 << 28);  // identity of the F prefix max

__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }
__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }

// DPP control words (GFX9 encoding).
constexpr int kQuadXor1 = 0xB1;       // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;       // quad_perm [2,3,0,1]
constexpr int kRowHalfMirror = 0x141;
constexpr int kRowMirror = 0x140;
constexpr int kRowShr = 0x110;        // row_shr:n = 0x110 + n

template <int CTRL>
__device__ __forceinline__ int dpp(int old, int src) {
    return __builtin_amdgcn_update_dpp(old, src, CTRL, 0xF, 0xF, false);
}

// Max over the L lanes of a group; every lane receives the result.
template <int L>
__device__ __forceinline__ int group_allmax(int x) {
    x = imax(x, dpp<kQuadXor1>(x, x));
    if constexpr (L >= 4) x = imax(x, dpp<kQuadXor2>(x, x));
    if constexpr (L >= 8) x = imax(x, dpp<kRowHalfMirror>(x, x));
    if constexpr (L >= 16) x = imax(x, dpp<kRowMirror>(x, x));
    return x;
}

// Packed two-half unsigned 16-bit max over the L lanes of a group.
__device__ __forceinline__ uint32_t pk_max_u16(uint32_t a, uint32_t b) {
    uint32_t lo = imax((int)(a & 0xFFFFu), (int)(b & 0xFFFFu));
    uint32_t hi = imax((int)(a >> 16), (int)(b >> 16));
    return lo | (hi << 16);
}
template <int L>
__device__ __forceinline__ uint32_t group_allmax_pk(uint32_t x) {
    x = pk_max_u16(x, (uint32_t)dpp<kQuadXor1>((int)x, (int)x));
    if constexpr (L >= 4) x = pk_max_u16(x, (uint32_t)dpp<kQuadXor2>((int)x, (int)x));
    if constexpr (L >= 8) x = pk_max_u16(x, (uint32_t)dpp<kRowHalfMirror>((int)x, (int)x));
    if constexpr (L >= 16) x = pk_max_u16(x, (uint32_t)dpp<kRowMirror>((int)x, (int)x));
    return x;
}

// Value of lane l-1 of the same group (lane 0 of a group receives `ident`).
template <int L, int K>
__device__ __forceinline__ int group_shr(int x, int l, int ident) {
    int y = dpp<kRowShr + K>(ident, x);
    if constexpr (L < 16) y = (l >= K) ? y : ident;
    return y;
}

// Exclusive max-scan over the lanes of a group.
template <int L>
__device__ __forceinline__ int group_excl_scan_max(int x, int l) {
    int y = group_shr<L, 1>(x, l, kNeg);
    if constexpr (L > 1) y = imax(y, group_shr<L, 1>(y, l, kNeg));
    if constexpr (L > 2) y = imax(y, group_shr<L, 2>(y, l, kNeg));
    if constexpr (L > 4) y = imax(y, group_shr<L, 4>(y, l, kNeg));
    if constexpr (L > 8) y = imax(y, group_shr<L, 8>(y, l, kNeg));
    return y;
}

__device__ __forceinline__ int decode_base(const uint32_t *seq2, const uint32_t *seqn,
                                           uint32_t off, int pos) {
    uint32_t w2 = seq2[off + (pos >> 4)];
    uint32_t wn = seqn[(off >> 1) + (pos >> 5)];
    int code = (int)((w2 >> ((pos & 15) * 2)) & 3u);
    return ((wn >> (pos & 31)) & 1u) ? 4 : code;
}

}  // namespace

// One wavefront processes up to 64/L pairs: order[slot0 .. slot0 + 64/L).
template <int L, int S>
__device__ __forceinline__ void extend_wave(const ExtArgs &a, uint32_t slot0, uint32_t nslots,
                                            uint8_t *lds_wave) {
    constexpr int G = 64 / L;
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / L;
    const int l = lane % L;
    const uint32_t slot = slot0 + (uint32_t)g;
    const bool valid = slot < nslots;
    const uint32_t pid = valid ? a.order[slot] : 0u;

    const int qlen = valid ? a.qlen[pid] : 1;
    const int tlen = valid ? a.tlen[pid] : 0;
    const int h0 = valid ? a.h0[pid] : 1;
    const int w = valid ? a.w[pid] : 0;
    const uint32_t qo = valid ? a.qo[pid] : 0u;
    const uint32_t to = valid ? a.to[pid] : 0u;

    const int o_del = a.o_del, e_del = a.e_del, o_ins = a.o_ins, e_ins = a.e_ins;
    const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;

    // Stage this pair's target as one code byte per row in LDS.
    uint8_t *tcodes = lds_wave + g * a.tpad;
    {
        const int nw = (tlen + 15) >> 4;
        for (int k = l; k < nw; k += L) {
            uint32_t w2 = a.seq2[to + k];
            uint32_t wn = (a.seqn[(to >> 1) + (k >> 1)] >> ((k & 1) * 16)) & 0xFFFFu;
            uint32_t packed[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                uint32_t v = 0;
#pragma unroll
                for (int x = 0; x < 4; ++x) {
                    int p = b * 4 + x;
                    uint32_t code = (w2 >> (p * 2)) & 3u;
                    code = ((wn >> p) & 1u) ? 4u : code;
                    v |= code << (x * 8);
                }
                packed[b] = v;
            }
            uint32_t *dst = reinterpret_cast<uint32_t *>(tcodes + k * 16);
            dst[0] = packed[0];
            dst[1] = packed[1];
            dst[2] = packed[2];
            dst[3] = packed[3];
        }
    }

    // Column state.  eh[c].h and eh[c].e of the reference for c = cb + s.
    const int cb = l * S;
    uint32_t prof[S];
    int Hh[S], Ee[S];
    const int v1 = h0 > oe_ins ? h0 - oe_ins : 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int c = cb + s;
        int code = 4;
        if (c < qlen) code = decode_base(a.seq2, a.seqn, qo, c);
        prof[s] = a.profw[code];
        // first row, ksw_affine_gap.rs:992-998 in closed form
        int hv = c == 0 ? h0 : imax(v1 - (c - 1) * e_ins, 0);
        Hh[s] = c <= qlen ? hv : 0;
        Ee[s] = 0;
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_s_waitcnt(0);

    int beg = 0, end = qlen;
    int best = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
    uint32_t ncell = 0;
    bool alive = valid && tlen > 0;
    const uint32_t nrow = a.nrow;

    int tb_next = alive ? tcodes[0] : 0;
    for (int i = 0;; ++i) {
        if (__builtin_amdgcn_ballot_w64(alive) == 0) break;
        const int tb = tb_next;
        tb_next = (alive && i + 1 < tlen) ? tcodes[i + 1] : 0;

        // band for this row, :1038-1046
        int rb = imax(beg, i - w);
        int re = imin(imin(end, i + w + 1), qlen);
        if (!alive) {
            rb = 1 << 20;
            re = -4;
        }
        int h1i = 0;  // :1049-1054
        if (rb == 0) h1i = imax(h0 - (o_del + e_del * (i + 1)), 0);
        const uint32_t sel = 0x0C0C0C00u | (uint32_t)tb;

        // pass 1: M and the lane-local exclusive prefix of G
        int M[S], pre[S];
        int run = kNeg;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = cb + s;
            const int sb = (int)__builtin_amdgcn_perm(nrow, prof[s], sel);
            const int mm = Hh[s] + sb - 128;
            M[s] = Hh[s] != 0 ? mm : 0;
            int gk = imax(M[s] - oe_ins, 0) + c * e_ins;
            gk = c >= rb ? gk : kNeg;
            pre[s] = run;
            run = imax(run, gk);
        }
        const int carry = group_excl_scan_max<L>(run, l);

        // pass 2: H, E, row-max key
        int H[S];
        int kmax = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = cb + s;
            const int F = imax(carry, pre[s]) - (c - 1) * e_ins;
            const int h = imax(imax(M[s], Ee[s]), F);
            H[s] = h;
            const int en = imax(imax(Ee[s] - e_del, M[s] - oe_del), 0);
            const bool inb = c >= rb && c < re;
            const int key = inb ? ((h << 10) | (1023 - c)) : 0;
            kmax = imax(kmax, key);
            Ee[s] = inb ? en : (c == re ? 0 : Ee[s]);
        }
        // eh[c].h <- H(i, c-1) on [beg, end]; eh[beg].h <- h1 of the first column
        const int hlast = dpp<kRowShr + 1>(0, H[S - 1]);
        int hend = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = cb + s;
            const int hp = s > 0 ? H[s - 1] : hlast;
            const int hv = c <= rb ? h1i : hp;
            const bool wr = (c >= rb && c < re) || c == re;
            Hh[s] = wr ? hv : Hh[s];
            hend = c == re ? Hh[s] : hend;
        }
        // next band, :1148-1158: first non-zero in [beg,end), last in [beg,end]
        uint32_t fl = 0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int c = cb + s;
            const bool nz = (Hh[s] | Ee[s]) != 0;
            const uint32_t f = (nz && c >= rb && c < re) ? (uint32_t)(1023 - c) : 0u;
            const uint32_t la = (nz && c >= rb && c <= re) ? (uint32_t)(c + 1) : 0u;
            fl = pk_max_u16(fl, f | (la << 16));
        }
        kmax = group_allmax<L>(kmax);
        hend = group_allmax<L>(hend);
        fl = group_allmax_pk<L>(fl);

        if (alive) {
            ncell += (uint32_t)imax(re - rb, 0);
            if (re == qlen && gscore < hend) {  // :1110-1120
                gscore = hend;
                max_ie = i;
            }
            const int m_row = kmax >> 10;
            const int mj = 1023 - (kmax & 1023);
            bool brk = m_row == 0;  // :1122
            if (!brk) {
                if (m_row > best) {  // :1126-1133
                    best = m_row;
                    max_i = i;
                    max_j = mj;
                    const int off = mj > i ? mj - i : i - mj;
                    max_off = imax(max_off, off);
                } else if (a.zdrop > 0) {  // :1134-1145
                    const int di = i - max_i, dj = mj - max_j;
                    if (di > dj) {
                        if (best - m_row - (di - dj) * e_del > a.zdrop) brk = true;
                    } else if (best - m_row - (dj - di) * e_ins > a.zdrop) {
                        brk = true;
                    }
                }
            }
            if (!brk) {
                const int fenc = (int)(fl & 0xFFFFu);
                const int nb = fenc ? 1023 - fenc : re;
                const int ne = imax(nb, (int)(fl >> 16) - 1);
                beg = nb;
                end = imin(ne + 2, qlen);
            }
            if (brk || i + 1 >= tlen) alive = false;
        }
    }

    if (valid && l == 0)  // score, tle, gtle, qle, global score, max offset
        store_result(a, pid, best, max_i + 1, max_ie + 1, max_j + 1, gscore, max_off, ncell);
}

template <int L, int S>
__global__ __launch_bounds__(256) void extend_plain_kernel(ExtArgs a, uint32_t first,
                                                           uint32_t count) {
    extern __shared__ uint8_t lds[];
    constexpr int G = 64 / L;
    const int wave = (int)(threadIdx.x >> 6);
    const uint32_t slot0 = (blockIdx.x * (blockDim.x >> 6) + (uint32_t)wave) * G;
    uint8_t *lds_wave = lds + wave * G * a.tpad;
    if (slot0 >= count) return;  // whole wave idle; no block barrier follows
    extend_wave<L, S>(a, first + slot0, first + count, lds_wave);
}

// Persistent form: a grid sized to the resident capacity of the GPU; every
// wavefront repeatedly claims `a.chunk` wave tasks (64/L consecutive pairs of
// the packed order each) from the device work counter a.queue[qidx] and
// exits once the counter passes the last task.  The counter only grows, so
// every wave reaches the exit.
template <int L, int S>
__global__ __launch_bounds__(256) void extend_persistent_kernel(ExtArgs a, uint32_t first,
                                                                uint32_t count, uint32_t qidx) {
    extern __shared__ uint8_t lds[];
    constexpr int G = 64 / L;
    const int wave = (int)(threadIdx.x >> 6);
    uint8_t *lds_wave = lds + wave * G * a.tpad;
    const uint32_t ntasks = (count + G - 1) / G;
    const uint32_t chunk = (uint32_t)a.chunk;
    uint32_t *counter = a.queue + qidx;
    for (;;) {
        uint32_t t0 = 0;
        if ((threadIdx.x & 63) == 0) t0 = atomicAdd(counter, chunk);
        t0 = __builtin_amdgcn_readfirstlane(t0);
        if (t0 >= ntasks) break;
        const uint32_t t1 = t0 + chunk < ntasks ? t0 + chunk : ntasks;
        for (uint32_t t = t0; t < t1; ++t) extend_wave<L, S>(a, first + t * G, first + count, lds_wave);
    }
}

// Clears the work counters of a run (a kernel, so that the sequence can be
// captured in a graph without memset nodes).
__global__ void queue_reset_kernel(uint32_t *q, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) q[i] = 0u;
}

// ---- host-side launch table ------------------------------------------------

// LDS a workgroup may take (the staged targets of its wavefronts)
constexpr size_t kLdsPerGroup = 65536;

template <int L, int S>
static hipError_t launch_plain(const ExtArgs &a, uint32_t first, uint32_t count,
                               hipStream_t st) {
    constexpr int G = 64 / L;
    // four wavefronts per workgroup, or one when four wavefronts' staged
    // targets (targets near 2048 bases) would not fit a workgroup's LDS
    const uint32_t waves = (size_t)4 * G * a.tpad <= kLdsPerGroup ? 4u : 1u;
    const uint32_t per_block = waves * G;
    const uint32_t blocks = (count + per_block - 1) / per_block;
    if (blocks == 0) return hipSuccess;
    const size_t lds = (size_t)waves * G * a.tpad;
    if (lds > kLdsPerGroup) return hipErrorInvalidValue;
    hipLaunchKernelGGL((extend_plain_kernel<L, S>), dim3(blocks), dim3(64 * waves), lds, st, a,
                       first, count);
    return hipGetLastError();
}

template <int L, int S>
static hipError_t launch_persistent(const ExtArgs &a, uint32_t first, uint32_t count,
                                    uint32_t qidx, int cus, int blocks_per_cu, hipStream_t st) {
    constexpr int G = 64 / L;
    if (count == 0) return hipSuccess;
    const size_t lds = (size_t)4 * G * a.tpad;
    // a persistent workgroup holds four wavefronts: when their staged targets
    // do not fit its LDS, the plain form runs the same work
    if (lds > kLdsPerGroup) return launch_plain<L, S>(a, first, count, st);
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void *>(&extend_persistent_kernel<L, S>), 256, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < occ) occ = blocks_per_cu;
    uint32_t blocks = (uint32_t)occ * (uint32_t)cus;
    const uint32_t need = (count + 4u * G - 1) / (4u * G);
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL((extend_persistent_kernel<L, S>), dim3(blocks), dim3(256), lds, st, a,
                       first, count, qidx);
    return hipGetLastError();
}

#define FA_SLOT_LIST_16(X) X(16, 1) X(16, 2) X(16, 3) X(16, 4) X(16, 6) X(16, 8) X(16, 10) \
    X(16, 12) X(16, 16)
#define FA_SLOT_LIST_8(X) X(8, 1) X(8, 2) X(8, 3) X(8, 4) X(8, 6) X(8, 8) X(8, 10) X(8, 12) \
    X(8, 16) X(8, 20) X(8, 24) X(8, 28) X(8, 32)
#define FA_SLOT_LIST_4(X) X(4, 2) X(4, 4) X(4, 6) X(4, 8) X(4, 12) X(4, 16) X(4, 20) X(4, 24) \
    X(4, 28) X(4, 32)

int extend_slot_count(int L, int qlen) {
    static const int s16[] = {1, 2, 3, 4, 6, 8, 10, 12, 16};
    static const int s8[] = {1, 2, 3, 4, 6, 8, 10, 12, 16, 20, 24, 28, 32};
    static const int s4[] = {2, 4, 6, 8, 12, 16, 20, 24, 28, 32};
    const int *sl = L == 16 ? s16 : (L == 8 ? s8 : s4);
    const int n = L == 16 ? 9 : (L == 8 ? 13 : 10);
    if (L != 16 && L != 8 && L != 4) return -1;
    for (int k = 0; k < n; ++k)
        if (L * sl[k] >= qlen + 1) return sl[k];
    return -1;
}

hipError_t extend_launch_plain(int L, int S, const ExtArgs &a, uint32_t first, uint32_t count,
                               hipStream_t st) {
#define FA_CASE(LL, SS) \
    if (L == LL && S == SS) return launch_plain<LL, SS>(a, first, count, st);
    FA_SLOT_LIST_16(FA_CASE)
    FA_SLOT_LIST_8(FA_CASE)
    FA_SLOT_LIST_4(FA_CASE)
#undef FA_CASE
    return hipErrorInvalidValue;
}

hipError_t extend_launch_persistent(int L, int S, const ExtArgs &a, uint32_t first,
                                    uint32_t count, uint32_t qidx, int cus, int blocks_per_cu,
                                    hipStream_t st) {
#define FA_CASE(LL, SS)                                                                   \
    if (L == LL && S == SS)                                                               \
        return launch_persistent<LL, SS>(a, first, count, qidx, cus, blocks_per_cu, st);
    FA_SLOT_LIST_16(FA_CASE)
    FA_SLOT_LIST_8(FA_CASE)
    FA_SLOT_LIST_4(FA_CASE)
#undef FA_CASE
    return hipErrorInvalidValue;
}

hipError_t queue_reset(uint32_t *q, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(queue_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, st, q, n);
    return hipGetLastError();
}

}  // namespace fa
