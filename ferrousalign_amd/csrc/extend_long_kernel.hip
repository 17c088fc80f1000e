// Long extension pairs: ksw_extend2 (src/core/alignment/ksw_affine_gap.rs:952-1169)
// for the pairs the lane-group kernels do not hold: a query longer than 255
// bases (their columns live in the registers of at most 16 lanes) or a
// target longer than 2048 bases (their rows are staged in LDS).
//
// One wavefront per pair.  The reference's eh[] row stays in HBM, in a slice
// of a per-launch scratch buffer (the band it touches stays in L2); the
// query's codes sit in LDS.  A row visits only its band [beg, end), in tiles
// of 64 x kLongS columns aligned to absolute column numbers, so column c is
// always read and written by the same lane (no cross-lane traffic through
// memory): lane l of a tile owns columns tile + l*kLongS + s.  F is the
// exclusive max-plus prefix of G_k = max(M_k - oe_ins, 0) + k*e_ins over the
// band (as in the lane-group kernels), carried from tile to tile; the row
// maximum and its first column, and the first and last non-zero column of
// the next band, are wave reductions at the end of the row.  32-bit
// arithmetic throughout; the rare pairs that come here are few, so the
// kernel keeps to plain shuffles.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "extend_kernel.h"

namespace fa {

namespace {

constexpr int kLongS = 8;                  // columns per lane per tile
constexpr int kLongTile = 64 * kLongS;     // columns per tile
constexpr int kNegL = -(1 << 30);          // identity of the F prefix max
constexpr int kBig = 0x7FFFFFFF;
constexpr uint32_t kLongBlocks = 1024;     // workgroups per launch (eh slices of the scratch)

__device__ __forceinline__ int wave_max(int x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = max(x, __shfl_xor(x, d, 64));
    return x;
}
__device__ __forceinline__ int wave_min(int x) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) x = min(x, __shfl_xor(x, d, 64));
    return x;
}
// max over the lanes below this one (kNegL for lane 0)
__device__ __forceinline__ int wave_excl_max(int x, int lane) {
    int y = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(y, d, 64);
        if (lane >= d) y = max(y, t);
    }
    const int e = __shfl_up(y, 1, 64);
    return lane == 0 ? kNegL : e;
}

}  // namespace

__global__ __launch_bounds__(64) void extend_long_kernel(ExtArgs a, uint32_t first, uint32_t count,
                                                         int2 *ehbuf, uint32_t ehstride) {
    extern __shared__ uint8_t qc[];  // the query's codes, one byte per base
    const int lane = (int)threadIdx.x;
    int2 *const eh = ehbuf + (size_t)blockIdx.x * ehstride;
    const int o_del = a.o_del, e_del = a.e_del, o_ins = a.o_ins, e_ins = a.e_ins;
    const int oe_del = o_del + e_del, oe_ins = o_ins + e_ins;

    for (uint32_t p = blockIdx.x; p < count; p += gridDim.x) {
        const uint32_t pid = a.order[first + p];
        const int qlen = a.qlen[pid], tlen = a.tlen[pid], h0 = a.h0[pid], w = a.w[pid];
        const uint32_t qo = a.qo[pid], to = a.to[pid];

        // query codes into LDS, 16 bases per 2-bit word
        const int nqw = (qlen + 15) >> 4;
        for (int k = lane; k < nqw; k += 64) {
            const uint32_t w2 = a.seq2[qo + k];
            const uint32_t wn = (a.seqn[(qo >> 1) + (k >> 1)] >> ((k & 1) * 16)) & 0xFFFFu;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                const uint32_t c = ((wn >> j) & 1u) ? 4u : ((w2 >> (2 * j)) & 3u);
                qc[k * 16 + j] = (uint8_t)c;
            }
        }
        // first row, :992-998 in closed form: eh[0].h = h0, eh[1].h = max(h0 -
        // oe_ins, 0), then e_ins less per column down to 0; e = 0
        const int v1 = h0 > oe_ins ? h0 - oe_ins : 0;
        for (int c = lane; c <= qlen; c += 64) {
            const int hv = c == 0 ? h0 : max(v1 - (c - 1) * e_ins, 0);
            eh[c] = make_int2(hv, 0);
        }
        __syncthreads();
        // a column is read and written by one lane from here on (tile-aligned
        // ownership), except eh[] above, written before the barrier

        int beg = 0, end = qlen;
        int best = h0, max_i = -1, max_j = -1, max_ie = -1, gscore = -1, max_off = 0;
        uint32_t ncell = 0;
        for (int i = 0; i < tlen; ++i) {
            // :1038-1046
            if (beg < i - w) beg = i - w;
            if (end > i + w + 1) end = i + w + 1;
            if (end > qlen) end = qlen;
            const int rb = beg, re = end;
            // :1049-1054
            const int h1 = rb == 0 ? max(h0 - (o_del + e_del * (i + 1)), 0) : 0;
            // the row's target code and its scores against query codes 0..4,
            // biased by 128 as bytes (an N target reads the uniform N row)
            const uint32_t tw = a.seq2[to + (i >> 4)];
            const uint32_t tn = a.seqn[(to >> 1) + (i >> 5)];
            const int t = ((tn >> (i & 31)) & 1u) ? 4 : (int)((tw >> ((i & 15) * 2)) & 3u);
            uint32_t rlo = 0, rhi = 0;
            if (t < 4) {
#pragma unroll
                for (int q = 0; q < 4; ++q) rlo |= ((a.profw[q] >> (8 * t)) & 0xFFu) << (8 * q);
                rhi = (a.profw[4] >> (8 * t)) & 0xFFu;
            } else {
                rlo = a.nrow * 0x01010101u;
                rhi = a.nrow;
            }

            int carryG = kNegL;  // max G over the band columns of the earlier tiles
            int hleft = h1;      // H(i, c-1) for the first column of the next tile
            int lm = 0, lj = kBig;         // this lane's row maximum and its first column
            int fnz = kBig, lnz = -1;      // first / last non-zero column after the row
            int hlast = h1;                // H(i, re-1), or h1 for an empty band
            const int t0 = (rb / kLongTile) * kLongTile;
            for (int tb = t0; rb < re && tb < re; tb += kLongTile) {
                const int cb = tb + lane * kLongS;
                int ee[kLongS], M[kLongS], pre[kLongS];
                int run = kNegL;
#pragma unroll
                for (int s = 0; s < kLongS; ++s) {
                    const int c = cb + s;
                    const bool inb = c >= rb && c < re;
                    int2 v = make_int2(0, 0);
                    if (inb) v = eh[c];
                    ee[s] = v.y;
                    const uint32_t q = inb ? qc[c] : 4u;
                    const int sc = (int)__builtin_amdgcn_perm(rhi, rlo, 0x0C0C0C00u | q) - 128;
                    M[s] = v.x != 0 ? v.x + sc : 0;  // :1061-1063
                    const int g = inb ? max(M[s] - oe_ins, 0) + c * e_ins : kNegL;
                    pre[s] = run;
                    run = max(run, g);
                }
                const int excl = max(carryG, wave_excl_max(run, lane));
                carryG = max(carryG, wave_max(run));
                int H[kLongS];
#pragma unroll
                for (int s = 0; s < kLongS; ++s) {
                    const int c = cb + s;
                    // F(c) = max over band columns k < c of G_k - (c-1)*e_ins (the
                    // reference's f, which starts at 0 on the first column; a
                    // negative F never wins against E >= 0)
                    const int F = max(excl, pre[s]) - (c - 1) * e_ins;
                    H[s] = max(max(M[s], ee[s]), F);  // :1064-1066
                }
                // eh[c].h <- H(i, c-1) (:1060), h1 at the band start
                const int hprev_lane = __shfl_up(H[kLongS - 1], 1, 64);
                const int hin = lane == 0 ? hleft : hprev_lane;
#pragma unroll
                for (int s = 0; s < kLongS; ++s) {
                    const int c = cb + s;
                    if (c < rb || c >= re) continue;
                    const int hl = c == rb ? h1 : (s > 0 ? H[s - 1] : hin);
                    // :1067-1075 row maximum, first column on ties
                    if (H[s] > lm) {
                        lm = H[s];
                        lj = c;
                    }
                    // :1076-1083 e for the next row
                    const int en = max(ee[s] - e_del, max(M[s] - oe_del, 0));
                    eh[c] = make_int2(hl, en);
                    if ((hl | en) != 0) {
                        fnz = min(fnz, c);
                        lnz = max(lnz, c);
                    }
                }
                // H of the tile's last band column: the next tile's hleft, or
                // hlast when this is the band's last tile
                const int lastc = min(tb + kLongTile, re) - 1;
                int mine = kNegL;
#pragma unroll
                for (int s = 0; s < kLongS; ++s)
                    if (cb + s == lastc) mine = H[s];
                hleft = wave_max(mine);
                hlast = hleft;
            }
            // :1106-1107 eh[end] <- (h1, 0), h1 being the band's last H
            if (re >= 0 && re <= qlen && (re / kLongTile) * kLongTile + lane * kLongS <= re &&
                re < (re / kLongTile) * kLongTile + (lane + 1) * kLongS) {
                eh[re] = make_int2(hlast, 0);
                if (hlast != 0) lnz = max(lnz, re);
            }
            const int m_row = wave_max(lm);
            const int mj = m_row > 0 ? wave_min(lm == m_row ? lj : kBig) : -1;
            const int first_nz = wave_min(fnz);
            const int last_nz = wave_max(lnz);

            if (re > rb) ncell += (uint32_t)(re - rb);
            // :1110-1120 ((beg..end) holding qlen-1 implies end == qlen)
            if (re == qlen && gscore < hlast) {
                gscore = hlast;
                max_ie = i;
            }
            if (m_row == 0) break;  // :1122-1124
            if (m_row > best) {     // :1126-1133
                best = m_row;
                max_i = i;
                max_j = mj;
                const int off = mj > i ? mj - i : i - mj;
                if (max_off < off) max_off = off;
            } else if (a.zdrop > 0) {  // :1134-1145
                const int di = i - max_i, dj = mj - max_j;
                if (di > dj) {
                    if (best - m_row - (di - dj) * e_del > a.zdrop) break;
                } else if (best - m_row - (dj - di) * e_ins > a.zdrop) {
                    break;
                }
            }
            // :1148-1158: first non-zero column in [beg, end), else end (beg when
            // the band is empty); the last non-zero one in (that, end], else
            // that (end when that is not below end); end two past it
            int nb, ne;
            if (rb >= re) {
                nb = rb;
                ne = re;
            } else {
                nb = first_nz < re ? first_nz : re;
                ne = re <= nb ? re : max(nb, last_nz);
            }
            beg = nb;
            end = ne + 2 < qlen ? ne + 2 : qlen;
        }
        if (lane == 0)  // score, tle, gtle, qle, global score, max offset
            store_result(a, pid, best, max_i + 1, max_ie + 1, max_j + 1, gscore, max_off, ncell);
        __syncthreads();  // the next pair reuses the LDS codes and the eh slice
    }
}

size_t extend_long_scratch(int qmax, uint32_t count) {
    const uint32_t blocks = count < kLongBlocks ? count : kLongBlocks;
    return (size_t)blocks * ((size_t)qmax + 1) * sizeof(int2);
}

hipError_t extend_long_launch(const ExtArgs &a, uint32_t first, uint32_t count, int qmax,
                              void *scratch, size_t scratch_bytes, hipStream_t st) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = count < kLongBlocks ? count : kLongBlocks;
    const uint32_t stride = (uint32_t)qmax + 1;
    if (!scratch || scratch_bytes < (size_t)blocks * stride * sizeof(int2)) return hipErrorInvalidValue;
    const size_t lds = ((size_t)qmax + 16) / 16 * 16;
    if (lds > 65536) return hipErrorInvalidValue;
    hipLaunchKernelGGL(extend_long_kernel, dim3(blocks), dim3(64), lds, st, a, first, count,
                       reinterpret_cast<int2 *>(scratch), stride);
    return hipGetLastError();
}

}  // namespace fa
