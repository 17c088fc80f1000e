// Mate-rescue local Smith-Waterman, packed: two pairs per lane group (kernel
// template; instantiated by mate16_sym.hip, mate16_kernel.hip and
// mate8_kernel.hip).
//
// Same execution model as the reference's 8-lane SSE ksw_i16
// (src/core/alignment/ksw_affine_gap.rs:503-735, :773-861; SimdEngine128,
// simd_abstraction/engine128.rs:247-326): GPU lane p of a group plays SSE
// lane p, query position k lives in lane k / slen, segment k % slen, and the
// lane walks its segments in order for every target row.  The reference's
// arithmetic is signed 16-bit saturating (adds/subs_epi16) with signed max,
// which is exactly v_pk_add_i16 / v_pk_sub_i16 with clamp and v_pk_max_i16,
// so every register holds two pairs, pair A in the low and pair B in the high
// half.  Per pair state (row maximum, best row, lazy-F flag, stopping) stays
// per half.  Pairs of one launch share slen (the host buckets by it); the
// reverse (XSTART) pass runs each half with its own query length, so its
// segments past a half's slen are masked: they neither feed the row maximum,
// nor the F chain, nor the lazy-F pass.
//
// 8-bit mode (KSW_XBYTE, ksw_u8_impl :267-494): the reference runs 16 unsigned
// byte lanes with slen = ceil(qlen / 16); GPU lane p of an 8-lane group plays
// SSE lanes 2p and 2p+1 one after the other, so its registers j = 0..2 slen-1
// hold query positions j + 2p slen (the 16-bit mapping with 2 slen registers):
// F restarts at 0 where SSE lane 2p+1 begins (register slen), and the lazy-F
// pass carries both lanes' F chains so that F still moves one SSE lane per
// pass, as the reference's byte shift does.  Its
// cell is H = subs(adds(H, mat + shift), shift) with shift = (u8)(256 - the
// smallest matrix entry), E and F decay by unsigned saturating subtraction
// (clamped at 0), and a best score reaching 255 - shift stops the pass and is
// reported as 255.  Here the halves hold the same values unclamped above: the
// profile adds mat itself (E, F >= 0 make the max at least 0), and a cell the
// reference would saturate at 255 already carries a row maximum >= 255 - shift,
// which ends that pass on the same row in both, so no saturated value is ever
// reported.  E and F use v_pk_sub_u16 with clamp.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mate_kernel.h"

namespace fa {

namespace {

typedef short s2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t bits(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2 vs(uint32_t v) { return __builtin_bit_cast(s2, v); }
__device__ __forceinline__ uint32_t pk(int lo, int hi) {
    return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}
__device__ __forceinline__ int lo16(uint32_t x) { return (int)(short)(x & 0xFFFFu); }
__device__ __forceinline__ int hi16(uint32_t x) { return (int)(short)(x >> 16); }
__device__ __forceinline__ uint32_t smax(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_max(vs(a), vs(b)));
}
__device__ __forceinline__ uint32_t addsat(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_add_sat(vs(a), vs(b)));
}
__device__ __forceinline__ uint32_t subsat(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_sub_sat(vs(a), vs(b)));
}
__device__ __forceinline__ uint32_t sub16(uint32_t a, uint32_t b) {
    return bits(vs(a) - vs(b));
}
typedef unsigned short u2 __attribute__((ext_vector_type(2)));
// unsigned saturating (clamp at 0) per half: subs_epu8 on values 0..~260
__device__ __forceinline__ uint32_t usubsat(uint32_t a, uint32_t b) {
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2, a),
                                                                      __builtin_bit_cast(u2, b)));
}
// the gap decay of the mode: signed saturating (16-bit) or clamped at 0 (8-bit)
template <bool U8>
__device__ __forceinline__ uint32_t gsub(uint32_t a, uint32_t b) {
    if constexpr (U8)
        return usubsat(a, b);
    else
        return subsat(a, b);
}
// 0xFFFF in each half that is negative
__device__ __forceinline__ uint32_t sgnmask(uint32_t d) {
    uint32_t r;
    asm("v_pk_ashrrev_i16 %0, 15, %1 op_sel_hi:[0,1]" : "=v"(r) : "v"(d));
    return r;
}
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) {
    return (m & a) | (~m & b);
}
// b = (m & a) | (~m & b) in b's own register, so a conditional update leaves
// the loop-carried value where it was
__device__ __forceinline__ void bfi_into(uint32_t &b, uint32_t m, uint32_t a) {
    asm("v_bfi_b32 %0, %1, %2, %0" : "+v"(b) : "v"(m), "v"(a));
}

constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowShr1 = 0x111;

// DPP move with bound_ctrl (a lane without a source lane reads 0): no old
// value, so no copy of the source, and 32-bit consumers fuse with it
template <int CTRL>
__device__ __forceinline__ uint32_t dmov(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
// per-half signed max over the 8 lanes of a group
__device__ __forceinline__ uint32_t grp8_smax2(uint32_t x) {
    x = smax(x, dmov<kQuadXor1>(x));
    x = smax(x, dmov<kQuadXor2>(x));
    x = smax(x, dmov<kRowHalfMirror>(x));
    return x;
}
__device__ __forceinline__ uint32_t grp8_or(uint32_t x) {
    x |= dmov<kQuadXor1>(x);
    x |= dmov<kQuadXor2>(x);
    x |= dmov<kRowHalfMirror>(x);
    return x;
}
__device__ __forceinline__ int grp8_imax(int x) {
    x = max(x, (int)dmov<kQuadXor1>((uint32_t)x));
    x = max(x, (int)dmov<kQuadXor2>((uint32_t)x));
    x = max(x, (int)dmov<kRowHalfMirror>((uint32_t)x));
    return x;
}
// value of lane p-1 of the group (lane 0: 0): the SSE shift by one lane
__device__ __forceinline__ uint32_t shift_up(uint32_t x, int p) {
    const uint32_t y = dmov<kRowShr1>(x);
    return p == 0 ? 0u : y;
}
// 0xFFFF in each half whose value equals j
__device__ __forceinline__ uint32_t eq16(int j, uint32_t vp) {
    const uint32_t d = sub16((uint32_t)j * 0x00010001u, vp);
    return ~(sgnmask(d) | sgnmask(sub16(0u, d)));
}

__device__ __forceinline__ int seq_code(const uint32_t *seq2, const uint32_t *seqn, uint32_t off,
                                        int pos) {
    const uint32_t c = (seq2[off + (pos >> 4)] >> ((pos & 15) * 2)) & 3u;
    return ((seqn[(off >> 1) + (pos >> 5)] >> (pos & 31)) & 1u) ? 4 : (int)c;
}

struct Half {
    bool active;
    uint32_t qo, to;
    int qlen, tlen, trev;
    uint32_t endsc;
};
struct Res {
    int score, te, qe, rows;
};

// Whether ksw_align2 runs the XSTART reverse pass after a forward result
// (:804-810): KSW_XSTART set, the score not below KSW_XSUBO's minimum, and a
// hit (qe, te >= 0; a saturated 8-bit pass reports qe = -1).
__device__ __forceinline__ bool xstart_needed(uint32_t xtra, int score, int te, int qe) {
    return (xtra & kKswXStart) && !((xtra & kKswXSubo) && score < (int)(xtra & 0xffffu)) &&
           qe >= 0 && te >= 0;
}

// Target rows staged per lane group at a time: byte r of a chunk holds row
// i0 + r of pair A in bits 0-3 and of pair B in bits 4-7.
constexpr int kTgtChunk = 512;  // the forward staging decodes 4 + 2 words per lane
static_assert(kTgtChunk == 512, "forward staging reads 64 rows per lane");

// wave-uniform largest j in [1, SLEN] with pred(j) true for some lane
template <int SLEN, class P>
__device__ __forceinline__ int wave_last(P pred) {
    int r = 0;
#pragma unroll
    for (int j = 1; j <= SLEN; ++j)
        if (__builtin_amdgcn_ballot_w64(pred(j)) != 0) r = j;
    return r;
}

// One ksw_i16 pass (U8: ksw_u8, SLEN = 2 x its slen) for the group's two pairs (`qrev`: the XSTART reverse
// pass, query position k = original qlen-1-k, target row i = original
// trev - i while i <= trev).
// UNIFORM: every pair of the wave has slen = SLEN (the host buckets the
// forward pass by slen and instantiates the kernel per slen), so no slot is
// conditional; otherwise pairs may have any slen <= SLEN and the slots past a
// pair's slen are masked.
// tc2: the group's kTgtChunk-byte target staging area.
// logp: lane 0 logs pair A's row maxima, lane 1 pair B's (nullptr: no log).
// hm: the wave's copy of H at the best row so far (Hmax), word j * 64 + lane.
// SYM: o_del == o_ins and e_del == e_ins, so E and F share the gap-open term.
template <int SLEN, bool UNIFORM, bool SYM, bool U8>
__device__ __forceinline__ void ksw_pass2(const MateArgs &a, int p, const Half (&hf)[2],
                                          uint8_t *tc2, uint16_t *logp, uint32_t *hm,
                                          bool qrev, Res (&res)[2]) {
    // R SSE lanes per GPU lane; slen = segments per SSE lane, ext = registers used
    constexpr int G = 8;  // GPU lanes per pair group
    constexpr int R = U8 ? 2 : 1, P16 = U8 ? 16 : 8;
    static_assert(SLEN % R == 0, "8-bit passes run an even register count");
    const int lane = (int)(threadIdx.x & 63);
    int slen[2], ext[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        slen[h] = hf[h].active ? (hf[h].qlen + P16 - 1) / P16 : 0;
        ext[h] = R * slen[h];
    }
    int smax_w = SLEN, sfull = SLEN;
    if constexpr (!UNIFORM) {
        smax_w = wave_last<SLEN>([&](int j) { return j <= ext[0] || j <= ext[1]; });
        // the smallest extent among the wave's active pairs: slots below it exist
        // for every pair
        sfull = SLEN;
#pragma unroll
        for (int j = SLEN; j >= 1; --j)
            if (__builtin_amdgcn_ballot_w64((hf[0].active && j > ext[0]) ||
                                            (hf[1].active && j > ext[1])) != 0)
                sfull = j - 1;
    }
    const uint32_t slenp = pk(ext[0], ext[1]);
    const uint32_t slm1p = pk(ext[0] - 1, ext[1] - 1);
    const uint32_t splitp = pk(slen[0], slen[1]);  // 8-bit: first register of SSE lane 2p+1

    uint32_t profA[SLEN], profB[SLEN], H[SLEN], H2[SLEN], E[SLEN];
#pragma unroll
    for (int j = 0; j < SLEN; ++j) {
        uint32_t pw[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int k = j + p * ext[h];
            pw[h] = 0x80808080u;  // past the query: score 0 for every target base
            if (hf[h].active && j < ext[h] && k < hf[h].qlen) {
                const int orig = qrev ? hf[h].qlen - 1 - k : k;
                pw[h] = a.profw[seq_code(a.seq2, a.seqn, hf[h].qo, orig)];
            }
        }
        profA[j] = pw[0];
        profB[j] = pw[1];
        H[j] = 0u;
        H2[j] = 0u;
        E[j] = 0u;
        hm[j * 64 + lane] = 0u;
    }
    const uint32_t edel = pk(a.e_del, a.e_del), eins = pk(a.e_ins, a.e_ins);
    const uint32_t oedel = pk(a.o_del + a.e_del, a.o_del + a.e_del);
    const uint32_t oeins = pk(a.o_ins + a.e_ins, a.o_ins + a.e_ins);
    // per-pair state, packed like the cells: 0xFFFF per live half, best score,
    // its row, rows run, and the stop score min(endsc, 0x7FFF) (a score of 0x7FFF
    // is saturated and stops too; 8-bit: min(endsc, 255 - shift))
    uint32_t alivem = (hf[0].active && hf[0].tlen > 0 ? 0xFFFFu : 0u) |
                      (hf[1].active && hf[1].tlen > 0 ? 0xFFFF0000u : 0u);
    uint32_t gmaxp = 0u, tep = 0xFFFFFFFFu, rowsp = 0u, ip = 0u;
    const uint32_t sat = U8 ? (uint32_t)a.ovf : 0x7FFFu;
    const uint32_t endscp = pk((int)min(hf[0].endsc, sat), (int)min(hf[1].endsc, sat));
    const uint32_t tlenp = pk(hf[0].tlen, hf[1].tlen);
    // stage target rows [i0, i0 + kTgtChunk) of both pairs; the reverse pass reads
    // the target reversed up to trev and forward past it (the reference reverses
    // the first te + 1 bases in place, :812-813)
    auto stage = [&](int i0) {
        if (!qrev) {
            // lane p decodes rows [i0 + RPL p, i0 + RPL (p + 1)) of both pairs from
            // RPL / 16 2-bit words and RPL / 32 N-bit words per pair (targets start
            // on a 32-base boundary: `to` counts 16-base words and is even)
            constexpr int RPL = kTgtChunk / G, NW2 = RPL / 16, NWN = RPL / 32;
            const int r0 = RPL * p, i = i0 + r0;
            uint32_t w[2][NW2], nw[2][NWN];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int nwords = hf[h].active ? (hf[h].tlen + 15) >> 4 : 0;
                const int nnw = hf[h].active ? (hf[h].tlen + 31) >> 5 : 0;
#pragma unroll
                for (int q = 0; q < NW2; ++q) {
                    const int x = (i >> 4) + q;
                    w[h][q] = x < nwords ? a.seq2[hf[h].to + x] : 0u;
                }
#pragma unroll
                for (int q = 0; q < NWN; ++q) {
                    const int x = (i >> 5) + q;
                    nw[h][q] = x < nnw ? a.seqn[(hf[h].to >> 1) + x] : 0u;
                }
            }
#pragma unroll
            for (int d = 0; d < RPL / 4; ++d) {
                uint32_t v = 0u;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int r = 4 * d + b;
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        uint32_t c = (w[h][r >> 4] >> ((r & 15) * 2)) & 3u;
                        if ((nw[h][r >> 5] >> (r & 31)) & 1u) c = 4u;
                        v |= c << (8 * b + 4 * h);
                    }
                }
                *(uint32_t *)(tc2 + r0 + 4 * d) = v;
            }
            __builtin_amdgcn_s_waitcnt(0);
            __builtin_amdgcn_wave_barrier();
            return;
        }
#pragma unroll 8
        for (int m = 0; m < kTgtChunk / G; ++m) {
            const int r = p + G * m, i = i0 + r;
            uint32_t c[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                c[h] = 0u;
                if (hf[h].active && i < hf[h].tlen) {
                    const int pos = (qrev && i <= hf[h].trev) ? hf[h].trev - i : i;
                    c[h] = (uint32_t)seq_code(a.seq2, a.seqn, hf[h].to, pos);
                }
            }
            tc2[r] = (uint8_t)(c[0] | (c[1] << 4));
        }
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
    };
    uint32_t hlast = 0u;  // H(i-1, slen-1) per half, carried to the next row's shift
    // one target row: H(i-1) is read from Hd, H(i) written to Hn; the row loop
    // alternates two register sets so that no row copies H back
    auto row = [&](const int i, const uint32_t (&Hd)[SLEN], uint32_t (&Hn)[SLEN]) {
        // target codes of this row
        const uint32_t tw = tc2[i & (kTgtChunk - 1)];
        // score byte mat[tb][q] + 128 of each pair's profile, bias 0x80; an N target
        // base scores 0 (the N row is zero, checked on the host): selector 0x0C
        // gives the byte 0x00 and its half's bias is 0
        const uint32_t nb = ((tw & 4u) >> 2) | ((tw & 0x40u) << 10);
        const uint32_t sel = (tw & 3u) | ((tw & 0x30u) << 12) | 0x0C040C00u | (nb * 0x0Cu);
        const uint32_t biasp = (nb ^ 0x00010001u) << 7;
        uint32_t f = 0u, mx = 0u, fA = 0u;  // fA: 8-bit, SSE lane 2p's F at its end
        {
            uint32_t h = shift_up(hlast, p);
#pragma unroll
            for (int j = 0; j < SLEN; ++j) {
                if (j < smax_w) {
                    if constexpr (U8) {  // SSE lane 2p+1 begins: its F enters at 0
                        if constexpr (UNIFORM) {
                            if (j == SLEN / 2) {
                                fA = f;
                                f = 0u;
                            }
                        } else if (j >= 1 && j <= SLEN / 2) {
                            const uint32_t eq = eq16(j, splitp);
                            fA = bfi(eq, f, fA);
                            f &= ~eq;
                        }
                    }
                    const uint32_t sb = sub16(__builtin_amdgcn_perm(profB[j], profA[j], sel), biasp);
                    uint32_t hv = addsat(h, sb);
                    const uint32_t e = E[j];
                    hv = smax(smax(hv, e), f);
                    uint32_t fn;
                    if constexpr (SYM) {
                        const uint32_t t = gsub<U8>(hv, oedel);
                        fn = smax(gsub<U8>(f, eins), t);
                        E[j] = smax(gsub<U8>(e, edel), t);
                    } else {
                        fn = smax(gsub<U8>(f, eins), gsub<U8>(hv, oeins));
                        E[j] = smax(gsub<U8>(e, edel), gsub<U8>(hv, oedel));
                    }
                    h = Hd[j];  // H(i-1, j): the next segment's diagonal
                    Hn[j] = hv;
                    if (UNIFORM || j < sfull) {
                        mx = smax(mx, hv);
                        f = fn;
                    } else {  // some pair has no segment j
                        const uint32_t vm = sgnmask(sub16((uint32_t)j * 0x00010001u, slenp));
                        mx = smax(mx, hv & vm);
                        f = bfi(vm, fn, f);
                    }
                }
            }
        }
        // the row maximum is taken before lazy F (:601, :635); materialize it here so
        // the row's H values need not stay live through the lazy-F pass
        asm volatile("" : "+v"(mx));
        // lazy F (:607-633): shift F one lane, raise H, decay F; a pair stops after a
        // pass without change.  Every lane's F entered its segment 0 at 0 and H >= F
        // along the chain, so a shifted F <= 0 in every lane of a pair changes no H:
        // the first pass is skipped when that holds for the whole wave.
        // 8-bit: SSE lane 2p takes lane 2p-1's F (GPU lane p-1's second chain) and
        // lane 2p+1 takes lane 2p's (this lane's first chain), both from the last pass
        uint32_t lazym = alivem;
        const uint32_t fs = shift_up(f, p);
        uint32_t fsB = fA;
        uint32_t pos = sgnmask(sub16(0u, fs));
        if constexpr (U8) pos |= sgnmask(sub16(0u, fsB));
        if (__builtin_amdgcn_ballot_w64((pos & lazym) != 0u) != 0) {
            f = fs;
            for (int pass = 0;;) {
                uint32_t chg = 0u, fAend = 0u;
#pragma unroll
                for (int j = 0; j < SLEN; ++j) {
                    if (j < smax_w) {
                        if constexpr (U8) {
                            if constexpr (UNIFORM) {
                                if (j == SLEN / 2) {
                                    fAend = f;
                                    f = fsB;
                                }
                            } else if (j >= 1 && j <= SLEN / 2) {
                                const uint32_t eq = eq16(j, splitp);
                                fAend = bfi(eq, f, fAend);
                                f = bfi(eq, fsB, f);
                            }
                        }
                        uint32_t hn = smax(Hn[j], f);
                        const uint32_t fn = gsub<U8>(f, eins);
                        if (UNIFORM || j < sfull) {
                            f = fn;
                        } else {  // a pair without segment j keeps H and F
                            const uint32_t vm = sgnmask(sub16((uint32_t)j * 0x00010001u, slenp));
                            hn = bfi(vm, hn, Hn[j]);
                            f = bfi(vm, fn, f);
                        }
                        chg |= hn ^ Hn[j];
                        bfi_into(Hn[j], lazym, hn);
                    }
                }
                chg = grp8_or(chg);
                if ((chg & 0xFFFFu) == 0u) lazym &= 0xFFFF0000u;
                if ((chg >> 16) == 0u) lazym &= 0x0000FFFFu;
                if (++pass == P16 || __builtin_amdgcn_ballot_w64(lazym != 0u) == 0) break;
                f = shift_up(f, p);
                fsB = fAend;
            }
        }
        // H(i, slen-1) per half for the next row's shift
#pragma unroll
        for (int j = 0; j < SLEN; ++j) {
            if constexpr (UNIFORM) {
                if (j == SLEN - 1) hlast = Hn[j];
            } else if (j < smax_w && j + 1 >= sfull) {
                hlast = bfi(eq16(j, slm1p), Hn[j], hlast);
            }
        }
        const uint32_t rmax = grp8_smax2(mx);  // >= 0 in both halves
        if (logp != nullptr && ((alivem >> (16 * (p & 1))) & 1u))
            logp[i] = (uint16_t)(rmax >> (16 * (p & 1)));
        rowsp += alivem & 0x00010001u;
        const uint32_t upd = sgnmask(sub16(gmaxp, rmax)) & alivem;  // rmax > gmax
        gmaxp = bfi(upd, rmax, gmaxp);
        tep = bfi(upd, ip, tep);
        ip += 0x00010001u;
        const uint32_t stop = (upd & ~sgnmask(sub16(gmaxp, endscp))) | ~sgnmask(sub16(ip, tlenp));
        alivem &= ~stop;
        if (__builtin_amdgcn_ballot_w64(upd != 0u) != 0) {
            uint16_t *const hm16 = (uint16_t *)hm;
            if (upd & 0xFFFFu) {
#pragma unroll
                for (int j = 0; j < SLEN; ++j)
                    if (j < smax_w) hm16[(j * 64 + lane) * 2] = (uint16_t)Hn[j];
            }
            if (upd >> 16) {
#pragma unroll
                for (int j = 0; j < SLEN; ++j)
                    if (j < smax_w) hm16[(j * 64 + lane) * 2 + 1] = (uint16_t)(Hn[j] >> 16);
            }
        }
        };
    for (int i = 0;; i += 2) {
        if (__builtin_amdgcn_ballot_w64(alivem != 0u) == 0) break;
        if ((i & (kTgtChunk - 1)) == 0) {
            __builtin_amdgcn_wave_barrier();  // the previous chunk's reads are issued
            stage(i);
        }
        row(i, H, H2);
        if (__builtin_amdgcn_ballot_w64(alivem != 0u) == 0) break;
        row(i + 1, H2, H);
    }
    const int gmax[2] = {lo16(gmaxp), hi16(gmaxp)};
    const int te[2] = {lo16(tep), hi16(tep)};
    const int rows[2] = {lo16(rowsp), hi16(rowsp)};
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
        // saturated: 0x7FFF (16-bit), or a best score >= 255 - shift reported as 255
        const bool satd = U8 ? gmax[hh] >= a.ovf : gmax[hh] == 0x7FFF;
        res[hh].score = U8 && satd ? 255 : gmax[hh];
        res[hh].te = te[hh];
        res[hh].rows = rows[hh];
        res[hh].qe = -1;
        // qe: largest Hmax over real positions, smallest position on ties
        int key = 0;
#pragma unroll
        for (int j = 0; j < SLEN; ++j) {
            const int k = j + p * ext[hh];
            const uint32_t hmj = hm[j * 64 + lane];
            const int hv = hh ? hi16(hmj) : lo16(hmj);
            if (j < ext[hh] && k < hf[hh].qlen) key = max(key, ((hv + 32769) << 12) | (4095 - k));
        }
        key = grp8_imax(key);
        if (!satd) res[hh].qe = key ? 4095 - (key & 4095) : -1;
    }
}

}  // namespace

// Stripe lengths 14..20 fit three waves per SIMD in 168 VGPRs with spills only
// outside the row loop; the shorter ones fit three or more by themselves, the
// longer ones would spill inside it.
constexpr int mate16_waves(int slen) { return slen >= 14 && slen <= 20 ? 3 : 1; }

// Two pairs per 8-lane group, 16 pairs per single-wave block; SLEN registers per
// lane (U8, the 8-bit mode: SLEN = 2 x its slen).
template <int SLEN, bool SYM, bool U8>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(mate16_waves(SLEN)))) void
mate16_kernel(MateArgs a, uint32_t first, uint32_t count) {
    constexpr int PPB = 16;
    extern __shared__ uint8_t lds[];
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane >> 3, p = lane & 7;
    Half hf[2];
    uint32_t pid[2], slot[2];
    bool valid[2];
    uint8_t *const tc2 = lds + g * kTgtChunk;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        slot[h] = blockIdx.x * (uint32_t)PPB + 2u * (uint32_t)g + (uint32_t)h;
        valid[h] = slot[h] < count;
        pid[h] = valid[h] ? a.order[first + slot[h]] : 0u;
        hf[h].qlen = valid[h] ? a.qlen[pid[h]] : 1;
        hf[h].tlen = valid[h] ? a.tlen[pid[h]] : 0;
        hf[h].qo = valid[h] ? a.qo[pid[h]] : 0u;
        hf[h].active = valid[h];
        hf[h].trev = 0;
        hf[h].to = valid[h] ? a.to[pid[h]] : 0u;
    }
    // row-maximum log of pair p (lanes 0 and 1 of the group)
    uint16_t *const logp =
        (p < 2 && valid[p & 1]) ? a.imax_log + (size_t)slot[p & 1] * (size_t)a.log_stride : nullptr;
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();

    const uint32_t xtra = a.xtra;
    const uint32_t minsc = (xtra & kKswXSubo) ? (xtra & 0xffffu) : 0x10000u;
    const uint32_t endsc = (xtra & kKswXStop) ? (xtra & 0xffffu) : 0x10000u;
    hf[0].endsc = hf[1].endsc = endsc;
    Res r[2];
    uint32_t *const hm = (uint32_t *)(lds + 8 * kTgtChunk);
    ksw_pass2<SLEN, true, SYM, U8>(a, p, hf, tc2, logp, hm, false, r);

    // suboptimal hits (:639-655, :717-731) from the row log: lane 0 of the group for
    // pair A, lane 1 for pair B (selects, not indexing, keep the state in registers)
    int score2 = -1, te2 = -1;
    const bool mineB = p == 1;
    const bool mine_valid = mineB ? valid[1] : valid[0];
    const Res mine = mineB ? r[1] : r[0];
    if (p < 2 && mine_valid && mine.score != (U8 ? 255 : 0x7FFF) && minsc <= 0xFFFFu) {
        const Res rr = mine;
        const uint16_t *log = logp;
        const int iv = (rr.score + a.qmax - 1) / a.qmax;
        const int low = rr.te - iv, high = rr.te + iv;
        int cs = -1, ce = -1;
        for (int i = 0; i < rr.rows; ++i) {
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

    if (a.split_rev) {
        // the XSTART pass runs in mate16_rev_kernel, over the pairs that need it
        if (p < 2 && mine_valid) {
            const uint32_t id = mineB ? pid[1] : pid[0];
            const int ql = mineB ? hf[1].qlen : hf[0].qlen;
            int32_t *o = a.out + (size_t)id * 7;
            o[0] = mine.score;
            o[1] = mine.te;
            o[2] = mine.qe;
            o[3] = score2;
            o[4] = te2;
            o[5] = -1;
            o[6] = -1;
            if (a.cells) a.cells[id] = (uint32_t)mine.rows * (uint32_t)ql;
        }
        return;
    }

    // XSTART reverse pass (:804-855), per pair
    Half hr[2];
    bool rev[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        rev[h] = valid[h] && xstart_needed(xtra, r[h].score, r[h].te, r[h].qe);
        hr[h] = hf[h];
        hr[h].active = rev[h];
        hr[h].qlen = rev[h] ? r[h].qe + 1 : 1;
        hr[h].trev = r[h].te;
        hr[h].endsc = (uint32_t)r[h].score & 0xffffu;
    }
    Res rr[2];
    ksw_pass2<SLEN, false, SYM, U8>(a, p, hr, tc2, nullptr, hm, true, rr);

    if (p < 2 && mine_valid) {
        const Res f = mine, b = mineB ? rr[1] : rr[0];
        const bool rv = mineB ? rev[1] : rev[0];
        const int ql = mineB ? hf[1].qlen : hf[0].qlen, ql2 = mineB ? hr[1].qlen : hr[0].qlen;
        const uint32_t id = mineB ? pid[1] : pid[0];
        int tb = -1, qb = -1;
        uint32_t cells = (uint32_t)f.rows * (uint32_t)ql;
        if (rv) {
            cells += (uint32_t)b.rows * (uint32_t)ql2;
            if (b.score >= f.score) {
                tb = f.te - b.te;
                qb = f.qe - b.qe;
            }
        }
        int32_t *o = a.out + (size_t)id * 7;
        o[0] = f.score;
        o[1] = f.te;
        o[2] = f.qe;
        o[3] = score2;
        o[4] = te2;
        o[5] = tb;
        o[6] = qb;
        if (a.cells) a.cells[id] = cells;
    }
}

// The XSTART reverse pass (:804-855) alone, two pairs per 8-lane group, for
// pairs whose reverse query length qe + 1 has the stripe length of SLEN (the
// host buckets them by it), so the pass runs unmasked.  The reference runs
// ksw_i16 / ksw_u8 on the first qe + 1 query bases and the first te + 1
// target bases, both reversed, with KSW_XSTOP at the forward score.
template <int SLEN, bool SYM, bool U8>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(mate16_waves(SLEN)))) void
mate16_rev_kernel(MateArgs a, uint32_t first, uint32_t count) {
    constexpr int PPB = 16;
    extern __shared__ uint8_t lds[];
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane >> 3, p = lane & 7;
    Half hr[2];
    Res f[2];
    uint32_t pid[2];
    bool valid[2];
    uint8_t *const tc2 = lds + g * kTgtChunk;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t slot = blockIdx.x * (uint32_t)PPB + 2u * (uint32_t)g + (uint32_t)h;
        valid[h] = slot < count;
        pid[h] = valid[h] ? a.rorder[first + slot] : 0u;
        const int32_t *o = a.out + (size_t)pid[h] * 7;
        f[h].score = valid[h] ? o[0] : 0;
        f[h].te = valid[h] ? o[1] : -1;
        f[h].qe = valid[h] ? o[2] : -1;
        f[h].rows = 0;
        hr[h].active = valid[h];
        hr[h].qlen = valid[h] ? f[h].qe + 1 : 1;
        hr[h].tlen = valid[h] ? a.tlen[pid[h]] : 0;
        hr[h].trev = f[h].te;
        hr[h].qo = valid[h] ? a.qo[pid[h]] : 0u;
        hr[h].to = valid[h] ? a.to[pid[h]] : 0u;
        hr[h].endsc = (uint32_t)f[h].score & 0xffffu;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    uint32_t *const hm = (uint32_t *)(lds + 8 * kTgtChunk);
    Res rr[2];
    ksw_pass2<SLEN, true, SYM, U8>(a, p, hr, tc2, nullptr, hm, true, rr);
    // lane 0 of the group finishes pair A, lane 1 pair B (:846-855)
    const bool mineB = p == 1;
    if (p < 2 && (mineB ? valid[1] : valid[0])) {
        const Res fw = mineB ? f[1] : f[0], b = mineB ? rr[1] : rr[0];
        const uint32_t id = mineB ? pid[1] : pid[0];
        int32_t *o = a.out + (size_t)id * 7;
        if (b.score >= fw.score) {
            o[5] = fw.te - b.te;
            o[6] = fw.qe - b.qe;
        }
        if (a.cells) a.cells[id] += (uint32_t)b.rows * (uint32_t)(fw.qe + 1);
    }
}

template <bool SYM, bool U8>
hipError_t mate16_rev_launch_set(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                                 hipStream_t st) {
    if (count == 0) return hipSuccess;
    const dim3 grid((count + 15) / 16);
    if constexpr (U8) {
        switch (slen) {
#define FA_M8(N)                                                                                 \
    case N:                                                                                      \
        hipLaunchKernelGGL((mate16_rev_kernel<2 * N, SYM, true>), grid, dim3(64),                \
                           (size_t)8 * kTgtChunk + (size_t)256 * 2 * N, st, a, first, count); \
        return hipGetLastError();
            FA_M8(1) FA_M8(2) FA_M8(3) FA_M8(4) FA_M8(5) FA_M8(6) FA_M8(7) FA_M8(8)
            FA_M8(9) FA_M8(10) FA_M8(11) FA_M8(12) FA_M8(13) FA_M8(14) FA_M8(15) FA_M8(16)
#undef FA_M8
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (slen) {
#define FA_M16(N)                                                                            \
    case N:                                                                                  \
        hipLaunchKernelGGL((mate16_rev_kernel<N, SYM, false>), grid, dim3(64),               \
                           (size_t)8 * kTgtChunk + (size_t)256 * N, st, a, first, count); \
        return hipGetLastError();
            FA_M16(1) FA_M16(2) FA_M16(3) FA_M16(4) FA_M16(5) FA_M16(6) FA_M16(7) FA_M16(8)
            FA_M16(9) FA_M16(10) FA_M16(11) FA_M16(12) FA_M16(13) FA_M16(14) FA_M16(15) FA_M16(16)
            FA_M16(17) FA_M16(18) FA_M16(19) FA_M16(20) FA_M16(21) FA_M16(22) FA_M16(23) FA_M16(24)
            FA_M16(25) FA_M16(26) FA_M16(27) FA_M16(28) FA_M16(29) FA_M16(30) FA_M16(31) FA_M16(32)
#undef FA_M16
            default: return hipErrorInvalidValue;
        }
    }
}

// Launch over every stripe length 1..kMateMaxSlen (8-bit: 1..kMate8MaxSlen) for
// one gap symmetry and mode (each instantiation set lives in its own
// translation unit so they build in parallel).  The 8-bit kernel of stripe
// length s runs 2 s registers per lane.
template <bool SYM, bool U8>
hipError_t mate16_launch_set(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st) {
    if (count == 0) return hipSuccess;
    const dim3 grid((count + 15) / 16);
    if constexpr (U8) {
        switch (slen) {
#define FA_M8(N)                                                                                 \
    case N:                                                                                      \
        hipLaunchKernelGGL((mate16_kernel<2 * N, SYM, true>), grid, dim3(64),                    \
                           (size_t)8 * kTgtChunk + (size_t)256 * 2 * N, st, a, first, count); \
        return hipGetLastError();
            FA_M8(1) FA_M8(2) FA_M8(3) FA_M8(4) FA_M8(5) FA_M8(6) FA_M8(7) FA_M8(8)
            FA_M8(9) FA_M8(10) FA_M8(11) FA_M8(12) FA_M8(13) FA_M8(14) FA_M8(15) FA_M8(16)
#undef FA_M8
            default: return hipErrorInvalidValue;
        }
    } else {
        switch (slen) {
#define FA_M16(N)                                                                            \
    case N:                                                                                  \
        hipLaunchKernelGGL((mate16_kernel<N, SYM, false>), grid, dim3(64),                   \
                           (size_t)8 * kTgtChunk + (size_t)256 * N, st, a, first, count); \
        return hipGetLastError();
            FA_M16(1) FA_M16(2) FA_M16(3) FA_M16(4) FA_M16(5) FA_M16(6) FA_M16(7) FA_M16(8)
            FA_M16(9) FA_M16(10) FA_M16(11) FA_M16(12) FA_M16(13) FA_M16(14) FA_M16(15) FA_M16(16)
            FA_M16(17) FA_M16(18) FA_M16(19) FA_M16(20) FA_M16(21) FA_M16(22) FA_M16(23) FA_M16(24)
            FA_M16(25) FA_M16(26) FA_M16(27) FA_M16(28) FA_M16(29) FA_M16(30) FA_M16(31) FA_M16(32)
#undef FA_M16
            default: return hipErrorInvalidValue;
        }
    }
}

}  // namespace fa
