// 16-bit banded seed extension for gfx950: two pairs per lane group.
//
// Same exact row-parallel formulation as extend_kernel.hip (ksw_extend2,
// src/core/alignment/ksw_affine_gap.rs:952-1169), but every column register
// holds two 16-bit values: pair A in the low half, pair B in the high half,
// and the arithmetic runs on packed v_pk_* instructions, so one lane group
// of L lanes advances two pairs per instruction.  Band tests are sign bits of
// packed differences (column - bound), selects are v_bfi on those masks.
//
// 16-bit range: the host routes a pair here only when every value provably
// fits: with Hmax = h0 + qlen*max(mat), Hmax < 4096 and Hmax*K < 32768
// (K = max(2, 1 + max(mat))), e_ins < 32 and o+e < 8192, so H, E, F, G stay
// in (-32768, 16384) and the row-argmax key (H << 4) | slot fits 16 bits;
// the wide slot counts (S = 18..24) take only pairs with Hmax < 2048, whose
// key (H << 5) | slot fits; other pairs run the 32-bit kernel.
//
// Differences from the 32-bit kernel that keep results identical:
//  * M = H ? H + s : 0 becomes min(H + s, H*K): equal when H > 0, and <= 0
//    when H == 0, which every later use (max with E >= 0, max with 0) treats
//    like 0;
//  * the F prefix identity is -16384 (F(beg) stays below E >= 0 without
//    wrapping); the cross-lane scan runs on values biased by +16384 as
//    unsigned, so its identity is 0;
//  * E' = max(E -sat e_del, M - oe_del): E >= 0, so the unsigned saturating
//    subtract equals max(E - e_del, 0).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "extend_kernel.h"

namespace fa {

namespace {

typedef short s2 __attribute__((ext_vector_type(2)));
typedef unsigned short u2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ uint32_t bits(s2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t bits(u2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ s2 vs(uint32_t v) { return __builtin_bit_cast(s2, v); }
__device__ __forceinline__ u2 vu(uint32_t v) { return __builtin_bit_cast(u2, v); }

__device__ __forceinline__ uint32_t pk(int lo, int hi) {
    return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}
__device__ __forceinline__ int lo16(uint32_t x) { return (int)(short)(x & 0xFFFFu); }
__device__ __forceinline__ int hi16(uint32_t x) { return (int)(short)(x >> 16); }
__device__ __forceinline__ uint32_t smax(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_max(vs(a), vs(b)));
}
__device__ __forceinline__ uint32_t smin(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_min(vs(a), vs(b)));
}
__device__ __forceinline__ uint32_t umax(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_max(vu(a), vu(b)));
}
__device__ __forceinline__ uint32_t umin(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_min(vu(a), vu(b)));
}
__device__ __forceinline__ uint32_t add16(uint32_t a, uint32_t b) { return bits(vu(a) + vu(b)); }
__device__ __forceinline__ uint32_t sub16(uint32_t a, uint32_t b) { return bits(vu(a) - vu(b)); }
__device__ __forceinline__ uint32_t mul16(uint32_t a, uint32_t b) { return bits(vu(a) * vu(b)); }
__device__ __forceinline__ uint32_t subsat(uint32_t a, uint32_t b) {
    return bits(__builtin_elementwise_sub_sat(vu(a), vu(b)));
}
// Sign of each 16-bit half spread over the half (0xFFFF if negative).  Kept
// as one VALU instruction: written in C, the compiler rewrites the
// mask-and-select idiom into per-half compares and selects.
__device__ __forceinline__ uint32_t sgnmask(uint32_t d) {
    uint32_t r;
    asm("v_pk_ashrrev_i16 %0, 15, %1 op_sel_hi:[0,1]" : "=v"(r) : "v"(d));
    return r;
}
// min(x, 1) per half: 1 where the half is non-zero
__device__ __forceinline__ uint32_t nz01(uint32_t x) {
    uint32_t r;
    asm("v_pk_min_u16 %0, %1, 1 op_sel_hi:[1,0]" : "=v"(r) : "v"(x));
    return r;
}
// a * 2 + b per half
__device__ __forceinline__ uint32_t dbl_add(uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_pk_mad_u16 %0, %1, 2, %2 op_sel_hi:[1,0,1]" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
// acc = max(acc, a * MUL + b) per half (unsigned), MUL an inline constant, b
// wave-uniform.  One statement, so the running maximum is taken slot by slot:
// with the multiply-add and the maximum apart, the compiler sinks the
// maximum to the end of the row and keeps every slot's term in a register.
// The packed maximum reads the multiply-add's result on the next
// instruction, which needs one wait state (s_nop 0) on gfx950.
template <int MUL>
__device__ __forceinline__ void key_acc(uint32_t &acc, uint32_t a, uint32_t b) {
    uint32_t t;
    asm("v_pk_mad_u16 %1, %2, %3, %4 op_sel_hi:[1,0,1]\n\ts_nop 0\n\tv_pk_max_u16 %0, %0, %1"
        : "+v"(acc), "=&v"(t)
        : "v"(a), "n"(MUL), "s"(b));
}
// a * m + b per half, m wave-uniform
__device__ __forceinline__ uint32_t mad_s(uint32_t a, uint32_t m, uint32_t b) {
    uint32_t r;
    asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(a), "s"(m), "v"(b));
    return r;
}
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b) {
    return (m & a) | (~m & b);
}
// bfi kept as one v_bfi_b32 where the compiler would split it
__device__ __forceinline__ uint32_t bfi1(uint32_t m, uint32_t a, uint32_t b) {
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}

constexpr int kQuadXor1 = 0xB1, kQuadXor2 = 0x4E, kRowHalfMirror = 0x141, kRowMirror = 0x140;
constexpr int kRowShr = 0x110;
// row forms of the packed kernel (see the row loop of extend16_rows)
constexpr int kFull = 0, kFast = 1;
constexpr int kRowNewBcast15 = 0x15F;  // every lane of a 16-lane row reads its lane 15
constexpr int kQuadBcast3 = 0xFF;      // quad_perm [3,3,3,3]: every lane of a quad reads its lane 3

// DPP move with bound_ctrl: a lane without a source lane (row_shr at the
// start of a 16-lane row) reads 0, and no old value is needed, so the
// compiler neither copies the source first nor keeps the move apart from a
// 32-bit consumer (v_max_i32_dpp)
template <int CTRL>
__device__ __forceinline__ uint32_t dmov(uint32_t x) {
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, CTRL, 0xF, 0xF, true);
}
template <int L>
__device__ __forceinline__ int allmax_i(int x) {
    x = max(x, (int)dmov<kQuadXor1>((uint32_t)x));
    if constexpr (L >= 4) x = max(x, (int)dmov<kQuadXor2>((uint32_t)x));
    if constexpr (L >= 8) x = max(x, (int)dmov<kRowHalfMirror>((uint32_t)x));
    if constexpr (L >= 16) x = max(x, (int)dmov<kRowMirror>((uint32_t)x));
    return x;
}
template <int L>
__device__ __forceinline__ uint32_t allmax_u2(uint32_t x) {
    x = umax(x, dmov<kQuadXor1>(x));
    if constexpr (L >= 4) x = umax(x, dmov<kQuadXor2>(x));
    if constexpr (L >= 8) x = umax(x, dmov<kRowHalfMirror>(x));
    if constexpr (L >= 16) x = umax(x, dmov<kRowMirror>(x));
    return x;
}
// y from lane l-K of the group for a max-scan step, 0 (the identity of the
// unsigned scan) where the group has no such lane
template <int L, int K>
__device__ __forceinline__ uint32_t shr0(uint32_t y, int l) {
    uint32_t z = dmov<kRowShr + K>(y);
    if constexpr (L < 16) z = (l >= K) ? z : 0u;
    return z;
}
// exclusive packed unsigned max-scan over the L lanes of a group (lane 0: 0)
template <int L>
__device__ __forceinline__ uint32_t excl_scan_u2(uint32_t x, int l) {
    uint32_t y = shr0<L, 1>(x, l);
    if constexpr (L > 1) y = umax(y, shr0<L, 1>(y, l));
    if constexpr (L > 2) y = umax(y, shr0<L, 2>(y, l));
    if constexpr (L > 4) y = umax(y, shr0<L, 4>(y, l));
    if constexpr (L > 8) y = umax(y, shr0<L, 8>(y, l));
    return y;
}

// nfree: the sequence has no N base, its N-mask word is not read
__device__ __forceinline__ int decode_base(const uint32_t *seq2, const uint32_t *seqn,
                                           uint32_t off, int pos, bool nfree) {
    const uint32_t w2 = seq2[off + (pos >> 4)];
    const uint32_t wn = nfree ? 0u : seqn[(off >> 1) + (pos >> 5)];
    const int code = (int)((w2 >> ((pos & 15) * 2)) & 3u);
    return ((wn >> (pos & 31)) & 1u) ? 4 : code;
}

// Stage the target codes of a lane group's two pairs: byte i = code of row i
// of pair A in bits 0-3 and of pair B in bits 4-7 (0 past a pair's target).
// Lane l of the group decodes 16-base words l, l+L, ...  A target without N
// (tfree) reads no N-mask words.
__device__ __forceinline__ void stage_targets2(const ExtArgs &a, uint8_t *tc2, const int (&tlen)[2],
                                               const uint32_t (&to)[2], const bool (&tfree)[2], int l,
                                               int L) {
    const int nw0 = (tlen[0] + 15) >> 4, nw1 = (tlen[1] + 15) >> 4;
    const int nw = max(nw0, nw1);
    for (int k = l; k < nw; k += L) {
        uint32_t w2[2], wn[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool in = k < (h ? nw1 : nw0);
            w2[h] = in ? a.seq2[to[h] + k] : 0u;
            wn[h] = in && !tfree[h] ? (a.seqn[(to[h] >> 1) + (k >> 1)] >> ((k & 1) * 16)) & 0xFFFFu : 0u;
        }
        uint32_t v[4];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            uint32_t x = 0;
#pragma unroll
            for (int y = 0; y < 4; ++y) {
                const int p = b * 4 + y;
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    uint32_t c = (w2[h] >> (p * 2)) & 3u;
                    c = ((wn[h] >> p) & 1u) ? 4u : c;
                    x |= c << (y * 8 + h * 4);
                }
            }
            v[b] = x;
        }
        uint32_t *dst = reinterpret_cast<uint32_t *>(tc2 + k * 16);
        dst[0] = v[0];
        dst[1] = v[1];
        dst[2] = v[2];
        dst[3] = v[3];
    }
}

// Per-pair state of the reference's row loop, both pairs of a lane group
// packed like the cells (pair A in the low, pair B in the high 16 bits) and
// replicated in every lane of the group.  Every field fits a signed 16-bit
// half: scores < 4096, rows < 2048, columns <= 256, w clamped to 4095 (any
// w >= 2304 gives the same band, since i < 2048 and qlen <= 256).
struct PairPack {
    uint32_t alive;                     // 0xFFFF per pair still running
    uint32_t qlen, tlen, w, w1, h0;     // w1 = w + 1
    uint32_t beg, end, best, max_i, max_j, max_ie, gscore, max_off;
    uint32_t nacc;                      // cells of the rows since the last flush, both pairs
    uint32_t ncell[2];                  // cells per pair (32-bit)
};

// A row adds at most 255 cells per pair, so nacc takes 256 rows without
// wrapping; it is folded into the 32-bit counts every 256 rows and at the end.
__device__ __forceinline__ void flush_cells(PairPack &P) {
    P.ncell[0] += P.nacc & 0xFFFFu;
    P.ncell[1] += P.nacc >> 16;
    P.nacc = 0u;
}

// 0xFFFF per half where a < b (signed, |a - b| < 32768)
__device__ __forceinline__ uint32_t lt16(uint32_t a, uint32_t b) { return sgnmask(sub16(a, b)); }

// End of row i for both pairs (:1110-1158), branch-free on packed halves.
// ip = i in both halves; rbp/rep the row's band; kA/kB the pairs' row-argmax
// keys (m_row << 10) | (1023 - mj); hend H(i, qlen-1) per half; fenc/lastp1
// the first (as 1023 - c) and last (as c + 1) non-zero column, 0 when none.
__device__ __forceinline__ void pair_row_end2(const ExtArgs &a, PairPack &P, uint32_t ip, uint32_t rbp,
                                              uint32_t rep, int kA, int kB, uint32_t hend,
                                              uint32_t fenc, uint32_t lastp1, uint32_t capdelp,
                                              uint32_t capinsp) {
    const uint32_t alive = P.alive;
    P.nacc = add16(P.nacc, smax(sub16(rep, rbp), 0u) & alive);
    // :1110-1120 (re <= qlen, so re == qlen is !(re < qlen))
    const uint32_t g = alive & ~lt16(rep, P.qlen) & lt16(P.gscore, hend);
    P.gscore = bfi1(g, hend, P.gscore);
    P.max_ie = bfi1(g, ip, P.max_ie);
    const uint32_t mrow = (uint32_t)(kA >> 10) | ((uint32_t)(kB >> 10) << 16);
    const uint32_t mj = sub16(0x03FF03FFu, ((uint32_t)kA & 1023u) | (((uint32_t)kB & 1023u) << 16));
    const uint32_t nz = lt16(0u, mrow);        // :1122 m_row != 0
    const uint32_t better = lt16(P.best, mrow);  // :1126 (best >= h0 > 0, so m_row != 0)
    // :1134-1145 z-drop: gap = (di - dj) * e_del if di > dj else (dj - di) * e_ins,
    // with |di - dj| clamped to cap = 32767 / e (a clamped gap exceeds any
    // best - m_row - zdrop, so the test is unchanged and the product fits).
    // The gap is never negative, so a pair can break only when
    // best - m_row > zdrop; the gap is computed only in rows where some pair
    // of the wavefront is there.
    // zdrop >= 4096 never breaks (best < 4096); clamped at 8191 it stays in range
    const int zd = min(a.zdrop, 8191);
    const uint32_t slack = sub16(sub16(P.best, mrow), pk(zd, zd));  // best - m_row - zdrop
    const uint32_t cand = zd > 0 ? alive & nz & ~better & lt16(0u, slack) : 0u;
    uint32_t zbrk = 0u;
    if (__builtin_amdgcn_ballot_w64(cand != 0u) != 0) {
        const uint32_t dd = sub16(sub16(ip, P.max_i), sub16(mj, P.max_j));
        const uint32_t pos = lt16(0u, dd);
        const uint32_t ad = smax(dd, sub16(0u, dd));
        const uint32_t gap = mul16(umin(ad, bfi1(pos, capdelp, capinsp)),
                                   bfi1(pos, pk(a.e_del, a.e_del), pk(a.e_ins, a.e_ins)));
        zbrk = cand & lt16(gap, slack);
    }
    const uint32_t upd = alive & better;
    P.best = bfi1(upd, mrow, P.best);
    P.max_i = bfi1(upd, ip, P.max_i);
    P.max_j = bfi1(upd, mj, P.max_j);
    const uint32_t od = sub16(mj, ip);
    P.max_off = bfi1(upd, smax(P.max_off, smax(od, sub16(0u, od))), P.max_off);
    // :1148-1158
    const uint32_t nb = bfi1(lt16(0u, fenc), sub16(0x03FF03FFu, fenc), rep);
    const uint32_t ne = smax(nb, sub16(lastp1, 0x00010001u));
    const uint32_t band = alive & nz & ~zbrk;
    P.beg = bfi1(band, nb, P.beg);
    P.end = bfi1(band, smin(add16(ne, 0x00020002u), P.qlen), P.end);
    P.alive = band & lt16(add16(ip, 0x00010001u), P.tlen);
}

}  // namespace

// One wavefront: 64/L lane groups, each running pairs order[slot0 + 2g] (low
// halves) and order[slot0 + 2g + 1] (high halves).
//
// Column layout: each pair is right-aligned in its group, column c sits in
// lane (c + pad) / S, slot (c + pad) % S with pad = L*S - 1 - qlen, so column
// qlen (whose eh.h after a row ending at qlen is H(i, qlen-1), the gscore
// candidate) is always the group's last slot.  Positions with c < 0 and
// columns left of the band are dead: the band start rb never decreases
// (rb = max(beg, i - w) with beg non-decreasing), so those registers may hold
// anything; every reader masks them (G in pass 1, the argmax, the non-zero
// scan) or overwrites them before use (eh.h at c = rb is h1).
// Slot count a wavefront moves to once every running pair's live columns
// [beg, qlen] fit it: one or two slots fewer, none below 3 (the per-row
// bookkeeping dominates there).  A restripe (two LDS instructions per slot
// and array) costs a few slot-rows of DP (about 30 instructions each), so
// the smallest step pays once the band stays inside the smaller stripe for
// a few rows; measured on uneven_mix_4m it beats steps of 2/3 and 3/4.
constexpr int next_slots16(int S) {
    return S > 12 ? S - 2 : S > 3 ? S - 1 : 0;
}

// Wavefront state carried from one slot count to the next.
struct WaveCtx {
    int i;          // next row
    uint32_t ip;    // i in both halves
    int l;          // lane in the group
    int ql[2];      // qlen per half
    uint8_t *tc2;   // the group's staged target codes
    uint32_t *xfer; // the wavefront's restripe buffer in LDS (64 x S words)
};

// Column position p = c + pad holds in lane p / S, slot p % S; with S' slots
// the pad shrinks by L (S - S') for both halves, so position p' of the new
// layout is position p' + L (S - S') of the old one.  The positions that fall
// off hold columns below every running pair's beg.
template <int L, int S, int NS>
__device__ __forceinline__ void restripe(const WaveCtx &w, const uint32_t (&src)[S], uint32_t (&dst)[NS]) {
    const int g = (int)(threadIdx.x & 63) / L;
    uint32_t *buf = w.xfer + g * (L * S);
#pragma unroll
    for (int s = 0; s < S; ++s) buf[w.l * S + s] = src[s];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    constexpr int D = L * (S - NS);
#pragma unroll
    for (int s = 0; s < NS; ++s) dst[s] = buf[w.l * NS + s + D];
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
}

// LDS bytes of a wavefront's restripe buffer at S slots.
constexpr size_t xfer_bytes16(int S) { return next_slots16(S) > 0 ? (size_t)256 * S : 0; }

// The row loop at S slots per lane, from row w.i until every pair of the
// wavefront has stopped (false) or, when next_slots16(S) > 0, until the next
// row's live columns fit that many slots (true, w at the next row).
template <int L, int S>
__device__ __forceinline__ bool extend16_rows(const ExtArgs &a, PairPack &P, WaveCtx &w,
                                              const uint32_t (&profA)[S], const uint32_t (&profB)[S],
                                              uint32_t (&Hh)[S], uint32_t (&Ee)[S]) {
    constexpr int NS = next_slots16(S);
    // one register per slot fewer for one instruction per slot more where that
    // decides the occupancy: S = 18 then fits 3 waves per SIMD instead of 2
    // (167 VGPRs; measured 3.5% faster); S = 20..24 stay at 2 either way
    constexpr bool GONLY = S == 18;
    // row-argmax key per half: (H << KB) | (KM - s); with S > 16 the host
    // routes only pairs whose scores stay below 2048 here, so the key fits
    constexpr int KB = S <= 8 ? 3 : (S <= 16 ? 4 : 5);
    constexpr uint32_t KM = (1u << KB) - 1u;
    const int l = w.l;
    const uint8_t *const tc2 = w.tc2;
    int cbh[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) cbh[h] = l * S - (L * S - 1 - w.ql[h]);  // column of slot 0 in this lane
    const int oe_ins = a.o_ins + a.e_ins, oe_del = a.o_del + a.e_del;
    const uint32_t einsp = pk(a.e_ins, a.e_ins), edelp = pk(a.e_del, a.e_del);
    const uint32_t oedelp = pk(oe_del, oe_del);
    const uint32_t kp = pk(a.mguard, a.mguard);
    const uint32_t cbp = pk(cbh[0], cbh[1]);
    // G(c) = max(M - oe_ins, 0) + c*e_ins is split into the lane's cb*e_ins
    // and a lane-local part max(M + (s*e_ins - oe_ins), s*e_ins) whose
    // constants are wave-uniform.
    const uint32_t cbe = mul16(cbp, einsp);  // cb * e_ins per half
    const uint32_t cbeb = add16(cbe, 0x40004000u);
    uint32_t se[S], seo[S];                  // s*e_ins, s*e_ins - oe_ins (wave-uniform)
#pragma unroll
    for (int s = 0; s < S; ++s) {
        se[s] = pk(s * a.e_ins, s * a.e_ins);
        seo[s] = pk(s * a.e_ins - oe_ins, s * a.e_ins - oe_ins);
    }
    const uint32_t se_m1 = pk(-a.e_ins, -a.e_ins);
    // identity of masked G: any value <= -e_ins leaves F(c >= rb) unchanged
    // (F <= 0 there while E >= 0); kept in a scalar register so the select
    // is one v_bfi_b32
    uint32_t negs;
    asm("s_mov_b32 %0, 0xf000f000" : "=s"(negs));

    // z-drop gap clamps 32767 / e per direction (e = 0: the gap is 0 anyway)
    const int capdel = a.e_del > 0 ? 32767 / a.e_del : 32767;
    const int capins = a.e_ins > 0 ? 32767 / a.e_ins : 32767;
    const uint32_t capdelp = pk(capdel, capdel), capinsp = pk(capins, capins);

    // per-lane constants of the unmasked row forms: column qlen (the gscore
    // column, never in a row's [rb, re)) is the last slot of the last lane for
    // both halves, and slot 0 of lane 0 is always dead there (c <= 0 < rb)
    const uint32_t notlast = l == L - 1 ? 0u : 0xFFFFFFFFu;
    const uint32_t notlane0 = l == 0 ? 0u : 0xFFFFFFFFu;

    // Rows of one form until the next row needs the other form (0), every
    // pair has stopped (1), or the next row's live columns fit NS slots (2).
    // One loop per form keeps the column arrays in their registers from row
    // to row (a loop body that joins both forms makes the compiler copy them
    // at every row).
    auto rows_in = [&](auto kind) -> int {
        constexpr bool FULLROWS = decltype(kind)::value == kFull;
        for (;; ++w.i, w.ip += 0x00010001u) {
            const int i = w.i;
            const uint32_t ip = w.ip;
            if (__builtin_amdgcn_ballot_w64(P.alive != 0u) == 0) return 1;
            // :1038-1046; a finished pair gets the empty band [4000, -3)
            const uint32_t rbp = bfi1(P.alive, smax(P.beg, sub16(ip, P.w)), 0x0FA00FA0u);
            const uint32_t rep = bfi1(P.alive, smin(smin(P.end, add16(ip, P.w1)), P.qlen), 0xFFFDFFFDu);
            // Row form, the same for the whole wavefront.  Every column left of the
            // band holds h = e = 0 at the start of a row: the band start moves past a
            // column by the beg rule only when its h and e are 0 (:1148-1152), a
            // column the i - w bound drops is zeroed before its row, and a masked
            // row writes 0 into every column it leaves behind.  A zero column
            // computed as if it were in the band stays 0 (M <= 0, E = 0, F <= 0 from
            // zero columns on its left) and adds nothing to the F prefix, the row
            // maximum or the non-zero scan.  So when every live pair has rb > 0 (no
            // first-column value h1 to place) and re == qlen (no stale columns on
            // the right to keep), no band mask is needed (kFast), after zeroing the
            // columns that the i - w bound drops this row; other rows keep every
            // mask (kFull).
            const uint32_t live = P.alive;
            const bool any_short = __builtin_amdgcn_ballot_w64((live & lt16(rep, P.qlen)) != 0u) != 0;
            const bool any_rb0 = __builtin_amdgcn_ballot_w64((live & lt16(rbp, 0x00010001u)) != 0u) != 0;
            const bool any_wk = __builtin_amdgcn_ballot_w64((live & lt16(P.beg, rbp)) != 0u) != 0;
            if ((any_short || any_rb0) != FULLROWS) return 0;
            // both pairs' target codes of this row (a row past a pair's target reads
            // its code 0, and a finished pair's half is never used)
            const uint32_t tw = tc2[i];
            // score byte mat[t][q] + 128 from each pair's profile word; an N target
            // (uniform N row) selects the byte 0x00 (selector 0x0C) and its half's
            // bias becomes 128 - (mat[20] + 128)
            const uint32_t nb = ((tw & 4u) >> 2) | ((tw & 0x40u) << 10);
            const uint32_t sel = (tw & 3u) | ((tw & 0x30u) << 12) | 0x0C040C00u | (nb * 0x0Cu);
            const uint32_t biasp = 0x00800080u - nb * (a.nrow & 0xFFu);  // (24-bit multiply)

            // Band tests are lane-relative: with t = bound - cb, column cb+s is
            // below the bound iff s - t < 0, and c <= bound at slot s is c < bound
            // at slot s-1.
            const uint32_t trb = sub16(rbp, cbp), tre = sub16(rep, cbp);

            uint32_t pkey = 0, acc = 0, accH = 0;
            auto row = [&](auto form) {
                constexpr int MODE = decltype(form)::value;
                constexpr bool FULL = MODE == kFull;   // band masks and h1
                // :1049-1054 h1 = max(h0 - (o_del + e_del*(i+1)), 0) when rb == 0 (h0 < 4096,
                // so clamping the subtrahend at 32767 keeps the difference in range)
                uint32_t h1ip = 0u;
                if constexpr (FULL) {
                    const int dsub = min(a.o_del + a.e_del * (i + 1), 32767);
                    h1ip = smax(sub16(P.h0, pk(dsub, dsub)), 0u) & lt16(rbp, 0x00010001u);
                }
                // pass 1: M, the band-start masks and the lane-local exclusive prefix of G.
                // With GONLY only G is kept per slot: pass 2 takes M back as G - seo
                // (a masked column's value is wrong, but every use of it is masked too)
                // and rebuilds the prefix from the carry, one more instruction per slot
                // for one register fewer.
                uint32_t Mv[GONLY ? 1 : S], pre[GONLY ? 1 : S], Gv[GONLY ? S : 1], mrb[S];
                uint32_t run = negs;
    #pragma unroll
                for (int s = 0; s < S; ++s) {
                    const uint32_t sb = __builtin_amdgcn_perm(profB[s], profA[s], sel);
                    const uint32_t M = smin(sub16(add16(Hh[s], sb), biasp), mul16(Hh[s], kp));
                    if constexpr (!GONLY) Mv[s] = M;
                    // G without the clamp of M - oe_ins at 0: that clamp only lifts F
                    // to >= 0, and H = max(M, E, F) with E >= 0 does not see it
                    const uint32_t G0 = add16(M, seo[s]);
                    uint32_t G = G0;
                    if constexpr (FULL) {
                        mrb[s] = sgnmask(sub16((uint32_t)s * 0x00010001u, trb));  // c < rb
                        G = bfi(mrb[s], negs, G0);
                    }
                    if constexpr (GONLY) {
                        Gv[s] = G;
                    } else {
                        pre[s] = run;
                    }
                    run = smax(run, G);
                }
                // carry relative to this lane's cb*e_ins; the scanned values run + cb*e_ins
                // lie in (-16384, 16384), so with 0x4000 added they scan as unsigned with
                // identity 0 (the empty prefix of lane 0 comes back as -16384 - cb*e_ins)
                const uint32_t carry = sub16(excl_scan_u2<L>(add16(run, cbeb), l), cbeb);

                // pass 2: H, E, row-max key, write-back, non-zero columns
                // acc: one bit per slot and half; with S > 16 the first S - 16 slots
                // go to accH (the high bits of a 32-bit mask per half)
                uint32_t Hprev = 0;
                uint32_t mle_rb0 = 0, mle_re0 = 0, wr0 = 0, mlt_rb0 = 0;
                uint32_t mlt_rb_prev = 0, mlt_re_prev = 0;
                if constexpr (FULL) {
                    mlt_rb_prev = sgnmask(sub16(0xFFFFFFFFu, trb));  // column cb-1
                    mlt_re_prev = sgnmask(sub16(0xFFFFFFFFu, tre));
                }
                // GONLY: max(carry, G of the slots left of s); it differs from
                // max(carry, pre[s]) only below the identity, where F < 0 <= E
                uint32_t runp = carry;
    #pragma unroll
                for (int s = 0; s < S; ++s) {
                    const uint32_t sp = (uint32_t)s * 0x00010001u;
                    uint32_t F, Ms;
                    if constexpr (GONLY) {
                        F = sub16(runp, s ? se[s - 1] : se_m1);  // - (c-1)*e_ins
                        runp = smax(runp, Gv[s]);
                        Ms = sub16(Gv[s], seo[s]);
                    } else {
                        F = sub16(smax(carry, pre[s]), s ? se[s - 1] : se_m1);
                        Ms = Mv[s];
                    }
                    const uint32_t me = smax(Ms, Ee[s]);
                    uint32_t x, H;
                    if constexpr (FULL) {
                        H = smax(me, F);
                        const uint32_t En = smax(subsat(Ee[s], edelp), sub16(Ms, oedelp));
                        const uint32_t mlt_rb = mrb[s], mle_rb = mlt_rb_prev;
                        const uint32_t mlt_re = sgnmask(sub16(sp, tre)), mle_re = mlt_re_prev;
                        mlt_rb_prev = mlt_rb;
                        mlt_re_prev = mlt_re;
                        const uint32_t Hm = H & mlt_re & ~mlt_rb;  // H inside [rb, re), else 0
                        key_acc<(1 << KB)>(pkey, Hm, (KM - (uint32_t)s) * 0x00010001u);
                        // E' in band, 0 at c = re (:1106-1107) and left of the band
                        Ee[s] = bfi1(mle_re, En & mlt_re & ~mlt_rb, Ee[s]);
                        const uint32_t wr = mle_re & ~mlt_rb;      // rb <= c <= re
                        if (s == 0) {
                            mle_rb0 = mle_rb;
                            mle_re0 = mle_re;
                            wr0 = wr;
                            mlt_rb0 = mlt_rb;
                            x = 0u;
                        } else {
                            // eh[c].h <- H(i, c-1), h1 at c = rb, 0 left of the band
                            const uint32_t hv = bfi1(mle_rb, h1ip, Hprev) & ~mlt_rb;
                            Hh[s] = bfi1(mle_re, hv, Hh[s]);
                            x = (Hh[s] | Ee[s]) & wr;
                        }
                    } else {
                        // eh[c+1].h <- H(i, c) (the next slot's old h was read in pass 1 only)
                        H = smax(me, F);
                        if (s < S - 1) Hh[s + 1] = H;
                        Ee[s] = smax(subsat(Ee[s], edelp), sub16(Ms, oedelp));
                        // re == qlen: every column but qlen is below re
                        uint32_t Hm = H;
                        if (s == S - 1) {
                            Hm = H & notlast;
                            Ee[s] = Ee[s] & notlast;
                        }
                        key_acc<(1 << KB)>(pkey, Hm, (KM - (uint32_t)s) * 0x00010001u);
                        // H(i, c-1) at s > 0; c <= rb takes h1 = 0 (rb > 0), which a
                        // zero column on the left already gives
                        x = s == 0 ? 0u : (Hh[s] | Ee[s]);
                    }
                    if (s > 0) {
                        if (S > 16 && s < S - 16)
                            accH = dbl_add(accH, nz01(x));
                        else
                            acc = dbl_add(acc, nz01(x));
                    }
                    Hprev = H;
                }
                {  // slot 0: its H(i, c-1) comes from the last slot of lane l-1
                    const uint32_t hl = dmov<kRowShr + 1>(Hprev);
                    uint32_t x0;
                    if constexpr (FULL) {
                        const uint32_t hv = bfi1(mle_rb0, h1ip, hl) & ~mlt_rb0;
                        Hh[0] = bfi1(mle_re0, hv, Hh[0]);
                        x0 = (Hh[0] | Ee[0]) & wr0;
                    } else {
                        // lane 0's slot 0 is dead; c <= rb takes 0
                        Hh[0] = hl & notlane0;
                        x0 = Hh[0] | Ee[0];
                    }
                    if constexpr (S > 16)
                        accH = mad_s(nz01(x0), (1u << (S - 17)) * 0x00010001u, accH);
                    else
                        acc = mad_s(nz01(x0), (1u << (S - 1)) * 0x00010001u, acc);
                }
            };
            if constexpr (FULLROWS) {
                row(kind);
            } else {
                if (any_wk) {
                    // columns [beg, rb) leave the band by the i - w bound: zero them
                    // (every column left of rb is below the band from now on)
    #pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const uint32_t m = sgnmask(sub16((uint32_t)s * 0x00010001u, trb));  // c < rb
                        Hh[s] = bfi1(m, 0u, Hh[s]);
                        Ee[s] = bfi1(m, 0u, Ee[s]);
                    }
                }
                row(kind);
            }

            // acc: bit S-1-s of each half set where column cb+s in [rb, re] has
            // eh.h or eh.e non-zero after the row.  First such column as 1023 - c,
            // last as c + 1, 0 when none (:1148-1158).
            uint32_t fenc[2], lenc[2];
            int kk[2];
    #pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint32_t bitsh = h ? (acc >> 16) : (acc & 0xFFFFu);
                if constexpr (S > 16) bitsh |= (h ? (accH >> 16) : (accH & 0xFFFFu)) << 16;
                const int clz = __builtin_clz(bitsh | 1u);
                const int ctz = __builtin_ctz(bitsh | (S > 16 ? 0x80000000u : 0x10000u));
                fenc[h] = bitsh ? (uint32_t)(1023 - (cbh[h] + S - 32 + clz)) : 0u;
                lenc[h] = bitsh ? (uint32_t)(cbh[h] + S - ctz) : 0u;
                const int key = (int)(h ? (pkey >> 16) : (pkey & 0xFFFFu));
                const int c = max(cbh[h] + (int)KM - (key & (int)KM), 0);
                kk[h] = ((key >> KB) << 10) | (1023 - c);
            }
            uint32_t ffirst = fenc[0] | (fenc[1] << 16);
            uint32_t flast = lenc[0] | (lenc[1] << 16);
            // column qlen (the group's last slot): H(i, qlen-1) when re == qlen
            uint32_t hend;
            if constexpr (L == 16) {
                hend = dmov<kRowNewBcast15>(Hh[S - 1]);  // lane 15 of the 16-lane row
            } else if constexpr (L == 8) {
                // lane 3 of each quad, then the other quad's through the half-row mirror
                const uint32_t x = dmov<kQuadBcast3>(Hh[S - 1]);
                const uint32_t y = dmov<kRowHalfMirror>(x);  // every lane, before the select
                hend = l < 4 ? y : x;
            } else {
                hend = dmov<kQuadBcast3>(Hh[S - 1]);  // lane 3 of the 4-lane group
            }
            // flush the packed cell counts before they could wrap
            if ((i & 255) == 255) flush_cells(P);
            const int kA = allmax_i<L>(kk[0]);
            const int kB = allmax_i<L>(kk[1]);
            ffirst = allmax_u2<L>(ffirst);
            flast = allmax_u2<L>(flast);
            pair_row_end2(a, P, ip, rbp, rep, kA, kB, hend, ffirst, flast, capdelp, capinsp);
            if constexpr (NS > 0) {
                // the next row's live columns [beg, qlen] of every running pair fit
                // NS slots per lane: move to that layout and go on there
                const uint32_t need = sub16(add16(P.qlen, 0x00010001u), P.beg);
                const uint32_t big = P.alive & ~lt16(need, (uint32_t)(L * NS + 1) * 0x00010001u);
                if (__builtin_amdgcn_ballot_w64(big != 0u) == 0) {
                    ++w.i;
                    w.ip += 0x00010001u;
                    return 2;
                }
            }
        }
    };
    for (;;) {
        int r = rows_in(std::integral_constant<int, kFast>());
        if (r == 0) r = rows_in(std::integral_constant<int, kFull>());
        if (r != 0) return r == 2;
    }
}

// Rows at S slots, then at the next slot counts as the band shrinks.
template <int L, int S>
__device__ __forceinline__ void extend16_phase(const ExtArgs &a, PairPack &P, WaveCtx &w,
                                               const uint32_t (&profA)[S], const uint32_t (&profB)[S],
                                               uint32_t (&Hh)[S], uint32_t (&Ee)[S]) {
    const bool shrink = extend16_rows<L, S>(a, P, w, profA, profB, Hh, Ee);
    constexpr int NS = next_slots16(S);
    if constexpr (NS > 0) {
        if (shrink) {
            uint32_t nA[NS], nB[NS], nH[NS], nE[NS];
            restripe<L, S, NS>(w, profA, nA);
            restripe<L, S, NS>(w, profB, nB);
            restripe<L, S, NS>(w, Hh, nH);
            restripe<L, S, NS>(w, Ee, nE);
            extend16_phase<L, NS>(a, P, w, nA, nB, nH, nE);
        }
    }
}

template <int L, int S>
__device__ __forceinline__ void extend16_wave(const ExtArgs &a, uint32_t slot0, uint32_t nslots,
                                              uint8_t *lds_wave) {
    // lds_wave: tpad bytes of staged target codes per lane group, then the
    // restripe buffer (64 x S words)
    const int lane = (int)(threadIdx.x & 63);
    const int g = lane / L;
    const int l = lane % L;
    PairPack P;
    bool valid[2];
    uint32_t pid[2], qo[2], to[2];
    int tl[2], ql[2], h0v[2];
    uint8_t *const tc2 = lds_wave + g * a.tpad;
    int cbh[2];
    int wv[2];
    bool qfree[2], tfree[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const uint32_t slot = slot0 + 2u * (uint32_t)g + (uint32_t)h;
        valid[h] = slot < nslots;
        uint4 m0 = {0u, 0u, 0u, 1u}, m1 = {0u, 1u, 0u, 0u};  // empty slot: qlen 1, tlen 0, h0 1
        if (valid[h]) {
            const uint4 *mp = reinterpret_cast<const uint4 *>(a.meta) + 2 * (size_t)slot;
            m0 = mp[0];
            m1 = mp[1];
        }
        pid[h] = m0.x;
        ql[h] = (int)m0.w;
        tl[h] = (int)m1.x;
        h0v[h] = (int)m1.y;
        wv[h] = min((int)m1.z, 4095);
        qo[h] = m0.y;
        to[h] = m0.z;
        qfree[h] = (m1.w & 1u) != 0u;
        tfree[h] = (m1.w & 2u) != 0u;
        cbh[h] = l * S - (L * S - 1 - ql[h]);  // column of slot 0 in this lane
    }
    P.alive = (valid[0] && tl[0] > 0 ? 0xFFFFu : 0u) | (valid[1] && tl[1] > 0 ? 0xFFFF0000u : 0u);
    P.qlen = pk(ql[0], ql[1]);
    P.tlen = pk(tl[0], tl[1]);
    P.w = pk(wv[0], wv[1]);
    P.w1 = add16(P.w, 0x00010001u);
    P.h0 = pk(h0v[0], h0v[1]);
    P.beg = 0u;
    P.end = P.qlen;
    P.best = P.h0;
    P.max_i = P.max_j = P.max_ie = P.gscore = 0xFFFFFFFFu;  // -1
    P.max_off = 0u;
    P.ncell[0] = P.ncell[1] = 0u;
    P.nacc = 0u;
    stage_targets2(a, tc2, tl, to, tfree, l, L);

    const int oe_ins = a.o_ins + a.e_ins;

    uint32_t profA[S], profB[S], Hh[S], Ee[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        int hv[2];
        uint32_t pw[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int c = cbh[h] + s;
            int code = 4;
            if (valid[h] && c >= 0 && c < ql[h]) code = decode_base(a.seq2, a.seqn, qo[h], c, qfree[h]);
            pw[h] = a.profw[code];
            const int h0 = h0v[h];
            const int v1 = h0 > oe_ins ? h0 - oe_ins : 0;
            const int x = c == 0 ? h0 : max(v1 - (c - 1) * a.e_ins, 0);  // :992-998
            hv[h] = c >= 0 ? x : 0;
        }
        profA[s] = pw[0];
        profB[s] = pw[1];
        Hh[s] = pk(hv[0], hv[1]);
        Ee[s] = 0u;
    }
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    WaveCtx w;
    w.i = 0;
    w.ip = 0u;
    w.l = l;
    w.ql[0] = ql[0];
    w.ql[1] = ql[1];
    w.tc2 = tc2;
    w.xfer = reinterpret_cast<uint32_t *>(lds_wave + (64 / L) * a.tpad);
    extend16_phase<L, S>(a, P, w, profA, profB, Hh, Ee);
    flush_cells(P);
    if (l == 0) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            if (!valid[h]) continue;
            const int sh = 16 * h;
            store_result(a, pid[h], (int)(short)(P.best >> sh), (int)(short)(P.max_i >> sh) + 1,
                         (int)(short)(P.max_ie >> sh) + 1, (int)(short)(P.max_j >> sh) + 1,
                         (int)(short)(P.gscore >> sh), (int)(short)(P.max_off >> sh), P.ncell[h]);
        }
    }
}

__global__ __launch_bounds__(256) void meta_gather_kernel(ExtArgs a, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t p = a.order[k];
    uint4 *mp = reinterpret_cast<uint4 *>(a.meta) + 2 * (size_t)k;
    if (a.params) {
        // the pair's 32-byte record: one line instead of six scattered words
        const uint4 *src = reinterpret_cast<const uint4 *>(a.params) + 2 * (size_t)p;
        mp[0] = src[0];
        mp[1] = src[1];
        return;
    }
    mp[0] = make_uint4(p, a.qo[p], a.to[p], (uint32_t)a.qlen[p]);
    mp[1] = make_uint4((uint32_t)a.tlen[p], (uint32_t)a.h0[p], (uint32_t)a.w[p], 0u);
}

// 1 when none of the first len bases of the sequence at word offset off is N.
__device__ __forceinline__ uint32_t n_free(const uint32_t *seqn, uint32_t off, int len) {
    uint32_t any = 0u;
    const int nw = (len + 31) >> 5;
    for (int j = 0; j < nw; ++j) {
        uint32_t m = seqn[(off >> 1) + j];
        if (j == nw - 1 && (len & 31)) m &= (1u << (len & 31)) - 1u;
        any |= m;
    }
    return any ? 0u : 1u;
}

__global__ __launch_bounds__(256) void params_pack_kernel(ExtArgs a, uint32_t n, uint32_t *params) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const uint32_t qo = a.qo[k], to = a.to[k];
    const int qlen = a.qlen[k], tlen = a.tlen[k];
    const uint32_t nfree = n_free(a.seqn, qo, qlen) | n_free(a.seqn, to, tlen) << 1;
    uint4 *mp = reinterpret_cast<uint4 *>(params) + 2 * (size_t)k;
    mp[0] = make_uint4(k, qo, to, (uint32_t)qlen);
    mp[1] = make_uint4((uint32_t)tlen, (uint32_t)a.h0[k], (uint32_t)a.w[k], nfree);
}

__global__ __launch_bounds__(256) void records_split_kernel(const int4 *__restrict__ rec, uint32_t n,
                                                            int2 *__restrict__ out,
                                                            uint32_t *__restrict__ cells) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int4 r0 = rec[2 * (size_t)k], r1 = rec[2 * (size_t)k + 1];
    out[3 * (size_t)k] = make_int2(r0.x, r0.y);
    out[3 * (size_t)k + 1] = make_int2(r0.z, r0.w);
    out[3 * (size_t)k + 2] = make_int2(r1.x, r1.y);
    cells[k] = (uint32_t)r1.z;
}

hipError_t meta_gather(const ExtArgs &a, uint32_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(meta_gather_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a, n);
    return hipGetLastError();
}

hipError_t params_pack(const ExtArgs &a, uint32_t n, uint32_t *params, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(params_pack_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a, n, params);
    return hipGetLastError();
}

hipError_t records_split(const int32_t *rec, uint32_t n, int32_t *out, uint32_t *cells, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(records_split_kernel, dim3((n + 255) / 256), dim3(256), 0, st,
                       reinterpret_cast<const int4 *>(rec), n, reinterpret_cast<int2 *>(out), cells);
    return hipGetLastError();
}

template <int L, int S>
__global__ __launch_bounds__(256) void extend16_plain_kernel(ExtArgs a, uint32_t first,
                                                             uint32_t count) {
    extern __shared__ uint8_t lds[];
    constexpr int G = 64 / L;
    const int wave = (int)(threadIdx.x >> 6);
    const uint32_t slot0 = (blockIdx.x * (blockDim.x >> 6) + (uint32_t)wave) * (2u * G);
    uint8_t *lds_wave = lds + wave * ((size_t)G * a.tpad + xfer_bytes16(S));
    if (slot0 >= count) return;
    extend16_wave<L, S>(a, first + slot0, first + count, lds_wave);
}

template <int L, int S>
__global__ __launch_bounds__(256) void extend16_persistent_kernel(ExtArgs a, uint32_t first,
                                                                  uint32_t count, uint32_t qidx) {
    extern __shared__ uint8_t lds[];
    constexpr int G = 64 / L;
    const int wave = (int)(threadIdx.x >> 6);
    uint8_t *lds_wave = lds + wave * ((size_t)G * a.tpad + xfer_bytes16(S));
    const uint32_t ntasks = (count + 2 * G - 1) / (2 * G);
    const uint32_t chunk = (uint32_t)a.chunk;
    uint32_t *counter = a.queue + qidx;
    for (;;) {
        uint32_t t0 = 0;
        if ((threadIdx.x & 63) == 0) t0 = atomicAdd(counter, chunk);
        t0 = __builtin_amdgcn_readfirstlane(t0);
        if (t0 >= ntasks) break;
        const uint32_t t1 = t0 + chunk < ntasks ? t0 + chunk : ntasks;
        for (uint32_t t = t0; t < t1; ++t)
            extend16_wave<L, S>(a, first + t * 2 * G, first + count, lds_wave);
    }
}

template <int L, int S>
static hipError_t launch16(const ExtArgs &a, uint32_t first, uint32_t count, bool persistent,
                           uint32_t qidx, int cus, int blocks_per_cu, hipStream_t st) {
    constexpr int G = 64 / L;
    if (count == 0) return hipSuccess;
    const size_t lds_wave = (size_t)G * a.tpad + xfer_bytes16(S);
    const size_t lds = (size_t)4 * lds_wave;  // persistent: 4 wavefronts per workgroup
    if (lds_wave > 65536) return hipErrorInvalidValue;
    // four wavefronts' staged targets (targets near 2048 bases) may not fit a
    // persistent workgroup's LDS: the plain form runs the same work then
    if (lds > 65536) persistent = false;
    const uint32_t per_block = 4u * 2u * G;
    const uint32_t need = (count + per_block - 1) / per_block;
    if (!persistent) {
        // one wavefront per workgroup: a workgroup's slot on a CU is reused
        // only when all of its wavefronts are done, so single-wave groups let
        // short wavefronts hand their slot on at once
        const uint32_t pw = 2u * G;
        hipLaunchKernelGGL((extend16_plain_kernel<L, S>), dim3((count + pw - 1) / pw), dim3(64),
                           lds_wave, st, a, first, count);
        return hipGetLastError();
    }
    int occ = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &occ, reinterpret_cast<const void *>(&extend16_persistent_kernel<L, S>), 256, lds);
    if (e != hipSuccess) return e;
    if (occ < 1) occ = 1;
    if (blocks_per_cu > 0 && blocks_per_cu < occ) occ = blocks_per_cu;
    uint32_t blocks = (uint32_t)occ * (uint32_t)cus;
    if (blocks > need) blocks = need;
    hipLaunchKernelGGL((extend16_persistent_kernel<L, S>), dim3(blocks), dim3(256), lds, st, a,
                       first, count, qidx);
    return hipGetLastError();
}

#define FA16_LIST(X, L) \
    X(L, 1) X(L, 2) X(L, 3) X(L, 4) X(L, 6) X(L, 8) X(L, 10) X(L, 12) X(L, 14) X(L, 16)
// wide slot counts (4 and 8 lanes): twice the pairs per wavefront for the same
// query length, so the per-row bookkeeping is shared by twice as many pairs
#define FA16_WIDE(X, L) X(L, 18) X(L, 20) X(L, 22) X(L, 24)

int extend16_slot_count(int L, int qlen, bool wide) {
    static const int sl[] = {1, 2, 3, 4, 6, 8, 10, 12, 14, 16, 18, 20, 22, 24};
    if (L != 4 && L != 8 && L != 16) return -1;
    const int n = wide && L <= 8 ? 14 : 10;
    for (int k = 0; k < n; ++k)
        if (L * sl[k] >= qlen + 1) return sl[k];
    return -1;
}

hipError_t extend16_launch(int L, int S, const ExtArgs &a, uint32_t first, uint32_t count,
                           bool persistent, uint32_t qidx, int cus, int blocks_per_cu,
                           hipStream_t st) {
#define FA_CASE(LL, SS)                                                                       \
    if (L == LL && S == SS)                                                                   \
        return launch16<LL, SS>(a, first, count, persistent, qidx, cus, blocks_per_cu, st);
    FA16_LIST(FA_CASE, 16)
    FA16_LIST(FA_CASE, 8)
    FA16_LIST(FA_CASE, 4)
    FA16_WIDE(FA_CASE, 8)
    FA16_WIDE(FA_CASE, 4)
#undef FA_CASE
    return hipErrorInvalidValue;
}

}  // namespace fa
