// Host-side base-code checks and 2-bit packing, eight codes per 64-bit word
// (plain host C++, built with the host compiler).
//
// Code check: a code is valid when it is at most 4; per byte,
// ((x & 0x7F) + 0x7B) sets bit 7 for 5..127 without a carry into the next
// byte, and x itself sets it for 128..255.
// Packing (pack layout of the extension kernels: base p at bits 2*(p%16) of
// word p/16, N mask bit p%32 of word p/32, N packed as code 0): with BMI2,
// pext of the low two bits of eight code bytes gives 16 packed bits and pext
// of bit 2 the eight N bits (codes are 0..4, so bit 2 is set only for N).
#include <immintrin.h>

#include <cstdint>
#include <cstring>

#include "pack_host.h"

namespace {

constexpr uint64_t kLow7 = 0x7F7F7F7F7F7F7F7Full;
constexpr uint64_t kAdd = 0x7B7B7B7B7B7B7B7Bull;
constexpr uint64_t kHigh = 0x8080808080808080ull;

inline uint64_t load8(const uint8_t *p) {
    uint64_t x;
    std::memcpy(&x, p, 8);
    return x;
}

void pack_scalar(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn) {
    const uint32_t nw = fa_host::seq_words(len);
    for (uint32_t k = 0; k < nw; ++k) w2[k] = 0;
    for (uint32_t k = 0; k < nw / 2; ++k) wn[k] = 0;
    for (int32_t p = 0; p < len; ++p) {
        uint32_t c = s[p];
        if (c == 4) {
            wn[p >> 5] |= 1u << (p & 31);
            c = 0;
        }
        w2[p >> 4] |= (c & 3u) << ((p & 15) * 2);
    }
}

template <bool CHECK>
__attribute__((target("bmi2"))) bool pack_bmi2(const uint8_t *s, int32_t len, uint32_t *w2,
                                               uint32_t *wn) {
    const int32_t full = len & ~31;
    uint8_t tail[32];
    uint64_t bad = 0;
    for (int32_t p0 = 0; p0 < len; p0 += 32) {
        const uint8_t *src = s + p0;
        if (p0 >= full) {  // last, partial block: zero padding (code 0, not N)
            std::memset(tail, 0, sizeof tail);
            std::memcpy(tail, src, (size_t)(len - p0));
            src = tail;
        }
        uint32_t n = 0;
        uint32_t lo = 0, hi = 0;
        for (int g = 0; g < 4; ++g) {
            const uint64_t x = load8(src + 8 * g);
            if constexpr (CHECK) bad |= (((x & kLow7) + kAdd) | x) & kHigh;
            const uint32_t c = (uint32_t)_pext_u64(x, 0x0303030303030303ull);
            n |= (uint32_t)_pext_u64(x, 0x0404040404040404ull) << (8 * g);
            if (g < 2)
                lo |= c << (16 * g);
            else
                hi |= c << (16 * (g - 2));
        }
        w2[p0 >> 4] = lo;
        w2[(p0 >> 4) + 1] = hi;
        wn[p0 >> 5] = n;
    }
    return bad == 0;
}

bool have_bmi2() {
    static const bool v = [] {
        __builtin_cpu_init();
        return __builtin_cpu_supports("bmi2") != 0;
    }();
    return v;
}

}  // namespace

namespace fa_host {

bool codes_ok(const uint8_t *p, int64_t len) {
    int64_t k = 0;
    uint64_t bad = 0;
    for (; k + 8 <= len; k += 8) {
        const uint64_t x = load8(p + k);
        bad |= (((x & kLow7) + kAdd) | x) & kHigh;
    }
    if (bad) return false;
    for (; k < len; ++k)
        if (p[k] > 4) return false;
    return true;
}

void pack_codes(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn) {
    if (have_bmi2())
        pack_bmi2<false>(s, len, w2, wn);
    else
        pack_scalar(s, len, w2, wn);
}

bool pack_codes_checked(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn) {
    if (have_bmi2()) return pack_bmi2<true>(s, len, w2, wn);
    if (!codes_ok(s, len)) return false;
    pack_scalar(s, len, w2, wn);
    return true;
}

}  // namespace fa_host
