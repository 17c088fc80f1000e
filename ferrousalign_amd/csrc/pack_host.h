// Host-side base-code checks and 2-bit packing (pack_host.cc, plain C++).
#pragma once
#include <cstdint>

namespace fa_host {

// 32-bit words of a packed sequence: 16 bases per word, every sequence
// starting on an even word.
inline uint32_t seq_words(int32_t len) { return (uint32_t)((len + 31) / 32) * 2u; }

// Every code of p[0 .. len) is at most 4.
bool codes_ok(const uint8_t *p, int64_t len);

// Pack len codes (0..4, checked by the caller) into seq_words(len) words of
// w2 and seq_words(len) / 2 words of N mask wn.
void pack_codes(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn);
// The same, checking the codes in the same pass: false when one exceeds 4
// (the words then hold garbage).
bool pack_codes_checked(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn);

}  // namespace fa_host
