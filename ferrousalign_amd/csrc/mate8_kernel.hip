// Packed mate-rescue kernels of the 8-bit mode (KSW_XBYTE): 16-lane groups,
// every stripe length 1..16, both gap symmetries.
#include "mate16_impl.h"

namespace fa {

hipError_t mate8_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                        hipStream_t st) {
    if (slen < 1 || slen > kMate8MaxSlen) return hipErrorInvalidValue;
    if (a.o_del == a.o_ins && a.e_del == a.e_ins)
        return mate16_launch_set<true, true>(a, first, count, slen, st);
    return mate16_launch_set<false, true>(a, first, count, slen, st);
}

hipError_t mate8_rev_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                            hipStream_t st) {
    if (slen < 1 || slen > kMate8MaxSlen) return hipErrorInvalidValue;
    if (a.o_del == a.o_ins && a.e_del == a.e_ins)
        return mate16_rev_launch_set<true, true>(a, first, count, slen, st);
    return mate16_rev_launch_set<false, true>(a, first, count, slen, st);
}

}  // namespace fa
