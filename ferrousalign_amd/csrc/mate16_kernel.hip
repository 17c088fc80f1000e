// Packed mate-rescue kernels for general gap penalties, every stripe length,
// and the launcher that picks the symmetric set when it applies.
#include "mate16_impl.h"

namespace fa {

hipError_t mate16_launch_sym(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st);
hipError_t mate8_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                        hipStream_t st);

hipError_t mate16_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                         hipStream_t st) {
    if (a.byte_mode) return mate8_launch(a, first, count, slen, st);
    if (slen < 1 || slen > kMateMaxSlen) return hipErrorInvalidValue;
    if (a.o_del == a.o_ins && a.e_del == a.e_ins) return mate16_launch_sym(a, first, count, slen, st);
    return mate16_launch_set<false, false>(a, first, count, slen, st);
}

}  // namespace fa
