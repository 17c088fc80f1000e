// Packed mate-rescue kernels for symmetric gap penalties (o_del == o_ins,
// e_del == e_ins), every stripe length.
#include "mate16_impl.h"

namespace fa {

hipError_t mate16_launch_sym(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st) {
    return mate16_launch_set<true, false>(a, first, count, slen, st);
}

hipError_t mate16_rev_launch_sym(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                                 hipStream_t st) {
    return mate16_rev_launch_set<true, false>(a, first, count, slen, st);
}

}  // namespace fa
