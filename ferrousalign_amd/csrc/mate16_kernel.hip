// Packed mate-rescue kernels for general gap penalties, every stripe length,
// the launchers that pick the symmetric set when it applies, and the sort keys
// of the separate XSTART reverse launches.
#include "mate16_impl.h"

namespace fa {

hipError_t mate16_launch_sym(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st);
hipError_t mate8_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                        hipStream_t st);

hipError_t mate16_rev_launch_sym(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                                 hipStream_t st);
hipError_t mate8_rev_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                            hipStream_t st);

hipError_t mate16_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                         hipStream_t st) {
    if (a.byte_mode) return mate8_launch(a, first, count, slen, st);
    if (slen < 1 || slen > kMateMaxSlen) return hipErrorInvalidValue;
    if (a.o_del == a.o_ins && a.e_del == a.e_ins) return mate16_launch_sym(a, first, count, slen, st);
    return mate16_launch_set<false, false>(a, first, count, slen, st);
}

hipError_t mate16_rev_launch(const MateArgs &a, uint32_t first, uint32_t count, int slen,
                             hipStream_t st) {
    if (a.byte_mode) return mate8_rev_launch(a, first, count, slen, st);
    if (slen < 1 || slen > kMateMaxSlen) return hipErrorInvalidValue;
    if (a.o_del == a.o_ins && a.e_del == a.e_ins)
        return mate16_rev_launch_sym(a, first, count, slen, st);
    return mate16_rev_launch_set<false, false>(a, first, count, slen, st);
}

namespace {

__global__ __launch_bounds__(256) void mate_rev_keys_kernel(MateArgs a, uint32_t n, uint32_t *keys,
                                                            uint32_t *ids, uint32_t *hist) {
    __shared__ uint32_t h[64];
    if (threadIdx.x < 64) h[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k < n) {
        const int32_t *o = a.out + (size_t)k * 7;
        const int score = o[0], te = o[1], qe = o[2];
        const int lanes = a.byte_mode ? 16 : 8;
        uint32_t bin = 63u;
        if (xstart_needed(a.xtra, score, te, qe)) bin = (uint32_t)(kMateMaxSlen - (qe + lanes) / lanes);
        keys[k] = bin << 15 | (uint32_t)(32767 - min(max(score, 0), 32767));
        ids[k] = k;
        atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    if (threadIdx.x < 64 && h[threadIdx.x] != 0u) atomicAdd(&hist[threadIdx.x], h[threadIdx.x]);
}

}  // namespace

hipError_t mate_rev_keys(const MateArgs &a, uint32_t n, uint32_t *keys, uint32_t *ids,
                         uint32_t *hist, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mate_rev_keys_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a, n, keys, ids,
                       hist);
    return hipGetLastError();
}

}  // namespace fa
