// Device probes for the roofline.  VALU: the integer add/max issue rate one
// device sustains,
// in the two forms the extension kernels are made of (packed 16-bit v_pk_*
// and 32-bit), to give their roofline a measured ceiling beside the
// published one.  Eight independent register chains per lane, adds and maxes
// alternating so that consecutive instructions never depend on each other.
// HBM: a plain vector copy of a known number of bytes, the measured bandwidth
// and the reference point that scales the HBM byte counters.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "valu_probe.h"

namespace fa {

namespace {

template <bool PACKED>
__global__ __launch_bounds__(256) void valu_probe_kernel(uint32_t *out, int iters, uint32_t c) {
    uint32_t x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = threadIdx.x * 8u + (uint32_t)k;
    for (int it = 0; it < iters; ++it) {
#pragma unroll
        for (int r = 0; r < kValuProbeUnroll; ++r) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if constexpr (PACKED)
                    asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                else
                    asm volatile("v_add_u32 %0, %0, %1" : "+v"(x[k]) : "v"(c));
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if constexpr (PACKED)
                    asm volatile("v_pk_max_i16 %0, %0, %1" : "+v"(x[k]) : "v"(c));
                else
                    asm volatile("v_max_i32 %0, %0, %1" : "+v"(x[k]) : "v"(c));
            }
        }
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) s ^= x[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

// 16-byte vector copy, grid-stride: reads and writes n16 * 16 bytes
__global__ __launch_bounds__(256) void hbm_copy_kernel(const uint4 *__restrict__ src,
                                                       uint4 *__restrict__ dst, size_t n16) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

}  // namespace

hipError_t hbm_copy_launch(const void *src, void *dst, size_t bytes, int blocks, hipStream_t st) {
    hipLaunchKernelGGL(hbm_copy_kernel, dim3(blocks), dim3(256), 0, st,
                       static_cast<const uint4 *>(src), static_cast<uint4 *>(dst), bytes / 16);
    return hipGetLastError();
}

hipError_t valu_probe_launch(bool packed16, uint32_t *out, int blocks, int iters, hipStream_t st) {
    if (packed16)
        hipLaunchKernelGGL(valu_probe_kernel<true>, dim3(blocks), dim3(256), 0, st, out, iters, 3u);
    else
        hipLaunchKernelGGL(valu_probe_kernel<false>, dim3(blocks), dim3(256), 0, st, out, iters, 3u);
    return hipGetLastError();
}

}  // namespace fa
