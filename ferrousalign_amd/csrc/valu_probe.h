// Device probes (internal; reported through fa_valu_peak and fa_hbm_copy).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

// instructions per lane per iteration: kValuProbeUnroll * 8 chains * 2 (add, max)
constexpr int kValuProbeUnroll = 16;
// out: blocks * 256 words
hipError_t valu_probe_launch(bool packed16, uint32_t *out, int blocks, int iters, hipStream_t st);
// bytes: multiple of 16
hipError_t hbm_copy_launch(const void *src, void *dst, size_t bytes, int blocks, hipStream_t st);

}  // namespace fa
