// Packing plan on the device: the sort keys that order a batch's pairs into
// lane groups (the host needs only a 768-bin histogram back to cut the sorted
// order into per-kernel launches).
//
// A pair's kernel follows from its 16-bit fit level f (0..2) and query
// length q, so the statistics are kept per (f, q) bin:
//   plan_hist_kernel  pairs, target rows and (cost packing) fixed-point cost
//                     sums per bin, block histograms in LDS, one atomic per
//                     bin and block;
//   plan_key_kernel   one key per pair, ascending = launch order: launch rank
//                     (the kernel's position among the kernels the batch
//                     uses, read from the histogram on the host; 32-bit keys
//                     for at most 32 ranks), then (cost packing) the early-stop flag and
//                     cost class of a pair whose cost is below half the mean
//                     of its (kernel, qlen) group, then qlen and tlen
//                     descending (cost-order packing: descending cost inside
//                     the kernels the host marks); long pairs (fit level 3) last, longest
//                     query first; the device radix sort of (key, pair id)
//                     is stable, so equal keys keep the input order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "plan_kernel.h"

namespace fa {

namespace {

constexpr int kBins = 3 * 256;

__global__ __launch_bounds__(1024) void plan_hist_kernel(PlanArgs a, PlanHist h) {
    __shared__ uint32_t cnt[kBins];
    __shared__ unsigned long long rows[kBins];
    __shared__ unsigned long long csum[kBins];
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) {
        cnt[b] = 0u;
        rows[b] = 0ull;
        csum[b] = 0ull;
    }
    __syncthreads();
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < a.n; k += gridDim.x * blockDim.x) {
        const int q = a.qlen[k];
        const int f = a.elig[k];
        if (f > 2 || q < 1 || q > 255) continue;  // long pairs: one launch of their own
        const int b = f * 256 + q;
        atomicAdd(&cnt[b], 1u);
        atomicAdd(&rows[b], (unsigned long long)max(a.tlen[k], 1));
        if (a.cost) atomicAdd(&csum[b], cost_fixed(a.cost[k]));
    }
    __syncthreads();
    for (int b = threadIdx.x; b < kBins; b += blockDim.x) {
        if (!cnt[b]) continue;
        atomicAdd(&h.cnt[b], cnt[b]);
        atomicAdd(&h.rows[b], rows[b]);
        if (a.cost) atomicAdd(&h.csum[b], csum[b]);
    }
}

template <class K>
__global__ __launch_bounds__(256) void plan_key_kernel(PlanArgs a, PlanHist h, PlanTable t, K *keys,
                                                       uint32_t *ids) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= a.n) return;
    ids[k] = k;
    const int q = a.qlen[k];
    const int f = a.elig[k];
    const uint32_t len20 = ((uint32_t)(255 - min(q, 255)) << 12) | (uint32_t)(4095 - min(a.tlen[k], 4095));
    const bool compact = a.shift == kRankShiftCompact;
    if (f > 2) {
        // long pair: after every lane-group kernel, longest query first
        keys[k] = compact ? ((K)a.long_rank << kRankShiftCompact) | ((K)(2047 - min(q >> 4, 2047)) << 8) |
                                (K)(255 - min(a.tlen[k] >> 8, 255))
                          : ((K)a.long_rank << kRankShift) | ((K)(65535 - min(q, 65535)) << 11) |
                                (K)(2047 - min(a.tlen[k] >> 5, 2047));
        return;
    }
    const int x = t.kidx[f][q];  // >= 0: the host checked the histogram
    uint32_t low = 0;
    if (a.mode == kPlanCost) {
        // the (kernel, qlen) group: every fit level whose kernel is x at this qlen
        unsigned long long gs = 0ull, gc = 0ull;
        for (int g = 0; g < 3; ++g)
            if (t.kidx[g][q] == x) {
                gs += h.csum[g * 256 + q];
                gc += h.cnt[g * 256 + q];
            }
        const float c = fmaxf(a.cost[k], 0.0f);
        low = len20;
        if ((unsigned long long)cost_fixed(c) * 2ull * gc < gs) {
            // an early stopper: after the others, by cost class (e^(1/2) apart,
            // at most 28 for 2^20 cells), highest first
            const int cls = (int)rint(2.0 * log1p((double)c));
            low |= (1u << 25) | ((uint32_t)(31 - min(cls, 31)) << 20);
        }
    } else if (a.mode == kPlanCostOrder) {
        if (t.cost_order[x]) {
            // descending cost as log2(1 + cells) in 1/50 octaves (a pair has at
            // most 256 x 2048 cells: 950 steps of 1.4%, finer than the
            // prediction's error), then qlen descending: pairs of one cost
            // step and one query length share a wavefront.  Measured on the
            // uneven mix, run only: 1.0% faster than exact cells (where equal
            // predictions are few and qlen rarely decides), with perfect costs
            // too; 1/100 to 1/12 octave all within 0.2%
            const float c = fmaxf(a.cost[k], 0.0f);
            const uint32_t lv = 2047u - (uint32_t)min((int)rintf(50.0f * __log2f(1.0f + c)), 2047);
            low = (lv << 8) | (uint32_t)(255 - q);
        } else {
            // narrow kernels: qlen, then tlen, descending
            low = ((uint32_t)(255 - min(q, 255)) << 11) | (uint32_t)(2047 - min(a.tlen[k], 2047));
        }
    } else if (a.mode != kPlanNone) {
        low = len20;
    }
    keys[k] = ((K)x << a.shift) | (K)low;
}

}  // namespace

hipError_t plan_hist(const PlanArgs &a, const PlanHist &h, int cus, hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    uint32_t blocks = (a.n + 1023) / 1024;
    const uint32_t cap = (uint32_t)(cus > 0 ? cus : 256);
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(plan_hist_kernel, dim3(blocks), dim3(1024), 0, st, a, h);
    return hipGetLastError();
}

hipError_t plan_keys(const PlanArgs &a, const PlanHist &h, const PlanTable &t, uint32_t *keys, uint32_t *ids,
                     hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(plan_key_kernel<uint32_t>, dim3((a.n + 255) / 256), dim3(256), 0, st, a, h, t, keys, ids);
    return hipGetLastError();
}

hipError_t plan_keys(const PlanArgs &a, const PlanHist &h, const PlanTable &t, unsigned long long *keys,
                     uint32_t *ids, hipStream_t st) {
    if (a.n == 0) return hipSuccess;
    hipLaunchKernelGGL(plan_key_kernel<unsigned long long>, dim3((a.n + 255) / 256), dim3(256), 0, st, a, h, t,
                       keys, ids);
    return hipGetLastError();
}

}  // namespace fa
