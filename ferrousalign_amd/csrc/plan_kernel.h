// Device-side packing plan of a batch (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

constexpr int kPlanLength = 0;  // (kernel, qlen desc, tlen desc)
constexpr int kPlanNone = 1;    // kernel only: input order inside a kernel's range
constexpr int kPlanCost = 2;    // length order with early stoppers apart, by cost class
constexpr int kPlanCostOrder = 3;  // descending cost inside the kernels marked in PlanTable::cost_order

struct PlanArgs {
    uint32_t n;
    int mode;
    const int32_t *qlen, *tlen;
    const uint8_t *elig;  // 16-bit fit level per pair (0..2), 3: long pair
    const float *cost;    // per-pair packing cost (kPlanCost, kPlanCostOrder), else null
    int long_rank;        // launch rank of the long pairs (after every lane-group kernel)
    int shift;            // bits of the order inside a launch (kRankShift, or kRankShiftCompact
                          // for kPlanCostOrder)
};

// Per (fit level, qlen) bin, 3 x 256 entries each; zero before plan_hist.
struct PlanHist {
    uint32_t *cnt;
    unsigned long long *rows;  // sum of max(tlen, 1)
    unsigned long long *csum;  // sum of cost_fixed(cost)
};

// Launch rank of each (fit level, qlen): the position of the pair's kernel
// among the kernels the batch uses, -1 where no kernel holds it.
struct PlanTable {
    int16_t kidx[3][256];
    uint8_t cost_order[64];  // kPlanCostOrder: rank x orders its pairs by descending cost
};

// Sort keys: launch rank << kRankShift | a 27-bit order inside the launch.
// kPlanCostOrder keeps 19 bits (the cost on a log scale in 1/50 octaves,
// then qlen), so that with up to 32 ranks the keys sort in three 8-bit radix
// passes instead of four.
constexpr int kRankShift = 27;
constexpr int kRankShiftCompact = 19;

// Packing cost in 1/16 cells, non-negative and saturating at 2^24 cells
// (a pair has at most 256 x 2048), so a sum over 2^31 pairs cannot wrap.
__host__ __device__ inline unsigned long long cost_fixed(float c) {
    c = c > 0.0f ? c : 0.0f;
    c = c < 16777216.0f ? c : 16777216.0f;
    return (unsigned long long)(c * 16.0f);
}

hipError_t plan_hist(const PlanArgs &a, const PlanHist &h, int cus, hipStream_t st);
// keys[k] and ids[k] = k for the stable sort (32-bit keys when the batch
// uses at most 32 launch ranks, else 64-bit).
hipError_t plan_keys(const PlanArgs &a, const PlanHist &h, const PlanTable &t, uint32_t *keys,
                     uint32_t *ids, hipStream_t st);
hipError_t plan_keys(const PlanArgs &a, const PlanHist &h, const PlanTable &t, unsigned long long *keys,
                     uint32_t *ids, hipStream_t st);

}  // namespace fa
