// Kernel launchers of the learned work predictor (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "extend_kernel.h"

namespace fa {

constexpr int kMlpParams = 1313;  // 6-32-32-1
constexpr int kMlpFeatures = 6;
constexpr int kMlpBlock = 64;     // pairs per backward block

hipError_t mlp_launch_features(const ExtArgs &a, uint32_t n, float *feat, hipStream_t st);
// Predicted cells of pairs [0, n): from the raw features in feat
// (from_feat), or computed from the batch's sequences in a (and written to
// feat when it is not null).
hipError_t mlp_launch_predict(const ExtArgs &a, const void *wh, const float *norm, float *feat,
                              bool from_feat, uint32_t n, float *cells_pred, hipStream_t st);
hipError_t mlp_launch_backward(const void *wh, const float *norm, const float *feat,
                               const uint32_t *cells, const uint32_t *perm, uint32_t first,
                               uint32_t count, float *partial, hipStream_t st);
hipError_t mlp_launch_reduce_adam(const float *partial, uint32_t nblocks, float *grad_out,
                                  float *loss_out, float *master, float *m, float *v, void *wh,
                                  float lr, float b1, float b2, float eps, float bc1, float bc2,
                                  int apply, hipStream_t st);
hipError_t mlp_launch_to_half(const float *src, void *dst, int n, hipStream_t st);

}  // namespace fa
