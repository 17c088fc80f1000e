// Learned work predictor: a 6-32-32-1 MLP that predicts log(1 + DP cells) of
// an extension pair from six features, used to pack pairs of similar cost
// into the lane groups of one wavefront.  Hand-written gfx950 kernels for the
// features, the forward pass, the backward pass (per-block partial
// gradients) and the gradient reduction + Adam step.  FP16 storage of
// weights and activations, FP32 accumulation.  The predictor has no
// counterpart in the reference; it never changes any alignment result.
//
// Parameter vector (1,313 floats): W1[32][6], b1[32], W2[32][32], b2[32],
// W3[32], b3.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "extend_kernel.h"
#include "predictor.h"

namespace fa {

namespace {

constexpr int NIN = 6, NH = 32;
constexpr int OFF_B1 = NH * NIN, OFF_W2 = OFF_B1 + NH, OFF_B2 = OFF_W2 + NH * NH,
              OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH;
static_assert(OFF_B3 + 1 == kMlpParams, "parameter layout");

__device__ __forceinline__ float r16(float v) { return __half2float(__float2half(v)); }

// Forward of one pair with weights `w` (fp16, LDS).  Fills x, h1, h2 (fp16
// values held in float) and the pre-activations' signs.
__device__ __forceinline__ float mlp_forward1(const __half *w, const float *raw, const float *norm,
                                              float *x, float *h1, float *h2, uint32_t &pos1,
                                              uint32_t &pos2) {
#pragma unroll
    for (int k = 0; k < NIN; ++k) x[k] = r16((raw[k] - norm[k]) * norm[NIN + k]);
    pos1 = 0;
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float a = __half2float(w[OFF_B1 + i]);
#pragma unroll
        for (int k = 0; k < NIN; ++k) a += __half2float(w[i * NIN + k]) * x[k];
        pos1 |= (a > 0.f ? 1u : 0u) << i;
        h1[i] = r16(a > 0.f ? a : 0.f);
    }
    pos2 = 0;
#pragma unroll
    for (int i = 0; i < NH; ++i) {
        float a = __half2float(w[OFF_B2 + i]);
#pragma unroll
        for (int k = 0; k < NH; ++k) a += __half2float(w[OFF_W2 + i * NH + k]) * h1[k];
        pos2 |= (a > 0.f ? 1u : 0u) << i;
        h2[i] = r16(a > 0.f ? a : 0.f);
    }
    float y = __half2float(w[OFF_B3]);
#pragma unroll
    for (int k = 0; k < NH; ++k) y += __half2float(w[OFF_W3 + k]) * h2[k];
    return y;
}

}  // namespace

// Even bits 0, 2, .., 30 of x moved to bits 0 .. 15.
__device__ __forceinline__ uint32_t even_bits(uint32_t x) {
    x &= 0x55555555u;
    x = (x | (x >> 1)) & 0x33333333u;
    x = (x | (x >> 2)) & 0x0F0F0F0Fu;
    x = (x | (x >> 4)) & 0x00FF00FFu;
    return (x | (x >> 8)) & 0x0000FFFFu;
}

// 32 bases of a sequence from base d (0..3) on: 2-bit words w0, w1, w2 (bases
// 0-15, 16-31, 32-47) as 64 bits, N-mask words m0, m1 (bases 0-31, 32-63)
// as 32 bits.
__device__ __forceinline__ uint64_t window2(uint32_t w0, uint32_t w1, uint32_t w2, int d) {
    const uint64_t x = ((uint64_t)w1 << 32) | w0;
    return d ? (x >> (2 * d)) | ((uint64_t)w2 << (64 - 2 * d)) : x;
}
__device__ __forceinline__ uint32_t window_n(uint32_t m0, uint32_t m1, int d) {
    return d ? (m0 >> d) | (m1 << (32 - d)) : m0;
}

// Raw features per pair from the batch's 2-bit sequences in HBM, word-wide:
// qlen, tlen, h0, w, mismatches among the first lim = min(32, qlen, tlen)
// bases of the best diagonal within three of the main one (query window from
// base max(0, -d), target window from base max(0, d), taken when both lie
// inside their sequences; an indel near the start leaves few mismatches on a
// neighbouring diagonal), query GC fraction.  Base j mismatches when exactly
// one of the two is N, or neither is and the 2-bit codes differ.
__device__ __forceinline__ void pair_features(const ExtArgs &a, uint32_t k, float *o) {
    int qlen, tlen, h0, w;
    uint32_t qo, to, nfree = 0u;
    if (a.params) {  // the pair's 32-byte record: one load, with its N-free bits
        const uint4 *r = reinterpret_cast<const uint4 *>(a.params) + 2 * (size_t)k;
        const uint4 r0 = r[0], r1 = r[1];
        qo = r0.y;
        to = r0.z;
        qlen = (int)r0.w;
        tlen = (int)r1.x;
        h0 = (int)r1.y;
        w = (int)r1.z;
        nfree = r1.w;
    } else {
        qlen = a.qlen[k];
        tlen = a.tlen[k];
        h0 = a.h0[k];
        w = a.w[k];
        qo = a.qo[k];
        to = a.to[k];
    }
    int lim = qlen < tlen ? qlen : tlen;
    if (lim > 32) lim = 32;
    int mm = 0;
    if (lim > 0) {  // both sequences have at least one 32-base block (two words)
        const uint2 q = *reinterpret_cast<const uint2 *>(a.seq2 + qo);
        const uint2 t = *reinterpret_cast<const uint2 *>(a.seq2 + to);
        // the next block only where the sequence has one (shifted windows reach
        // at most base 34; below a sequence's end the missing bits are masked)
        const uint32_t q2 = qlen > 32 ? a.seq2[qo + 2] : 0u, t2 = tlen > 32 ? a.seq2[to + 2] : 0u;
        // N-mask words only of a sequence that has an N
        const bool qn = !(nfree & 1u), tn = !(nfree & 2u);
        const uint32_t qn0 = qn ? a.seqn[qo >> 1] : 0u, tn0 = tn ? a.seqn[to >> 1] : 0u;
        const uint32_t qn1 = qn && qlen > 32 ? a.seqn[(qo >> 1) + 1] : 0u;
        const uint32_t tn1 = tn && tlen > 32 ? a.seqn[(to >> 1) + 1] : 0u;
        const uint32_t lm = lim == 32 ? 0xFFFFFFFFu : ((1u << lim) - 1u);
        mm = 32;
#pragma unroll
        for (int d = -3; d <= 3; ++d) {
            const int qs = d < 0 ? -d : 0, ts = d > 0 ? d : 0;
            if (qs + lim > qlen || ts + lim > tlen) continue;
            const uint64_t x = window2(q.x, q.y, q2, qs) ^ window2(t.x, t.y, t2, ts);
            const uint32_t lo = (uint32_t)x, hi = (uint32_t)(x >> 32);
            const uint32_t code = even_bits(lo | (lo >> 1)) | (even_bits(hi | (hi >> 1)) << 16);
            const uint32_t qn = window_n(qn0, qn1, qs), tn = window_n(tn0, tn1, ts);
            const uint32_t mism = (qn ^ tn) | (~(qn | tn) & code);
            mm = min(mm, __popc(mism & lm));
        }
    }
    // C (01) and G (10) are the 2-bit codes whose two bits differ; N is 00
    int gc = 0;
    const int nw = (qlen + 15) >> 4;
    for (int x = 0; x < nw; ++x) {
        const uint32_t w2 = a.seq2[qo + x];
        gc += __popc((w2 ^ (w2 >> 1)) & 0x55555555u);
    }
    o[0] = (float)qlen;
    o[1] = (float)tlen;
    o[2] = (float)h0;
    o[3] = (float)w;
    o[4] = (float)mm;
    o[5] = qlen > 0 ? (float)gc / (float)qlen : 0.f;
}

__global__ __launch_bounds__(256) void mlp_features_kernel(ExtArgs a, uint32_t n, float *feat) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    pair_features(a, k, feat + (size_t)k * NIN);
}

typedef _Float16 half4v __attribute__((ext_vector_type(4)));
typedef float float16v __attribute__((ext_vector_type(16)));

// Row of accumulator r (0..15) of lane l in the 32 x 32 output of
// v_mfma_f32_32x32x8_f16; its column is l % 32.  The A operand (32 x 8) of
// lane l is row l % 32, columns 4 (l / 32) .. +3; the B operand (8 x 32) is
// rows 4 (l / 32) .. +3 of column l % 32.  So accumulators 4s .. 4s+3 of a
// 32 x 32 output are exactly the B operand of K-step s of the next product.
__device__ __forceinline__ int mfma_row(int r, int l) { return 8 * (r >> 2) + 4 * (l >> 5) + (r & 3); }

__device__ __forceinline__ uint32_t pack_h2(float a, float b) {
    const _Float16 ha = (_Float16)a, hb = (_Float16)b;
    return (uint32_t)__builtin_bit_cast(uint16_t, ha) | ((uint32_t)__builtin_bit_cast(uint16_t, hb) << 16);
}
__device__ __forceinline__ half4v as_half4(uint32_t lo, uint32_t hi) {
    const uint64_t v = (uint64_t)lo | ((uint64_t)hi << 32);
    return __builtin_bit_cast(half4v, v);
}

// Predicted cells (expm1 of the network output) of 64 pairs per wavefront,
// the 6-32-32-1 network on the matrix cores: with the pairs as the N
// dimension of v_mfma_f32_32x32x8_f16, layer 1 is one product (K = 6 padded
// to 8), layer 2 four (K = 32), and layer 1's ReLU'd, fp16-rounded
// accumulators are layer 2's B operand in place; layer 3 is a 16-term dot
// per lane and one cross-half add.  Two passes of 32 pairs per wavefront.
// FROM_FEAT: the raw features come from feat; otherwise they are computed
// here from the batch's sequences (and written to feat when it is set).
template <bool FROM_FEAT>
__global__ __launch_bounds__(256) void mlp_forward_kernel(ExtArgs a, const __half *__restrict__ wg,
                                                          const float *__restrict__ normg, float *feat,
                                                          uint32_t n, float *cells_pred) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t tile = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
    if (tile >= n) return;  // whole wavefront idle
    const int m = lane & 31, kh = 4 * (lane >> 5);
    // this lane's operands of the weights (same for every pass)
    half4v A1;
#pragma unroll
    for (int j = 0; j < 4; ++j) A1[j] = kh + j < NIN ? (_Float16)wg[m * NIN + kh + j] : (_Float16)0.0f;
    half4v A2[4];
#pragma unroll
    for (int st = 0; st < 4; ++st)
#pragma unroll
        for (int j = 0; j < 4; ++j) A2[st][j] = (_Float16)wg[OFF_W2 + m * NH + 8 * st + kh + j];
    float b1r[16], b2r[16], w3r[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = mfma_row(r, lane);
        b1r[r] = __half2float(wg[OFF_B1 + row]);
        b2r[r] = __half2float(wg[OFF_B2 + row]);
        w3r[r] = __half2float(wg[OFF_W3 + row]);
    }
    const float b3 = __half2float(wg[OFF_B3]);

    // this lane's pair: raw features, normalised, fp16, two halves per word
    const uint32_t k = tile + (uint32_t)lane;
    float f[NIN] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (k < n) {
        if constexpr (FROM_FEAT) {
#pragma unroll
            for (int i = 0; i < NIN; ++i) f[i] = feat[(size_t)k * NIN + i];
        } else {
            pair_features(a, k, f);
            if (feat) {
#pragma unroll
                for (int i = 0; i < NIN; ++i) feat[(size_t)k * NIN + i] = f[i];
            }
        }
    }
    float xn[NIN];
#pragma unroll
    for (int i = 0; i < NIN; ++i) xn[i] = (f[i] - normg[i]) * normg[NIN + i];
    const uint32_t x01 = pack_h2(xn[0], xn[1]), x23 = pack_h2(xn[2], xn[3]), x45 = pack_h2(xn[4], xn[5]);

#pragma unroll
    for (int pass = 0; pass < 2; ++pass) {
        // B operand of layer 1: features 0-3 (lanes 0-31) or 4-7 (lanes
        // 32-63) of pair 32 pass + lane % 32
        const int src = 32 * pass + m;
        const uint32_t lo0 = (uint32_t)__shfl((int)x01, src, 64), lo1 = (uint32_t)__shfl((int)x23, src, 64);
        const uint32_t hi0 = (uint32_t)__shfl((int)x45, src, 64);
        const half4v B1 = lane < 32 ? as_half4(lo0, lo1) : as_half4(hi0, 0u);
        float16v acc = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x8f16(A1, B1, acc, 0, 0, 0);
        // layer 1 bias and ReLU, rounded to fp16: layer 2's B operand in place
        half4v H1[4];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float v = acc[r] + b1r[r];
            H1[r >> 2][r & 3] = (_Float16)(v > 0.f ? v : 0.f);
        }
        float16v acc2 = {};
#pragma unroll
        for (int st = 0; st < 4; ++st) acc2 = __builtin_amdgcn_mfma_f32_32x32x8f16(A2[st], H1[st], acc2, 0, 0, 0);
        // layer 2 bias, ReLU and fp16 rounding, then layer 3 over this lane's
        // 16 neurons and the other half's 16
        float y = 0.f;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const float v = acc2[r] + b2r[r];
            y += w3r[r] * (float)(_Float16)(v > 0.f ? v : 0.f);
        }
        y += __shfl_xor(y, 32, 64);
        y += b3;
        const uint32_t pk = tile + (uint32_t)src;
        if (lane < 32 && pk < n) {
            const float c = expm1f(y);
            cells_pred[pk] = c > 0.f ? c : 0.f;
        }
    }
}

// One minibatch slice of kMlpBlock pairs per block: forward, backward, and the
// block's partial gradient sum (already divided by the minibatch size) in
// partial[blockIdx.x][0 .. kMlpParams), block loss sum in partial[..][kMlpParams].
__global__ __launch_bounds__(kMlpBlock) void mlp_backward_kernel(
    const __half *wg, const float *normg, const float *feat, const uint32_t *cells,
    const uint32_t *perm, uint32_t first, uint32_t count, float inv_count, float *partial) {
    __shared__ __half w[kMlpParams];
    __shared__ float norm[2 * NIN];
    // per pair: x[6], h1[32], h2[32], d1[32], d2[32], d3  (fp32)
    constexpr int REC = NIN + 4 * NH + 1;
    __shared__ float rec[kMlpBlock * REC];
    __shared__ float lsum[kMlpBlock];
    for (int i = threadIdx.x; i < kMlpParams; i += blockDim.x) w[i] = wg[i];
    if (threadIdx.x < 2 * NIN) norm[threadIdx.x] = normg[threadIdx.x];
    __syncthreads();

    const uint32_t slot = blockIdx.x * kMlpBlock + threadIdx.x;
    float *r = rec + threadIdx.x * REC;
    float loss = 0.f;
    if (slot < count) {
        const uint32_t k = perm[first + slot];
        float raw[NIN], x[NIN], h1[NH], h2[NH];
        uint32_t p1, p2;
#pragma unroll
        for (int i = 0; i < NIN; ++i) raw[i] = feat[(size_t)k * NIN + i];
        const float y = mlp_forward1(w, raw, norm, x, h1, h2, p1, p2);
        const float t = log1pf((float)cells[k]);
        const float d = y - t;
        loss = 0.5f * d * d;
        const float d3 = d * inv_count;
        float d2[NH];
#pragma unroll
        for (int j = 0; j < NH; ++j)
            d2[j] = ((p2 >> j) & 1u) ? d3 * __half2float(w[OFF_W3 + j]) : 0.f;
#pragma unroll
        for (int kk = 0; kk < NH; ++kk) {
            float acc = 0.f;
#pragma unroll
            for (int j = 0; j < NH; ++j) acc += __half2float(w[OFF_W2 + j * NH + kk]) * d2[j];
            r[NIN + 2 * NH + kk] = ((p1 >> kk) & 1u) ? acc : 0.f;  // d1
        }
#pragma unroll
        for (int i = 0; i < NIN; ++i) r[i] = x[i];
#pragma unroll
        for (int j = 0; j < NH; ++j) {
            r[NIN + j] = h1[j];
            r[NIN + NH + j] = h2[j];
            r[NIN + 3 * NH + j] = d2[j];
        }
        r[NIN + 4 * NH] = d3;
    } else {
        for (int i = 0; i < REC; ++i) r[i] = 0.f;
    }
    lsum[threadIdx.x] = loss;
    __syncthreads();

    float *out = partial + (size_t)blockIdx.x * (kMlpParams + 1);
    for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) {
        float g = 0.f;
        if (p < OFF_B1) {  // W1[i][m] : d1[i] * x[m]
            const int i = p / NIN, m = p % NIN;
            for (int q = 0; q < kMlpBlock; ++q) g += rec[q * REC + NIN + 2 * NH + i] * rec[q * REC + m];
        } else if (p < OFF_W2) {  // b1
            const int i = p - OFF_B1;
            for (int q = 0; q < kMlpBlock; ++q) g += rec[q * REC + NIN + 2 * NH + i];
        } else if (p < OFF_B2) {  // W2[j][k] : d2[j] * h1[k]
            const int j = (p - OFF_W2) / NH, kk = (p - OFF_W2) % NH;
            for (int q = 0; q < kMlpBlock; ++q)
                g += rec[q * REC + NIN + 3 * NH + j] * rec[q * REC + NIN + kk];
        } else if (p < OFF_W3) {  // b2
            const int j = p - OFF_B2;
            for (int q = 0; q < kMlpBlock; ++q) g += rec[q * REC + NIN + 3 * NH + j];
        } else if (p < OFF_B3) {  // W3[j] : d3 * h2[j]
            const int j = p - OFF_W3;
            for (int q = 0; q < kMlpBlock; ++q) g += rec[q * REC + NIN + 4 * NH] * rec[q * REC + NIN + NH + j];
        } else {  // b3
            for (int q = 0; q < kMlpBlock; ++q) g += rec[q * REC + NIN + 4 * NH];
        }
        out[p] = g;
    }
    if (threadIdx.x == 0) {
        float s = 0.f;
        for (int q = 0; q < kMlpBlock; ++q) s += lsum[q];
        out[kMlpParams] = s;
    }
}

// Sums the per-block partial gradients in block order (deterministic), then
// optionally takes one Adam step on the fp32 master weights and refreshes the
// fp16 copy.  grad_out / loss_out receive the summed gradient and loss sum.
__global__ __launch_bounds__(256) void mlp_reduce_adam_kernel(const float *partial, uint32_t nblocks,
                                                              float *grad_out, float *loss_out,
                                                              float *master, float *m, float *v,
                                                              __half *wh, float lr, float b1,
                                                              float b2, float eps, float bc1,
                                                              float bc2, int apply) {
    // 16 parameters per block, 16 threads per parameter: thread j sums the
    // block partials j, j + 16, ... and the 16 sums are added in j order, a
    // fixed order, so the result does not depend on timing
    __shared__ float part[16][16];
    const int pl = (int)(threadIdx.x & 15), j = (int)(threadIdx.x >> 4);
    const int p = (int)blockIdx.x * 16 + pl;
    float s = 0.f;
    if (p <= kMlpParams)
        for (uint32_t b = (uint32_t)j; b < nblocks; b += 16u) s += partial[(size_t)b * (kMlpParams + 1) + p];
    part[j][pl] = s;
    __syncthreads();
    if (j != 0 || p > kMlpParams) return;
    float g = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) g += part[k][pl];
    if (p == kMlpParams) {
        if (loss_out) *loss_out = g;
        return;
    }
    if (grad_out) grad_out[p] = g;
    if (!apply) return;
    const float mm = b1 * m[p] + (1.f - b1) * g;
    const float vv = b2 * v[p] + (1.f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float upd = lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    const float nw = master[p] - upd;
    master[p] = nw;
    wh[p] = __float2half(nw);
}

__global__ void mlp_to_half_kernel(const float *src, __half *dst, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = __float2half(src[i]);
}

// ---- host launchers ----------------------------------------------------------

hipError_t mlp_launch_features(const ExtArgs &a, uint32_t n, float *feat, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(mlp_features_kernel, dim3((n + 255) / 256), dim3(256), 0, st, a, n, feat);
    return hipGetLastError();
}

hipError_t mlp_launch_predict(const ExtArgs &a, const void *wh, const float *norm, float *feat,
                              bool from_feat, uint32_t n, float *cells_pred, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const dim3 grid((n + 255) / 256), block(256);  // 64 pairs per wavefront, 4 wavefronts
    if (from_feat)
        hipLaunchKernelGGL(mlp_forward_kernel<true>, grid, block, 0, st, a, (const __half *)wh, norm, feat, n,
                           cells_pred);
    else
        hipLaunchKernelGGL(mlp_forward_kernel<false>, grid, block, 0, st, a, (const __half *)wh, norm, feat, n,
                           cells_pred);
    return hipGetLastError();
}

hipError_t mlp_launch_backward(const void *wh, const float *norm, const float *feat,
                               const uint32_t *cells, const uint32_t *perm, uint32_t first,
                               uint32_t count, float *partial, hipStream_t st) {
    if (count == 0) return hipSuccess;
    const uint32_t blocks = (count + kMlpBlock - 1) / kMlpBlock;
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(blocks), dim3(kMlpBlock), 0, st,
                       (const __half *)wh, norm, feat, cells, perm, first, count,
                       1.0f / (float)count, partial);
    return hipGetLastError();
}

hipError_t mlp_launch_reduce_adam(const float *partial, uint32_t nblocks, float *grad_out,
                                  float *loss_out, float *master, float *m, float *v, void *wh,
                                  float lr, float b1, float b2, float eps, float bc1, float bc2,
                                  int apply, hipStream_t st) {
    hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((kMlpParams + 1 + 15) / 16), dim3(256), 0,
                       st, partial, nblocks, grad_out, loss_out, master, m, v, (__half *)wh, lr,
                       b1, b2, eps, bc1, bc2, apply);
    return hipGetLastError();
}

hipError_t mlp_launch_to_half(const float *src, void *dst, int n, hipStream_t st) {
    hipLaunchKernelGGL(mlp_to_half_kernel, dim3((n + 255) / 256), dim3(256), 0, st, src,
                       (__half *)dst, n);
    return hipGetLastError();
}

}  // namespace fa
