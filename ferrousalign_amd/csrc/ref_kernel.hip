// Target extraction from a device-resident 2-bit reference.
//
// Layout of the reference's .pac (src/pipelines/linear/index/bntseq.rs:559-575):
// base p sits in byte p/4, first base of a byte in its two most significant
// bits, codes A=0 C=1 G=2 T=3.  Coordinates p >= l_pac address the reverse
// complement (get_reference_segment, :451-522): base p there is 3 minus forward
// base 2*l_pac - 1 - p.  A left-extension job reads its target segment
// backwards (collect_soa.rs:156-166).  Each output word packs 16 bases, base k
// of the target at bits 2*(k%16) of word k/16, the layout the extension
// kernels read.  One wavefront per job, one lane per output word: consecutive
// lanes read consecutive reference bytes (forward or backward) and write
// consecutive words.
#include "ref_kernel.h"

namespace fa {

namespace {

__device__ __forceinline__ uint32_t pac_base(const uint8_t *pac, int64_t p) {
    return (pac[p >> 2] >> ((~(uint32_t)p & 3u) << 1)) & 3u;
}

__global__ __launch_bounds__(256) void pac_gather_kernel(const uint8_t *__restrict__ pac,
                                                         int64_t l_pac, const int64_t *beg,
                                                         const int32_t *len, const uint8_t *rev,
                                                         const uint32_t *to, uint32_t *seq2,
                                                         uint32_t n) {
    const uint32_t job = blockIdx.x * 4u + (threadIdx.x >> 6);
    if (job >= n) return;
    const int lane = (int)(threadIdx.x & 63);
    const int64_t b = beg[job];
    const int L = len[job];
    const bool r = rev[job] != 0;
    const int64_t two = l_pac << 1;
    const uint32_t nw = (uint32_t)(L + 15) / 16u;
    for (uint32_t w = (uint32_t)lane; w < nw; w += 64u) {
        uint32_t word = 0;
        for (int j = 0; j < 16; ++j) {
            const int k = (int)w * 16 + j;
            if (k >= L) break;
            const int64_t p = r ? b + L - 1 - k : b + k;
            const uint32_t c = p < l_pac ? pac_base(pac, p) : 3u - pac_base(pac, two - 1 - p);
            word |= c << (2 * j);
        }
        seq2[to[job] + w] = word;
    }
}

}  // namespace

hipError_t pac_gather(const uint8_t *pac, int64_t l_pac, const int64_t *beg, const int32_t *len,
                      const uint8_t *rev, const uint32_t *to, uint32_t *seq2, uint32_t n,
                      hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(pac_gather_kernel, dim3((n + 3) / 4), dim3(256), 0, st, pac, l_pac, beg,
                       len, rev, to, seq2, n);
    return hipGetLastError();
}

}  // namespace fa
