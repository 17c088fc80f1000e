// Targets read from a device-resident reference in the .pac layout
// (internal to the library; reached through fa_ref_upload and
// fa_batch_upload_ref).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fa {

// For job j < n: target base k < len[j] is base p = rev[j] ? beg[j] + len[j] - 1 - k
// : beg[j] + k of the bidirectional coordinate space [0, 2*l_pac) (p >= l_pac: the
// complement of forward base 2*l_pac - 1 - p), written as 2-bit words at
// seq2[to[j] ..].  One wavefront per job.  The caller guarantees
// 0 <= beg, beg + len <= 2*l_pac, and no segment containing both l_pac - 1 and l_pac.
hipError_t pac_gather(const uint8_t *pac, int64_t l_pac, const int64_t *beg, const int32_t *len,
                      const uint8_t *rev, const uint32_t *to, uint32_t *seq2, uint32_t n,
                      hipStream_t st);

}  // namespace fa
