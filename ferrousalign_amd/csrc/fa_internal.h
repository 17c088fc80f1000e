// Library-internal types shared by the host translation units (fa_api.cpp,
// fa_pipeline.cpp): the context, the device reference, a batch of extension
// pairs with its launch plan, and the host helpers both sides use.  Not part
// of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ferrousalign_amd.h"
#include "extend_kernel.h"
#include "mate_kernel.h"
#include "pack_host.h"

struct fa_ctx {
    int device = 0;
    hipStream_t stream = nullptr;           // main stream of the context
    hipStream_t side[3] = {nullptr, nullptr, nullptr};  // fork/join streams (default priority)
    hipEvent_t ev_fork = nullptr, ev_join[3] = {nullptr, nullptr, nullptr};
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    hipEvent_t ev_part[8] = {};  // pieces of a download the host consumes as they land
    std::string arch;
    int cus = 0;
    double last_dp_ms = 0.0, last_tb_ms = 0.0;  // kernels of the last fa_scalar_banded_swa_batch
    void *chain_ws = nullptr;  // device workspace of the chain-extension path (chain_api.cpp)
    void *pin = nullptr;       // pinned host staging for uploads and downloads (grow-only)
    size_t pin_bytes = 0;
    void *slot[32] = {};       // grow-only device buffers of the CIGAR-regeneration calls and of
    size_t slot_bytes[32] = {};  // the one-call batches of fa_execute_batch_simd_scoring
    void *hist_pin = nullptr;  // pinned packing histogram of those batches
    void *hscr = nullptr;      // pinned per-pair arrays of the CIGAR-regeneration calls (grow-only)
    size_t hscr_bytes = 0;
};

namespace fa_host {

struct Bucket {
    int S;
    uint32_t first, count;
    double work;   // estimated work: target rows x VALU instructions per row / pairs per wavefront
    int kind = 0;  // 0: 32-bit kernel, 1: packed 16-bit kernel (two pairs per lane group),
                   // 4: long pairs (one wavefront per pair)
    int L = 16;    // lanes per pair of this bucket
    int slen = 0;  // mate rescue: segments per lane ceil(qlen / 8) of every pair
};

// fits16 level of a pair that only the long-pair kernel holds (query over
// FA_MAX_QLEN_PACKED or target over FA_MAX_TLEN_PACKED bases)
constexpr uint8_t kEligLong = 3;

}  // namespace fa_host

// Device-resident reference in the .pac layout (4 bases per byte, first base
// in the two most significant bits).
struct fa_ref {
    int device = 0;
    uint8_t *d_pac = nullptr;
    int64_t l_pac = 0;
};

struct fa_batch {
    int64_t n = 0;
    // host copies kept for re-packing
    std::vector<int32_t> qlen, tlen, h0, w;
    std::vector<uint8_t> elig16;  // fits16 of each pair
    bool cost_set = false;         // d_pred holds the packing cost (FA_PACK_COST)
    bool plan_dirty = false;       // the cost changed since the last plan
    std::vector<uint32_t> qo_host, to_host;  // word offsets of each query / target in d_seq2
    // device
    uint32_t *d_seq2 = nullptr, *d_seqn = nullptr, *d_qo = nullptr, *d_to = nullptr;
    int32_t *d_qlen = nullptr, *d_tlen = nullptr, *d_h0 = nullptr, *d_w = nullptr;
    uint8_t *d_elig = nullptr;   // fits16 of each pair (device planner)
    uint32_t *d_order = nullptr;
    uint32_t *d_meta = nullptr;  // per packed slot: pid, qo, to, qlen, tlen, h0, w, 0
    uint32_t *d_params = nullptr;  // the same 8 words per pair in job order (meta_gather's source)
    int32_t *d_rec = nullptr;    // the extension kernels' 32-byte records (ExtArgs::orec)
    bool rec_pending = false;    // d_rec is newer than d_out / d_cells
    int32_t *d_out = nullptr;    // OutScore per pair and cells per pair (job order), split from
    uint32_t *d_cells = nullptr; // d_rec on demand (download, predictor training)
    uint32_t *d_queue = nullptr;
    void *d_sort = nullptr;      // packing sort: keys in/out, pair ids, scratch
    size_t sort_bytes = 0;
    void *h_hist = nullptr;      // pinned host copy of the packing histogram
    // kernel table of the packing plan for (lanes, FA_SCORE_I32, merged),
    // kept between plans: kernel key per (fit level, qlen) and the distinct
    // keys in launch order
    uint32_t tab_key = 0xFFFFFFFFu;
    int tab[3][256];
    std::vector<int> kinds;
    fa::ExtArgs args{};
    bool in_arena = false;      // arrays, sort scratch and histogram belong to the context
    // launch plan for the current packing
    uint32_t packed_flags = 0xFFFFFFFFu;
    std::vector<fa_host::Bucket> buckets;
    int lanes = 0;  // 0: narrowest lane group per query length
    int chunk = 0;
    int waves_per_cu = 0;
    int streams = 4;
    int tmax = 0;
    int64_t n_fallback32 = 0;   // pairs whose scores may leave the 16-bit range (fits16 == 0)
    int64_t nlong = 0;          // pairs for the long-pair kernel (fits16 level kEligLong)
    int long_qmax = 0;          // their longest query
    double long_work = 0.0;     // their band-free cell bound (sum of qlen x tlen), for stream balance
    void *d_long = nullptr;     // eh[] slices of the long-pair kernel
    size_t long_bytes = 0;
    int64_t last_overflow = 0;  // of the last extension run: pairs that ran the 32-bit kernel for that
    // mate-rescue local SW (FA_RUN_ALIGN2)
    bool align2_set = false;
    fa::MateArgs margs{};
    int32_t *d_mout = nullptr;     // 7 int32 per pair
    uint32_t *d_mcells = nullptr;  // cells per pair of the mate-rescue run
    uint16_t *d_imax_log = nullptr;
    uint32_t log_slots = 0;
    void *d_mrev = nullptr;        // reverse launches: keys in/out, ids, order, histogram, sort scratch
    size_t mrev_bytes = 0;
    std::vector<fa_host::Bucket> mbuckets;
    float *d_feat = nullptr;       // predictor features [n][6]
    float *d_pred = nullptr;       // predicted cells [n], or the caller's cost (fa_batch_set_cost)
    hipGraphExec_t graph_exec = nullptr;
    uint32_t graph_key = 0;
    int graph_wpc = 0;
};

namespace fa_host {

// FA_OK, or FA_ERR_DEVICE after printing what failed.
int hip_fail(hipError_t e, const char *what);
#define FA_HIP(call)                                        \
    do {                                                    \
        int _rc = fa_host::hip_fail((call), #call);         \
        if (_rc != FA_OK) return _rc;                       \
    } while (0)

int check_scoring(const fa_banded_pairwise_sw *sw);

// Kernel choice of one pair: kind (1: packed 16-bit, 0: 32-bit), lanes per
// pair and slot count, for the configured lane count L (0: narrowest group
// that holds the query), the FA_SCORE_I32 request `wide` and the pair's
// fits16 level; false when no instantiated kernel holds qlen.
bool pair_kernel(int L, bool wide, int elig16, int qlen, int *kind, int *lanes, int *S);

// 0: the pair needs the 32-bit kernel; 1: its scores provably fit the packed
// 16-bit kernel (Hmax = h0 + qlen * max(mat) < 4096, Hmax * K < 32768, gap
// costs in range); 2: also Hmax < 2048, which the wide slot counts need.
// matmax = max(0, max(mat)).
int fits16(const fa_banded_pairwise_sw &sw, int matmax, int64_t h0, int64_t qlen);

// Scoring fields of the kernel argument block (profile bytes, gap costs,
// z-drop, 16-bit guard).
void set_scoring(fa::ExtArgs &a, const fa_banded_pairwise_sw &sw);

// Enqueue one run of a planned batch (buckets over up to 4 streams).
int enqueue_run(fa_ctx *ctx, fa_batch *b, uint32_t flags);

// Releases the chain-extension workspace of a context (chain_api.cpp).
void chain_workspace_free(fa_ctx *ctx);

// Looks up, with `device` current, the device address of the fixed scratch
// array of fa_execute_batch_device_noalloc (noalloc.hip); once per device.
int noalloc_init(int device);

// Host-to-device copies of pageable host arrays through the context's pinned
// staging buffer on stream st: the host threads fill one half of it while the
// other half's copy is in flight.  Returns once every copy is enqueued; the
// staging buffer is in use until the stream reaches them.
struct HostSeg {
    void *dst;        // device
    const void *src;  // host
    size_t bytes;
};
int staged_upload(fa_ctx *ctx, const HostSeg *segs, int nseg, hipStream_t st);
// Device-to-host copy into pageable memory through the staging buffer, in
// pieces that the host threads copy out as they land; synchronous.
int staged_download(fa_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st);


}  // namespace fa_host
