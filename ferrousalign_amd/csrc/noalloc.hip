// Allocation-free form of fa_execute_batch_simd_scoring for callers whose
// inputs are already in device memory and that may not allocate (HIP graph
// capture, tight serving loops): every buffer beyond the caller's inputs and
// outputs lives in one file-scope device array, `scratch`, of a fixed size.
//
// Steps, all on the context's stream:
//   1. noalloc_clear_kernel: zero the header (histogram, fill counters, max
//      target length, first-error key);
//   2. noalloc_prep_kernel (one wavefront per job): the checks of
//      validate_job in fa_api.cpp, in the same order; 2-bit packing of the
//      query and target into fixed per-job slots; end-bonus width capping
//      (ksw_affine_gap.rs:1001-1018) and the 16-bit fit level; a histogram of
//      (fit level, qlen); an invalid job gets its code, a zeroed record and 0
//      cells;
//   3. the host reads the header back (3 KB), orders the (level, qlen) bins
//      by kernel choice as plan_packing does, and writes the bin bases;
//   4. noalloc_scatter_kernel: the pair order (bins in that order, query
//      length descending inside a kernel's range);
//   5. meta_gather and one extension launch per kernel choice (the same
//      kernels as every other entry point).
// The scratch array is shared by every call on a device, so calls that
// overlap on one device race; each call rewrites everything it reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>

#include "../../include/ferrousalign_amd.h"
#include "extend_kernel.h"
#include "fa_internal.h"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
constexpr int kQWords = 16;  // 2-bit words of a query slot (seq_words(255))

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

// Device address of `scratch` per device, looked up once with that device
// current (each device holds its own copy of the array).
std::mutex g_mu;
void *g_ptr[kMaxDevices] = {};
int g_rc[kMaxDevices] = {};

struct Header {
    uint32_t hist[3][256];   // valid jobs per (16-bit fit level, qlen)
    uint32_t base[3][256];   // first packed position of each bin (host-written)
    uint32_t fill[3][256];   // scatter cursors
    int32_t tmax;            // longest valid target
    uint32_t nvalid;
    unsigned long long errkey;  // min over invalid jobs of (job << 8) | -code
};
constexpr size_t kHeaderBytes = (sizeof(Header) + 255) & ~(size_t)255;

struct PrepArgs {
    int64_t n;
    const uint8_t *qseq, *tseq;
    const int64_t *qoff, *toff;
    const int32_t *qlen, *tlen, *h0, *w;
    int32_t max_tlen;
    uint32_t stride;  // words per job slot: kQWords + seq_words(max_tlen)
    int32_t matmax, end_bonus, o_del, e_del, o_ins, e_ins;
    int32_t gaps16;   // the gap costs alone allow the 16-bit kernel
    Header *hdr;
    uint32_t *seq2, *seqn, *qo, *to;
    int32_t *wcap;
    int32_t *out;
    uint32_t *cells;
    int32_t *job_rc;
};

__global__ __launch_bounds__(256) void noalloc_clear_kernel(Header *h) {
    // every word before errkey to 0 (errkey is written by one thread only)
    uint32_t *p = reinterpret_cast<uint32_t *>(h);
    const uint32_t n = (uint32_t)(offsetof(Header, errkey) / 4);
    for (uint32_t k = blockIdx.x * 256u + threadIdx.x; k < n; k += gridDim.x * 256u) p[k] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) h->errkey = ~0ull;
}

__device__ int32_t f64_as_i32_dev(double x) {
    if (isnan(x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (int32_t)0x80000000u;
    return (int32_t)x;
}

// 2-bit word j (bases 16j .. 16j+15) of a sequence and its 16 N-mask bits;
// bad: a code above 4 among them
__device__ __forceinline__ void pack_word(const uint8_t *s, int32_t len, int j, uint32_t *w2,
                                          uint32_t *nm, bool *bad) {
    uint32_t w = 0, m = 0;
    bool b = false;
    for (int y = 0; y < 16; ++y) {
        const int p = 16 * j + y;
        if (p >= len) break;
        const uint32_t c = s[p];
        b |= c > 4u;
        if (c == 4u)
            m |= 1u << y;
        else
            w |= (c & 3u) << (2 * y);
    }
    *w2 = w;
    *nm = m;
    *bad = b;
}

__global__ __launch_bounds__(64) void noalloc_prep_kernel(PrepArgs a) {
    const int64_t k = blockIdx.x;
    const int lane = (int)threadIdx.x;
    if (k >= a.n) return;
    const int32_t ql = a.qlen[k], tl = a.tlen[k], h0 = a.h0[k], w = a.w[k];
    // validate_job (fa_api.cpp), same order of checks
    int rc = FA_OK;
    if (h0 <= 0) rc = FA_ERR_H0;
    else if (ql <= 0 || ql > FA_MAX_QLEN_PACKED) rc = FA_ERR_QLEN;
    else if (tl < 0 || tl > a.max_tlen) rc = FA_ERR_TLEN;
    else if (w < 0) rc = FA_ERR_BAND;
    else if (h0 > (1 << 19)) rc = FA_ERR_SCORING;
    const uint32_t qo = (uint32_t)k * a.stride, to = qo + kQWords;
    if (rc == FA_OK) {
        bool bad = false;
        // query: lanes 0..15, one 2-bit word each; the N words pair lanes 2m, 2m+1
        {
            uint32_t w2 = 0, nm = 0;
            bool b = false;
            if (lane < kQWords) pack_word(a.qseq + a.qoff[k], ql, lane, &w2, &nm, &b);
            const uint32_t nm1 = (uint32_t)__shfl_xor((int)nm, 1);
            if (lane < kQWords) {
                a.seq2[qo + lane] = w2;
                if ((lane & 1) == 0) a.seqn[qo / 2 + lane / 2] = nm | (nm1 << 16);
            }
            bad |= b;
        }
        const int tw = (int)(a.stride - kQWords);
        for (int j0 = 0; j0 < tw; j0 += 64) {
            const int j = j0 + lane;
            uint32_t w2 = 0, nm = 0;
            bool b = false;
            if (j < tw) pack_word(a.tseq + a.toff[k], tl, j, &w2, &nm, &b);
            const uint32_t nm1 = (uint32_t)__shfl_xor((int)nm, 1);
            if (j < tw) {
                a.seq2[to + j] = w2;
                if ((j & 1) == 0) a.seqn[to / 2 + j / 2] = nm | (nm1 << 16);
            }
            bad |= b;
        }
        if (__ballot(bad) != 0) rc = FA_ERR_CODE;
    }
    if (lane != 0) return;
    if (a.job_rc) a.job_rc[k] = rc;
    if (rc != FA_OK) {
        int32_t *o = a.out + (size_t)k * 6;
        for (int f = 0; f < 6; ++f) o[f] = 0;
        if (a.cells) a.cells[k] = 0u;
        atomicMin(&a.hdr->errkey, ((unsigned long long)k << 8) | (unsigned long long)(-rc));
        return;
    }
    // capped_w (fa_api.cpp): ksw_affine_gap.rs:1001-1018
    int32_t mi = f64_as_i32_dev(((double)ql * a.matmax + a.end_bonus - a.o_ins) / (double)a.e_ins + 1.0);
    mi = max(mi, 1);
    int32_t wc = min(w, mi);
    int32_t md = f64_as_i32_dev(((double)ql * a.matmax + a.end_bonus - a.o_del) / (double)a.e_del + 1.0);
    md = max(md, 1);
    wc = min(wc, md);
    // fits16 (fa_api.cpp)
    const int64_t bound = (int64_t)h0 + (int64_t)ql * a.matmax;
    const int64_t K = max(2, 1 + a.matmax);
    const int lvl = (a.gaps16 && bound < 4096 && bound * K < 32768) ? (bound < 2048 ? 2 : 1) : 0;
    a.qo[k] = qo;
    a.to[k] = to;
    a.wcap[k] = wc;
    atomicAdd(&a.hdr->hist[lvl][ql], 1u);
    atomicAdd(&a.hdr->nvalid, 1u);
    atomicMax(&a.hdr->tmax, tl);
}

__global__ __launch_bounds__(256) void noalloc_scatter_kernel(PrepArgs a, uint32_t *order) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= a.n) return;
    if (a.job_rc[k] != FA_OK) return;
    const int32_t ql = a.qlen[k];
    // recompute the level exactly as the prep kernel did
    const int64_t bound = (int64_t)a.h0[k] + (int64_t)ql * a.matmax;
    const int64_t K = max(2, 1 + a.matmax);
    const int lvl = (a.gaps16 && bound < 4096 && bound * K < 32768) ? (bound < 2048 ? 2 : 1) : 0;
    const uint32_t pos = a.hdr->base[lvl][ql] + atomicAdd(&a.hdr->fill[lvl][ql], 1u);
    order[pos] = (uint32_t)k;
}

}  // namespace

namespace fa_host {

// Called by fa_ctx_create with `device` current.
int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return FA_ERR_DEVICE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device] || g_rc[device]) return g_rc[device];
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        g_rc[device] = FA_ERR_DEVICE;
        return FA_ERR_DEVICE;
    }
    g_ptr[device] = p;
    return FA_OK;
}

}  // namespace fa_host

extern "C" int fa_execute_batch_device_noalloc(
    fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs, const uint8_t *d_query_seqs,
    const int64_t *d_query_offset, const int32_t *d_query_len, const uint8_t *d_ref_seqs,
    const int64_t *d_ref_offset, const int32_t *d_ref_len, const int32_t *d_h0,
    const int32_t *d_band_width, int32_t max_ref_len, fa_out_score *d_out, uint32_t *d_cells,
    int32_t *d_job_rc) {
    if (!ctx || njobs < 0 || max_ref_len < 0 || max_ref_len > FA_MAX_TLEN_PACKED) return FA_ERR_ARG;
    int rc = fa_host::check_scoring(sw);
    if (rc) return rc;
    if (njobs == 0) return FA_OK;
    if (!d_query_seqs || !d_query_offset || !d_query_len || !d_ref_seqs || !d_ref_offset ||
        !d_ref_len || !d_h0 || !d_band_width || !d_out)
        return FA_ERR_ARG;
    if (ctx->device < 0 || ctx->device >= kMaxDevices) return FA_ERR_DEVICE;
    uint8_t *base;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        base = (uint8_t *)g_ptr[ctx->device];
    }
    if (!base) return FA_ERR_DEVICE;
    // scratch layout: header | seq2 | seqn | qo | to | wcap | rc | order | meta
    const uint32_t stride = (uint32_t)kQWords + fa_host::seq_words(max_ref_len);
    const uint64_t n = (uint64_t)njobs;
    auto al = [](uint64_t x) { return (x + 255) & ~(uint64_t)255; };
    const uint64_t o_seq2 = kHeaderBytes;
    const uint64_t o_seqn = al(o_seq2 + n * stride * 4);
    const uint64_t o_qo = al(o_seqn + n * stride * 2);
    const uint64_t o_to = al(o_qo + n * 4);
    const uint64_t o_w = al(o_to + n * 4);
    const uint64_t o_rc = al(o_w + n * 4);
    const uint64_t o_order = al(o_rc + n * 4);
    const uint64_t o_meta = al(o_order + n * 4);
    const uint64_t need = o_meta + n * 32;
    if (need > kScratchBytes || n > 0x7FFFFFFFull) return FA_ERR_SCRATCH;

    FA_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    Header *hdr = (Header *)base;
    PrepArgs pa{};
    pa.n = njobs;
    pa.qseq = d_query_seqs;
    pa.tseq = d_ref_seqs;
    pa.qoff = d_query_offset;
    pa.toff = d_ref_offset;
    pa.qlen = d_query_len;
    pa.tlen = d_ref_len;
    pa.h0 = d_h0;
    pa.w = d_band_width;
    pa.max_tlen = max_ref_len;
    pa.stride = stride;
    int matmax = 0;
    for (int x = 0; x < 25; ++x) matmax = std::max(matmax, (int)sw->mat[x]);
    pa.matmax = matmax;
    pa.end_bonus = sw->end_bonus;
    pa.o_del = sw->o_del;
    pa.e_del = sw->e_del;
    pa.o_ins = sw->o_ins;
    pa.e_ins = sw->e_ins;
    pa.gaps16 = (sw->e_ins < 32 && sw->o_ins + sw->e_ins < 8192 && sw->o_del + sw->e_del < 8192) ? 1 : 0;
    pa.hdr = hdr;
    pa.seq2 = (uint32_t *)(base + o_seq2);
    pa.seqn = (uint32_t *)(base + o_seqn);
    pa.qo = (uint32_t *)(base + o_qo);
    pa.to = (uint32_t *)(base + o_to);
    pa.wcap = (int32_t *)(base + o_w);
    pa.out = (int32_t *)d_out;
    pa.cells = d_cells;
    // per-job codes: the caller's array, or a scratch one (the scatter reads them)
    pa.job_rc = d_job_rc ? d_job_rc : (int32_t *)(base + o_rc);
    uint32_t *order = (uint32_t *)(base + o_order);

    hipLaunchKernelGGL(noalloc_clear_kernel, dim3(4), dim3(256), 0, st, hdr);
    FA_HIP(hipGetLastError());
    hipLaunchKernelGGL(noalloc_prep_kernel, dim3((uint32_t)n), dim3(64), 0, st, pa);
    FA_HIP(hipGetLastError());
    // the histogram, the longest target and the first error
    uint32_t hist[3][256];
    int32_t tmax_nvalid[2];
    unsigned long long errkey = 0;
    FA_HIP(hipMemcpyAsync(hist, hdr->hist, sizeof hist, hipMemcpyDeviceToHost, st));
    FA_HIP(hipMemcpyAsync(tmax_nvalid, &hdr->tmax, sizeof tmax_nvalid, hipMemcpyDeviceToHost, st));
    FA_HIP(hipMemcpyAsync(&errkey, &hdr->errkey, sizeof errkey, hipMemcpyDeviceToHost, st));
    FA_HIP(hipStreamSynchronize(st));
    const int first_err = errkey == ~0ull ? FA_OK : -(int)(errkey & 0xFFu);
    const uint32_t nvalid = (uint32_t)tmax_nvalid[1];
    if (nvalid == 0) return first_err;

    // bins (level, qlen) ordered like plan_packing's buckets: kernel key
    // (kind << 16 | lanes << 8 | slots) descending, then qlen descending
    struct Bin {
        int key, q, lvl;
    };
    Bin bins[3 * 256];
    int nb = 0;
    for (int lvl = 2; lvl >= 0; --lvl)
        for (int q = FA_MAX_QLEN_PACKED; q >= 1; --q) {
            if (!hist[lvl][q]) continue;
            int kind, lanes, S;
            if (!fa_host::pair_kernel(0, false, lvl, q, &kind, &lanes, &S)) return FA_ERR_QLEN;
            bins[nb++] = Bin{(kind << 16) | (lanes << 8) | S, q, lvl};
        }
    // insertion sort (at most 765 bins), stable on the (level, qlen) order above
    for (int x = 1; x < nb; ++x) {
        const Bin v = bins[x];
        int y = x - 1;
        while (y >= 0 && (bins[y].key < v.key || (bins[y].key == v.key && bins[y].q < v.q))) {
            bins[y + 1] = bins[y];
            --y;
        }
        bins[y + 1] = v;
    }
    uint32_t bases[3][256];
    std::memset(bases, 0, sizeof bases);
    struct Launch {
        int key;
        uint32_t first, count;
    };
    Launch launches[3 * 256];
    int nl = 0;
    uint32_t pos = 0;
    for (int x = 0; x < nb; ++x) {
        const uint32_t c = hist[bins[x].lvl][bins[x].q];
        bases[bins[x].lvl][bins[x].q] = pos;
        if (nl > 0 && launches[nl - 1].key == bins[x].key)
            launches[nl - 1].count += c;
        else
            launches[nl++] = Launch{bins[x].key, pos, c};
        pos += c;
    }
    FA_HIP(hipMemcpyAsync(hdr->base, bases, sizeof bases, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(noalloc_scatter_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, st, pa,
                       order);
    FA_HIP(hipGetLastError());

    fa::ExtArgs a{};
    a.seq2 = pa.seq2;
    a.seqn = pa.seqn;
    a.qo = pa.qo;
    a.to = pa.to;
    a.qlen = d_query_len;
    a.tlen = d_ref_len;
    a.h0 = d_h0;
    a.w = pa.wcap;
    a.order = order;
    a.meta = (uint32_t *)(base + o_meta);
    a.out = (int32_t *)d_out;
    a.cells = d_cells;
    a.queue = nullptr;
    a.ovf = nullptr;
    fa_host::set_scoring(a, *sw);
    a.tpad = ((tmax_nvalid[0] + 16) / 16) * 16;
    a.chunk = 0;
    FA_HIP(fa::meta_gather(a, nvalid, st));
    for (int x = 0; x < nl; ++x) {
        const int key = launches[x].key, kind = key >> 16, L = (key >> 8) & 0xFF, S = key & 0xFF;
        if (kind == 1)
            FA_HIP(fa::extend16_launch(L, S, a, launches[x].first, launches[x].count, false, 0, ctx->cus, 0, st));
        else
            FA_HIP(fa::extend_launch_plain(L, S, a, launches[x].first, launches[x].count, st));
    }
    FA_HIP(hipStreamSynchronize(st));
    return first_err;
}
