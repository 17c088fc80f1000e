// Host side of the C ABI (include/ferrousalign_amd.h): validation, 2-bit
// packing, length/cost packing into lane groups, device upload, kernel
// launches and result download.  Replaces the reference's CPU dispatch
// (src/pipelines/linear/batch_extension/dispatch.rs:31-268 and
// src/core/alignment/banded_swa/shared.rs:161-389) for one GPU.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ferrousalign_amd.h"
#include "fa_internal.h"
#include "extend_kernel.h"
#include "predictor.h"
#include "mate_kernel.h"
#include "valu_probe.h"
#include "ref_kernel.h"
#include "swtb_kernel.h"
#include "plan_kernel.h"

using fa_host::Bucket;

namespace fa_host {

int hip_fail(hipError_t e, const char *what) {
    if (e == hipSuccess) return FA_OK;
    std::fprintf(stderr, "ferrousalign_amd: %s failed: %s\n", what, hipGetErrorString(e));
    return FA_ERR_DEVICE;
}

}  // namespace fa_host

namespace {

using fa_host::hip_fail;

// Rust `f64 as i32`: truncation toward zero, saturating, NaN -> 0.
int32_t f64_as_i32(double x) {
    if (std::isnan(x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

// Band width after end-bonus capping, ksw_affine_gap.rs:1001-1018
// (max_score = max(0, max(mat))).
int32_t capped_w(const fa_banded_pairwise_sw &sw, int32_t max_score, int32_t qlen, int32_t w) {
    int32_t max_ins = f64_as_i32(((double)qlen * max_score + sw.end_bonus - sw.o_ins) /
                                     (double)sw.e_ins + 1.0);
    max_ins = std::max(max_ins, 1);
    w = std::min(w, max_ins);
    int32_t max_del = f64_as_i32(((double)qlen * max_score + sw.end_bonus - sw.o_del) /
                                     (double)sw.e_del + 1.0);
    max_del = std::max(max_del, 1);
    return std::min(w, max_del);
}

}  // namespace

namespace fa_host {

int check_scoring(const fa_banded_pairwise_sw *sw) {
    if (!sw) return FA_ERR_ARG;
    if (sw->m != 5) return FA_ERR_SCORING;
    if (sw->o_del < 0 || sw->e_del < 0 || sw->o_ins < 0 || sw->e_ins < 0) return FA_ERR_SCORING;
    if (sw->o_del > 4096 || sw->e_del > 4096 || sw->o_ins > 4096 || sw->e_ins > 4096)
        return FA_ERR_SCORING;
    for (int q = 1; q < 5; ++q)
        if (sw->mat[20 + q] != sw->mat[20]) return FA_ERR_SCORING;
    return FA_OK;
}

bool pair_kernel(int L, bool wide, int elig16, int q, int *kind, int *lanes, int *S) {
    // lanes per pair of a query length: the configured L when its slot
    // counts reach, else the next wider group; L == 0 picks the narrowest
    // group that holds the query (the fewest lanes per pair amortise the
    // per-row reductions over the most pairs per wavefront)
    if (!wide && elig16 > 0) {
        const bool wok = elig16 > 1;
        int l16 = 16;
        if (L == 0) {
            for (int c : {4, 8, 16})
                if (fa::extend16_slot_count(c, q, wok) > 0) {
                    l16 = c;
                    break;
                }
        } else {
            l16 = fa::extend16_slot_count(L, q, wok) > 0 ? L : 16;
        }
        const int s16 = fa::extend16_slot_count(l16, q, wok);
        if (s16 > 0) {
            *kind = 1;
            *lanes = l16;
            *S = s16;
            return true;
        }
    }
    const int c = L ? L : 8;
    const int l32 = fa::extend_slot_count(c, q) > 0 ? c : 8;
    *kind = 0;
    *lanes = l32;
    *S = fa::extend_slot_count(l32, q);
    return *S > 0;
}

int fits16(const fa_banded_pairwise_sw &sw, int matmax, int64_t h0, int64_t qlen) {
    const int64_t bound = h0 + qlen * matmax;
    const int64_t K = std::max(2, 1 + matmax);
    if (!(bound < 4096 && bound * K < 32768 && sw.e_ins < 32 && sw.o_ins + sw.e_ins < 8192 &&
          sw.o_del + sw.e_del < 8192))
        return 0;
    return bound < 2048 ? 2 : 1;
}

void set_scoring(fa::ExtArgs &a, const fa_banded_pairwise_sw &sw) {
    for (int q = 0; q < 5; ++q) {
        uint32_t v = 0;
        for (int t = 0; t < 4; ++t) v |= (uint32_t)(uint8_t)(sw.mat[t * 5 + q] + 128) << (8 * t);
        a.profw[q] = v;
    }
    a.nrow = (uint32_t)(uint8_t)(sw.mat[20] + 128);
    int mx = 0;
    for (int x = 0; x < 25; ++x) mx = std::max(mx, (int)sw.mat[x]);
    a.mguard = std::max(2, 1 + mx);
    a.o_del = sw.o_del;
    a.e_del = sw.e_del;
    a.o_ins = sw.o_ins;
    a.e_ins = sw.e_ins;
    a.zdrop = sw.zdrop;
}

}  // namespace fa_host

namespace {

using fa_host::check_scoring;

// f(lo, hi, thread) over up to 16 contiguous ranges of [0, n) on threads of
// their own; serially below `serial_below` items.
template <class F>
void parallel_for(int64_t n, F &&f, int64_t serial_below = 65536) {
    unsigned nt = std::thread::hardware_concurrency();
    if (nt < 1) nt = 1;
    if (nt > 16) nt = 16;
    if (n < serial_below || nt == 1) {
        f(0, n, 0u);
        return;
    }
    std::vector<std::thread> th;
    int64_t per = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; ++t) {
        int64_t lo = per * t, hi = std::min<int64_t>(lo + per, n);
        if (lo >= hi) break;
        th.emplace_back([=, &f] { f(lo, hi, t); });
    }
    for (auto &x : th) x.join();
}

using fa_host::seq_words;

// t == nullptr: the target comes from a device reference (codes 0-3 by construction)
// The checks of validate_job that read no sequence.
int validate_scalar(int32_t ql, int32_t tl, int32_t h0, int32_t w) {
    if (h0 <= 0) return FA_ERR_H0;
    if (ql <= 0 || ql > FA_MAX_QLEN) return FA_ERR_QLEN;
    if (tl < 0 || tl > FA_MAX_TLEN) return FA_ERR_TLEN;
    if (w < 0) return FA_ERR_BAND;
    if (h0 > (1 << 19)) return FA_ERR_SCORING;
    return FA_OK;
}

int validate_job(int32_t ql, int32_t tl, int32_t h0, int32_t w, const uint8_t *q,
                 const uint8_t *t) {
    const int r = validate_scalar(ql, tl, h0, w);
    if (r != FA_OK) return r;
    if (!fa_host::codes_ok(q, ql)) return FA_ERR_CODE;
    if (t && !fa_host::codes_ok(t, tl)) return FA_ERR_CODE;
    return FA_OK;
}

inline void pack_seq(const uint8_t *s, int32_t len, uint32_t *w2, uint32_t *wn) {
    fa_host::pack_codes(s, len, w2, wn);
}

template <class T>
int dev_alloc(T **p, size_t count) {
    *p = nullptr;
    if (count == 0) count = 1;
    return hip_fail(hipMalloc((void **)p, count * sizeof(T)), "hipMalloc");
}

// context slots of the one-call batches (the CIGAR path uses 0-11)
constexpr int kSlotBatchArena = 28, kSlotBatchSort = 29;

// Grow-only device buffer slot k of the context (CIGAR-regeneration calls,
// one-call batches); the previous contents are not kept.
template <class T>
int ctx_alloc(fa_ctx *ctx, int k, T **p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (ctx->slot_bytes[k] < bytes) {
        if (ctx->slot[k]) (void)hipFree(ctx->slot[k]);
        ctx->slot[k] = nullptr;
        ctx->slot_bytes[k] = 0;
        const size_t want = bytes + bytes / 8;
        FA_HIP(hipMalloc(&ctx->slot[k], want));
        ctx->slot_bytes[k] = want;
    }
    *p = (T *)ctx->slot[k];
    return FA_OK;
}

template <class T>
int dev_upload(T **p, const T *src, size_t count, hipStream_t st) {
    int rc = dev_alloc(p, count);
    if (rc) return rc;
    if (count) FA_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
    return FA_OK;
}

void drop_graph(fa_batch *b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    b->graph_exec = nullptr;
}

// The extension records of the last run as OutScore and cell arrays (one
// coalesced pass on the context's stream, where every run ends).
int ensure_records(fa_ctx *ctx, fa_batch *b) {
    if (!b->rec_pending) return FA_OK;
    FA_HIP(fa::records_split(b->d_rec, (uint32_t)b->n, b->d_out, b->d_cells, ctx->stream));
    b->rec_pending = false;
    return FA_OK;
}

void batch_release(fa_batch *b) {
    if (!b) return;
    drop_graph(b);
    if (b->in_arena) {  // the context keeps these for the next one-call batch
        void **carved[] = {(void **)&b->d_seq2, (void **)&b->d_seqn, (void **)&b->d_qo, (void **)&b->d_to,
                           (void **)&b->d_qlen, (void **)&b->d_tlen, (void **)&b->d_h0, (void **)&b->d_w,
                           (void **)&b->d_elig, (void **)&b->d_order, (void **)&b->d_meta, (void **)&b->d_params,
                           (void **)&b->d_out, (void **)&b->d_rec, (void **)&b->d_cells, (void **)&b->d_queue,
                           (void **)&b->d_long, (void **)&b->d_sort, (void **)&b->h_hist};
        for (void **p : carved) *p = nullptr;
    }
    if (b->margs.order) (void)hipFree(const_cast<uint32_t *>(b->margs.order));
    void *ptrs[] = {b->d_seq2, b->d_seqn, b->d_qo, b->d_to, b->d_qlen, b->d_tlen, b->d_h0,
                    b->d_w, b->d_order, b->d_meta, b->d_out, b->d_cells, b->d_queue, b->d_feat, b->d_pred,
                    b->d_mout, b->d_imax_log, b->d_sort, b->d_elig, b->d_mrev, b->d_long, b->d_params,
                    b->d_rec, b->d_mcells};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (b->h_hist) (void)hipHostFree(b->h_hist);
    delete b;
}

// Build the pair order and the per-slot-count buckets for `flags`, on the
// device (plan_kernel.hip): a (fit level, qlen) histogram, one sort key per
// pair, a stable radix sort of (key, pair id) and the launch metadata; the
// host reads the 9 KB histogram back and cuts the order into buckets.
// Buckets run widest first; inside a bucket pairs run from the most to the
// least expensive (by (qlen, tlen), or by the cost), so that lane groups of
// one wavefront hold pairs of similar cost and the long pairs start first.
int plan_packing(fa_ctx *ctx, fa_batch *b, uint32_t flags) {
    const uint32_t pack = flags & 0xF0u;
    const uint32_t single = flags & FA_LAUNCH_SINGLE;
    const uint32_t wide = flags & FA_SCORE_I32;
    const int L = b->lanes;
    const uint32_t key = pack | single | wide | (uint32_t)L << 16;
    if (b->packed_flags == key && !(flags & FA_PLAN_FRESH) && !b->plan_dirty) return FA_OK;
    const int64_t n = b->n;
    const bool by_cost = pack == FA_PACK_COST || pack == FA_PACK_COST_ORDER;
    if (by_cost && !b->cost_set) return FA_ERR_ARG;
    // kernel of a pair, (kind << 16) | (lanes << 8) | slots, as a function of
    // its fits16 level and qlen (-1: no kernel holds it); merged buckets (one
    // per kernel kind) take the kind alone.  Kept between plans.
    const bool merged = pack == FA_PACK_NONE || single;
    const uint32_t tkey = (uint32_t)L << 8 | (wide ? 2u : 0u) | (merged ? 1u : 0u);
    if (b->tab_key != tkey) {
        std::vector<int> kinds;  // distinct keys, descending: the bucket order
        for (int f = 0; f < 3; ++f)
            for (int q = 0; q < 256; ++q) {
                int kind, lanes, S;
                const bool ok = q > 0 && fa_host::pair_kernel(L, wide != 0, f, q, &kind, &lanes, &S);
                b->tab[f][q] = !ok ? -1 : (merged ? kind << 16 : (kind << 16) | (lanes << 8) | S);
                if (b->tab[f][q] >= 0) kinds.push_back(b->tab[f][q]);
            }
        std::sort(kinds.begin(), kinds.end(), [](int x, int y) { return x > y; });
        kinds.erase(std::unique(kinds.begin(), kinds.end()), kinds.end());
        for (int f = 0; f < 3; ++f)
            for (int q = 0; q < 256; ++q)
                if (b->tab[f][q] >= 0)
                    b->tab[f][q] = (int)(std::find(kinds.begin(), kinds.end(), b->tab[f][q]) - kinds.begin());
        b->kinds.swap(kinds);
        b->tab_key = tkey;
    }
    const std::vector<int> &kinds = b->kinds;  // b->tab now holds indices into kinds
    const size_t nk = kinds.size();
    // device buffers (grow-only): histogram, keys in/out (64-bit room), ids, sort scratch
    constexpr size_t kBins = 768;
    size_t tb32 = 0, tb64 = 0;
    FA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb32, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (uint32_t *)nullptr,
                                              (int)std::max<int64_t>(n, 1), 0, 32, ctx->stream));
    FA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb64, (unsigned long long *)nullptr,
                                              (unsigned long long *)nullptr, (uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (int)std::max<int64_t>(n, 1), 0, 34,
                                              ctx->stream));
    size_t tb = std::max(tb32, tb64);
    const size_t hist_bytes = kBins * (4 + 8 + 8) + 256;
    const size_t need = hist_bytes + 16 * (size_t)n + 4 * (size_t)n + tb + 512;
    if (b->in_arena) {  // a one-call batch: the context's grow-only scratch
        uint8_t *p = nullptr;
        int rc0 = ctx_alloc(ctx, kSlotBatchSort, &p, need);
        if (rc0) return rc0;
        b->d_sort = p;
        b->sort_bytes = need;
        if (!ctx->hist_pin) FA_HIP(hipHostMalloc(&ctx->hist_pin, kBins * 12, hipHostMallocDefault));
        b->h_hist = ctx->hist_pin;
    } else if (b->sort_bytes < need) {
        if (b->d_sort) (void)hipFree(b->d_sort);
        b->d_sort = nullptr;
        b->sort_bytes = 0;
        FA_HIP(hipMalloc(&b->d_sort, need));
        b->sort_bytes = need;
    }
    if (!b->h_hist) FA_HIP(hipHostMalloc(&b->h_hist, kBins * 12, hipHostMallocDefault));
    uint8_t *base = (uint8_t *)b->d_sort;
    fa::PlanHist h;
    h.rows = (unsigned long long *)base;
    h.csum = h.rows + kBins;
    h.cnt = (uint32_t *)(h.csum + kBins);
    uint8_t *kbase = base + hist_bytes;
    uint32_t *vin = (uint32_t *)(kbase + 16 * (size_t)n);
    void *temp = (void *)(((uintptr_t)(vin + n) + 255) & ~(uintptr_t)255);
    fa::PlanArgs pa;
    pa.n = (uint32_t)n;
    pa.mode = pack == FA_PACK_COST         ? fa::kPlanCost
              : pack == FA_PACK_COST_ORDER ? fa::kPlanCostOrder
              : pack == FA_PACK_NONE       ? fa::kPlanNone
                                           : fa::kPlanLength;
    pa.qlen = b->d_qlen;
    pa.tlen = b->d_tlen;
    pa.elig = b->d_elig;
    pa.cost = by_cost ? b->d_pred : nullptr;
    pa.shift = pa.mode == fa::kPlanCostOrder ? fa::kRankShiftCompact : fa::kRankShift;
    // the histogram first: the launches, their ranks and the key width come
    // from it, so the rest of the plan needs no further readback
    FA_HIP(hipMemsetAsync(base, 0, hist_bytes, ctx->stream));
    FA_HIP(fa::plan_hist(pa, h, ctx->cus, ctx->stream));
    uint32_t *cnt = (uint32_t *)b->h_hist;
    unsigned long long *rows = (unsigned long long *)(cnt + kBins);
    FA_HIP(hipMemcpyAsync(cnt, h.cnt, kBins * 4, hipMemcpyDeviceToHost, ctx->stream));
    FA_HIP(hipMemcpyAsync(rows, h.rows, kBins * 8, hipMemcpyDeviceToHost, ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    // launch ranks of the kernels the batch uses, in kernel order
    std::vector<int> rank(nk, -1);
    std::vector<uint64_t> kcount(nk, 0), krows(nk, 0);
    std::vector<int> kmaxq(nk, 1), kminf(nk, 2);
    for (int f = 0; f < 3; ++f)
        for (int q = 1; q < 256; ++q) {
            const uint32_t c = cnt[(size_t)f * 256 + (size_t)q];
            if (!c) continue;
            const int x = b->tab[f][q];
            if (x < 0) return FA_ERR_QLEN;  // a pair no kernel holds
            kcount[(size_t)x] += c;
            krows[(size_t)x] += rows[(size_t)f * 256 + (size_t)q];
            kmaxq[(size_t)x] = std::max(kmaxq[(size_t)x], q);
            kminf[(size_t)x] = std::min(kminf[(size_t)x], f);
        }
    int nr = 0;
    for (size_t x = 0; x < nk; ++x)
        if (kcount[x]) rank[x] = nr++;
    pa.long_rank = nr;
    const int nranks = nr + (b->nlong > 0 ? 1 : 0);
    if (nranks > 64) return FA_ERR_ARG;
    fa::PlanTable t;
    for (int f = 0; f < 3; ++f)
        for (int q = 0; q < 256; ++q) t.kidx[f][q] = (int16_t)(b->tab[f][q] < 0 ? -1 : rank[(size_t)b->tab[f][q]]);
    // cost order where a lane group holds 64 or more columns: there a
    // wavefront of pairs with equal cost ends together; in the narrow kernels
    // (short queries, a few hundred cells) the prediction's error is too large
    // a part of the cost and length order does better
    std::memset(t.cost_order, 0, sizeof(t.cost_order));
    for (size_t x = 0; x < nk; ++x)
        if (rank[x] >= 0 && (kinds[x] & 0xFFFF) != 0) {
            const int kl = (kinds[x] >> 8) & 0xFF, ks = kinds[x] & 0xFF;
            t.cost_order[rank[x]] = (uint8_t)(kl * ks >= 64 ? 1 : 0);
        }
    // keys: rank << shift | order inside the launch
    int rbits = 0;
    while ((1 << rbits) < nranks) ++rbits;
    const int end_bit = pa.shift + rbits;
    if (n > 0) {
        if (end_bit <= 32) {
            uint32_t *kin = (uint32_t *)kbase, *kout = kin + n;
            FA_HIP(fa::plan_keys(pa, h, t, kin, vin, ctx->stream));
            FA_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, kin, kout, vin, b->d_order, (int)n, 0, end_bit,
                                                      ctx->stream));
        } else {
            unsigned long long *kin = (unsigned long long *)kbase, *kout = kin + n;
            FA_HIP(fa::plan_keys(pa, h, t, kin, vin, ctx->stream));
            FA_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, kin, kout, vin, b->d_order, (int)n, 0, end_bit,
                                                      ctx->stream));
        }
    }
    // the launch metadata; the launches that follow on this stream wait for it
    FA_HIP(fa::meta_gather(b->args, (uint32_t)n, ctx->stream));
    // buckets: the used kernels in rank order with their pair counts
    std::vector<Bucket> buckets;
    uint32_t first = 0;
    for (size_t x = 0; x < nk; ++x) {
        if (rank[x] < 0) continue;
        const uint64_t count = kcount[x], nrows = krows[x];
        const int maxq = kmaxq[x], minf = kminf[x];
        const int key0 = kinds[x];
        Bucket bk;
        bk.kind = key0 >> 16;
        bk.first = first;
        bk.count = (uint32_t)count;
        first += (uint32_t)count;
        if ((key0 & 0xFFFF) == 0) {
            // merged bucket: the slot count of its longest query, wide slots only
            // when every pair of the bucket is under their bound (fits16 level 2)
            int kind, lanes, S;
            if (!fa_host::pair_kernel(L, bk.kind == 0, bk.kind == 0 ? 0 : minf, maxq, &kind, &lanes, &S))
                return FA_ERR_QLEN;
            bk.L = lanes;
            bk.S = S;
        } else {
            bk.L = (key0 >> 8) & 0xFF;
            bk.S = key0 & 0xFF;
        }
        // work: rows x VALU instructions per wavefront-row / pairs per
        // wavefront (31 per slot for the packed kernel's two pairs, ~40 for
        // the 32-bit kernel's one, ~175 per row of bookkeeping)
        bk.work = (double)nrows *
                  (bk.kind == 1 ? (31.0 * bk.S + 175.0) * bk.L / 128.0 : (40.0 * bk.S + 175.0) * bk.L / 64.0);
        buckets.push_back(bk);
    }
    if (b->nlong > 0) {
        // long pairs sort after every lane-group kernel: one launch of their own
        Bucket bk;
        bk.kind = 4;
        bk.first = first;
        bk.count = (uint32_t)b->nlong;
        bk.L = 64;
        bk.S = 0;
        bk.work = b->long_work;
        buckets.push_back(bk);
    }
    if (buckets.size() > 64) return FA_ERR_ARG;
    // a re-plan that gives the same launches keeps the captured graph (it
    // reads the order and metadata from device memory at replay)
    bool same = key == b->packed_flags && buckets.size() == b->buckets.size();
    for (size_t x = 0; same && x < buckets.size(); ++x)
        same = buckets[x].first == b->buckets[x].first && buckets[x].count == b->buckets[x].count &&
               buckets[x].L == b->buckets[x].L && buckets[x].S == b->buckets[x].S &&
               buckets[x].kind == b->buckets[x].kind;
    b->buckets.swap(buckets);
    b->packed_flags = key;
    b->plan_dirty = false;
    if (!same) drop_graph(b);
    return FA_OK;
}

}  // namespace

namespace fa_host {

hipError_t launch_bucket(fa_ctx *ctx, fa_batch *b, uint32_t flags, size_t k, hipStream_t st) {
    const Bucket &bk = b->buckets[k];
    fa::ExtArgs a = b->args;
    if (a.orec) b->rec_pending = true;
    if (bk.kind == 4)
        return fa::extend_long_launch(a, bk.first, bk.count, b->long_qmax, b->d_long, b->long_bytes, st);
    a.chunk = std::max(1, b->chunk / (64 / bk.L) / (bk.kind == 1 ? 2 : 1));
    if (bk.kind == 1)
        return fa::extend16_launch(bk.L, bk.S, a, bk.first, bk.count,
                                   (flags & FA_LAUNCH_PERSISTENT) != 0, (uint32_t)k, ctx->cus,
                                   b->waves_per_cu / 4, st);
    if (flags & FA_LAUNCH_PERSISTENT)
        return fa::extend_launch_persistent(bk.L, bk.S, a, bk.first, bk.count, (uint32_t)k,
                                            ctx->cus, b->waves_per_cu / 4, st);
    return fa::extend_launch_plain(bk.L, bk.S, a, bk.first, bk.count, st);
}

// Enqueue one run of the batch (no synchronisation).  Buckets are spread over
// up to 4 streams (the main stream and three side streams forked from it and
// joined back), heaviest bucket first on the least-loaded stream, so that
// the drain of one bucket overlaps the start of the next.  Under graph
// capture the fork/join becomes parallel graph branches.
int enqueue_run(fa_ctx *ctx, fa_batch *b, uint32_t flags) {
    if (flags & FA_LAUNCH_PERSISTENT)
        FA_HIP(fa::queue_reset(b->d_queue, (uint32_t)b->buckets.size(), ctx->stream));
    const int ns = std::max(1, std::min(4, b->streams));
    const size_t nb = b->buckets.size();
    if (ns == 1 || nb <= 1) {
        for (size_t k = 0; k < nb; ++k) FA_HIP(launch_bucket(ctx, b, flags, k, ctx->stream));
        return FA_OK;
    }
    hipStream_t st[4] = {ctx->stream, ctx->side[0], ctx->side[1], ctx->side[2]};
    FA_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    for (int j = 1; j < ns; ++j) FA_HIP(hipStreamWaitEvent(st[j], ctx->ev_fork, 0));
    std::vector<size_t> idx(nb);
    for (size_t k = 0; k < nb; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(),
                     [&](size_t x, size_t y) { return b->buckets[x].work > b->buckets[y].work; });
    double load[4] = {0, 0, 0, 0};
    for (size_t k : idx) {
        int j = 0;
        for (int q = 1; q < ns; ++q)
            if (load[q] < load[j]) j = q;
        load[j] += b->buckets[k].work;
        FA_HIP(launch_bucket(ctx, b, flags, k, st[j]));
    }
    for (int j = 1; j < ns; ++j) {
        FA_HIP(hipEventRecord(ctx->ev_join[j - 1], st[j]));
        FA_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join[j - 1], 0));
    }
    return FA_OK;
}

}  // namespace fa_host

namespace {

using fa_host::enqueue_run;
using fa_host::launch_bucket;

constexpr uint32_t kMateChunk = 262144;  // pairs per mate launch (row-log scratch reuse)

// One launch of a mate-rescue run, for fa_batch_profile_launches: kernel
// (2 << 8: forward, 3 << 8: XSTART reverse; | SSE lanes per pair), stripe
// length, range in its order and duration.
struct MateLaunch {
    int kernel, slen;
    uint32_t first, count;
    double ms;
};

// One mate-rescue run: the forward launches over the slen buckets (chunks of
// kMateChunk pairs, the row log's size), then, when KSW_XSTART is set and
// FA_ALIGN2_FUSED is not, the reverse pass as launches of its own over only
// the pairs that need it: sort keys from the forward records (bin = stripe
// length of qe + 1, then forward score descending), a stable radix sort, the
// bins' counts read back, one unmasked launch per stripe length.  A pair's
// reverse pass in the forward kernel would keep its wavefront's 16 pairs
// until the slowest pass ends, masked to the longest stripe among them, and
// 20-25% of the pairs of a mate window need none.  With `rec`, every launch
// is timed alone (a synchronisation after each).
int run_align2(fa_ctx *ctx, fa_batch *b, uint32_t flags, int *nlaunch,
               std::vector<MateLaunch> *rec = nullptr) {
    if (!b->align2_set) return FA_ERR_ARG;
    const bool split = !(flags & FA_ALIGN2_FUSED) && (b->margs.xtra & fa::kKswXStart) && b->n > 0;
    fa::MateArgs a = b->margs;
    a.split_rev = split ? 1 : 0;
    const int lanes = a.byte_mode ? 16 : 8;
    int launches = 0;
    auto timed = [&](int kernel, int slen, uint32_t first, uint32_t count, auto &&go) -> int {
        if (rec) FA_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        FA_HIP(go());
        ++launches;
        if (rec) {
            FA_HIP(hipEventRecord(ctx->ev1, ctx->stream));
            FA_HIP(hipEventSynchronize(ctx->ev1));
            float t = 0.f;
            FA_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
            rec->push_back(MateLaunch{kernel, slen, first, count, (double)t});
        }
        return FA_OK;
    };
    int rc;
    for (const Bucket &bk : b->mbuckets) {
        for (uint32_t off = 0; off < bk.count; off += kMateChunk) {
            const uint32_t cnt = std::min<uint32_t>(kMateChunk, bk.count - off);
            rc = timed(2 << 8 | lanes, bk.slen, bk.first + off, cnt,
                       [&] { return fa::mate16_launch(a, bk.first + off, cnt, bk.slen, ctx->stream); });
            if (rc) return rc;
        }
    }
    if (split) {
        const int64_t n = b->n;
        size_t tb = 0;
        FA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0,
                                                  fa::kMateRevKeyBits, ctx->stream));
        const size_t need = 16 * (size_t)n + 64 * 4 + tb + 512;
        if (b->mrev_bytes < need) {
            if (b->d_mrev) (void)hipFree(b->d_mrev);
            b->d_mrev = nullptr;
            b->mrev_bytes = 0;
            FA_HIP(hipMalloc(&b->d_mrev, need));
            b->mrev_bytes = need;
        }
        uint32_t *kin = (uint32_t *)b->d_mrev;
        uint32_t *kout = kin + n, *ids = kout + n, *rorder = ids + n, *hist = rorder + n;
        void *temp = (void *)(((uintptr_t)(hist + 64) + 255) & ~(uintptr_t)255);
        FA_HIP(hipMemsetAsync(hist, 0, 64 * 4, ctx->stream));
        FA_HIP(fa::mate_rev_keys(a, (uint32_t)n, kin, ids, hist, ctx->stream));
        FA_HIP(hipcub::DeviceRadixSort::SortPairs(temp, tb, kin, kout, ids, rorder, (int)n, 0,
                                                  fa::kMateRevKeyBits, ctx->stream));
        uint32_t cnt[64];
        FA_HIP(hipMemcpyAsync(cnt, hist, sizeof(cnt), hipMemcpyDeviceToHost, ctx->stream));
        FA_HIP(hipStreamSynchronize(ctx->stream));
        a.rorder = rorder;
        // the bins run on the context's four streams, largest first on the
        // least-loaded one, so that the small bins (a single wavefront's pass
        // each) overlap the large ones (timed runs stay on one stream)
        hipStream_t st[4] = {ctx->stream, ctx->side[0], ctx->side[1], ctx->side[2]};
        const int ns = rec ? 1 : 4;
        if (ns > 1) {
            FA_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
            for (int j = 1; j < ns; ++j) FA_HIP(hipStreamWaitEvent(st[j], ctx->ev_fork, 0));
        }
        double load[4] = {0, 0, 0, 0};
        uint32_t start = 0;
        for (int bin = 0; bin < fa::kMateMaxSlen; ++bin) {
            const uint32_t c = cnt[bin];
            const int slen = fa::kMateMaxSlen - bin;
            if (c == 0) continue;
            int j = 0;
            for (int q = 1; q < ns; ++q)
                if (load[q] < load[j]) j = q;
            load[j] += (double)c * slen;
            rc = timed(3 << 8 | lanes, slen, start, c,
                       [&] { return fa::mate16_rev_launch(a, start, c, slen, st[j]); });
            if (rc) return rc;
            start += c;
        }
        for (int j = 1; j < ns; ++j) {
            FA_HIP(hipEventRecord(ctx->ev_join[j - 1], st[j]));
            FA_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join[j - 1], 0));
        }
    }
    if (nlaunch) *nlaunch = launches;
    return FA_OK;
}

int run_batch(fa_ctx *ctx, fa_batch *b, uint32_t flags, int *nlaunch) {
    if (flags & FA_RUN_ALIGN2) {
        b->last_overflow = 0;
        return run_align2(ctx, b, flags, nlaunch);
    }
    int rc = plan_packing(ctx, b, flags);
    if (rc) return rc;
    b->last_overflow = b->n_fallback32;
    if (b->args.orec) b->rec_pending = true;  // also for a graph replay
    if (nlaunch) *nlaunch = (int)b->buckets.size();
    if (!(flags & FA_LAUNCH_GRAPH)) return enqueue_run(ctx, b, flags);
    const uint32_t key = (flags & 0xFF4u) | (uint32_t)b->lanes << 16 | (uint32_t)b->chunk << 21;
    if (!b->graph_exec || b->graph_key != key || b->graph_wpc != b->waves_per_cu * 8 + b->streams) {
        drop_graph(b);
        FA_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_run(ctx, b, flags & ~FA_LAUNCH_GRAPH);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(ctx->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        FA_HIP(e);
        e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FA_HIP(e);
        b->graph_key = key;
        b->graph_wpc = b->waves_per_cu * 8 + b->streams;
    }
    FA_HIP(hipGraphLaunch(b->graph_exec, ctx->stream));
    return FA_OK;
}

}  // namespace

extern "C" {

const char *fa_strerror(int code) {
    switch (code) {
        case FA_OK: return "ok";
        case FA_ERR_ARG: return "invalid argument";
        case FA_ERR_H0: return "h0 must be positive";
        case FA_ERR_QLEN: return "query length out of range (1..FA_MAX_QLEN; 1..FA_MAX_QLEN_PACKED on the paths without long pairs)";
        case FA_ERR_CODE: return "base code out of range (0..4)";
        case FA_ERR_TLEN: return "target length out of range (0..FA_MAX_TLEN; 0..FA_MAX_TLEN_PACKED on the paths without long pairs)";
        case FA_ERR_BAND: return "band width must be non-negative";
        case FA_ERR_SCORING: return "unsupported scoring parameters";
        case FA_ERR_DEVICE: return "HIP device error";
        case FA_ERR_NODEVICE: return "no usable gfx950 device";
        case FA_ERR_OVERFLOW: return "8-bit score overflow";
        case FA_ERR_SEGMENT: return "target segment outside the reference or across the strand boundary";
        case FA_ERR_SCRATCH: return "batch larger than the allocation-free entry point's fixed scratch";
        default: return "unknown error";
    }
}

int fa_pack_2bit(const uint8_t *codes, int32_t len, uint32_t *w2, uint32_t *wn) {
    if (len < 0 || (len > 0 && (!codes || !w2 || !wn))) return FA_ERR_ARG;
    // the check runs in the packing pass, as in the batch uploads
    return fa_host::pack_codes_checked(codes, len, w2, wn) ? FA_OK : FA_ERR_CODE;
}

int fa_unpack_2bit(const uint32_t *w2, const uint32_t *wn, int32_t len, uint8_t *codes) {
    if (len < 0 || (len > 0 && (!codes || !w2 || !wn))) return FA_ERR_ARG;
    for (int32_t p = 0; p < len; ++p) {
        uint32_t c = (w2[p >> 4] >> ((p & 15) * 2)) & 3u;
        codes[p] = ((wn[p >> 5] >> (p & 31)) & 1u) ? 4 : (uint8_t)c;
    }
    return FA_OK;
}

int fa_ctx_create(int device, fa_ctx **out) {
    if (!out) return FA_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return FA_ERR_NODEVICE;
    if (device < 0 || device >= ndev) return FA_ERR_NODEVICE;
    hipDeviceProp_t prop;
    FA_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return FA_ERR_NODEVICE;
    FA_HIP(hipSetDevice(device));
    fa_ctx *c = new (std::nothrow) fa_ctx();
    if (!c) return FA_ERR_ARG;
    c->device = device;
    c->arch = prop.gcnArchName;
    c->cus = prop.multiProcessorCount;
    bool ok = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&c->ev0) == hipSuccess && hipEventCreate(&c->ev1) == hipSuccess &&
              hipEventCreate(&c->ev2) == hipSuccess &&
              hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int j = 0; ok && j < 3; ++j)
        ok = hipStreamCreateWithFlags(&c->side[j], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->ev_join[j], hipEventDisableTiming) == hipSuccess;
    for (hipEvent_t &ev : c->ev_part)
        ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        fa_ctx_destroy(c);
        return FA_ERR_DEVICE;
    }
    // device address of the allocation-free entry point's scratch (once per
    // device; a failure there only disables that entry point)
    (void)fa_host::noalloc_init(device);
    *out = c;
    return FA_OK;
}

int fa_ctx_destroy(fa_ctx *ctx) {
    if (!ctx) return FA_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    fa_host::chain_workspace_free(ctx);
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    if (ctx->hist_pin) (void)hipHostFree(ctx->hist_pin);
    if (ctx->hscr) (void)hipHostFree(ctx->hscr);
    for (void *p : ctx->slot)
        if (p) (void)hipFree(p);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    for (int j = 0; j < 3; ++j) {
        if (ctx->side[j]) (void)hipStreamDestroy(ctx->side[j]);
        if (ctx->ev_join[j]) (void)hipEventDestroy(ctx->ev_join[j]);
    }
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (hipEvent_t ev : ctx->ev_part)
        if (ev) (void)hipEventDestroy(ev);
    delete ctx;
    return FA_OK;
}

int fa_ctx_describe(fa_ctx *ctx, char *buf, size_t cap) {
    if (!ctx || !buf || cap == 0) return FA_ERR_ARG;
    std::snprintf(buf, cap, "device=%d arch=%s cus=%d", ctx->device, ctx->arch.c_str(), ctx->cus);
    return FA_OK;
}

}  // extern "C"

namespace {

// Where the targets of a batch come from: host codes (seqs + offset) or
// segments of a device reference (ref, seg_beg, seg_rev).
struct TargetSource {
    const uint8_t *seqs = nullptr;
    const int64_t *offset = nullptr;
    const fa_ref *ref = nullptr;
    const int64_t *seg_beg = nullptr;
    const uint8_t *seg_rev = nullptr;
    bool validated = false;  // the caller already ran validate_job on every job
};
// Without `validated`, batch_upload runs the checks that read no sequence
// first and checks the codes while packing (one pass over the bases instead
// of two); any error then makes it validate every job in full, so that the
// error reported is the same: the first invalid job's, in job order.

// Grow-only pinned host buffer of the context (staging for uploads; the
// caller synchronises before the next use).
int pinned_staging(fa_ctx *ctx, size_t bytes) {
    if (ctx->pin_bytes >= bytes) return FA_OK;
    if (ctx->pin) (void)hipHostFree(ctx->pin);
    ctx->pin = nullptr;
    ctx->pin_bytes = 0;
    const size_t want = bytes + bytes / 8 + 4096;
    FA_HIP(hipHostMalloc(&ctx->pin, want));
    ctx->pin_bytes = want;
    return FA_OK;
}


template <class T>
int ctx_upload(fa_ctx *ctx, int k, T **p, const T *src, size_t count) {
    int rc = ctx_alloc(ctx, k, p, count);
    if (rc) return rc;
    if (count) FA_HIP(hipMemcpyAsync(*p, src, count * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
    return FA_OK;
}

// in_arena: a batch that lives for one call (fa_execute_batch_simd_scoring):
// its arrays come from the context's grow-only arena rather than from a
// hipMalloc each, and are not freed with it.
int batch_upload(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                 const uint8_t *query_seqs, const int64_t *query_offset, const int32_t *query_len,
                 const TargetSource &ts, const int32_t *ref_len, const int32_t *h0,
                 const int32_t *band_width, fa_batch **out, bool in_arena = false) {
    if (!ctx || !out || njobs < 0) return FA_ERR_ARG;
    *out = nullptr;
    int rc = check_scoring(sw);
    if (rc) return rc;
    const bool from_ref = ts.ref != nullptr;
    if (njobs > 0 && (!query_seqs || !query_offset || !query_len || !ref_len || !h0 ||
                      !band_width || (from_ref ? (!ts.seg_beg || !ts.seg_rev)
                                               : (!ts.seqs || !ts.offset))))
        return FA_ERR_ARG;
    if (njobs > (int64_t)0x7FFFFFFF) return FA_ERR_ARG;
    if (from_ref && ts.ref->device != ctx->device) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));

    fa_batch *b = new (std::nothrow) fa_batch();
    if (!b) return FA_ERR_ARG;
    const int64_t n = njobs;
    b->n = n;
    b->qlen.assign(query_len, query_len + n);
    b->tlen.assign(ref_len, ref_len + n);
    b->h0.assign(h0, h0 + n);
    b->w.resize((size_t)n);
    b->elig16.resize((size_t)n);

    // validate + offsets
    std::vector<uint32_t> qo((size_t)n), to((size_t)n);
    uint64_t words = 0;
    int first_err = FA_OK;
    int tmax = 0;
    std::vector<int8_t> vrc((size_t)n);
    int matmax = 0;  // max(0, max(mat))
    for (int x = 0; x < 25; ++x) matmax = std::max(matmax, (int)sw->mat[x]);
    std::vector<int> ttmax(16, 0), tqmax(16, 0);
    std::vector<int64_t> tnlong(16, 0);
    std::vector<double> twork(16, 0.0);
    parallel_for(n, [&](int64_t lo, int64_t hi, unsigned t) {
        int tm = 0, qm = 0;
        int64_t nl = 0;
        double wk = 0.0;
        for (int64_t k = lo; k < hi; ++k) {
            int r = ts.validated ? FA_OK : validate_scalar(query_len[k], ref_len[k], h0[k], band_width[k]);
            if (r == FA_OK && from_ref) {
                // inside [0, 2 l_pac) and on one strand (get_reference_segment
                // returns an empty segment for one that bridges, bntseq.rs:461-469)
                const int64_t b = ts.seg_beg[k], e = b + ref_len[k], lp = ts.ref->l_pac;
                if (b < 0 || e > 2 * lp || (b < lp && e > lp)) r = FA_ERR_SEGMENT;
            }
            vrc[(size_t)k] = (int8_t)r;
            if (r == FA_OK) {
                b->w[(size_t)k] = capped_w(*sw, matmax, query_len[k], band_width[k]);
                if (query_len[k] > FA_MAX_QLEN_PACKED || ref_len[k] > FA_MAX_TLEN_PACKED) {
                    // beyond the lane-group kernels: the long-pair kernel
                    b->elig16[(size_t)k] = fa_host::kEligLong;
                    qm = std::max(qm, query_len[k]);
                    ++nl;
                    wk += (double)query_len[k] * (double)ref_len[k];
                } else {
                    b->elig16[(size_t)k] = (uint8_t)fa_host::fits16(*sw, matmax, h0[k], query_len[k]);
                    tm = std::max(tm, ref_len[k]);
                }
            }
        }
        ttmax[t] = std::max(ttmax[t], tm);
        tqmax[t] = std::max(tqmax[t], qm);
        tnlong[t] += nl;
        twork[t] += wk;
    });
    for (int v : ttmax) tmax = std::max(tmax, v);
    if (!ts.validated && std::any_of(vrc.begin(), vrc.end(), [](int8_t r) { return r != FA_OK; })) {
        // an invalid job: the batch is rejected with the first invalid job's
        // error, which the full checks (codes included) decide
        parallel_for(n, [&](int64_t lo, int64_t hi, unsigned) {
            for (int64_t k = lo; k < hi; ++k) {
                int r = validate_job(query_len[k], ref_len[k], h0[k], band_width[k], query_seqs + query_offset[k],
                                     from_ref ? nullptr : ts.seqs + ts.offset[k]);
                if (r == FA_OK && from_ref) {
                    const int64_t b0 = ts.seg_beg[k], e = b0 + ref_len[k], lp = ts.ref->l_pac;
                    if (b0 < 0 || e > 2 * lp || (b0 < lp && e > lp)) r = FA_ERR_SEGMENT;
                }
                vrc[(size_t)k] = (int8_t)r;
            }
        });
    }
    for (int x = 0; x < 16; ++x) {
        b->long_qmax = std::max(b->long_qmax, tqmax[(size_t)x]);
        b->nlong += tnlong[(size_t)x];
        b->long_work += twork[(size_t)x];
    }
    for (int64_t k = 0; k < n; ++k)
        if (vrc[(size_t)k] != FA_OK) {
            first_err = vrc[(size_t)k];
            break;
        }
    if (first_err == FA_OK && n > 0) {
        // a pair's query and target words, back to back, start where they
        // touch the fewest 128-byte lines (32 words): the kernels read each
        // pair's words alone, in packed order, and HBM is read in whole
        // lines.  Placed per chunk of jobs in parallel, each chunk from a line
        // start, then shifted by the prefix of the chunks' sizes.
        const int64_t nch = std::min<int64_t>(256, (n + 16383) / 16384);
        std::vector<uint64_t> base((size_t)nch + 1, 0);
        parallel_for(
            nch,
            [&](int64_t c0, int64_t c1, unsigned) {
                for (int64_t c = c0; c < c1; ++c) {
                    uint64_t wds = 0;
                    for (int64_t k = n * c / nch; k < n * (c + 1) / nch; ++k) {
                        const uint64_t blk = (uint64_t)seq_words(query_len[k]) + seq_words(ref_len[k]);
                        if (((wds & 31u) + blk + 31u) / 32u > (blk + 31u) / 32u)
                            wds = (wds + 31u) & ~(uint64_t)31u;
                        wds = std::min<uint64_t>(wds, 0xFFFFFFFFull);  // a total this big is rejected below
                        qo[(size_t)k] = (uint32_t)wds;
                        wds += seq_words(query_len[k]);
                        to[(size_t)k] = (uint32_t)std::min<uint64_t>(wds, 0xFFFFFFFFull);
                        wds += seq_words(ref_len[k]);
                    }
                    base[(size_t)c + 1] = (wds + 31u) & ~(uint64_t)31u;
                }
            },
            2);
        for (int64_t c = 0; c < nch; ++c) base[(size_t)c + 1] += base[(size_t)c];
        words = base[(size_t)nch];
        if (words > 0xFFFFFFF0ull) {
            first_err = FA_ERR_ARG;
        } else {
            parallel_for(
                nch,
                [&](int64_t c0, int64_t c1, unsigned) {
                    for (int64_t c = c0; c < c1; ++c)
                        for (int64_t k = n * c / nch; k < n * (c + 1) / nch; ++k) {
                            qo[(size_t)k] += (uint32_t)base[(size_t)c];
                            to[(size_t)k] += (uint32_t)base[(size_t)c];
                        }
                },
                2);
        }
    }
    if (first_err != FA_OK) {
        delete b;
        return first_err;
    }
    // packed sequences staged in the context's pinned buffer (every word is
    // written: sequences are packed back to back, and a device-reference
    // target, which pac_gather writes on the device, is zeroed here)
    if ((rc = pinned_staging(ctx, (size_t)words * 4 + (size_t)(words / 2) * 4))) {
        delete b;
        return rc;
    }
    uint32_t *seq2 = (uint32_t *)ctx->pin, *seqn = seq2 + words;
    // every job passed the checks that read no sequence; unless the caller
    // validated the codes too, they are checked while packing
    const bool check = !ts.validated;
    std::atomic<bool> bad_code{false};
    parallel_for(n, [&](int64_t lo, int64_t hi, unsigned) {
        bool bad = false;
        for (int64_t k = lo; k < hi; ++k) {
            // the alignment gap before the pair (never read; zeroed so that
            // uploads are deterministic)
            const uint32_t gap0 = k ? to[(size_t)k - 1] + seq_words(ref_len[k - 1]) : 0u;
            if (gap0 < qo[(size_t)k]) {
                std::memset(&seq2[gap0], 0, (qo[(size_t)k] - gap0) * sizeof(uint32_t));
                std::memset(&seqn[gap0 / 2], 0, (qo[(size_t)k] - gap0) / 2 * sizeof(uint32_t));
            }
            const uint8_t *qs = query_seqs + query_offset[k];
            uint32_t *qw = &seq2[qo[(size_t)k]], *qn = &seqn[qo[(size_t)k] / 2];
            if (check)
                bad |= !fa_host::pack_codes_checked(qs, query_len[k], qw, qn);
            else
                pack_seq(qs, query_len[k], qw, qn);
            if (!from_ref) {
                const uint8_t *tsq = ts.seqs + ts.offset[k];
                uint32_t *tw = &seq2[to[(size_t)k]], *tn = &seqn[to[(size_t)k] / 2];
                if (check)
                    bad |= !fa_host::pack_codes_checked(tsq, ref_len[k], tw, tn);
                else
                    pack_seq(tsq, ref_len[k], tw, tn);
            } else {
                const uint32_t nw = seq_words(ref_len[k]);
                std::memset(&seq2[to[(size_t)k]], 0, nw * sizeof(uint32_t));
                std::memset(&seqn[to[(size_t)k] / 2], 0, nw / 2 * sizeof(uint32_t));
            }
        }
        if (bad) bad_code.store(true, std::memory_order_relaxed);
    });
    if (n > 0) {  // the last chunk's padding to its line
        const uint64_t end = (uint64_t)to[(size_t)n - 1] + seq_words(ref_len[n - 1]);
        std::memset(&seq2[end], 0, (size_t)(words - end) * sizeof(uint32_t));
        std::memset(&seqn[end / 2], 0, (size_t)(words - end) / 2 * sizeof(uint32_t));
    }
    if (bad_code.load()) {
        // every job passed the other checks, so the first invalid one is a
        // code error
        delete b;
        return FA_ERR_CODE;
    }

    hipStream_t st = ctx->stream;
    // the batch's device arrays, and the host data the first nine start from
    const size_t un = (size_t)n;
    b->long_bytes = b->nlong > 0 ? fa::extend_long_scratch(b->long_qmax, (uint32_t)b->nlong) : 0;
    struct DevArray {
        void **p;
        size_t bytes;
        const void *src;
    } arrays[] = {
        {(void **)&b->d_seq2, (size_t)words * 4, seq2},     {(void **)&b->d_seqn, (size_t)(words / 2) * 4, seqn},
        {(void **)&b->d_qo, un * 4, qo.data()},             {(void **)&b->d_to, un * 4, to.data()},
        {(void **)&b->d_qlen, un * 4, b->qlen.data()},      {(void **)&b->d_tlen, un * 4, b->tlen.data()},
        {(void **)&b->d_h0, un * 4, b->h0.data()},          {(void **)&b->d_w, un * 4, b->w.data()},
        {(void **)&b->d_elig, un, b->elig16.data()},        {(void **)&b->d_order, un * 4, nullptr},
        {(void **)&b->d_meta, un * 32, nullptr},            {(void **)&b->d_params, un * 32, nullptr},
        {(void **)&b->d_out, un * 24, nullptr},             {(void **)&b->d_rec, un * 32, nullptr},
        {(void **)&b->d_cells, un * 4, nullptr},            {(void **)&b->d_queue, 256, nullptr},
        {(void **)&b->d_long, b->long_bytes, nullptr},
    };
    auto al = [](size_t x) { return (std::max<size_t>(x, 1) + 255) & ~(size_t)255; };
    if (in_arena) {
        size_t total = 0;
        for (const DevArray &d : arrays) total += al(d.bytes);
        uint8_t *base = nullptr;
        if ((rc = ctx_alloc(ctx, kSlotBatchArena, &base, total))) {
            delete b;
            return rc;
        }
        b->in_arena = true;
        for (const DevArray &d : arrays) {
            *d.p = base;
            base += al(d.bytes);
        }
    } else {
        for (const DevArray &d : arrays) {
            if (d.p == (void **)&b->d_long && !d.bytes) continue;
            if ((rc = hip_fail(hipMalloc(d.p, std::max<size_t>(d.bytes, 1)), "hipMalloc"))) {
                batch_release(b);
                return rc;
            }
        }
    }
    for (const DevArray &d : arrays)
        if (d.src && d.bytes &&
            (rc = hip_fail(hipMemcpyAsync(*d.p, d.src, d.bytes, hipMemcpyHostToDevice, st), "hipMemcpyAsync"))) {
            (void)hipStreamSynchronize(st);  // copies from the staging buffer may be in flight
            batch_release(b);
            return rc;
        }
    if (in_arena && !b->nlong) b->d_long = nullptr;
    if (from_ref && n > 0) {
        int64_t *d_beg = nullptr;
        uint8_t *d_rev = nullptr;
        rc = dev_upload(&d_beg, ts.seg_beg, (size_t)n, st);
        if (!rc) rc = dev_upload(&d_rev, ts.seg_rev, (size_t)n, st);
        if (!rc && fa::pac_gather(ts.ref->d_pac, ts.ref->l_pac, d_beg, b->d_tlen, d_rev, b->d_to,
                                  b->d_seq2, (uint32_t)n, st) != hipSuccess)
            rc = FA_ERR_DEVICE;
        if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = FA_ERR_DEVICE;
        if (d_beg) (void)hipFree(d_beg);
        if (d_rev) (void)hipFree(d_rev);
        if (rc) {
            batch_release(b);
            return rc;
        }
    }
    if (hipStreamSynchronize(st) != hipSuccess) {
        batch_release(b);
        return FA_ERR_DEVICE;
    }
    b->qo_host.swap(qo);
    b->to_host.swap(to);
    for (int64_t k = 0; k < n; ++k) b->n_fallback32 += b->elig16[(size_t)k] == 0;

    fa::ExtArgs &a = b->args;
    a.seq2 = b->d_seq2;
    a.seqn = b->d_seqn;
    a.qo = b->d_qo;
    a.to = b->d_to;
    a.qlen = b->d_qlen;
    a.tlen = b->d_tlen;
    a.h0 = b->d_h0;
    a.w = b->d_w;
    a.order = b->d_order;
    a.meta = b->d_meta;
    // the kernels write one 32-byte record per pair (a single HBM sector),
    // split into the OutScore and cell arrays when they are read
    a.out = b->d_rec;
    a.orec = 1;
    a.cells = b->d_cells;
    a.queue = b->d_queue;
    a.ovf = nullptr;
    fa_host::set_scoring(a, *sw);
    b->tmax = tmax;
    a.tpad = ((tmax + 16) / 16) * 16;
    a.chunk = 0;
    // the per-pair records meta_gather reads, one 32-byte line per pair
    a.params = nullptr;
    if (fa::params_pack(a, (uint32_t)n, b->d_params, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        batch_release(b);
        return FA_ERR_DEVICE;
    }
    a.params = b->d_params;
    *out = b;
    return FA_OK;
}

}  // namespace

namespace fa_host {

constexpr size_t kStagedMin = (size_t)4 << 20;  // below: plain copies

int staged_upload(fa_ctx *ctx, const HostSeg *segs, int nseg, hipStream_t st) {
    size_t total = 0;
    for (int k = 0; k < nseg; ++k) total += segs[k].bytes;
    if (total == 0) return FA_OK;
    if (total < kStagedMin) {  // small: the runtime's own staging, no host threads
        for (int k = 0; k < nseg; ++k)
            if (segs[k].bytes) FA_HIP(hipMemcpyAsync(segs[k].dst, segs[k].src, segs[k].bytes, hipMemcpyHostToDevice, st));
        return FA_OK;
    }
    // pieces of the arrays laid end to end, alternately in the two halves
    const size_t piece = std::min<size_t>(std::max<size_t>(total / 8, (size_t)4 << 20), (size_t)64 << 20);
    int rc = pinned_staging(ctx, 2 * piece);
    if (rc) return rc;
    uint8_t *const half[2] = {(uint8_t *)ctx->pin, (uint8_t *)ctx->pin + piece};
    bool pending[2] = {false, false};
    struct Run {
        const uint8_t *src;
        uint8_t *dst;
        size_t bytes, at;  // at: offset in the piece
    };
    std::vector<Run> runs;
    hipError_t e = hipSuccess;
    int k = 0;
    size_t koff = 0;  // position in segment k
    for (size_t x = 0, done = 0; done < total && e == hipSuccess; ++x, done += piece) {
        const int h = (int)(x & 1);
        const size_t len = std::min(piece, total - done);
        if (pending[h] && (e = hipEventSynchronize(ctx->ev_part[h])) != hipSuccess) break;
        runs.clear();
        for (size_t got = 0; got < len;) {
            const size_t take = std::min(segs[k].bytes - koff, len - got);
            if (take > 0)
                runs.push_back({(const uint8_t *)segs[k].src + koff, (uint8_t *)segs[k].dst + koff, take, got});
            got += take;
            koff += take;
            if (koff == segs[k].bytes) {
                ++k;
                koff = 0;
            }
        }
        parallel_for(16, [&](int64_t lo, int64_t hi, unsigned) {
            for (int64_t t = lo; t < hi; ++t) {
                const size_t b0 = len * (size_t)t / 16, b1 = len * (size_t)(t + 1) / 16;
                for (const Run &r : runs) {
                    const size_t s0 = std::max(b0, r.at), s1 = std::min(b1, r.at + r.bytes);
                    if (s1 > s0) std::memcpy(half[h] + s0, r.src + (s0 - r.at), s1 - s0);
                }
            }
        }, 2);
        for (const Run &r : runs)
            if (e == hipSuccess) e = hipMemcpyAsync(r.dst, half[h] + r.at, r.bytes, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_part[h], st);
        pending[h] = true;
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);  // copies from the staging buffer may be in flight
        return hip_fail(e, "staged_upload");
    }
    return FA_OK;
}

int staged_download(fa_ctx *ctx, void *dst, const void *src, size_t bytes, hipStream_t st) {
    if (bytes == 0) return FA_OK;
    if (bytes < kStagedMin) {
        FA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st));
        FA_HIP(hipStreamSynchronize(st));
        return FA_OK;
    }
    int rc = pinned_staging(ctx, bytes);
    if (rc) return rc;
    const int parts = (int)std::max<size_t>(1, std::min<size_t>(8, bytes / ((size_t)8 << 20)));
    uint8_t *const pin = (uint8_t *)ctx->pin;
    hipError_t e = hipSuccess;
    for (int x = 0; x < parts && e == hipSuccess; ++x) {
        const size_t b0 = bytes * (size_t)x / parts, b1 = bytes * (size_t)(x + 1) / parts;
        e = hipMemcpyAsync(pin + b0, (const uint8_t *)src + b0, b1 - b0, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_part[x], st);
    }
    for (int x = 0; x < parts && e == hipSuccess; ++x) {
        if ((e = hipEventSynchronize(ctx->ev_part[x])) != hipSuccess) break;
        const size_t b0 = bytes * (size_t)x / parts, b1 = bytes * (size_t)(x + 1) / parts;
        parallel_for(16, [&](int64_t lo, int64_t hi, unsigned) {
            for (int64_t t = lo; t < hi; ++t) {
                const size_t c0 = b0 + (b1 - b0) * (size_t)t / 16, c1 = b0 + (b1 - b0) * (size_t)(t + 1) / 16;
                if (c1 > c0) std::memcpy((uint8_t *)dst + c0, pin + c0, c1 - c0);
            }
        }, 2);
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        return hip_fail(e, "staged_download");
    }
    return FA_OK;
}

}  // namespace fa_host

extern "C" {

int fa_batch_upload(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                    const uint8_t *query_seqs, const int64_t *query_offset,
                    const int32_t *query_len, const uint8_t *ref_seqs, const int64_t *ref_offset,
                    const int32_t *ref_len, const int32_t *h0, const int32_t *band_width,
                    fa_batch **out) {
    TargetSource ts;
    ts.seqs = ref_seqs;
    ts.offset = ref_offset;
    return batch_upload(ctx, sw, njobs, query_seqs, query_offset, query_len, ts, ref_len, h0,
                        band_width, out);
}

int fa_ref_length(const fa_ref *ref, int64_t *l_pac) {
    if (!ref || !l_pac) return FA_ERR_ARG;
    *l_pac = ref->l_pac;
    return FA_OK;
}

int fa_ref_upload(fa_ctx *ctx, const uint8_t *pac, int64_t l_pac, fa_ref **out) {
    if (!ctx || !out || !pac || l_pac <= 0 || l_pac > ((int64_t)1 << 40)) return FA_ERR_ARG;
    *out = nullptr;
    FA_HIP(hipSetDevice(ctx->device));
    fa_ref *r = new (std::nothrow) fa_ref();
    if (!r) return FA_ERR_ARG;
    r->device = ctx->device;
    r->l_pac = l_pac;
    const size_t bytes = (size_t)((l_pac + 3) / 4);
    // 16 zero bytes after the last base: kernels read the .pac in aligned words
    int rc = dev_alloc(&r->d_pac, bytes + 16);
    if (!rc && (hipMemcpyAsync(r->d_pac, pac, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
                hipMemsetAsync(r->d_pac + bytes, 0, 16, ctx->stream) != hipSuccess))
        rc = FA_ERR_DEVICE;
    if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FA_ERR_DEVICE;
    if (rc) {
        if (r->d_pac) (void)hipFree(r->d_pac);
        delete r;
        return rc;
    }
    *out = r;
    return FA_OK;
}

int fa_ref_free(fa_ctx *ctx, fa_ref *ref) {
    if (!ctx || !ref) return FA_ERR_ARG;
    (void)hipSetDevice(ref->device);
    if (ref->d_pac) (void)hipFree(ref->d_pac);
    delete ref;
    return FA_OK;
}

int fa_batch_upload_ref(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                        const uint8_t *query_seqs, const int64_t *query_offset,
                        const int32_t *query_len, const fa_ref *ref, const int64_t *seg_beg,
                        const int32_t *seg_len, const uint8_t *seg_rev, const int32_t *h0,
                        const int32_t *band_width, fa_batch **out) {
    if (!ref) return FA_ERR_ARG;
    TargetSource ts;
    ts.ref = ref;
    ts.seg_beg = seg_beg;
    ts.seg_rev = seg_rev;
    return batch_upload(ctx, sw, njobs, query_seqs, query_offset, query_len, ts, seg_len, h0,
                        band_width, out);
}

int fa_batch_download_targets(fa_ctx *ctx, fa_batch *b, uint8_t *codes, const int64_t *offset) {
    if (!ctx || !b || (b->n > 0 && (!codes || !offset))) return FA_ERR_ARG;
    if (b->n == 0) return FA_OK;
    FA_HIP(hipSetDevice(ctx->device));
    const uint32_t last = b->to_host[(size_t)b->n - 1];
    const size_t words = (size_t)last + seq_words(b->tlen[(size_t)b->n - 1]);
    std::vector<uint32_t> seq2(words), seqn(words / 2);
    FA_HIP(hipMemcpy(seq2.data(), b->d_seq2, words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    FA_HIP(hipMemcpy(seqn.data(), b->d_seqn, words / 2 * sizeof(uint32_t), hipMemcpyDeviceToHost));
    parallel_for(b->n, [&](int64_t lo, int64_t hi, unsigned) {
        for (int64_t k = lo; k < hi; ++k) {
            const uint32_t o = b->to_host[(size_t)k];
            uint8_t *dst = codes + offset[k];
            for (int32_t p = 0; p < b->tlen[(size_t)k]; ++p) {
                const uint32_t c = (seq2[o + (uint32_t)(p >> 4)] >> ((p & 15) * 2)) & 3u;
                dst[p] = ((seqn[o / 2 + (uint32_t)(p >> 5)] >> (p & 31)) & 1u) ? 4 : (uint8_t)c;
            }
        }
    });
    return FA_OK;
}

namespace {

// Where fa_scalar_banded_swa_batch{,_dense} put the CIGARs and aligned bases:
// caller-given slots with counts, or back to back with offsets written.
struct SwtbOut {
    uint32_t *cigar;
    uint8_t *ref_aligned, *query_aligned;
    bool dense;
    // slots
    const int64_t *cigar_offset, *aligned_offset;
    int32_t *n_cigar, *n_ref_aligned, *n_query_aligned;
    // dense: capacities (entries) and njobs + 1 offsets
    int64_t cigar_cap, aligned_cap;
    int64_t *cigar_off_out, *ref_off_out, *query_off_out;
};

// Context slots of the CIGAR-regeneration calls.
enum SwtbSlot {
    kSwtbSeq2, kSwtbSeqN, kSwtbBlock, kSwtbSort, kSwtbTrace, kSwtbCigar, kSwtbRefAln,
    kSwtbQryAln, kSwtbDense, kSwtbOutCigar, kSwtbOutRef, kSwtbOutQry
};
static_assert(kSwtbOutQry < kSlotBatchArena, "slots of the CIGAR calls and the one-call batches");

// Per-pair arrays of a CIGAR-regeneration call in one block, laid out alike
// in the context's pinned host scratch and in one device slot: what the
// kernels read ([0, up), one upload), the per-pair results ([up, down), one
// download) and the host's own arrays ([down, end)).  256-byte aligned.
struct SwtbBlock {
    size_t qo, to, ql, tl, ww, hh, lsl, tbo, cgo, alo, key, idx, up;
    size_t out, cells, nc, nr, nq, down;
    size_t dense, job, end;
    SwtbBlock(size_t m, bool with_job) {
        size_t o = 0;
        auto take = [&o](size_t bytes) {
            const size_t at = o;
            o += (bytes + 255) & ~(size_t)255;
            return at;
        };
        qo = take(4 * m);
        to = take(4 * m);
        ql = take(4 * m);
        tl = take(4 * m);
        ww = take(4 * m);
        hh = take(4 * m);
        lsl = take(4 * m);
        tbo = take(8 * m);
        cgo = take(8 * m);
        alo = take(8 * m);
        key = take(4 * m);
        idx = take(4 * m);
        up = o;
        out = take(24 * m);
        cells = take(4 * m);
        nc = take(4 * m);
        nr = take(4 * m);
        nq = take(4 * m);
        down = o;
        dense = take(24 * m);
        job = take(with_job ? 8 * m : 0);
        end = o;
    }
};

// Grow-only pinned host scratch of the context for those arrays (a call
// synchronises its copies before it returns, so the next call may reuse it).
int host_scratch(fa_ctx *ctx, size_t bytes) {
    if (ctx->hscr_bytes >= bytes) return FA_OK;
    if (ctx->hscr) (void)hipHostFree(ctx->hscr);
    ctx->hscr = nullptr;
    ctx->hscr_bytes = 0;
    const size_t want = bytes + bytes / 8 + 4096;
    FA_HIP(hipHostMalloc(&ctx->hscr, want));
    ctx->hscr_bytes = want;
    return FA_OK;
}

// Bucket code of a pair's kernel (L in {4, 8, 16}, S < 64): band flag,
// log2(L) - 2 and S, ordered as lsl = kSwtbBand | L << 8 | S is.
constexpr int kSwtbCodes = 512;
constexpr int kSwtbKeyBits = 29;  // code (9 bits) << 20 | 255 - qlen << 12 | 4095 - tlen
constexpr int kSwtbParts = 8;     // pieces of the output download (the context's ev_part)
inline uint32_t swtb_code(uint32_t lsl) {
    const uint32_t L = (lsl >> 8) & 0xFFu, lg = L == 4 ? 0u : (L == 8 ? 1u : 2u);
    return ((lsl & fa::kSwtbBand) ? 256u : 0u) | (lg << 6) | (lsl & 63u);
}

int swtb_batch(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
               const uint8_t *query_seqs, const int64_t *query_offset, const int32_t *query_len,
               const uint8_t *ref_seqs, const int64_t *ref_offset, const int32_t *ref_len,
               const int32_t *band_width, const int32_t *h0, fa_out_score *out, uint32_t *cells,
               const SwtbOut &so) {
    int rc;
    if (njobs > (int64_t)0x7FFFFFFF) return FA_ERR_ARG;
    // validation (h0 = 0 is allowed here: region.rs:1058 passes 0) and the
    // count of non-empty jobs, in parallel; the code returned is the one of the
    // first invalid job
    std::vector<int64_t> filled(16, 0);  // non-empty jobs of each thread's range
    {
        std::vector<std::pair<int64_t, int>> first_bad(16, {njobs, FA_OK});
        parallel_for(njobs, [&](int64_t lo, int64_t hi, unsigned t) {
            int64_t nonempty = 0;
            for (int64_t k = lo; k < hi; ++k) {
                const int32_t ql = query_len[k], tl = ref_len[k];
                int c = FA_OK;
                if (ql < 0 || ql > FA_MAX_QLEN_PACKED) c = FA_ERR_QLEN;
                else if (tl < 0 || tl > FA_MAX_TLEN_PACKED) c = FA_ERR_TLEN;
                else if (band_width[k] < 0) c = FA_ERR_BAND;
                else if (h0[k] < 0 || h0[k] > (1 << 19)) c = FA_ERR_H0;
                else if (!fa_host::codes_ok(query_seqs + query_offset[k], ql) ||
                         !fa_host::codes_ok(ref_seqs + ref_offset[k], tl))
                    c = FA_ERR_CODE;
                if (c != FA_OK) {
                    first_bad[t] = {k, c};
                    return;
                }
                nonempty += ql > 0 && tl > 0;
            }
            filled[t] = nonempty;
        });
        std::pair<int64_t, int> bad{njobs, FA_OK};
        for (const auto &b : first_bad)
            if (b.first < bad.first) bad = b;
        if (bad.second != FA_OK) return bad.second;
    }
    FA_HIP(hipSetDevice(ctx->device));
    size_t m = 0;
    for (int64_t x : filled) m += (size_t)x;
    const bool all = m == (size_t)njobs;  // no empty job: pair p is job p
    // per-pair arrays in the context's pinned scratch, reused from call to call
    const SwtbBlock B(m, !all);
    if ((rc = host_scratch(ctx, B.end))) return rc;
    uint8_t *const hb = (uint8_t *)ctx->hscr;
    uint32_t *const qo = (uint32_t *)(hb + B.qo), *const to = (uint32_t *)(hb + B.to);
    int32_t *const ql = (int32_t *)(hb + B.ql), *const tl = (int32_t *)(hb + B.tl);
    int32_t *const ww = (int32_t *)(hb + B.ww), *const hh = (int32_t *)(hb + B.hh);
    uint32_t *const lsl = (uint32_t *)(hb + B.lsl), *const key = (uint32_t *)(hb + B.key);
    uint32_t *const idx = (uint32_t *)(hb + B.idx);
    uint64_t *const tbo = (uint64_t *)(hb + B.tbo), *const cgo = (uint64_t *)(hb + B.cgo);
    uint64_t *const alo = (uint64_t *)(hb + B.alo), *const dense = (uint64_t *)(hb + B.dense);
    const int32_t *const hout = (const int32_t *)(hb + B.out);
    const uint32_t *const hcells = (const uint32_t *)(hb + B.cells);
    const int32_t *const hnc = (const int32_t *)(hb + B.nc), *const hnr = (const int32_t *)(hb + B.nr);
    const int32_t *const hnq = (const int32_t *)(hb + B.nq);
    int64_t *const job = all ? nullptr : (int64_t *)(hb + B.job);
    if (!all) {
        // non-empty jobs run on the device, listed in job order (each thread's
        // from the counts of the ranges before it); empty ones get the
        // reference's zero record and an empty CIGAR (implementation.rs:17-30)
        parallel_for(njobs, [&](int64_t lo, int64_t hi, unsigned t) {
            int64_t p = 0;
            for (unsigned u = 0; u < t; ++u) p += filled[u];
            for (int64_t k = lo; k < hi; ++k) {
                if (query_len[k] > 0 && ref_len[k] > 0) {
                    job[p++] = k;
                    continue;
                }
                std::memset(&out[k], 0, sizeof(fa_out_score));
                if (cells) cells[k] = 0;
                if (!so.dense) so.n_cigar[k] = so.n_ref_aligned[k] = so.n_query_aligned[k] = 0;
            }
        });
    }
    if (m == 0) {
        if (so.dense)
            for (int64_t k = 0; k <= njobs; ++k)
                so.cigar_off_out[k] = so.ref_off_out[k] = so.query_off_out[k] = 0;
        return FA_OK;
    }
    auto jobof = [&](size_t p) -> int64_t { return all ? (int64_t)p : job[p]; };
    // per pair: kernel, sort key, sizes of its sequence words, traceback block,
    // CIGAR and aligned-base slots; per thread range: their sums, the longest
    // target and band-window query, and the pairs of each kernel
    struct Range {
        uint64_t words = 0, tbw = 0, cgw = 0, alw = 0;
        int tmax = 0, qmax = 0;
        uint32_t hist[kSwtbCodes] = {};
    };
    std::vector<Range> part(16);
    parallel_for((int64_t)m, [&](int64_t lo, int64_t hi, unsigned t) {
        Range r;
        for (int64_t p = lo; p < hi; ++p) {
            const int64_t k = jobof((size_t)p);
            const int32_t q = query_len[k], tt = ref_len[k];
            ql[p] = q;
            tl[p] = tt;
            ww[p] = band_width[k];
            hh[p] = h0[k];
            int L = 16;
            for (int c : {4, 8, 16})
                if (fa::swtb_slot_count(c, q) > 0) {
                    L = c;
                    break;
                }
            int S = fa::swtb_slot_count(L, q);
            uint32_t ls = ((uint32_t)L << 8) | (uint32_t)S;
            int bl, bs;
            if (fa::swtb_band_config(q, ww[p], L * S, &bl, &bs)) {  // band-window kernel
                L = bl;
                S = bs;
                ls = fa::kSwtbBand | ((uint32_t)L << 8) | (uint32_t)S;
                r.qmax = std::max(r.qmax, q);
            }
            r.tmax = std::max(r.tmax, tt);
            lsl[p] = ls;
            const uint32_t code = swtb_code(ls);
            ++r.hist[code];
            // ascending keys: kernels by descending lsl, then qlen and tlen descending
            key[p] = ((uint32_t)(kSwtbCodes - 1 - code) << 20) | ((uint32_t)(255 - q) << 12) |
                     (uint32_t)(4095 - tt);
            idx[p] = (uint32_t)p;
            qo[p] = seq_words(q);  // sizes here, offsets in the packing pass
            to[p] = seq_words(tt);
            tbo[p] = (uint64_t)tt * (uint64_t)L * (uint64_t)((S + 15) / 16);
            r.words += (uint64_t)qo[p] + to[p];
            r.tbw += tbo[p];
            r.cgw += (uint64_t)(q + tt + 4);
            r.alw += (uint64_t)(q + tt + 2);
        }
        part[t] = r;
    });
    uint64_t words = 0, tbw = 0, cgw = 0, alw = 0;
    int tmax = 0, qmax = 0;
    std::vector<std::array<uint64_t, 4>> start(16);  // each range's first offsets
    for (size_t t = 0; t < part.size(); ++t) {
        start[t] = {words, tbw, cgw, alw};
        words += part[t].words;
        tbw += part[t].tbw;
        cgw += part[t].cgw;
        alw += part[t].alw;
        tmax = std::max(tmax, part[t].tmax);
        qmax = std::max(qmax, part[t].qmax);
    }
    if (words > 0xFFFFFFF0ull) return FA_ERR_ARG;
    // offsets, then the packed sequences staged in the context's pinned buffer
    // (every word is written: sequences are packed back to back)
    if ((rc = pinned_staging(ctx, (size_t)words * 4 + (size_t)(words / 2) * 4))) return rc;
    uint32_t *seq2 = (uint32_t *)ctx->pin, *seqn = seq2 + words;
    parallel_for((int64_t)m, [&](int64_t lo, int64_t hi, unsigned t) {
        uint64_t wo = start[t][0], bo = start[t][1], co = start[t][2], ao = start[t][3];
        for (int64_t p = lo; p < hi; ++p) {
            const uint32_t qw = qo[p], tw = to[p];
            qo[p] = (uint32_t)wo;
            to[p] = (uint32_t)(wo + qw);
            wo += (uint64_t)qw + tw;
            const uint64_t tbs = tbo[p];
            tbo[p] = bo;
            bo += tbs;
            cgo[p] = co;
            co += (uint64_t)(ql[p] + tl[p] + 4);
            alo[p] = ao;
            ao += (uint64_t)(ql[p] + tl[p] + 2);
            const int64_t k = jobof((size_t)p);
            pack_seq(query_seqs + query_offset[k], ql[p], &seq2[qo[p]], &seqn[qo[p] / 2]);
            pack_seq(ref_seqs + ref_offset[k], tl[p], &seq2[to[p]], &seqn[to[p] / 2]);
        }
    });
    hipStream_t st = ctx->stream;
    // device buffers: grow-only slots of the context (the traceback block
    // alone is ~2.5 KB per 150-base pair), reused by the next call; the
    // per-pair block goes up in one copy from the pinned scratch, and the
    // order of the pairs is sorted on the device
    uint32_t *d_seq2, *d_seqn, *d_tb, *d_cig;
    uint8_t *db, *ds, *d_ra, *d_qa;
    size_t sort_tmp = 0;
    const size_t s_order = (4 * m + 255) & ~(size_t)255, s_tmp = 2 * s_order;
    FA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_tmp, (const uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (const uint32_t *)nullptr,
                                              (uint32_t *)nullptr, (int)m, 0, kSwtbKeyBits, st));
    if ((rc = ctx_upload(ctx, kSwtbSeq2, &d_seq2, seq2, (size_t)words)) ||
        (rc = ctx_upload(ctx, kSwtbSeqN, &d_seqn, seqn, (size_t)(words / 2))) ||
        (rc = ctx_alloc(ctx, kSwtbBlock, &db, B.down)) ||
        (rc = hip_fail(hipMemcpyAsync(db, hb, B.up, hipMemcpyHostToDevice, st), "hipMemcpyAsync")) ||
        (rc = ctx_alloc(ctx, kSwtbSort, &ds, s_tmp + sort_tmp)) ||
        (rc = hip_fail(hipcub::DeviceRadixSort::SortPairs(
                           ds + s_tmp, sort_tmp, (const uint32_t *)(db + B.key), (uint32_t *)ds,
                           (const uint32_t *)(db + B.idx), (uint32_t *)(ds + s_order), (int)m, 0,
                           kSwtbKeyBits, st),
                       "hipcub::DeviceRadixSort::SortPairs")) ||
        (rc = ctx_alloc(ctx, kSwtbTrace, &d_tb, (size_t)tbw + 32)) ||  // + a line: whole-line reads
        (rc = ctx_alloc(ctx, kSwtbCigar, &d_cig, (size_t)cgw)) ||
        (rc = ctx_alloc(ctx, kSwtbRefAln, &d_ra, (size_t)alw)) ||
        (rc = ctx_alloc(ctx, kSwtbQryAln, &d_qa, (size_t)alw))) {
        (void)hipStreamSynchronize(st);  // copies from the staging buffers may be in flight
        return rc;
    }
    fa::SwtbArgs a{};
    a.seq2 = d_seq2;
    a.seqn = d_seqn;
    a.qo = (const uint32_t *)(db + B.qo);
    a.to = (const uint32_t *)(db + B.to);
    a.qlen = (const int32_t *)(db + B.ql);
    a.tlen = (const int32_t *)(db + B.tl);
    a.w = (const int32_t *)(db + B.ww);
    a.h0 = (const int32_t *)(db + B.hh);
    a.order = (const uint32_t *)(ds + s_order);
    a.tbo = (const uint64_t *)(db + B.tbo);
    a.lsl = (const uint32_t *)(db + B.lsl);
    a.tb = d_tb;
    a.out = (int32_t *)(db + B.out);
    a.cells = (uint32_t *)(db + B.cells);
    a.cig_off = (const uint64_t *)(db + B.cgo);
    a.cigar = d_cig;
    a.n_cigar = (int32_t *)(db + B.nc);
    a.aln_off = (const uint64_t *)(db + B.alo);
    a.ref_aln = d_ra;
    a.qry_aln = d_qa;
    a.n_ref = (int32_t *)(db + B.nr);
    a.n_qry = (int32_t *)(db + B.nq);
    for (int q = 0; q < 5; ++q) {
        uint32_t v = 0;
        for (int t = 0; t < 4; ++t) v |= (uint32_t)(uint8_t)(sw->mat[t * 5 + q] + 128) << (8 * t);
        a.profw[q] = v;
    }
    a.nrow = (uint32_t)(uint8_t)(sw->mat[20] + 128);
    a.o_del = sw->o_del;
    a.e_del = sw->e_del;
    a.o_ins = sw->o_ins;
    a.e_ins = sw->e_ins;
    a.zdrop = sw->zdrop;
    a.pen_clip3 = sw->pen_clip3;
    a.tpad = ((tmax + 16) / 16) * 16;
    a.qpad = ((qmax + 16) / 16) * 16;
    hipError_t e = hipEventRecord(ctx->ev0, st);
    {  // one launch per (kernel, L, S) bucket, in the sorted order: descending codes
        uint32_t first = 0;
        for (int code = kSwtbCodes - 1; code >= 0 && e == hipSuccess; --code) {
            uint32_t cnt = 0;
            for (const Range &r : part) cnt += r.hist[code];
            if (cnt == 0) continue;
            e = fa::swtb_dp_launch(4 << ((code >> 6) & 3), code & 63, (code & 256) != 0, a, first,
                                   cnt, st);
            first += cnt;
        }
    }
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, st);
    if (e == hipSuccess) e = fa::swtb_traceback_launch(a, (uint32_t)m, st);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev2, st);
    // records and counts first (one copy into the pinned scratch); then the
    // CIGARs and aligned bases are gathered on the device into dense buffers
    // (prefix sums of the counts) so only the used entries cross PCIe
    if (e == hipSuccess)
        e = hipMemcpyAsync(hb + B.out, db + B.out, B.down - B.out, hipMemcpyDeviceToHost, st);
    {
        const hipError_t es = hipStreamSynchronize(st);  // uploads from the scratch end here
        if (e == hipSuccess) e = es;
    }
    float t_dp = 0.f, t_tb = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t_dp, ctx->ev0, ctx->ev1);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_tb, ctx->ev1, ctx->ev2);
    if (e != hipSuccess) return hip_fail(e, "fa_scalar_banded_swa_batch");
    // dense[3p..3p+2]: pair p's CIGAR, reference-base and query-base offsets
    // (per-range sums, then each range from its start)
    std::vector<std::array<uint64_t, 3>> dsum(16, {0, 0, 0});
    parallel_for((int64_t)m, [&](int64_t lo, int64_t hi, unsigned t) {
        uint64_t c = 0, r = 0, q = 0;
        for (int64_t p = lo; p < hi; ++p) {
            c += (uint64_t)hnc[p];
            r += (uint64_t)hnr[p];
            q += (uint64_t)hnq[p];
        }
        dsum[t] = {c, r, q};
    });
    uint64_t nc_all = 0, nr_all = 0, nq_all = 0;
    for (auto &s : dsum) {
        const std::array<uint64_t, 3> n3 = s;
        s = {nc_all, nr_all, nq_all};
        nc_all += n3[0];
        nr_all += n3[1];
        nq_all += n3[2];
    }
    if (so.dense && (nc_all > (uint64_t)so.cigar_cap || nr_all > (uint64_t)so.aligned_cap ||
                     nq_all > (uint64_t)so.aligned_cap))
        return FA_ERR_ARG;  // the capacities are too small for this batch
    parallel_for((int64_t)m, [&](int64_t lo, int64_t hi, unsigned t) {
        uint64_t c = dsum[t][0], r = dsum[t][1], q = dsum[t][2];
        for (int64_t p = lo; p < hi; ++p) {
            dense[3 * p] = c;
            dense[3 * p + 1] = r;
            dense[3 * p + 2] = q;
            c += (uint64_t)hnc[p];
            r += (uint64_t)hnr[p];
            q += (uint64_t)hnq[p];
        }
    });
    uint64_t *d_dense;
    uint32_t *d_ocig;
    uint8_t *d_ora, *d_oqa;
    if ((rc = ctx_upload(ctx, kSwtbDense, &d_dense, dense, 3 * m)) ||
        (rc = ctx_alloc(ctx, kSwtbOutCigar, &d_ocig, (size_t)nc_all + 1)) ||
        (rc = ctx_alloc(ctx, kSwtbOutRef, &d_ora, (size_t)nr_all + 1)) ||
        (rc = ctx_alloc(ctx, kSwtbOutQry, &d_oqa, (size_t)nq_all + 1))) {
        (void)hipStreamSynchronize(st);
        return rc;
    }
    // downloads land in the pinned staging buffer (the uploads from it are
    // complete: the stream was synchronised above)
    if ((rc = pinned_staging(ctx, (size_t)nc_all * 4 + (size_t)nr_all + (size_t)nq_all + 64))) {
        (void)hipStreamSynchronize(st);  // the offsets' upload from the scratch may be in flight
        return rc;
    }
    uint32_t *hcig = (uint32_t *)ctx->pin;
    uint8_t *hra = (uint8_t *)(hcig + nc_all), *hqa = hra + nr_all;
    // the gathered CIGARs and bases come back in pieces of consecutive pairs
    // (consecutive in the dense buffers too); the host threads copy each out
    // as soon as it has landed, while the next ones are still in flight (the
    // dense call only: the slot call's pair-by-pair copies measured slower so)
    const int parts =
        so.dense ? (int)std::max<size_t>(1, std::min<size_t>(kSwtbParts, m / 65536)) : 1;
    auto first_pair = [&](int x) { return (size_t)((uint64_t)m * (uint64_t)x / (uint64_t)parts); };
    auto cut = [&](size_t p, int f) -> uint64_t {  // entry of pair p in dense buffer f
        return p < m ? dense[3 * p + f] : (f == 0 ? nc_all : (f == 1 ? nr_all : nq_all));
    };
    e = fa::swtb_compact_launch(a, (uint32_t)m, d_dense, d_ocig, d_ora, d_oqa, st);
    for (int x = 0; x < parts && e == hipSuccess; ++x) {
        const size_t p0 = first_pair(x), p1 = first_pair(x + 1);
        const uint64_t c0 = cut(p0, 0), c1 = cut(p1, 0), r0 = cut(p0, 1), r1 = cut(p1, 1);
        const uint64_t q0 = cut(p0, 2), q1 = cut(p1, 2);
        if (c1 > c0) e = hipMemcpyAsync(hcig + c0, d_ocig + c0, (c1 - c0) * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && r1 > r0) e = hipMemcpyAsync(hra + r0, d_ora + r0, r1 - r0, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && q1 > q0) e = hipMemcpyAsync(hqa + q0, d_oqa + q0, q1 - q0, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipEventRecord(ctx->ev_part[x], st);
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);
        return hip_fail(e, "fa_scalar_banded_swa_batch");
    }
    ctx->last_dp_ms = t_dp;
    ctx->last_tb_ms = t_tb;
    if (so.dense) {
        // offsets in job order (an empty job adds nothing; each thread range
        // starts at the first pair of a job not before it)
        parallel_for(njobs, [&](int64_t lo, int64_t hi, unsigned) {
            size_t p = all ? (size_t)lo : (size_t)(std::lower_bound(job, job + m, lo) - job);
            for (int64_t k = lo; k < hi; ++k) {
                const bool here = p < m && jobof(p) == k;
                so.cigar_off_out[k] = (int64_t)cut(p, 0);
                so.ref_off_out[k] = (int64_t)cut(p, 1);
                so.query_off_out[k] = (int64_t)cut(p, 2);
                p += here;
            }
        });
        so.cigar_off_out[njobs] = (int64_t)nc_all;
        so.ref_off_out[njobs] = (int64_t)nr_all;
        so.query_off_out[njobs] = (int64_t)nq_all;
    }
    for (int x = 0; x < parts; ++x) {
        e = hipEventSynchronize(ctx->ev_part[x]);
        if (e != hipSuccess) {
            (void)hipStreamSynchronize(st);
            return hip_fail(e, "fa_scalar_banded_swa_batch");
        }
        // records of the piece's pairs, then their CIGARs and bases: into the
        // caller's slots pair by pair, or (dense) the thread's run of pairs
        // in one copy per buffer, so only the pages they fill are touched
        const size_t p0 = first_pair(x), p1 = first_pair(x + 1);
        parallel_for((int64_t)(p1 - p0), [&](int64_t lo, int64_t hi, unsigned) {
            const size_t a0 = p0 + (size_t)lo, a1 = p0 + (size_t)hi;
            for (size_t p = a0; p < a1; ++p) {
                const int64_t k = jobof(p);
                const int32_t *o = &hout[p * 6];
                out[k].score = o[0];
                out[k].target_end_pos = o[1];
                out[k].gtarget_end_pos = o[2];
                out[k].query_end_pos = o[3];
                out[k].global_score = o[4];
                out[k].max_offset = o[5];
                if (cells) cells[k] = hcells[p];
                if (so.dense) continue;
                so.n_cigar[k] = hnc[p];
                so.n_ref_aligned[k] = hnr[p];
                so.n_query_aligned[k] = hnq[p];
                std::memcpy(so.cigar + so.cigar_offset[k], hcig + dense[3 * p], (size_t)hnc[p] * 4);
                std::memcpy(so.ref_aligned + so.aligned_offset[k], hra + dense[3 * p + 1], (size_t)hnr[p]);
                std::memcpy(so.query_aligned + so.aligned_offset[k], hqa + dense[3 * p + 2],
                            (size_t)hnq[p]);
            }
            if (!so.dense) return;
            const uint64_t c0 = cut(a0, 0), c1 = cut(a1, 0), r0 = cut(a0, 1), r1 = cut(a1, 1);
            const uint64_t q0 = cut(a0, 2), q1 = cut(a1, 2);
            if (c1 > c0) std::memcpy(so.cigar + c0, hcig + c0, (size_t)(c1 - c0) * 4);
            if (r1 > r0) std::memcpy(so.ref_aligned + r0, hra + r0, (size_t)(r1 - r0));
            if (q1 > q0) std::memcpy(so.query_aligned + q0, hqa + q0, (size_t)(q1 - q0));
        });
    }
    return FA_OK;
}

}  // namespace

int fa_scalar_banded_swa_batch(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                               const uint8_t *query_seqs, const int64_t *query_offset,
                               const int32_t *query_len, const uint8_t *ref_seqs,
                               const int64_t *ref_offset, const int32_t *ref_len,
                               const int32_t *band_width, const int32_t *h0, fa_out_score *out,
                               uint32_t *cells, uint32_t *cigar, const int64_t *cigar_offset,
                               int32_t *n_cigar, uint8_t *ref_aligned, uint8_t *query_aligned,
                               const int64_t *aligned_offset, int32_t *n_ref_aligned,
                               int32_t *n_query_aligned) {
    if (!ctx || njobs < 0) return FA_ERR_ARG;
    int rc = check_scoring(sw);
    if (rc) return rc;
    if (njobs == 0) return FA_OK;
    if (!query_seqs || !query_offset || !query_len || !ref_seqs || !ref_offset || !ref_len ||
        !band_width || !h0 || !out || !cigar || !cigar_offset || !n_cigar || !ref_aligned ||
        !query_aligned || !aligned_offset || !n_ref_aligned || !n_query_aligned)
        return FA_ERR_ARG;
    SwtbOut so{};
    so.cigar = cigar;
    so.ref_aligned = ref_aligned;
    so.query_aligned = query_aligned;
    so.cigar_offset = cigar_offset;
    so.aligned_offset = aligned_offset;
    so.n_cigar = n_cigar;
    so.n_ref_aligned = n_ref_aligned;
    so.n_query_aligned = n_query_aligned;
    return swtb_batch(ctx, sw, njobs, query_seqs, query_offset, query_len, ref_seqs, ref_offset,
                      ref_len, band_width, h0, out, cells, so);
}

int fa_scalar_banded_swa_batch_dense(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                                     const uint8_t *query_seqs, const int64_t *query_offset,
                                     const int32_t *query_len, const uint8_t *ref_seqs,
                                     const int64_t *ref_offset, const int32_t *ref_len,
                                     const int32_t *band_width, const int32_t *h0,
                                     fa_out_score *out, uint32_t *cells, uint32_t *cigar,
                                     int64_t cigar_cap, int64_t *cigar_offset,
                                     uint8_t *ref_aligned, int64_t *ref_aligned_offset,
                                     uint8_t *query_aligned, int64_t *query_aligned_offset,
                                     int64_t aligned_cap) {
    if (!ctx || njobs < 0 || cigar_cap < 0 || aligned_cap < 0) return FA_ERR_ARG;
    int rc = check_scoring(sw);
    if (rc) return rc;
    if (!cigar_offset || !ref_aligned_offset || !query_aligned_offset) return FA_ERR_ARG;
    if (njobs == 0) {
        cigar_offset[0] = ref_aligned_offset[0] = query_aligned_offset[0] = 0;
        return FA_OK;
    }
    if (!query_seqs || !query_offset || !query_len || !ref_seqs || !ref_offset || !ref_len ||
        !band_width || !h0 || !out || !cigar || !ref_aligned || !query_aligned)
        return FA_ERR_ARG;
    SwtbOut so{};
    so.cigar = cigar;
    so.ref_aligned = ref_aligned;
    so.query_aligned = query_aligned;
    so.dense = true;
    so.cigar_cap = cigar_cap;
    so.aligned_cap = aligned_cap;
    so.cigar_off_out = cigar_offset;
    so.ref_off_out = ref_aligned_offset;
    so.query_off_out = query_aligned_offset;
    return swtb_batch(ctx, sw, njobs, query_seqs, query_offset, query_len, ref_seqs, ref_offset,
                      ref_len, band_width, h0, out, cells, so);
}

int fa_scalar_banded_swa_last_ms(fa_ctx *ctx, double *dp_ms, double *traceback_ms) {
    if (!ctx) return FA_ERR_ARG;
    if (dp_ms) *dp_ms = ctx->last_dp_ms;
    if (traceback_ms) *traceback_ms = ctx->last_tb_ms;
    return FA_OK;
}

int fa_batch_set_cost(fa_ctx *ctx, fa_batch *b, const float *cost) {
    if (!ctx || !b || (!cost && b->n > 0)) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc;
    if (!b->d_pred && (rc = dev_alloc(&b->d_pred, (size_t)b->n))) return rc;
    if (b->n)
        FA_HIP(hipMemcpyAsync(b->d_pred, cost, (size_t)b->n * sizeof(float), hipMemcpyHostToDevice,
                              ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    b->cost_set = true;
    b->plan_dirty = true;  // the next run re-plans
    return FA_OK;
}

int fa_batch_set_launch(fa_ctx *ctx, fa_batch *b, int32_t lanes_per_pair, int32_t chunk_pairs,
                        int32_t waves_per_cu) {
    if (!ctx || !b) return FA_ERR_ARG;
    if (lanes_per_pair != 0 && lanes_per_pair != 4 && lanes_per_pair != 8 && lanes_per_pair != 16)
        return FA_ERR_ARG;
    if (chunk_pairs < 0 || waves_per_cu < 0) return FA_ERR_ARG;
    if (b->lanes != lanes_per_pair) b->packed_flags = 0xFFFFFFFFu;
    drop_graph(b);
    b->lanes = lanes_per_pair;
    b->chunk = chunk_pairs;
    b->waves_per_cu = waves_per_cu;
    return FA_OK;
}

int fa_batch_set_streams(fa_ctx *ctx, fa_batch *b, int32_t streams) {
    if (!ctx || !b || streams < 1 || streams > 4) return FA_ERR_ARG;
    b->streams = streams;
    drop_graph(b);
    return FA_OK;
}

int fa_batch_run(fa_ctx *ctx, fa_batch *b, uint32_t flags) {
    if (!ctx || !b) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc = run_batch(ctx, b, flags, nullptr);
    if (rc) return rc;
    FA_HIP(hipStreamSynchronize(ctx->stream));
    return FA_OK;
}

int fa_batch_time(fa_ctx *ctx, fa_batch *b, uint32_t flags, int32_t reps, double *ms_per_run,
                  double *ms_dominant_launch, int32_t *launches_per_run) {
    if (!ctx || !b || reps <= 0) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc = (flags & FA_RUN_ALIGN2) ? FA_OK : plan_packing(ctx, b, flags);
    if (rc) return rc;
    FA_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    int nl = 0;
    for (int r = 0; r < reps; ++r) {
        rc = run_batch(ctx, b, flags, &nl);
        if (rc) return rc;
    }
    FA_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    FA_HIP(hipEventSynchronize(ctx->ev1));
    float ms = 0.f;
    FA_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (ms_per_run) *ms_per_run = ms / reps;
    if (launches_per_run) *launches_per_run = nl;
    if (ms_dominant_launch && (flags & FA_RUN_ALIGN2)) {
        *ms_dominant_launch = ms / reps;
    } else if (ms_dominant_launch) {
        int32_t nl2 = 0;
        double ms_l[64];
        rc = fa_batch_profile_launches(ctx, b, flags, 64, &nl2, nullptr, nullptr, nullptr, nullptr,
                                       ms_l);
        if (rc) return rc;
        double best = 0.0;
        for (int k = 0; k < nl2 && k < 64; ++k) best = std::max(best, ms_l[k]);
        *ms_dominant_launch = best;
    }
    return FA_OK;
}

int fa_batch_profile_launches(fa_ctx *ctx, fa_batch *b, uint32_t flags, int32_t cap,
                              int32_t *nlaunch, int32_t *slots, int32_t *kernel, int64_t *first,
                              int64_t *count, double *ms) {
    if (!ctx || !b || !nlaunch || cap < 0) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    if (flags & FA_RUN_ALIGN2) {
        std::vector<MateLaunch> rec;
        int nl = 0;
        const int rc = run_align2(ctx, b, flags, &nl, &rec);
        if (rc) return rc;
        for (size_t k = 0; k < rec.size() && (int)k < cap; ++k) {
            if (slots) slots[k] = rec[k].slen;
            if (kernel) kernel[k] = rec[k].kernel;
            if (first) first[k] = rec[k].first;
            if (count) count[k] = rec[k].count;
            if (ms) ms[k] = rec[k].ms;
        }
        *nlaunch = (int)rec.size();
        return FA_OK;
    }
    int rc = plan_packing(ctx, b, flags);
    if (rc) return rc;
    int k = 0;
    if (flags & FA_LAUNCH_PERSISTENT)
        FA_HIP(fa::queue_reset(b->d_queue, (uint32_t)b->buckets.size(), ctx->stream));
    for (const Bucket &bk : b->buckets) {
        FA_HIP(hipEventRecord(ctx->ev0, ctx->stream));
        FA_HIP(launch_bucket(ctx, b, flags, (size_t)k, ctx->stream));
        FA_HIP(hipEventRecord(ctx->ev1, ctx->stream));
        FA_HIP(hipEventSynchronize(ctx->ev1));
        float t = 0.f;
        FA_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
        if (k < cap) {
            if (slots) slots[k] = bk.S;
            if (kernel) kernel[k] = bk.kind << 8 | bk.L;
            if (first) first[k] = bk.first;
            if (count) count[k] = bk.count;
            if (ms) ms[k] = t;
        }
        ++k;
    }
    *nlaunch = k;
    return FA_OK;
}

int fa_valu_peak(fa_ctx *ctx, int32_t packed16, double *tera_ops) {
    if (!ctx || !tera_ops) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    const int blocks = ctx->cus * 8;  // 32 waves per CU
    const int iters = 512;
    uint32_t *out = nullptr;
    FA_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint32_t)));
    hipError_t e = fa::valu_probe_launch(packed16 != 0, out, blocks, 8, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) e = fa::valu_probe_launch(packed16 != 0, out, blocks, iters, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(out);
    if (e != hipSuccess) return FA_ERR_DEVICE;
    const double instr = (double)blocks * 256.0 * iters * fa::kValuProbeUnroll * 16.0;
    *tera_ops = instr * (packed16 ? 2.0 : 1.0) / (ms * 1e-3) / 1e12;
    return FA_OK;
}

int fa_hbm_copy(fa_ctx *ctx, int64_t bytes, double *gbytes_per_s) {
    if (!ctx || !gbytes_per_s || bytes < 16 || bytes % 16) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    void *src = nullptr, *dst = nullptr;
    FA_HIP(hipMalloc(&src, (size_t)bytes));
    hipError_t e = hipMalloc(&dst, (size_t)bytes);
    const int blocks = ctx->cus * 16;
    if (e == hipSuccess) e = hipMemsetAsync(src, 1, (size_t)bytes, ctx->stream);
    if (e == hipSuccess) e = fa::hbm_copy_launch(src, dst, (size_t)bytes, blocks, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, ctx->stream);
    if (e == hipSuccess) e = fa::hbm_copy_launch(src, dst, (size_t)bytes, blocks, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(ctx->ev1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
    (void)hipFree(src);
    if (dst) (void)hipFree(dst);
    if (e != hipSuccess) return FA_ERR_DEVICE;
    *gbytes_per_s = 2.0 * (double)bytes / (ms * 1e-3) / 1e9;
    return FA_OK;
}

int fa_batch_get_order(fa_ctx *ctx, fa_batch *b, uint32_t flags, uint32_t *order) {
    if (!ctx || !b || (!order && b->n)) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    if (flags & FA_RUN_ALIGN2) {
        if (!b->align2_set) return FA_ERR_ARG;
        if (b->n)
            FA_HIP(hipMemcpy(order, b->margs.order, (size_t)b->n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
        return FA_OK;
    }
    int rc = plan_packing(ctx, b, flags);
    if (rc) return rc;
    // the plan's sort runs on the context's stream (non-blocking): copy there
    if (b->n)
        FA_HIP(hipMemcpyAsync(order, b->d_order, (size_t)b->n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                              ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    return FA_OK;
}

int fa_batch_download(fa_ctx *ctx, fa_batch *b, fa_out_score *out, uint32_t *cells) {
    if (!ctx || !b) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    static_assert(sizeof(fa_out_score) == 24, "OutScore layout");
    int rc = ensure_records(ctx, b);
    if (rc) return rc;
    if (out && b->n &&
        (rc = fa_host::staged_download(ctx, out, b->d_out, (size_t)b->n * sizeof(fa_out_score), ctx->stream)))
        return rc;
    if (cells && b->n &&
        (rc = fa_host::staged_download(ctx, cells, b->d_cells, (size_t)b->n * sizeof(uint32_t), ctx->stream)))
        return rc;
    return FA_OK;
}

int fa_batch_overflow_count(fa_ctx *ctx, fa_batch *b, int64_t *count) {
    if (!ctx || !b || !count) return FA_ERR_ARG;
    *count = b->last_overflow;
    return FA_OK;
}

int fa_batch_free(fa_ctx *ctx, fa_batch *b) {
    if (!ctx || !b) return FA_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    batch_release(b);
    return FA_OK;
}

int fa_execute_batch_simd_scoring(fa_ctx *ctx, const fa_banded_pairwise_sw *sw, int64_t njobs,
                                  const uint8_t *query_seqs, const int64_t *query_offset,
                                  const int32_t *query_len, const uint8_t *ref_seqs,
                                  const int64_t *ref_offset, const int32_t *ref_len,
                                  const int32_t *h0, const int32_t *band_width, uint32_t flags,
                                  fa_out_score *out, uint32_t *cells, int32_t *per_job_rc) {
    if (!ctx || njobs < 0 || (njobs > 0 && !out)) return FA_ERR_ARG;
    if (njobs == 0) return check_scoring(sw);
    int rc = check_scoring(sw);
    if (rc) return rc;
    if (!query_seqs || !query_offset || !query_len || !ref_seqs || !ref_offset || !ref_len ||
        !h0 || !band_width)
        return FA_ERR_ARG;
    // per-job validation (on the host threads): invalid jobs are reported and
    // left out of the launch.  First the checks that read no sequence; the
    // codes are checked while the upload packs them, and only when one is
    // out of range are the jobs validated again in full, to find which.
    std::vector<int8_t> vrc((size_t)njobs);
    auto validate = [&](bool codes) {
        parallel_for(njobs, [&](int64_t lo, int64_t hi, unsigned) {
            for (int64_t k = lo; k < hi; ++k)
                vrc[(size_t)k] = (int8_t)(codes ? validate_job(query_len[k], ref_len[k], h0[k], band_width[k],
                                                               query_seqs + query_offset[k], ref_seqs + ref_offset[k])
                                                : validate_scalar(query_len[k], ref_len[k], h0[k], band_width[k]));
        });
    };
    validate(false);
    std::vector<int64_t> keep;
    int first_err = FA_OK;
    auto apply = [&]() {
        keep.clear();
        keep.reserve((size_t)njobs);
        first_err = FA_OK;
        for (int64_t k = 0; k < njobs; ++k) {
            const int r = vrc[(size_t)k];
            if (per_job_rc) per_job_rc[k] = r;
            if (r == FA_OK)
                keep.push_back(k);
            else if (first_err == FA_OK)
                first_err = r;
            if (r != FA_OK) {
                std::memset(&out[k], 0, sizeof(fa_out_score));
                if (cells) cells[k] = 0;
            }
        }
    };
    apply();
    for (int attempt = 0; attempt < 2 && !keep.empty(); ++attempt) {
        const bool all = (int64_t)keep.size() == njobs;
        std::vector<int64_t> qo2, to2;
        std::vector<int32_t> ql2, tl2, h02, w2;
        const int64_t *qo = query_offset, *to = ref_offset;
        const int32_t *ql = query_len, *tl = ref_len, *hh = h0, *ww = band_width;
        if (!all) {
            for (int64_t k : keep) {
                qo2.push_back(query_offset[k]);
                to2.push_back(ref_offset[k]);
                ql2.push_back(query_len[k]);
                tl2.push_back(ref_len[k]);
                h02.push_back(h0[k]);
                w2.push_back(band_width[k]);
            }
            qo = qo2.data();
            to = to2.data();
            ql = ql2.data();
            tl = tl2.data();
            hh = h02.data();
            ww = w2.data();
        }
        fa_batch *b = nullptr;
        TargetSource ts;
        ts.seqs = ref_seqs;
        ts.offset = to;
        ts.validated = attempt == 1;  // the first upload checks the codes while packing
        rc = batch_upload(ctx, sw, (int64_t)keep.size(), query_seqs, qo, ql, ts, tl, hh, ww, &b, true);
        if (rc == FA_ERR_CODE && attempt == 0) {
            validate(true);  // which jobs: the full checks, then the upload again without them
            apply();
            continue;
        }
        if (rc) return rc;
        rc = fa_batch_run(ctx, b, flags);
        std::vector<fa_out_score> o2;
        std::vector<uint32_t> c2;
        if (!rc) {
            if (all) {
                rc = fa_batch_download(ctx, b, out, cells);
            } else {
                o2.resize(keep.size());
                c2.resize(keep.size());
                rc = fa_batch_download(ctx, b, o2.data(), c2.data());
                if (!rc)
                    for (size_t x = 0; x < keep.size(); ++x) {
                        out[keep[x]] = o2[x];
                        if (cells) cells[keep[x]] = c2[x];
                    }
            }
        }
        fa_batch_free(ctx, b);
        if (rc) return rc;
        break;
    }
    return first_err;
}

static int soa_common(fa_ctx *ctx, int32_t lanes, int32_t num_jobs, const int32_t *qlen,
                      const int32_t *tlen, const int32_t *h0, const int32_t *w, int32_t o_del,
                      int32_t e_del, int32_t o_ins, int32_t e_ins, int32_t zdrop,
                      int32_t end_bonus, const int8_t *mat, int32_t m, fa_out_score *out,
                      std::vector<uint8_t> &q, std::vector<uint8_t> &t) {
    fa_banded_pairwise_sw sw{};
    sw.m = m;
    sw.end_bonus = end_bonus;
    sw.zdrop = zdrop;
    sw.o_del = o_del;
    sw.e_del = e_del;
    sw.o_ins = o_ins;
    sw.e_ins = e_ins;
    std::memcpy(sw.mat, mat, 25);
    std::vector<int64_t> qo(num_jobs), to(num_jobs);
    int64_t qacc = 0, tacc = 0;
    for (int32_t k = 0; k < num_jobs; ++k) {
        qo[k] = qacc;
        qacc += std::max(qlen[k], 0);
        to[k] = tacc;
        tacc += std::max(tlen[k], 0);
    }
    (void)lanes;
    q.push_back(0);  // keep data() non-null for empty jobs; offsets never reach it
    t.push_back(0);
    return fa_execute_batch_simd_scoring(ctx, &sw, num_jobs, q.data(), qo.data(), qlen, t.data(),
                                         to.data(), tlen, h0, w, FA_SCORE_I16 | FA_PACK_LENGTH,
                                         out, nullptr, nullptr);
}

int fa_simd_banded_swa_batch_soa(fa_ctx *ctx, int32_t lanes, const uint8_t *query_soa,
                                 const uint8_t *target_soa, const int32_t *qlen,
                                 const int32_t *tlen, const int32_t *h0, const int32_t *w,
                                 int32_t num_jobs, int32_t o_del, int32_t e_del, int32_t o_ins,
                                 int32_t e_ins, int32_t zdrop, int32_t end_bonus,
                                 const int8_t *mat, int32_t m, fa_out_score *out) {
    if (!ctx || lanes <= 0 || num_jobs < 0 || num_jobs > lanes || !mat) return FA_ERR_ARG;
    if (num_jobs > 0 && (!query_soa || !target_soa || !qlen || !tlen || !h0 || !w || !out))
        return FA_ERR_ARG;
    std::vector<uint8_t> q(1, 0), t(1, 0);
    q.clear();
    t.clear();
    for (int32_t k = 0; k < num_jobs; ++k) {
        for (int32_t p = 0; p < qlen[k]; ++p) q.push_back(query_soa[(size_t)p * lanes + k]);
        for (int32_t p = 0; p < tlen[k]; ++p) t.push_back(target_soa[(size_t)p * lanes + k]);
    }
    return soa_common(ctx, lanes, num_jobs, qlen, tlen, h0, w, o_del, e_del, o_ins, e_ins, zdrop,
                      end_bonus, mat, m, out, q, t);
}

int fa_simd_banded_swa_batch_int16_soa(fa_ctx *ctx, int32_t lanes, const int16_t *query_soa,
                                       const int16_t *target_soa, const int32_t *qlen,
                                       const int32_t *tlen, const int32_t *h0, const int32_t *w,
                                       int32_t num_jobs, int32_t o_del, int32_t e_del,
                                       int32_t o_ins, int32_t e_ins, int32_t zdrop,
                                       int32_t end_bonus, const int8_t *mat, int32_t m,
                                       fa_out_score *out) {
    if (!ctx || lanes <= 0 || num_jobs < 0 || num_jobs > lanes || !mat) return FA_ERR_ARG;
    if (num_jobs > 0 && (!query_soa || !target_soa || !qlen || !tlen || !h0 || !w || !out))
        return FA_ERR_ARG;
    std::vector<uint8_t> q, t;
    for (int32_t k = 0; k < num_jobs; ++k) {
        for (int32_t p = 0; p < qlen[k]; ++p) {
            int16_t v = query_soa[(size_t)p * lanes + k];
            q.push_back(v < 0 || v > 255 ? 255 : (uint8_t)v);
        }
        for (int32_t p = 0; p < tlen[k]; ++p) {
            int16_t v = target_soa[(size_t)p * lanes + k];
            t.push_back(v < 0 || v > 255 ? 255 : (uint8_t)v);
        }
    }
    return soa_common(ctx, lanes, num_jobs, qlen, tlen, h0, w, o_del, e_del, o_ins, e_ins, zdrop,
                      end_bonus, mat, m, out, q, t);
}

// ---- mate-rescue local SW (ksw_align2, 16-bit and 8-bit modes) ---------------------

int fa_batch_set_align2(fa_ctx *ctx, fa_batch *b, const int8_t *mat, int32_t m, int32_t o_del,
                        int32_t e_del, int32_t o_ins, int32_t e_ins, uint32_t xtra) {
    if (!ctx || !b || !mat) return FA_ERR_ARG;
    if (m != 5) return FA_ERR_SCORING;
    const bool byte_mode = (xtra & fa::kKswXByte) != 0;
    if (o_del < 0 || e_del < 0 || o_ins < 0 || e_ins < 0 || o_del > 4096 || o_ins > 4096 ||
        e_del > 4096 || e_ins > 4096)
        return FA_ERR_SCORING;
    for (int q = 0; q < 5; ++q)
        if (mat[20 + q] != 0) return FA_ERR_SCORING;  // the N target row must score 0
    int qmax = 0, qmin = 127;  // ksw_qinit's scan (ksw_affine_gap.rs:87-99)
    for (int k = 0; k < 25; ++k) {
        qmax = std::max(qmax, (int)mat[k]);
        qmin = std::min(qmin, (int)mat[k]);
    }
    if (qmax < 1) return FA_ERR_SCORING;
    // 8-bit mode: profile bytes mat + shift with shift = (u8)(256 - qmin); the zero N
    // row makes qmin <= 0, so every byte is mat - qmin in 0..255 and never wraps
    const int shift = (256 - qmin) & 0xFF;
    if (b->tmax > 2032) return FA_ERR_TLEN;
    FA_HIP(hipSetDevice(ctx->device));
    const int64_t n = b->n;
    // bucket by segments per lane slen = ceil(qlen / lanes), longest first: the
    // packed kernel runs two pairs per 8-lane group in the reference's stripe
    // geometry (8 lanes; 16 byte lanes in the 8-bit mode, two per GPU lane)
    const int lanes = byte_mode ? 16 : 8;
    std::vector<int> sk((size_t)n);
    for (int64_t k = 0; k < n; ++k) {
        const int ql = b->qlen[(size_t)k];
        if (ql < 1 || ql > FA_MAX_QLEN_PACKED) return FA_ERR_QLEN;
        sk[(size_t)k] = (ql + lanes - 1) / lanes;
    }
    // stable counting sort by slen, longest first
    std::vector<uint32_t> order((size_t)n);
    {
        std::vector<int64_t> start(fa::kMateMaxSlen + 2, 0);
        for (int64_t k = 0; k < n; ++k) ++start[(size_t)(fa::kMateMaxSlen - sk[(size_t)k] + 1)];
        for (int x = 0; x <= fa::kMateMaxSlen; ++x) start[(size_t)x + 1] += start[(size_t)x];
        for (int64_t k = 0; k < n; ++k)
            order[(size_t)(start[(size_t)(fa::kMateMaxSlen - sk[(size_t)k])]++)] = (uint32_t)k;
    }
    std::vector<Bucket> bk;
    for (int64_t k = 0; k < n;) {
        const int sl = sk[order[(size_t)k]];
        int64_t e = k;
        while (e < n && sk[order[(size_t)e]] == sl) ++e;
        Bucket mb{sl, (uint32_t)k, (uint32_t)(e - k), (double)(e - k)};
        mb.slen = sl;
        bk.push_back(mb);
        k = e;
    }
    int rc;
    if (!b->d_mout && (rc = dev_alloc(&b->d_mout, (size_t)n * 7))) return rc;
    if (!b->d_mcells && (rc = dev_alloc(&b->d_mcells, (size_t)n))) return rc;
    const uint32_t stride = (uint32_t)std::max(16, ((b->tmax + 15) / 16) * 16);
    const uint32_t slots = (uint32_t)std::min<int64_t>(n, kMateChunk);
    if (b->log_slots < slots || !b->d_imax_log) {
        if (b->d_imax_log) (void)hipFree(b->d_imax_log);
        b->d_imax_log = nullptr;
        if ((rc = dev_alloc(&b->d_imax_log, (size_t)slots * stride + 64))) return rc;
        b->log_slots = slots;
    }
    // the order buffer of the extension path is reused: the mate order is its own upload
    uint32_t *d_morder = nullptr;
    if ((rc = dev_upload(&d_morder, order.data(), order.size(), ctx->stream))) return rc;
    FA_HIP(hipStreamSynchronize(ctx->stream));
    if (b->margs.order) (void)hipFree(const_cast<uint32_t *>(b->margs.order));
    fa::MateArgs &a = b->margs;
    a.seq2 = b->d_seq2;
    a.seqn = b->d_seqn;
    a.qo = b->d_qo;
    a.to = b->d_to;
    a.qlen = b->d_qlen;
    a.tlen = b->d_tlen;
    a.order = d_morder;
    a.out = b->d_mout;
    a.cells = b->d_mcells;
    a.imax_log = b->d_imax_log;
    a.log_stride = stride;
    for (int q = 0; q < 5; ++q) {
        uint32_t v = 0;
        for (int t = 0; t < 4; ++t) v |= (uint32_t)(uint8_t)(mat[t * 5 + q] + 128) << (8 * t);
        a.profw[q] = v;
    }
    a.o_del = o_del;
    a.e_del = e_del;
    a.o_ins = o_ins;
    a.e_ins = e_ins;
    a.xtra = xtra;
    a.qmax = qmax;
    a.byte_mode = byte_mode ? 1 : 0;
    a.ovf = 255 - shift;
    b->mbuckets.swap(bk);
    b->align2_set = true;
    return FA_OK;
}

int fa_batch_download_align2(fa_ctx *ctx, fa_batch *b, fa_kswr *out, uint32_t *cells) {
    if (!ctx || !b || !b->align2_set) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    static_assert(sizeof(fa_kswr) == 28, "Kswr layout");
    int rc;
    if (out && b->n &&
        (rc = fa_host::staged_download(ctx, out, b->d_mout, (size_t)b->n * sizeof(fa_kswr), ctx->stream)))
        return rc;
    if (cells && b->n &&
        (rc = fa_host::staged_download(ctx, cells, b->d_mcells, (size_t)b->n * sizeof(uint32_t), ctx->stream)))
        return rc;
    return FA_OK;
}

int fa_ksw_align2_batch(fa_ctx *ctx, int64_t n, const uint8_t *query_seqs,
                        const int64_t *query_offset, const int32_t *query_len,
                        const uint8_t *target_seqs, const int64_t *target_offset,
                        const int32_t *target_len, int32_t m, const int8_t *mat, int32_t o_del,
                        int32_t e_del, int32_t o_ins, int32_t e_ins, uint32_t xtra, fa_kswr *out,
                        uint32_t *cells) {
    if (!ctx || n < 0 || (n > 0 && !out)) return FA_ERR_ARG;
    if (n == 0) return FA_OK;
    fa_banded_pairwise_sw sw{};
    sw.m = 5;
    sw.e_del = sw.e_ins = 1;
    std::vector<int32_t> ones((size_t)n, 1), zeros((size_t)n, 0);
    fa_batch *b = nullptr;
    // a batch for this call only: its arrays come from the context's arena
    TargetSource ts;
    ts.seqs = target_seqs;
    ts.offset = target_offset;
    int rc = batch_upload(ctx, &sw, n, query_seqs, query_offset, query_len, ts, target_len,
                          ones.data(), zeros.data(), &b, true);
    if (rc) return rc;
    rc = fa_batch_set_align2(ctx, b, mat, m, o_del, e_del, o_ins, e_ins, xtra);
    if (!rc) rc = fa_batch_run(ctx, b, FA_RUN_ALIGN2);
    if (!rc) rc = fa_batch_download_align2(ctx, b, out, cells);
    fa_batch_free(ctx, b);
    return rc;
}

// ---- learned work predictor ---------------------------------------------------

struct fa_mlp {
    float *master = nullptr, *m = nullptr, *v = nullptr, *norm = nullptr;
    void *wh = nullptr;  // fp16 weights
    float *partial = nullptr;
    uint32_t partial_blocks = 0;
    float *grad = nullptr, *loss = nullptr;
    float *losses = nullptr;  // per-minibatch loss sums of one fa_mlp_train call
    int64_t losses_n = 0;
    uint32_t *perm = nullptr;
    int64_t perm_n = 0;
    int64_t step = 0;
};

static void mlp_release(fa_mlp *m) {
    void *ptrs[] = {m->master, m->m, m->v, m->norm, m->wh, m->partial, m->grad, m->loss, m->perm,
                    m->losses};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    delete m;
}

int fa_mlp_create(fa_ctx *ctx, fa_mlp **out) {
    if (!ctx || !out) return FA_ERR_ARG;
    *out = nullptr;
    FA_HIP(hipSetDevice(ctx->device));
    fa_mlp *m = new (std::nothrow) fa_mlp();
    if (!m) return FA_ERR_ARG;
    int rc;
    if ((rc = dev_alloc(&m->master, fa::kMlpParams)) || (rc = dev_alloc(&m->m, fa::kMlpParams)) ||
        (rc = dev_alloc(&m->v, fa::kMlpParams)) || (rc = dev_alloc(&m->norm, 2 * fa::kMlpFeatures)) ||
        (rc = dev_alloc((uint16_t **)&m->wh, fa::kMlpParams)) ||
        (rc = dev_alloc(&m->grad, fa::kMlpParams)) || (rc = dev_alloc(&m->loss, 1))) {
        mlp_release(m);
        return rc;
    }
    *out = m;
    return FA_OK;
}

int fa_mlp_destroy(fa_ctx *ctx, fa_mlp *m) {
    if (!ctx || !m) return FA_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    mlp_release(m);
    return FA_OK;
}

int fa_mlp_set_params(fa_ctx *ctx, fa_mlp *m, const float *params, const float *norm) {
    if (!ctx || !m || !params || !norm) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    FA_HIP(hipMemcpyAsync(m->master, params, fa::kMlpParams * sizeof(float), hipMemcpyHostToDevice,
                          ctx->stream));
    FA_HIP(hipMemcpyAsync(m->norm, norm, 2 * fa::kMlpFeatures * sizeof(float),
                          hipMemcpyHostToDevice, ctx->stream));
    FA_HIP(hipMemsetAsync(m->m, 0, fa::kMlpParams * sizeof(float), ctx->stream));
    FA_HIP(hipMemsetAsync(m->v, 0, fa::kMlpParams * sizeof(float), ctx->stream));
    FA_HIP(fa::mlp_launch_to_half(m->master, m->wh, fa::kMlpParams, ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    m->step = 0;
    return FA_OK;
}

int fa_mlp_get_params(fa_ctx *ctx, fa_mlp *m, float *params) {
    if (!ctx || !m || !params) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    FA_HIP(hipMemcpy(params, m->master, fa::kMlpParams * sizeof(float), hipMemcpyDeviceToHost));
    return FA_OK;
}

static int ensure_features(fa_ctx *ctx, fa_batch *b) {
    if (!b->d_feat) {
        int rc = dev_alloc(&b->d_feat, (size_t)b->n * fa::kMlpFeatures);
        if (rc) return rc;
        FA_HIP(fa::mlp_launch_features(b->args, (uint32_t)b->n, b->d_feat, ctx->stream));
    }
    return FA_OK;
}

int fa_batch_features(fa_ctx *ctx, fa_batch *b, float *feat) {
    if (!ctx || !b) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc = ensure_features(ctx, b);
    if (rc) return rc;
    if (feat && b->n)
        FA_HIP(hipMemcpyAsync(feat, b->d_feat, (size_t)b->n * fa::kMlpFeatures * sizeof(float),
                              hipMemcpyDeviceToHost, ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    return FA_OK;
}

int fa_mlp_predict(fa_ctx *ctx, fa_mlp *m, fa_batch *b, float *cells_pred, int32_t set_cost) {
    if (!ctx || !m || !b) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc;
    if (!b->d_pred && (rc = dev_alloc(&b->d_pred, (size_t)b->n))) return rc;
    // features from the batch's feature buffer once training or
    // fa_batch_features has filled it, unless the caller asks for a new
    // batch's cost (set_cost & 2); otherwise the forward kernel computes them
    // itself and stores nothing (only training reads them)
    const bool cached = b->d_feat && !(set_cost & 2);
    FA_HIP(fa::mlp_launch_predict(b->args, m->wh, m->norm, cached ? b->d_feat : nullptr, cached,
                                  (uint32_t)b->n, b->d_pred, ctx->stream));
    if (cells_pred && b->n)
        FA_HIP(hipMemcpyAsync(cells_pred, b->d_pred, (size_t)b->n * sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    if (set_cost & 1) {  // the predictions stay on the device as the packing cost
        b->cost_set = true;
        b->plan_dirty = true;
    }
    return FA_OK;
}

static int mlp_ensure_partial(fa_mlp *m, uint32_t blocks) {
    if (m->partial_blocks >= blocks) return FA_OK;
    if (m->partial) (void)hipFree(m->partial);
    m->partial = nullptr;
    m->partial_blocks = 0;
    int rc = dev_alloc(&m->partial, (size_t)blocks * (fa::kMlpParams + 1));
    if (rc) return rc;
    m->partial_blocks = blocks;
    return FA_OK;
}

static int mlp_set_perm(fa_ctx *ctx, fa_mlp *m, const uint32_t *perm, int64_t n) {
    if (m->perm_n < n) {
        if (m->perm) (void)hipFree(m->perm);
        m->perm = nullptr;
        m->perm_n = 0;
        int rc = dev_alloc(&m->perm, (size_t)n);
        if (rc) return rc;
        m->perm_n = n;
    }
    FA_HIP(hipMemcpyAsync(m->perm, perm, (size_t)n * sizeof(uint32_t), hipMemcpyHostToDevice,
                          ctx->stream));
    return FA_OK;
}

int fa_mlp_gradient(fa_ctx *ctx, fa_mlp *m, fa_batch *b, const uint32_t *pairs, int64_t count,
                    float *grad, float *mean_loss) {
    if (!ctx || !m || !b || !pairs || count <= 0 || count > b->n) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc = ensure_features(ctx, b);
    if (!rc) rc = ensure_records(ctx, b);  // the cell counts are the labels
    if (rc) return rc;
    const uint32_t blocks = (uint32_t)((count + fa::kMlpBlock - 1) / fa::kMlpBlock);
    if ((rc = mlp_ensure_partial(m, blocks)) || (rc = mlp_set_perm(ctx, m, pairs, count))) return rc;
    FA_HIP(fa::mlp_launch_backward(m->wh, m->norm, b->d_feat, b->d_cells, m->perm, 0,
                                   (uint32_t)count, m->partial, ctx->stream));
    FA_HIP(fa::mlp_launch_reduce_adam(m->partial, blocks, m->grad, m->loss, m->master, m->m, m->v,
                                      m->wh, 0.f, 0.f, 0.f, 0.f, 1.f, 1.f, 0, ctx->stream));
    float lsum = 0.f;
    if (grad)
        FA_HIP(hipMemcpyAsync(grad, m->grad, fa::kMlpParams * sizeof(float), hipMemcpyDeviceToHost,
                              ctx->stream));
    FA_HIP(hipMemcpyAsync(&lsum, m->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    if (mean_loss) *mean_loss = lsum / (float)count;
    return FA_OK;
}

int fa_mlp_train(fa_ctx *ctx, fa_mlp *m, fa_batch *b, const uint32_t *order, int64_t count,
                 int32_t minibatch, float lr, float *mean_loss) {
    if (!ctx || !m || !b || !order || count <= 0 || count > b->n || minibatch <= 0) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    int rc = ensure_features(ctx, b);
    if (!rc) rc = ensure_records(ctx, b);  // the cell counts are the labels
    if (rc) return rc;
    const uint32_t mb_blocks = (uint32_t)((minibatch + fa::kMlpBlock - 1) / fa::kMlpBlock);
    if ((rc = mlp_ensure_partial(m, mb_blocks)) || (rc = mlp_set_perm(ctx, m, order, count)))
        return rc;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    // each minibatch's loss sum goes to its own slot; one read-back at the end
    const int64_t nmb = (count + minibatch - 1) / minibatch;
    if (m->losses_n < nmb) {
        if (m->losses) (void)hipFree(m->losses);
        m->losses = nullptr;
        m->losses_n = 0;
        if ((rc = dev_alloc(&m->losses, (size_t)nmb))) return rc;
        m->losses_n = nmb;
    }
    int64_t mb = 0;
    for (int64_t first = 0; first < count; first += minibatch, ++mb) {
        const uint32_t cnt = (uint32_t)std::min<int64_t>(minibatch, count - first);
        const uint32_t blocks = (cnt + fa::kMlpBlock - 1) / fa::kMlpBlock;
        m->step += 1;
        const float bc1 = 1.f - std::pow(b1, (float)m->step);
        const float bc2 = 1.f - std::pow(b2, (float)m->step);
        FA_HIP(fa::mlp_launch_backward(m->wh, m->norm, b->d_feat, b->d_cells, m->perm,
                                       (uint32_t)first, cnt, m->partial, ctx->stream));
        FA_HIP(fa::mlp_launch_reduce_adam(m->partial, blocks, nullptr, m->losses + mb, m->master, m->m,
                                          m->v, m->wh, lr, b1, b2, eps, bc1, bc2, 1, ctx->stream));
    }
    std::vector<float> lsum((size_t)nmb);
    FA_HIP(hipMemcpyAsync(lsum.data(), m->losses, (size_t)nmb * sizeof(float), hipMemcpyDeviceToHost,
                          ctx->stream));
    FA_HIP(hipStreamSynchronize(ctx->stream));
    double total = 0.0;
    for (float x : lsum) total += x;
    if (mean_loss) *mean_loss = (float)(total / (double)count);
    return FA_OK;
}

}  // extern "C"
