// Coalescing of concurrent small extension batches (fa_agg_*).
//
// The reference calls its batch extension path from every rayon worker at
// once, each call synchronous and small (single_end.rs:92-110; a chunk of
// one read batch, or 16-64 jobs at the kernel-level SoA entry points), and
// the GPU pays off only on large batches.  An aggregator owns one context and
// one worker thread.  A caller's jobs join a pending list and the caller
// sleeps; the worker takes everything pending (once at least min_jobs jobs
// are there, or max_wait_us after the oldest arrived), runs it as one batch
// per (scoring, flags) through fa_execute_batch_simd_scoring, copies each
// caller's records into that caller's outputs and wakes it.  Results are the
// per-job results of the direct call: only the grouping differs.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "fa_internal.h"

namespace {

struct AggCall {
    const fa_banded_pairwise_sw *sw;
    int64_t njobs;
    const uint8_t *query_seqs, *ref_seqs;
    const int64_t *query_offset, *ref_offset;
    const int32_t *query_len, *ref_len, *h0, *band_width;
    uint32_t flags;
    fa_out_score *out;
    uint32_t *cells;
    int32_t *per_job_rc;
    int rc = FA_OK;
    bool done = false;
};

bool same_group(const AggCall &a, const AggCall &b) {
    return a.flags == b.flags && std::memcmp(a.sw, b.sw, sizeof(fa_banded_pairwise_sw)) == 0;
}

}  // namespace

struct fa_agg {
    fa_ctx *ctx = nullptr;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<AggCall *> pending;
    int64_t pending_jobs = 0;
    std::chrono::steady_clock::time_point oldest;
    bool stop = false;
    int64_t min_jobs = 1;
    std::chrono::microseconds max_wait{0};
    int64_t n_calls = 0, n_runs = 0, n_jobs = 0;
};

namespace {

// One run over the calls of one group: their jobs back to back in call
// order, the sequences copied into one buffer per side so that one offset
// array addresses them.
void run_group(fa_agg *g, const std::vector<AggCall *> &calls) {
    int64_t n = 0, qbytes = 0, tbytes = 0;
    for (const AggCall *c : calls) {
        n += c->njobs;
        for (int64_t k = 0; k < c->njobs; ++k) {
            qbytes += c->query_len[k] > 0 ? c->query_len[k] : 0;
            tbytes += c->ref_len[k] > 0 ? c->ref_len[k] : 0;
        }
    }
    std::vector<uint8_t> q((size_t)qbytes + 1), t((size_t)tbytes + 1);
    std::vector<int64_t> qo((size_t)n), to((size_t)n);
    std::vector<int32_t> ql((size_t)n), tl((size_t)n), h0((size_t)n), w((size_t)n);
    int64_t j = 0, qa = 0, ta = 0;
    for (const AggCall *c : calls) {
        for (int64_t k = 0; k < c->njobs; ++k, ++j) {
            const int32_t a = c->query_len[k] > 0 ? c->query_len[k] : 0;
            const int32_t b = c->ref_len[k] > 0 ? c->ref_len[k] : 0;
            if (a) std::memcpy(&q[(size_t)qa], c->query_seqs + c->query_offset[k], (size_t)a);
            if (b) std::memcpy(&t[(size_t)ta], c->ref_seqs + c->ref_offset[k], (size_t)b);
            qo[(size_t)j] = qa;
            to[(size_t)j] = ta;
            qa += a;
            ta += b;
            ql[(size_t)j] = c->query_len[k];
            tl[(size_t)j] = c->ref_len[k];
            h0[(size_t)j] = c->h0[k];
            w[(size_t)j] = c->band_width[k];
        }
    }
    std::vector<fa_out_score> out((size_t)n);
    std::vector<uint32_t> cells((size_t)n);
    std::vector<int32_t> jrc((size_t)n);
    const int rc = fa_execute_batch_simd_scoring(g->ctx, calls[0]->sw, n, q.data(), qo.data(), ql.data(),
                                                 t.data(), to.data(), tl.data(), h0.data(), w.data(),
                                                 calls[0]->flags, out.data(), cells.data(), jrc.data());
    // a device or scoring error belongs to every call of the group; a job's
    // own validation code only to its call (the first invalid job in the
    // call's order, as the direct call returns it)
    int first_job_rc = FA_OK;
    for (int32_t x : jrc)
        if (x != FA_OK) {
            first_job_rc = x;
            break;
        }
    const bool per_job = rc == FA_OK || (first_job_rc != FA_OK && rc == first_job_rc);
    j = 0;
    for (AggCall *c : calls) {
        int crc = FA_OK;
        for (int64_t k = 0; k < c->njobs; ++k, ++j) {
            c->out[k] = out[(size_t)j];
            if (c->cells) c->cells[k] = cells[(size_t)j];
            if (c->per_job_rc) c->per_job_rc[k] = jrc[(size_t)j];
            if (crc == FA_OK && jrc[(size_t)j] != FA_OK) crc = jrc[(size_t)j];
        }
        c->rc = per_job ? crc : rc;
    }
}

void worker_main(fa_agg *g) {
    (void)hipSetDevice(g->ctx->device);
    std::unique_lock<std::mutex> lk(g->mu);
    for (;;) {
        g->cv_work.wait(lk, [&] { return g->stop || !g->pending.empty(); });
        if (g->pending.empty()) return;  // stop, nothing left
        // more calls may join until the batch is big enough or the oldest has
        // waited long enough
        g->cv_work.wait_until(lk, g->oldest + g->max_wait,
                              [&] { return g->stop || g->pending_jobs >= g->min_jobs; });
        std::vector<AggCall *> take;
        take.swap(g->pending);
        g->pending_jobs = 0;
        lk.unlock();
        // one run per (scoring, flags), calls in arrival order
        std::vector<bool> used(take.size(), false);
        int64_t runs = 0, jobs = 0;
        for (size_t x = 0; x < take.size(); ++x) {
            if (used[x]) continue;
            std::vector<AggCall *> grp;
            for (size_t y = x; y < take.size(); ++y)
                if (!used[y] && same_group(*take[x], *take[y])) {
                    used[y] = true;
                    grp.push_back(take[y]);
                    jobs += take[y]->njobs;
                }
            try {
                run_group(g, grp);
            } catch (...) {  // host allocation failure: the group's calls get the error
                for (AggCall *c : grp) c->rc = FA_ERR_ARG;
            }
            ++runs;
        }
        lk.lock();
        g->n_calls += (int64_t)take.size();
        g->n_runs += runs;
        g->n_jobs += jobs;
        for (AggCall *c : take) c->done = true;
        g->cv_done.notify_all();
    }
}

}  // namespace

extern "C" {

int fa_agg_create(int device, int32_t min_jobs, int32_t max_wait_us, fa_agg **out) {
    if (!out || min_jobs < 1 || max_wait_us < 0) return FA_ERR_ARG;
    *out = nullptr;
    fa_ctx *ctx = nullptr;
    const int rc = fa_ctx_create(device, &ctx);
    if (rc) return rc;
    fa_agg *g = new (std::nothrow) fa_agg();
    if (!g) {
        fa_ctx_destroy(ctx);
        return FA_ERR_ARG;
    }
    g->ctx = ctx;
    g->min_jobs = min_jobs;
    g->max_wait = std::chrono::microseconds(max_wait_us);
    try {
        g->worker = std::thread(worker_main, g);
    } catch (...) {
        fa_ctx_destroy(ctx);
        delete g;
        return FA_ERR_ARG;
    }
    *out = g;
    return FA_OK;
}

int fa_agg_destroy(fa_agg *g) {
    if (!g) return FA_ERR_ARG;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        g->stop = true;
    }
    g->cv_work.notify_all();
    if (g->worker.joinable()) g->worker.join();
    const int rc = fa_ctx_destroy(g->ctx);
    delete g;
    return rc;
}

int fa_agg_execute(fa_agg *g, const fa_banded_pairwise_sw *sw, int64_t njobs, const uint8_t *query_seqs,
                   const int64_t *query_offset, const int32_t *query_len, const uint8_t *ref_seqs,
                   const int64_t *ref_offset, const int32_t *ref_len, const int32_t *h0,
                   const int32_t *band_width, uint32_t flags, fa_out_score *out, uint32_t *cells,
                   int32_t *per_job_rc) {
    if (!g || njobs < 0 || (njobs > 0 && !out)) return FA_ERR_ARG;
    const int src = fa_host::check_scoring(sw);
    if (src || njobs == 0) return src;
    if (!query_seqs || !query_offset || !query_len || !ref_seqs || !ref_offset || !ref_len || !h0 ||
        !band_width)
        return FA_ERR_ARG;
    AggCall c{sw, njobs, query_seqs, ref_seqs, query_offset, ref_offset, query_len, ref_len, h0,
              band_width, flags, out, cells, per_job_rc};
    std::unique_lock<std::mutex> lk(g->mu);
    if (g->stop) return FA_ERR_ARG;
    if (g->pending.empty()) g->oldest = std::chrono::steady_clock::now();
    g->pending.push_back(&c);
    g->pending_jobs += njobs;
    g->cv_work.notify_one();
    g->cv_done.wait(lk, [&] { return c.done; });
    return c.rc;
}

int fa_agg_stats(fa_agg *g, int64_t *calls, int64_t *runs, int64_t *jobs) {
    if (!g) return FA_ERR_ARG;
    std::lock_guard<std::mutex> lk(g->mu);
    if (calls) *calls = g->n_calls;
    if (runs) *runs = g->n_runs;
    if (jobs) *jobs = g->n_jobs;
    return FA_OK;
}

}  // extern "C"
