// Chain extension on the device: fa_chains_upload / fa_chains_extend /
// fa_chains_download / fa_chains_free, and fa_extend_chains over host
// buffers.  The job producer, the packing of queries and targets, both
// extension batches and the region merge run on the GPU (chain_kernel.hip);
// the host reads back a few hundred bytes of per-kernel job counts between
// the steps to plan the extension launches.  Reference counterpart: the
// producer / batch / consumer sequence of src/pipelines/linear/
// batch_extension/collect_soa.rs:28-228, dispatch.rs:31-268 and
// region.rs:737-890, one chain batch at a time.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <new>
#include <vector>

#include "chain_kernel.h"
#include "fa_internal.h"

namespace {

using fa_host::Bucket;

struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

// Grow-only device buffer (an eighth of headroom, so that batches of
// slightly varying size reuse it).
int grow(DBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 1;
    if (b.bytes >= bytes) return FA_OK;
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
    const size_t want = bytes + bytes / 8;
    FA_HIP(hipMalloc(&b.p, want));
    b.bytes = want;
    return FA_OK;
}

void release(DBuf &b) {
    if (b.p) (void)hipFree(b.p);
    b = DBuf{};
}

template <class T>
T *P(const DBuf &b) {
    return (T *)b.p;
}

}  // namespace

struct fa_chain_batch {
    int device = 0;
    int64_t nreads = 0, nchains = 0, ncseeds = 0, nseeds = 0, qbytes = 0, nslots = 0;
    bool has_frac = false;
    DBuf qseq, qoff, qlen, rcs, rcc, kept, frac, cstart, ccount, slot_base, cseeds, seeds, regions;
    std::vector<int64_t> slot_base_host;

    void release_all() {
        for (DBuf *b : {&qseq, &qoff, &qlen, &rcs, &rcc, &kept, &frac, &cstart, &ccount, &slot_base,
                        &cseeds, &seeds, &regions})
            release(*b);
    }
};

namespace {

struct SideBufs {
    DBuf qlen, tlen, h0, w, qsrc, seg, okey, err, words, wend, qo, to, seq2, seqn, key, slot,
        key_out, order, out;
    void release_all() {
        for (DBuf *b : {&qlen, &tlen, &h0, &w, &qsrc, &seg, &okey, &err, &words, &wend, &qo,
                        &to, &seq2, &seqn, &key, &slot, &key_out, &order, &out})
            release(*b);
    }
};

struct ChainWork {
    SideBufs side[2];
    DBuf chain_read, chain_state, status, temp, meta, cells, queue;
    fa::ChainStatus *host_status = nullptr;  // pinned
    uint64_t *host_words = nullptr;          // pinned, 2 entries
    hipEvent_t ev[8] = {};
    fa_chain_batch staged;  // the inputs of fa_extend_chains
};

int workspace(fa_ctx *ctx, ChainWork **out) {
    if (!ctx->chain_ws) {
        ChainWork *w = new (std::nothrow) ChainWork();
        if (!w) return FA_ERR_ARG;
        bool ok = hipHostMalloc((void **)&w->host_status, sizeof(fa::ChainStatus)) == hipSuccess &&
                  hipHostMalloc((void **)&w->host_words, 2 * sizeof(uint64_t)) == hipSuccess;
        for (auto &e : w->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
        ctx->chain_ws = w;
        if (!ok) return FA_ERR_DEVICE;
    }
    *out = (ChainWork *)ctx->chain_ws;
    return FA_OK;
}

fa::ChainSide side_view(SideBufs &b) {
    fa::ChainSide s;
    s.qlen = P<int32_t>(b.qlen);
    s.tlen = P<int32_t>(b.tlen);
    s.h0 = P<int32_t>(b.h0);
    s.w = P<int32_t>(b.w);
    s.qsrc = P<int64_t>(b.qsrc);
    s.seg = P<int64_t>(b.seg);
    s.okey = P<uint64_t>(b.okey);
    s.err = P<uint8_t>(b.err);
    s.words = P<uint64_t>(b.words);
    s.wend = P<uint64_t>(b.wend);
    s.qo = P<uint32_t>(b.qo);
    s.to = P<uint32_t>(b.to);
    s.seq2 = P<uint32_t>(b.seq2);
    s.seqn = P<uint32_t>(b.seqn);
    s.key = P<uint32_t>(b.key);
    s.slot = P<uint32_t>(b.slot);
    s.out = P<int32_t>(b.out);
    return s;
}

int grow_side(SideBufs &b, size_t n) {
    int rc;
    if ((rc = grow(b.qlen, 4 * n)) || (rc = grow(b.tlen, 4 * n)) ||
        (rc = grow(b.h0, 4 * n)) || (rc = grow(b.w, 4 * n)) || (rc = grow(b.qsrc, 8 * n)) ||
        (rc = grow(b.seg, 8 * n)) || (rc = grow(b.okey, 8 * n)) || (rc = grow(b.err, n)) ||
        (rc = grow(b.words, 8 * n)) || (rc = grow(b.wend, 8 * n)) || (rc = grow(b.qo, 4 * n)) ||
        (rc = grow(b.to, 4 * n)) || (rc = grow(b.key, 4 * n)) || (rc = grow(b.slot, 4 * n)) ||
        (rc = grow(b.key_out, 4 * n)) || (rc = grow(b.order, 4 * n)) || (rc = grow(b.out, 24 * n)))
        return rc;
    return FA_OK;
}

int upload_into(fa_ctx *ctx, fa_chain_batch *cb, int64_t nreads, const uint8_t *query_seqs,
                int64_t query_bytes, const int64_t *query_offset, const int32_t *query_len,
                const int64_t *read_chain_start, const int32_t *read_chain_count, int64_t nchains,
                const uint8_t *chain_kept, const float *chain_frac_rep,
                const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                int64_t n_chain_seeds, const int64_t *chain_seeds, int64_t nseeds,
                const fa_seed *seeds) {
    if (nreads < 0 || nchains < 0 || query_bytes < 0 || n_chain_seeds < 0 || nseeds < 0)
        return FA_ERR_ARG;
    if (nreads >= ((int64_t)1 << 31) || nchains >= ((int64_t)1 << 31)) return FA_ERR_ARG;
    if ((query_bytes > 0 && !query_seqs) ||
        (nreads > 0 && (!query_offset || !query_len || !read_chain_start || !read_chain_count)) ||
        (nchains > 0 && (!chain_kept || !chain_seed_start || !chain_seed_count)) ||
        (n_chain_seeds > 0 && !chain_seeds) || (nseeds > 0 && !seeds))
        return FA_ERR_ARG;
    // job slots: chain c's seeds own [slot_base[c], slot_base[c] + chain_seed_count[c])
    cb->slot_base_host.resize((size_t)nchains + 1);
    int64_t total = 0;
    for (int64_t c = 0; c < nchains; ++c) {
        if (chain_seed_count[c] < 0) return FA_ERR_ARG;
        cb->slot_base_host[(size_t)c] = total;
        total += chain_seed_count[c];
    }
    cb->slot_base_host[(size_t)nchains] = total;
    if (total >= ((int64_t)1 << 31)) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    int rc;
    // reads with 64 bytes of padding on both sides (word-wide loads of the job
    // windows); every array goes up through the context's pinned staging
    // buffer, which the host threads fill while earlier pieces are in flight
    std::vector<fa_host::HostSeg> segs;
    auto put = [&](DBuf &d, const auto *src, int64_t count, size_t at = 0, size_t extra = 0) {
        const size_t bytes = (size_t)count * sizeof(*src);
        const int r = grow(d, bytes + extra);
        if (r == FA_OK && bytes > 0) segs.push_back({(uint8_t *)d.p + at, src, bytes});
        return r;
    };
    cb->has_frac = chain_frac_rep != nullptr;
    if ((rc = put(cb->qseq, query_seqs, query_bytes, 64, 128)) ||
        (rc = put(cb->qoff, query_offset, nreads)) || (rc = put(cb->qlen, query_len, nreads)) ||
        (rc = put(cb->rcs, read_chain_start, nreads)) || (rc = put(cb->rcc, read_chain_count, nreads)) ||
        (rc = put(cb->kept, chain_kept, nchains)) || (rc = put(cb->cstart, chain_seed_start, nchains)) ||
        (rc = put(cb->ccount, chain_seed_count, nchains)) ||
        (rc = put(cb->slot_base, cb->slot_base_host.data(), nchains + 1)) ||
        (rc = put(cb->cseeds, chain_seeds, n_chain_seeds)) || (rc = put(cb->seeds, seeds, nseeds)) ||
        (cb->has_frac && (rc = put(cb->frac, chain_frac_rep, nchains))) ||
        (rc = grow(cb->regions, (size_t)nchains * sizeof(fa_region))))
        return rc;
    if ((rc = fa_host::staged_upload(ctx, segs.data(), (int)segs.size(), st))) return rc;
    FA_HIP(hipStreamSynchronize(st));
    cb->device = ctx->device;
    cb->nreads = nreads;
    cb->nchains = nchains;
    cb->ncseeds = n_chain_seeds;
    cb->nseeds = nseeds;
    cb->qbytes = query_bytes;
    cb->nslots = total;
    return FA_OK;
}

// Launch plan of one side from its per-kernel job counts: buckets in rank
// order (widest kernel first), jobs inside a bucket by (qlen, tlen)
// descending, as the length packing of the batch entry points.
int run_side(fa_ctx *ctx, ChainWork *ws, int sd, const fa::ChainStatus &st,
             const fa_banded_pairwise_sw &sw, const std::vector<int> &rank_skey, int64_t n) {
    if (n == 0) return FA_OK;
    SideBufs &b = ws->side[sd];
    fa_batch tmp;
    tmp.n = n;
    uint32_t first = 0;
    for (int rk = 0; rk < fa::kChainRanks; ++rk) {
        const uint64_t cnt = st.count[sd][rk];
        if (!cnt) continue;
        if (rk >= (int)rank_skey.size()) return FA_ERR_ARG;
        const int skey = rank_skey[(size_t)rk];
        Bucket bk;
        bk.kind = skey >> 16;
        bk.L = (skey >> 8) & 0xFF;
        bk.S = skey & 0xFF;
        bk.first = first;
        bk.count = (uint32_t)cnt;
        bk.work = (double)st.work[sd][rk];
        first += (uint32_t)cnt;
        tmp.buckets.push_back(bk);
    }
    if ((int64_t)first != n) return FA_ERR_ARG;
    tmp.streams = 4;
    tmp.packed_flags = 0;
    fa::ExtArgs &a = tmp.args;
    a.seq2 = P<uint32_t>(b.seq2);
    a.seqn = P<uint32_t>(b.seqn);
    a.qo = P<uint32_t>(b.qo);
    a.to = P<uint32_t>(b.to);
    a.qlen = P<int32_t>(b.qlen);
    a.tlen = P<int32_t>(b.tlen);
    a.h0 = P<int32_t>(b.h0);
    a.w = P<int32_t>(b.w);
    a.order = P<uint32_t>(b.order);
    a.meta = P<uint32_t>(ws->meta);
    a.out = P<int32_t>(b.out);
    a.cells = P<uint32_t>(ws->cells);
    a.queue = P<uint32_t>(ws->queue);
    a.ovf = nullptr;
    fa_host::set_scoring(a, sw);
    a.tpad = (int32_t)((st.tmax[sd] + 16) / 16 * 16);
    a.chunk = 0;
    FA_HIP(fa::meta_gather(a, (uint32_t)n, ctx->stream));
    return fa_host::enqueue_run(ctx, &tmp, 0);
}

int read_status(fa_ctx *ctx, ChainWork *ws, bool words, int64_t nslots) {
    hipStream_t st = ctx->stream;
    FA_HIP(hipMemcpyAsync(ws->host_status, ws->status.p, sizeof(fa::ChainStatus), hipMemcpyDeviceToHost, st));
    if (words) {
        ws->host_words[0] = ws->host_words[1] = 0;
        if (nslots > 0)
            for (int sd = 0; sd < 2; ++sd)
                FA_HIP(hipMemcpyAsync(ws->host_words + sd, P<uint64_t>(ws->side[sd].wend) + nslots - 1,
                                      sizeof(uint64_t), hipMemcpyDeviceToHost, st));
    }
    FA_HIP(hipStreamSynchronize(st));
    return FA_OK;
}

}  // namespace

namespace fa_host {

void chain_workspace_free(fa_ctx *ctx) {
    if (!ctx || !ctx->chain_ws) return;
    ChainWork *w = (ChainWork *)ctx->chain_ws;
    for (auto &s : w->side) s.release_all();
    for (DBuf *b : {&w->chain_read, &w->chain_state, &w->status, &w->temp, &w->meta, &w->cells, &w->queue})
        release(*b);
    w->staged.release_all();
    if (w->host_status) (void)hipHostFree(w->host_status);
    if (w->host_words) (void)hipHostFree(w->host_words);
    for (auto &e : w->ev)
        if (e) (void)hipEventDestroy(e);
    delete w;
    ctx->chain_ws = nullptr;
}

}  // namespace fa_host

extern "C" {

int fa_chains_upload(fa_ctx *ctx, int64_t nreads, const uint8_t *query_seqs, int64_t query_bytes,
                     const int64_t *query_offset, const int32_t *query_len,
                     const int64_t *read_chain_start, const int32_t *read_chain_count,
                     int64_t nchains, const uint8_t *chain_kept, const float *chain_frac_rep,
                     const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                     int64_t n_chain_seeds, const int64_t *chain_seeds, int64_t nseeds,
                     const fa_seed *seeds, fa_chain_batch **out) {
    if (!ctx || !out) return FA_ERR_ARG;
    *out = nullptr;
    fa_chain_batch *cb = new (std::nothrow) fa_chain_batch();
    if (!cb) return FA_ERR_ARG;
    int rc = upload_into(ctx, cb, nreads, query_seqs, query_bytes, query_offset, query_len,
                         read_chain_start, read_chain_count, nchains, chain_kept, chain_frac_rep,
                         chain_seed_start, chain_seed_count, n_chain_seeds, chain_seeds, nseeds, seeds);
    if (rc) {
        cb->release_all();
        delete cb;
        return rc;
    }
    *out = cb;
    return FA_OK;
}

int fa_chains_free(fa_ctx *ctx, fa_chain_batch *cb) {
    if (!ctx || !cb) return FA_ERR_ARG;
    (void)hipSetDevice(cb->device);
    cb->release_all();
    delete cb;
    return FA_OK;
}

int fa_chains_download(fa_ctx *ctx, fa_chain_batch *cb, fa_region *regions) {
    if (!ctx || !cb || (cb->nchains > 0 && !regions)) return FA_ERR_ARG;
    if (cb->nchains == 0) return FA_OK;
    FA_HIP(hipSetDevice(ctx->device));
    return fa_host::staged_download(ctx, regions, cb->regions.p, (size_t)cb->nchains * sizeof(fa_region),
                                    ctx->stream);
}

int fa_chains_extend(fa_ctx *ctx, fa_chain_batch *cb, const fa_ref *ref,
                     const fa_banded_pairwise_sw *sw, const fa_mem_opt *opt, int64_t *n_left,
                     int64_t *n_right, double *stage_ms) {
    if (!ctx || !cb || !ref || !sw || !opt) return FA_ERR_ARG;
    if (cb->device != ctx->device || ref->device != ctx->device) return FA_ERR_ARG;
    if (opt->e_del <= 0 || opt->e_ins <= 0 || opt->a <= 0) return FA_ERR_SCORING;
    FA_HIP(hipSetDevice(ctx->device));
    ChainWork *ws = nullptr;
    int rc = workspace(ctx, &ws);
    if (rc) return rc;
    const int sw_rc = fa_host::check_scoring(sw);
    const int64_t N = cb->nslots, C = cb->nchains;
    hipStream_t st = ctx->stream;
    for (int sd = 0; sd < 2; ++sd)
        if ((rc = grow_side(ws->side[sd], (size_t)N))) return rc;
    if ((rc = grow(ws->chain_read, 4 * (size_t)C)) || (rc = grow(ws->chain_state, (size_t)C)) ||
        (rc = grow(ws->status, sizeof(fa::ChainStatus))) || (rc = grow(ws->meta, 32 * (size_t)N)) ||
        (rc = grow(ws->cells, 4 * (size_t)N)) || (rc = grow(ws->queue, 64 * sizeof(uint32_t))))
        return rc;
    // scratch of the scans and sorts
    {
        size_t a = 0, b = 0;
        FA_HIP(hipcub::DeviceScan::InclusiveSum(nullptr, a, (uint64_t *)nullptr, (uint64_t *)nullptr,
                                                (int)std::max<int64_t>(N, 1), st));
        FA_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, b, (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (uint32_t *)nullptr, (uint32_t *)nullptr,
                                                  (int)std::max<int64_t>(N, 1), 0, 26, st));
        if ((rc = grow(ws->temp, std::max(a, b)))) return rc;
    }

    fa::ChainIn in;
    in.qseq = P<uint8_t>(cb->qseq) + 64;
    in.qbytes = cb->qbytes;
    in.qoff = P<int64_t>(cb->qoff);
    in.qlen = P<int32_t>(cb->qlen);
    in.nreads = cb->nreads;
    in.rcs = P<int64_t>(cb->rcs);
    in.rcc = P<int32_t>(cb->rcc);
    in.nchains = C;
    in.kept = P<uint8_t>(cb->kept);
    in.frac = cb->has_frac ? P<float>(cb->frac) : nullptr;
    in.cstart = P<int64_t>(cb->cstart);
    in.ccount = P<int32_t>(cb->ccount);
    in.slot_base = P<int64_t>(cb->slot_base);
    in.cseeds = P<int64_t>(cb->cseeds);
    in.ncseeds = cb->ncseeds;
    in.seeds = P<fa_seed>(cb->seeds);
    in.nseeds = cb->nseeds;

    fa::ChainOpt o;
    o.a = opt->a;
    o.o_del = opt->o_del;
    o.e_del = opt->e_del;
    o.o_ins = opt->o_ins;
    o.e_ins = opt->e_ins;
    o.w = opt->w;
    o.pen_clip5 = opt->pen_clip5;
    o.pen_clip3 = opt->pen_clip3;
    o.l_pac = ref->l_pac;
    int mx = 0;
    for (int x = 0; x < 25; ++x) mx = std::max(mx, (int)sw->mat[x]);
    o.max_score = mx;
    o.end_bonus = sw->end_bonus;
    o.sw_o_del = sw->o_del;
    o.sw_e_del = sw->e_del;
    o.sw_o_ins = sw->o_ins;
    o.sw_e_ins = sw->e_ins;
    o.k16 = std::max(2, 1 + mx);
    o.sw16ok = sw->e_ins < 32 && sw->o_ins + sw->e_ins < 8192 && sw->o_del + sw->e_del < 8192;

    // kernel choice of every (fits16, qlen) and its rank (widest kernel first)
    fa::ChainRanks rk{};
    std::vector<int> rank_skey;
    {
        int sk[3][256] = {};
        for (int f = 0; f < 3; ++f)
            for (int q = 1; q < 256; ++q) {
                int kind = 0, lanes = 0, S = 0;
                if (!fa_host::pair_kernel(0, false, f, q, &kind, &lanes, &S)) return FA_ERR_QLEN;
                sk[f][q] = (kind << 16) | (lanes << 8) | S;
                rank_skey.push_back(sk[f][q]);
            }
        std::sort(rank_skey.begin(), rank_skey.end(), [](int x, int y) { return x > y; });
        rank_skey.erase(std::unique(rank_skey.begin(), rank_skey.end()), rank_skey.end());
        if ((int)rank_skey.size() > fa::kChainRanks) return FA_ERR_ARG;
        for (int f = 0; f < 3; ++f)
            for (int q = 1; q < 256; ++q)
                rk.r[f][q] = (uint8_t)(std::find(rank_skey.begin(), rank_skey.end(), sk[f][q]) -
                                       rank_skey.begin());
    }

    fa::ChainSide L = side_view(ws->side[0]), R = side_view(ws->side[1]);
    fa::ChainStatus *dst = P<fa::ChainStatus>(ws->status);
    int32_t *chain_read = P<int32_t>(ws->chain_read);
    uint8_t *chain_state = P<uint8_t>(ws->chain_state);
    const fa::ChainStatus &hs = *ws->host_status;
    hipEvent_t *ev = ws->ev;

    // producer
    FA_HIP(hipEventRecord(ev[0], st));
    FA_HIP(fa::chain_init(dst, chain_read, C, st));
    FA_HIP(fa::chain_owner(in, chain_read, dst, st));
    FA_HIP(fa::chain_jobs(in, o, chain_read, chain_state, L, R, dst, st));
    for (int sd = 0; sd < 2 && N > 0; ++sd) {
        size_t tb = ws->temp.bytes;
        FA_HIP(hipcub::DeviceScan::InclusiveSum(ws->temp.p, tb, P<uint64_t>(ws->side[sd].words),
                                                P<uint64_t>(ws->side[sd].wend), (int)N, st));
    }
    if ((rc = read_status(ctx, ws, true, N))) return rc;
    if (hs.arg_err) return FA_ERR_ARG;
    if (sw_rc != FA_OK && hs.any_job) return sw_rc;
    for (int sd = 0; sd < 2; ++sd) {
        const uint64_t words = ws->host_words[sd];
        if (words > 0xFFFFFFF0ull) return FA_ERR_ARG;
        if ((rc = grow(ws->side[sd].seq2, 4 * (size_t)words)) ||
            (rc = grow(ws->side[sd].seqn, 4 * (size_t)(words / 2))))
            return rc;
    }
    L = side_view(ws->side[0]);
    R = side_view(ws->side[1]);

    // left batch: sequences, keys, sort, extension
    FA_HIP(hipEventRecord(ev[1], st));
    FA_HIP(fa::chain_fill(in, ref->d_pac, ref->l_pac, L, R, N, dst, st));
    FA_HIP(fa::chain_keys(o, rk, L, L, 0, N, dst, st));
    if (N > 0) {
        size_t tb = ws->temp.bytes;
        FA_HIP(hipcub::DeviceRadixSort::SortPairs(ws->temp.p, tb, L.key, P<uint32_t>(ws->side[0].key_out),
                                                  L.slot, P<uint32_t>(ws->side[0].order), (int)N, 0, 26, st));
    }
    if ((rc = read_status(ctx, ws, false, N))) return rc;
    if (hs.left_key != ~0ull) return -(int)(hs.left_key & 15u);
    int64_t nl = 0;
    for (int k = 0; k < fa::kChainRanks; ++k) nl += (int64_t)hs.count[0][k];
    FA_HIP(hipEventRecord(ev[2], st));
    if ((rc = run_side(ctx, ws, 0, hs, *sw, rank_skey, nl))) return rc;

    // right batch: h0 from the left scores, keys, sort, extension
    FA_HIP(hipEventRecord(ev[3], st));
    FA_HIP(fa::chain_keys(o, rk, R, L, 1, N, dst, st));
    if (N > 0) {
        size_t tb = ws->temp.bytes;
        FA_HIP(hipcub::DeviceRadixSort::SortPairs(ws->temp.p, tb, R.key, P<uint32_t>(ws->side[1].key_out),
                                                  R.slot, P<uint32_t>(ws->side[1].order), (int)N, 0, 26, st));
    }
    if ((rc = read_status(ctx, ws, false, N))) return rc;
    if (hs.right_key != ~0ull) return -(int)(hs.right_key & 15u);
    int64_t nr = 0;
    for (int k = 0; k < fa::kChainRanks; ++k) nr += (int64_t)hs.count[1][k];
    FA_HIP(hipEventRecord(ev[4], st));
    if ((rc = run_side(ctx, ws, 1, hs, *sw, rank_skey, nr))) return rc;

    // merge
    FA_HIP(hipEventRecord(ev[5], st));
    FA_HIP(fa::chain_merge(in, o, chain_read, chain_state, L, R, P<fa_region>(cb->regions), st));
    FA_HIP(hipEventRecord(ev[6], st));
    FA_HIP(hipEventSynchronize(ev[6]));
    if (stage_ms) {
        // producer, left preparation, left run, right preparation, right run, merge
        for (int k = 0; k < 6; ++k) {
            float t = 0.f;
            FA_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
            stage_ms[k + 1] = t;
        }
    }
    if (n_left) *n_left = nl;
    if (n_right) *n_right = nr;
    return FA_OK;
}

int fa_extend_chains(fa_ctx *ctx, const fa_ref *ref, const fa_banded_pairwise_sw *sw,
                     const fa_mem_opt *opt, int64_t nreads, const uint8_t *query_seqs,
                     int64_t query_bytes, const int64_t *query_offset, const int32_t *query_len,
                     const int64_t *read_chain_start, const int32_t *read_chain_count,
                     int64_t nchains, const uint8_t *chain_kept, const float *chain_frac_rep,
                     const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                     int64_t n_chain_seeds, const int64_t *chain_seeds, int64_t nseeds,
                     const fa_seed *seeds, fa_region *regions, int64_t *n_left, int64_t *n_right,
                     double *stage_ms) {
    if (!ctx || !ref || !sw || !opt || (nchains > 0 && !regions)) return FA_ERR_ARG;
    FA_HIP(hipSetDevice(ctx->device));
    ChainWork *ws = nullptr;
    int rc = workspace(ctx, &ws);
    if (rc) return rc;
    double st[FA_CHAIN_STAGES] = {};
    auto t0 = std::chrono::steady_clock::now();
    rc = upload_into(ctx, &ws->staged, nreads, query_seqs, query_bytes, query_offset, query_len,
                     read_chain_start, read_chain_count, nchains, chain_kept, chain_frac_rep,
                     chain_seed_start, chain_seed_count, n_chain_seeds, chain_seeds, nseeds, seeds);
    if (rc) return rc;
    auto t1 = std::chrono::steady_clock::now();
    rc = fa_chains_extend(ctx, &ws->staged, ref, sw, opt, n_left, n_right, st);
    if (rc) return rc;
    auto t2 = std::chrono::steady_clock::now();
    rc = fa_chains_download(ctx, &ws->staged, regions);
    if (rc) return rc;
    auto t3 = std::chrono::steady_clock::now();
    if (stage_ms) {
        st[0] = std::chrono::duration<double, std::milli>(t1 - t0).count();
        st[7] = std::chrono::duration<double, std::milli>(t3 - t2).count();
        std::memcpy(stage_ms, st, sizeof st);
    }
    return FA_OK;
}

}  // extern "C"
