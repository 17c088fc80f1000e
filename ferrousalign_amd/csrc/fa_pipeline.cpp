// The callers either side of the extension batch on the host: the job
// producer (collect_extension_jobs_batch_soa, src/pipelines/linear/
// batch_extension/collect_soa.rs:28-228) and the consumer
// (merge_scores_to_regions, src/pipelines/linear/region.rs:737-890).  The
// device path that runs both around the GPU batches is chain_api.cpp.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/ferrousalign_amd.h"

namespace {

// chaining.rs:1108-1116: the longest gap a query stretch of qlen can pay for,
// capped at 2w (Rust `as i32` truncates toward zero and saturates)
int32_t f64_to_i32(double x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0) return INT32_MAX;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}
int32_t cal_max_gap(const fa_mem_opt &opt, int32_t qlen) {
    const int32_t l_del = f64_to_i32((double)(qlen * opt.a - opt.o_del) / (double)opt.e_del + 1.0);
    const int32_t l_ins = f64_to_i32((double)(qlen * opt.a - opt.o_ins) / (double)opt.e_ins + 1.0);
    int32_t l = l_del > l_ins ? l_del : l_ins;
    l = l > 1 ? l : 1;
    return l < (opt.w << 1) ? l : (opt.w << 1);
}

}  // namespace

extern "C" {

int fa_collect_extension_jobs(const fa_mem_opt *opt, int64_t l_pac, int64_t nreads,
                              const int32_t *query_len, const int64_t *read_chain_start,
                              const int32_t *read_chain_count, const uint8_t *chain_kept,
                              const int64_t *chain_seed_start, const int32_t *chain_seed_count,
                              const int64_t *chain_seeds, const fa_seed *seeds, int64_t capacity,
                              fa_ext_job *left, int64_t *n_left, fa_ext_job *right,
                              int64_t *n_right, int64_t *map_start, int32_t *map_count,
                              int64_t *map_seed, int64_t *map_left, int64_t *map_right,
                              int64_t *rmax) {
    if (!opt || nreads < 0 || l_pac <= 0 || capacity < 0 || !n_left || !n_right) return FA_ERR_ARG;
    if (nreads > 0 && (!query_len || !read_chain_start || !read_chain_count || !chain_kept ||
                       !chain_seed_start || !chain_seed_count || !chain_seeds || !seeds ||
                       !left || !right || !map_start || !map_count || !map_seed || !map_left ||
                       !map_right || !rmax))
        return FA_ERR_ARG;
    if (opt->e_del <= 0 || opt->e_ins <= 0 || opt->a <= 0) return FA_ERR_SCORING;
    const uint64_t lp = (uint64_t)l_pac;
    int64_t nl = 0, nr = 0, nm = 0;
    for (int64_t r = 0; r < nreads; ++r) {
        const int32_t qlen = query_len[r];
        for (int32_t lc = 0; lc < read_chain_count[r]; ++lc) {
            const int64_t c = read_chain_start[r] + lc;
            map_start[c] = nm;
            map_count[c] = 0;
            rmax[2 * c] = rmax[2 * c + 1] = -1;
            const int32_t ns = chain_seed_count[c];
            if (!chain_kept[c] || ns == 0) continue;  // :73-91
            const int64_t *si = chain_seeds + chain_seed_start[c];
            // :94-117 window over all seeds, kept on one strand
            uint64_t r0 = lp << 1, r1 = 0;
            for (int32_t k = 0; k < ns; ++k) {
                const fa_seed &s = seeds[si[k]];
                // a seed lies inside its read (the reference's seeds always do)
                if (s.query_pos < 0 || s.len < 0 || s.query_pos > qlen - s.len) return FA_ERR_ARG;
                const int32_t lm = s.query_pos + cal_max_gap(*opt, s.query_pos);
                const uint64_t b = (uint64_t)s.ref_pos > (uint64_t)lm ? (uint64_t)s.ref_pos - (uint64_t)lm : 0;
                const int32_t rem = qlen - s.query_pos - s.len;
                const int32_t rm = rem + cal_max_gap(*opt, rem);
                const uint64_t e = (uint64_t)s.ref_pos + (uint64_t)s.len + (uint64_t)rm;
                r0 = std::min(r0, b);
                r1 = std::max(r1, e);
            }
            r1 = std::min(r1, lp << 1);
            if (r0 < lp && lp < r1) {
                if ((uint64_t)seeds[si[0]].ref_pos < lp)
                    r1 = lp;
                else
                    r0 = lp;
            }
            // get_reference_segment(r0, r1 - r0) (bntseq.rs:451-522) cannot bridge
            // after the adjustment; an empty window (r1 <= r0) gives no jobs
            const int64_t wlen = r1 > r0 ? (int64_t)(r1 - r0) : 0;
            rmax[2 * c] = (int64_t)r0;
            rmax[2 * c + 1] = (int64_t)r1;
            // :126-219 seeds in reverse order
            for (int32_t k = ns - 1; k >= 0; --k) {
                if (nm >= capacity) return FA_ERR_ARG;
                const int64_t sid = si[k];
                const fa_seed &s = seeds[sid];
                int64_t lj = -1, rj = -1;
                if (s.query_pos > 0) {  // left (:135-176)
                    const int64_t tmp = s.ref_pos - (int64_t)r0;
                    if (tmp > 0 && tmp <= wlen) {
                        fa_ext_job &j = left[nl];
                        j.read_idx = (int32_t)r;
                        j.chain_idx = lc;
                        j.seed_idx = sid;
                        j.qbeg = 0;
                        j.qlen = s.query_pos;
                        j.h0 = s.len * opt->a;
                        j.w = opt->w;
                        j.seg_beg = (int64_t)r0;
                        j.seg_len = (int32_t)tmp;
                        j.seg_rev = 1;
                        lj = nl++;
                    }
                }
                const int32_t qend = s.query_pos + s.len;
                if (qend < qlen) {  // right (:178-205)
                    const int64_t tmp = s.ref_pos + s.len - (int64_t)r0;
                    if (tmp >= 0 && tmp < wlen) {
                        fa_ext_job &j = right[nr];
                        j.read_idx = (int32_t)r;
                        j.chain_idx = lc;
                        j.seed_idx = sid;
                        j.qbeg = qend;
                        j.qlen = qlen - qend;
                        j.h0 = 0;
                        j.w = opt->w;
                        j.seg_beg = (int64_t)r0 + tmp;
                        j.seg_len = (int32_t)(wlen - tmp);
                        j.seg_rev = 0;
                        rj = nr++;
                    }
                }
                map_seed[nm] = sid;
                map_left[nm] = lj;
                map_right[nm] = rj;
                ++nm;
                ++map_count[c];
            }
        }
    }
    *n_left = nl;
    *n_right = nr;
    return FA_OK;
}

int fa_merge_scores_to_regions(const fa_mem_opt *opt, int64_t nreads, const int32_t *query_len,
                               const int64_t *read_chain_start, const int32_t *read_chain_count,
                               const float *chain_frac_rep, const int64_t *chain_seed_start,
                               const int32_t *chain_seed_count, const int64_t *chain_seeds,
                               const fa_seed *seeds, const int64_t *map_start,
                               const int32_t *map_count, const int64_t *map_seed,
                               const int64_t *map_left, const int64_t *map_right,
                               const int64_t *rmax, const fa_out_score *left_scores,
                               const fa_out_score *right_scores, int32_t right_from_left,
                               fa_region *regions) {
    if (!opt || nreads < 0) return FA_ERR_ARG;
    if (nreads > 0 && (!query_len || !read_chain_start || !read_chain_count || !chain_seed_start ||
                       !chain_seed_count || !chain_seeds || !seeds || !map_start || !map_count ||
                       !map_seed || !map_left || !map_right || !rmax || !regions))
        return FA_ERR_ARG;
    for (int64_t r = 0; r < nreads; ++r) {
        const int32_t qlen = query_len[r];
        for (int32_t lc = 0; lc < read_chain_count[r]; ++lc) {
            const int64_t c = read_chain_start[r] + lc;
            fa_region best{};
            best.seed_idx = -1;
            regions[c] = best;
            if (chain_seed_count[c] == 0 || rmax[2 * c] < 0) continue;  // :753-762
            int32_t best_score = INT32_MIN;
            for (int32_t m = 0; m < map_count[c]; ++m) {
                const int64_t mi = map_start[c] + m;
                const fa_seed &s = seeds[map_seed[mi]];
                fa_region g{};
                g.chain_idx = lc;
                g.seed_idx = map_seed[mi];
                g.qb = s.query_pos;  // :771-779
                g.qe = s.query_pos + s.len;
                g.rb = s.ref_pos;
                g.re = s.ref_pos + s.len;
                g.seedlen0 = s.len;
                g.frac_rep = chain_frac_rep ? chain_frac_rep[c] : 0.f;
                g.w = opt->w;
                int32_t total = s.len * opt->a;
                if (map_left[mi] >= 0) {  // :783-804
                    const fa_out_score &L = left_scores[map_left[mi]];
                    total = L.score;
                    if (L.global_score <= 0 || L.global_score <= L.score - opt->pen_clip5) {
                        g.qb = s.query_pos - L.query_end_pos;
                        g.rb = s.ref_pos - L.target_end_pos;
                        g.truesc = L.score;
                    } else {
                        g.qb = 0;
                        g.rb = s.ref_pos - L.gtarget_end_pos;
                        g.truesc = L.global_score;
                    }
                } else if (s.query_pos == 0) {
                    g.truesc = s.len * opt->a;
                }
                if (map_right[mi] >= 0) {  // :807-831
                    const fa_out_score &R = right_scores[map_right[mi]];
                    total = right_from_left ? R.score : total + R.score;
                    if (R.global_score <= 0 || R.global_score <= R.score - opt->pen_clip3) {
                        g.qe = s.query_pos + s.len + R.query_end_pos;
                        g.re = s.ref_pos + s.len + R.target_end_pos;
                    } else {
                        g.qe = qlen;
                        g.re = s.ref_pos + s.len + R.gtarget_end_pos;
                    }
                }
                g.score = total;  // :837-840
                if (g.truesc == 0) g.truesc = total;
                g.seedcov = 0;  // :843-853
                const int64_t *si = chain_seeds + chain_seed_start[c];
                for (int32_t k = 0; k < chain_seed_count[c]; ++k) {
                    const fa_seed &o = seeds[si[k]];
                    if (o.query_pos >= g.qb && o.query_pos + o.len <= g.qe && o.ref_pos >= g.rb &&
                        o.ref_pos + o.len <= g.re)
                        g.seedcov += o.len;
                }
                if (total > best_score) {  // :862-866
                    best_score = total;
                    best = g;
                }
            }
            regions[c] = best;
        }
    }
    return FA_OK;
}

}  // extern "C"
