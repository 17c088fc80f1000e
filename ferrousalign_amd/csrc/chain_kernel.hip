// Device side of the chain-extension path.
//
// Job slots: chain c's seed k owns slot slot_base[c] + k of each side, so the
// producer needs no compaction: a slot holds a job when its qlen is > 0, and
// the length sort that packs jobs into lane groups puts empty slots last.
//
//   chain_owner  one thread per read: which read owns each chain (a chain in
//                two reads' ranges is an argument error)
//   chain_jobs   one thread per chain: the reference window over the chain's
//                seeds and, per seed, the left and the right job
//                (collect_soa.rs:28-228; cal_max_gap chaining.rs:1108-1116)
//   chain_fill   16 lanes per job, both sides in one launch: the query
//                (read backwards for a left job) and the target (.pac
//                segment, backwards for a left job, bntseq.rs:451-522) packed
//                16 bases per word, as the extension kernels read them; query
//                base codes checked
//   chain_keys   one thread per job: right jobs' h0 from the left score, the
//                rest of the validation, the length-packing sort key and the
//                per-kernel job counts
//   chain_merge  one thread per chain: one region per seed from its left and
//                right scores, clipping and seed coverage, the chain's first
//                highest-scoring region (region.rs:737-890)
//
// A job's position in the reference's job order (read, chain, seeds in
// reverse) is kept as okey; errors go to the status block through atomicMin
// on okey | -code, so the error reported is the one of the first invalid job,
// as a sequential producer and batch validation would report it.
#include "chain_kernel.h"

namespace fa {

namespace {

__device__ __forceinline__ int32_t as_i32(double x) {
    if (!(x == x)) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return INT32_MIN;
    return (int32_t)x;
}

__device__ __forceinline__ int32_t cal_max_gap(const ChainOpt &o, int32_t qlen) {
    const int32_t l_del = as_i32((double)(qlen * o.a - o.o_del) / (double)o.e_del + 1.0);
    const int32_t l_ins = as_i32((double)(qlen * o.a - o.o_ins) / (double)o.e_ins + 1.0);
    int32_t l = l_del > l_ins ? l_del : l_ins;
    l = l > 1 ? l : 1;
    return l < (o.w << 1) ? l : (o.w << 1);
}

// ksw_affine_gap.rs:1001-1018: the band after end-bonus capping
__device__ __forceinline__ int32_t capped_w(const ChainOpt &o, int32_t qlen, int32_t w) {
    int32_t mi = as_i32(((double)qlen * o.max_score + o.end_bonus - o.sw_o_ins) / (double)o.sw_e_ins + 1.0);
    mi = mi > 1 ? mi : 1;
    w = w < mi ? w : mi;
    int32_t md = as_i32(((double)qlen * o.max_score + o.end_bonus - o.sw_o_del) / (double)o.sw_e_del + 1.0);
    md = md > 1 ? md : 1;
    return w < md ? w : md;
}

__device__ __forceinline__ uint32_t seq_words(int32_t len) { return (uint32_t)((len + 31) / 32) * 2u; }

// validate_job's order for everything but the base codes: h0, qlen, tlen, band, h0 range
__device__ __forceinline__ int job_error(int32_t h0, int32_t ql, int32_t tl, int32_t w) {
    if (h0 <= 0) return FA_ERR_H0;
    if (ql <= 0 || ql > FA_MAX_QLEN_PACKED) return FA_ERR_QLEN;
    if (tl < 0 || tl > FA_MAX_TLEN_PACKED) return FA_ERR_TLEN;
    if (w < 0) return FA_ERR_BAND;
    if (h0 > (1 << 19)) return FA_ERR_SCORING;
    return FA_OK;
}

__device__ __forceinline__ uint32_t pac_base(const uint8_t *pac, int64_t p) {
    return (pac[p >> 2] >> ((~(uint32_t)p & 3u) << 1)) & 3u;
}

__global__ __launch_bounds__(256) void chain_init_kernel(ChainStatus *st, int32_t *chain_read,
                                                         int64_t nchains) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c < nchains) chain_read[c] = -1;
    if (c == 0) {
        st->left_key = ~0ull;
        st->right_key = ~0ull;
        st->arg_err = 0;
        st->tmax[0] = st->tmax[1] = 0;
        st->any_job = 0;
        for (int k = 0; k < kChainRanks; ++k) {
            st->count[0][k] = st->count[1][k] = 0;
            st->work[0][k] = st->work[1][k] = 0;
        }
    }
}

__global__ __launch_bounds__(256) void chain_owner_kernel(ChainIn in, int32_t *chain_read,
                                                          ChainStatus *st) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= in.nreads) return;
    const int64_t qo = in.qoff[r];
    const int32_t ql = in.qlen[r];
    const int64_t c0 = in.rcs[r];
    const int32_t n = in.rcc[r];
    if (qo < 0 || ql < 0 || qo > in.qbytes - ql || n < 0 || c0 < 0 || c0 > in.nchains - n) {
        atomicOr(&st->arg_err, 1u);
        return;
    }
    // positions of a read's jobs are kept in 29 bits
    if (n > 0 && in.slot_base[c0 + n] - in.slot_base[c0] >= ((int64_t)1 << 29)) {
        atomicOr(&st->arg_err, 1u);
        return;
    }
    for (int32_t k = 0; k < n; ++k)
        if (atomicCAS(&chain_read[c0 + k], -1, (int32_t)r) != -1) atomicOr(&st->arg_err, 1u);
}

__device__ __forceinline__ void clear_slots(ChainSide &L, ChainSide &R, int64_t base, int32_t cnt) {
    for (int32_t k = 0; k < cnt; ++k) {
        L.qlen[base + k] = 0;
        L.words[base + k] = 0;
        L.err[base + k] = 0;
        R.qlen[base + k] = 0;
        R.words[base + k] = 0;
        R.err[base + k] = 0;
    }
}

__global__ __launch_bounds__(256) void chain_jobs_kernel(ChainIn in, ChainOpt o,
                                                         const int32_t *chain_read,
                                                         uint8_t *chain_state, ChainSide L,
                                                         ChainSide R, ChainStatus *st) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= in.nchains) return;
    const int64_t base = in.slot_base[c];
    const int32_t cnt = in.ccount[c];
    const int32_t r = chain_read[c];
    const bool visited = r >= 0 && in.kept[c] && cnt > 0;  // collect_soa.rs:73-91
    chain_state[c] = r < 0 ? 0 : (visited ? 2 : 1);
    if (!visited) {
        clear_slots(L, R, base, cnt);
        return;
    }
    const int64_t cs = in.cstart[c];
    const int32_t qlen = in.qlen[r];
    bool bad = cs < 0 || cs > in.ncseeds - cnt;
    // :94-117 window over all seeds, kept on one strand
    const uint64_t lp = (uint64_t)o.l_pac;
    uint64_t r0 = lp << 1, r1 = 0;
    for (int32_t k = 0; k < cnt && !bad; ++k) {
        const int64_t sid = in.cseeds[cs + k];
        if (sid < 0 || sid >= in.nseeds) {
            bad = true;
            break;
        }
        const fa_seed s = in.seeds[sid];
        if (s.query_pos < 0 || s.len < 0 || s.query_pos > qlen - s.len) {
            bad = true;
            break;
        }
        const int32_t lm = s.query_pos + cal_max_gap(o, s.query_pos);
        const uint64_t b = (uint64_t)s.ref_pos > (uint64_t)lm ? (uint64_t)s.ref_pos - (uint64_t)lm : 0;
        const int32_t rem = qlen - s.query_pos - s.len;
        const int32_t rm = rem + cal_max_gap(o, rem);
        const uint64_t e = (uint64_t)s.ref_pos + (uint64_t)s.len + (uint64_t)rm;
        r0 = r0 < b ? r0 : b;
        r1 = r1 > e ? r1 : e;
    }
    if (bad) {
        atomicOr(&st->arg_err, 1u);
        chain_state[c] = 1;
        clear_slots(L, R, base, cnt);
        return;
    }
    r1 = r1 < (lp << 1) ? r1 : (lp << 1);
    if (r0 < lp && lp < r1) {
        if ((uint64_t)in.seeds[in.cseeds[cs]].ref_pos < lp)
            r1 = lp;
        else
            r0 = lp;
    }
    const int64_t wlen = r1 > r0 ? (int64_t)(r1 - r0) : 0;
    const int64_t pos0 = base - in.slot_base[in.rcs[r]];
    bool had_job = false;
    for (int32_t k = 0; k < cnt; ++k) {
        const int64_t slot = base + k;
        const fa_seed s = in.seeds[in.cseeds[cs + k]];
        const uint64_t okey = ((uint64_t)(uint32_t)r << 33) | ((uint64_t)(pos0 + (cnt - 1 - k)) << 4);
        // left (:135-176)
        int32_t lq = 0;
        L.err[slot] = 0;
        if (s.query_pos > 0) {
            const int64_t tmp = s.ref_pos - (int64_t)r0;
            if (tmp > 0 && tmp <= wlen) {
                const int32_t tl = (int32_t)tmp, h0 = s.len * o.a;
                const int e = job_error(h0, s.query_pos, tl, o.w);
                L.okey[slot] = okey;
                had_job = true;
                if (e != FA_OK) {
                    atomicMin((unsigned long long *)&st->left_key, okey | (uint64_t)(-e));
                } else {
                    lq = s.query_pos;
                    L.qsrc[slot] = in.qoff[r];
                    L.tlen[slot] = tl;
                    L.seg[slot] = (int64_t)r0;
                    L.h0[slot] = h0;
                    L.w[slot] = capped_w(o, lq, o.w);
                }
            }
        }
        L.qlen[slot] = lq;
        L.words[slot] = lq ? (uint64_t)seq_words(lq) + seq_words(L.tlen[slot]) : 0;
        // right (:178-205); its h0 is the left score, known after the left batch
        int32_t rq = 0;
        uint8_t rerr = 0;
        const int32_t qend = s.query_pos + s.len;
        if (qend < qlen) {
            const int64_t tmp = s.ref_pos + s.len - (int64_t)r0;
            if (tmp >= 0 && tmp < wlen) {
                rq = qlen - qend;
                const int32_t tl = (int32_t)(wlen - tmp);
                int e = FA_OK;
                if (rq > FA_MAX_QLEN_PACKED) e = FA_ERR_QLEN;
                else if (tl < 0 || tl > FA_MAX_TLEN_PACKED) e = FA_ERR_TLEN;
                else if (o.w < 0) e = FA_ERR_BAND;
                rerr = (uint8_t)(-e);
                R.okey[slot] = okey;
                R.qsrc[slot] = in.qoff[r] + qend;
                R.tlen[slot] = tl;
                R.seg[slot] = (int64_t)r0 + tmp;
                R.h0[slot] = s.len * o.a;
                R.w[slot] = capped_w(o, rq, o.w);
            }
        }
        R.qlen[slot] = rq;
        R.err[slot] = rerr;
        R.words[slot] = (rq && !rerr) ? (uint64_t)seq_words(rq) + seq_words(R.tlen[slot]) : 0;
        if (rq) had_job = true;
    }
    const uint64_t m = __ballot(had_job);  // one flag write per wavefront
    if (had_job && (int)__lane_id() == __ffsll((unsigned long long)m) - 1) atomicOr(&st->any_job, 1u);
}

// 16 lanes per job, lane w builds 2-bit words w, w + 16, ... of the query,
// then of the target.  A query word takes its 16 codes from five aligned
// 32-bit loads funnelled to the word's first byte (the codes of a left job
// sit backwards in the 16 bytes that end at its first base); a target word
// takes its 16 bases from two aligned 32-bit loads of the .pac (16 bases span
// at most 5 bytes).
__device__ __forceinline__ void window16(const uint8_t *base, int64_t at, uint32_t b[4]) {
    const int64_t w0 = at >> 2;
    const uint32_t sh = (uint32_t)(at & 3);
    const uint32_t *p = reinterpret_cast<const uint32_t *>(base) + w0;
    const uint32_t x0 = p[0], x1 = p[1], x2 = p[2], x3 = p[3], x4 = p[4];
    b[0] = __builtin_amdgcn_alignbyte(x1, x0, sh);
    b[1] = __builtin_amdgcn_alignbyte(x2, x1, sh);
    b[2] = __builtin_amdgcn_alignbyte(x3, x2, sh);
    b[3] = __builtin_amdgcn_alignbyte(x4, x3, sh);
}

__global__ __launch_bounds__(256) void chain_fill_kernel(ChainIn in, const uint8_t *__restrict__ pac,
                                                         int64_t l_pac, ChainSide L, ChainSide R,
                                                         int64_t nslots, ChainStatus *st) {
    const int64_t gs = (int64_t)blockIdx.x * 16 + (threadIdx.x >> 4);
    const int lane = (int)(threadIdx.x & 15);
    if (gs >= 2 * nslots) return;
    const int sd = gs >= nslots;  // uniform over the slot's 16 lanes
    const int64_t slot = sd ? gs - nslots : gs;
    const ChainSide &S = sd ? R : L;
    const uint64_t nw = S.words[slot];
    if (nw == 0) return;
    const int32_t ql = S.qlen[slot], tl = S.tlen[slot];
    const int64_t qsrc = S.qsrc[slot], seg = S.seg[slot];
    const uint32_t qo = (uint32_t)(S.wend[slot] - nw);
    const uint32_t nqw = seq_words(ql), ntw = seq_words(tl);
    const uint32_t to = qo + nqw;
    if (lane == 0) {
        S.qo[slot] = qo;
        S.to[slot] = to;
    }
    // the 64-byte padding before qseq keeps a left job's first window in the allocation
    const uint8_t *qbase = in.qseq - 64;
    bool bad = false;
    for (uint32_t b0 = 0; b0 < nqw; b0 += 16) {
        const uint32_t w = b0 + (uint32_t)lane;
        uint32_t word = 0, nb = 0;
        if (w < nqw) {
            const int32_t p0 = (int32_t)w * 16;
            const int n = ql - p0 < 16 ? ql - p0 : 16;  // <= 0 past the query (padding words)
            if (n > 0) {
                uint32_t b[4];
                window16(qbase, 64 + (sd == 0 ? qsrc + ql - p0 - 16 : qsrc + p0), b);
#pragma unroll
                for (int t = 0; t < 16; ++t) {
                    const int x = sd == 0 ? 15 - t : t;
                    uint32_t c = (b[x >> 2] >> (8 * (x & 3))) & 0xFFu;
                    if (t < n) {
                        if (c > 4) bad = true;
                        if (c >= 4) {
                            nb |= 1u << t;
                            c = 0;
                        }
                        word |= (c & 3u) << (2 * t);
                    }
                }
            }
            S.seq2[qo + w] = word;
        }
        const uint32_t nb_hi = __shfl_down(nb, 1, 16);
        if (w < nqw && (w & 1u) == 0) S.seqn[qo / 2 + w / 2] = nb | (nb_hi << 16);
    }
    // target base k is bidirectional position seg + k (right) or seg + tl - 1 - k
    // (left); on the reverse strand it is the complement of forward base
    // 2 l_pac - 1 - p.  Forward positions f of one word are consecutive.
    const bool rstrand = seg >= l_pac;
    const int fdir = (sd == 0) != rstrand ? -1 : 1;
    const uint32_t *pac32 = reinterpret_cast<const uint32_t *>(pac);
    for (uint32_t b0 = 0; b0 < ntw; b0 += 16) {
        const uint32_t w = b0 + (uint32_t)lane;
        if (w >= ntw) break;
        uint32_t word = 0;
        const int32_t k0 = (int32_t)w * 16;
        const int n = tl - k0 < 16 ? tl - k0 : 16;
        if (n > 0) {
            const int64_t p0 = sd == 0 ? seg + tl - 1 - k0 : seg + k0;
            const int64_t f0 = rstrand ? 2 * l_pac - 1 - p0 : p0;
            const int64_t fmin = fdir > 0 ? f0 : f0 - (n - 1);
            const int64_t wi = fmin >> 4;  // 32-bit word holding the byte of fmin
            const uint64_t v = (uint64_t)pac32[wi] | ((uint64_t)pac32[wi + 1] << 32);
            const int32_t g0 = (int32_t)(f0 - (wi << 4));  // base offset inside the 8 bytes
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                if (t < n) {
                    const int32_t g = g0 + fdir * t;
                    uint32_t c = (uint32_t)(v >> (8 * (g >> 2) + 6 - 2 * (g & 3))) & 3u;
                    if (rstrand) c = 3u - c;
                    word |= c << (2 * t);
                }
            }
        }
        S.seq2[to + w] = word;
        if ((w & 1u) == 0) S.seqn[to / 2 + w / 2] = 0;
    }
    // base codes above 4 anywhere in the query (validate_job)
    const uint64_t any = __ballot(bad);
    const uint64_t grp = 0xFFFFull << (threadIdx.x & 48u);
    if (lane == 0 && (any & grp)) {
        S.err[slot] = (uint8_t)(-FA_ERR_CODE);
        if (sd == 0) atomicMin((unsigned long long *)&st->left_key, S.okey[slot] | (uint64_t)(-FA_ERR_CODE));
    }
}

__global__ __launch_bounds__(256) void chain_keys_kernel(ChainOpt o, ChainRanks rk, ChainSide S,
                                                         ChainSide Lft, int sd, int64_t nslots,
                                                         ChainStatus *st) {
    __shared__ unsigned long long cnt[kChainRanks], wk[kChainRanks];
    __shared__ uint32_t tm;
    if (threadIdx.x < kChainRanks) {
        cnt[threadIdx.x] = 0;
        wk[threadIdx.x] = 0;
    }
    if (threadIdx.x == 0) tm = 0;
    __syncthreads();
    const int64_t slot = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (slot < nslots) {
        const int32_t ql = S.qlen[slot];
        bool valid = ql > 0;
        const uint8_t err = valid ? S.err[slot] : 0;
        int32_t h0 = valid ? S.h0[slot] : 0;
        if (valid && sd == 1) {
            // right job: h0 = the seed's left-extended score, else seed_len * a
            if (Lft.qlen[slot] > 0) h0 = Lft.out[slot * 6];
            S.h0[slot] = h0;
            int e = FA_OK;
            if (h0 <= 0) e = FA_ERR_H0;
            else if (err && err != (uint8_t)(-FA_ERR_CODE)) e = -(int)err;
            else if (h0 > (1 << 19)) e = FA_ERR_SCORING;
            else if (err) e = FA_ERR_CODE;
            if (e != FA_OK) {
                atomicMin((unsigned long long *)&st->right_key, S.okey[slot] | (uint64_t)(-e));
                valid = false;
            }
        } else if (valid && err) {
            valid = false;  // left job with a bad base code, already reported
        }
        uint32_t key = 0xFFFFFFFFu;
        if (valid) {
            const int32_t tl = S.tlen[slot];
            const int64_t bound = (int64_t)h0 + (int64_t)ql * o.max_score;
            const int fits = (o.sw16ok && bound < 4096 && bound * o.k16 < 32768) ? (bound < 2048 ? 2 : 1) : 0;
            const uint32_t rank = rk.r[fits][ql];
            key = (rank << 20) | ((uint32_t)(255 - ql) << 12) | (uint32_t)(4095 - tl);
            atomicAdd(&cnt[rank], 1ull);
            atomicAdd(&wk[rank], (unsigned long long)ql * (unsigned long long)(tl > 1 ? tl : 1));
            atomicMax(&tm, (uint32_t)tl);
        }
        S.key[slot] = key;
        S.slot[slot] = (uint32_t)slot;
    }
    __syncthreads();
    if (threadIdx.x < kChainRanks && cnt[threadIdx.x]) {
        atomicAdd(&st->count[sd][threadIdx.x], cnt[threadIdx.x]);
        atomicAdd(&st->work[sd][threadIdx.x], wk[threadIdx.x]);
    }
    if (threadIdx.x == 0 && tm) atomicMax(&st->tmax[sd], tm);
}

__global__ __launch_bounds__(256) void chain_merge_kernel(ChainIn in, ChainOpt o,
                                                          const int32_t *chain_read,
                                                          const uint8_t *chain_state, ChainSide L,
                                                          ChainSide R, fa_region *regions) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= in.nchains) return;
    fa_region best{};
    best.seed_idx = -1;
    if (chain_state[c] == 2) {
        const int32_t r = chain_read[c];
        const int32_t qlen = in.qlen[r];
        const int32_t lc = (int32_t)(c - in.rcs[r]);
        const int64_t cs = in.cstart[c], base = in.slot_base[c];
        const int32_t cnt = in.ccount[c];
        const float frac = in.frac ? in.frac[c] : 0.f;
        int32_t best_score = INT32_MIN;
        for (int32_t k = cnt - 1; k >= 0; --k) {  // the producer's seed order
            const int64_t slot = base + k;
            const int64_t sid = in.cseeds[cs + k];
            const fa_seed s = in.seeds[sid];
            fa_region g{};
            g.chain_idx = lc;
            g.seed_idx = sid;
            g.qb = s.query_pos;  // :771-779
            g.qe = s.query_pos + s.len;
            g.rb = s.ref_pos;
            g.re = s.ref_pos + s.len;
            g.seedlen0 = s.len;
            g.frac_rep = frac;
            g.w = o.w;
            int32_t total = s.len * o.a;
            if (L.qlen[slot] > 0) {  // :783-804
                const int32_t *x = L.out + slot * 6;
                const int32_t sc = x[0], tle = x[1], gtle = x[2], qle = x[3], gsc = x[4];
                total = sc;
                if (gsc <= 0 || gsc <= sc - o.pen_clip5) {
                    g.qb = s.query_pos - qle;
                    g.rb = s.ref_pos - tle;
                    g.truesc = sc;
                } else {
                    g.qb = 0;
                    g.rb = s.ref_pos - gtle;
                    g.truesc = gsc;
                }
            } else if (s.query_pos == 0) {
                g.truesc = s.len * o.a;
            }
            if (R.qlen[slot] > 0) {  // :807-831, right h0 chained from the left score
                const int32_t *x = R.out + slot * 6;
                const int32_t sc = x[0], tle = x[1], gtle = x[2], qle = x[3], gsc = x[4];
                total = sc;
                if (gsc <= 0 || gsc <= sc - o.pen_clip3) {
                    g.qe = s.query_pos + s.len + qle;
                    g.re = s.ref_pos + s.len + tle;
                } else {
                    g.qe = qlen;
                    g.re = s.ref_pos + s.len + gtle;
                }
            }
            g.score = total;  // :837-840
            if (g.truesc == 0) g.truesc = total;
            int32_t cov = 0;  // :843-853
            for (int32_t j = 0; j < cnt; ++j) {
                const fa_seed t = in.seeds[in.cseeds[cs + j]];
                if (t.query_pos >= g.qb && t.query_pos + t.len <= g.qe && t.ref_pos >= g.rb &&
                    t.ref_pos + t.len <= g.re)
                    cov += t.len;
            }
            g.seedcov = cov;
            if (total > best_score) {  // :862-866
                best_score = total;
                best = g;
            }
        }
    }
    regions[c] = best;
}

inline unsigned grid(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

hipError_t chain_init(ChainStatus *st, int32_t *chain_read, int64_t nchains, hipStream_t s) {
    hipLaunchKernelGGL(chain_init_kernel, dim3(grid(nchains > 0 ? nchains : 1, 256)), dim3(256), 0, s,
                       st, chain_read, nchains);
    return hipGetLastError();
}

hipError_t chain_owner(const ChainIn &in, int32_t *chain_read, ChainStatus *st, hipStream_t s) {
    if (in.nreads == 0) return hipSuccess;
    hipLaunchKernelGGL(chain_owner_kernel, dim3(grid(in.nreads, 256)), dim3(256), 0, s, in,
                       chain_read, st);
    return hipGetLastError();
}

hipError_t chain_jobs(const ChainIn &in, const ChainOpt &o, const int32_t *chain_read,
                      uint8_t *chain_state, ChainSide L, ChainSide R, ChainStatus *st,
                      hipStream_t s) {
    if (in.nchains == 0) return hipSuccess;
    hipLaunchKernelGGL(chain_jobs_kernel, dim3(grid(in.nchains, 256)), dim3(256), 0, s, in, o,
                       chain_read, chain_state, L, R, st);
    return hipGetLastError();
}

hipError_t chain_fill(const ChainIn &in, const uint8_t *pac, int64_t l_pac, ChainSide L, ChainSide R,
                      int64_t nslots, ChainStatus *st, hipStream_t s) {
    if (nslots == 0) return hipSuccess;
    hipLaunchKernelGGL(chain_fill_kernel, dim3(grid(2 * nslots, 16)), dim3(256), 0, s, in, pac, l_pac,
                       L, R, nslots, st);
    return hipGetLastError();
}

hipError_t chain_keys(const ChainOpt &o, const ChainRanks &rk, ChainSide S, const ChainSide &Lft,
                      int sd, int64_t nslots, ChainStatus *st, hipStream_t s) {
    if (nslots == 0) return hipSuccess;
    hipLaunchKernelGGL(chain_keys_kernel, dim3(grid(nslots, 256)), dim3(256), 0, s, o, rk, S, Lft,
                       sd, nslots, st);
    return hipGetLastError();
}

hipError_t chain_merge(const ChainIn &in, const ChainOpt &o, const int32_t *chain_read,
                       const uint8_t *chain_state, const ChainSide &L, const ChainSide &R,
                       fa_region *regions, hipStream_t s) {
    if (in.nchains == 0) return hipSuccess;
    hipLaunchKernelGGL(chain_merge_kernel, dim3(grid(in.nchains, 256)), dim3(256), 0, s, in, o,
                       chain_read, chain_state, L, R, regions);
    return hipGetLastError();
}

}  // namespace fa
