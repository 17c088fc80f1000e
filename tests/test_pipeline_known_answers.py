"""Known answers for the extension-job producer and the region merge, worked
by hand from the reference (collect_soa.rs:28-228, chaining.rs:1108-1116,
region.rs:737-890), for both the Python restatement (tests/pipeline_oracle.py)
and the library's host code (fa_collect_extension_jobs,
fa_merge_scores_to_regions).  Default MemOpt: a = 1, o_del = o_ins = 6,
e_del = e_ins = 1, w = 100, pen_clip5 = pen_clip3 = 5.

cal_max_gap(q) = min(max((q*a - 6)/1 + 1 truncated, 1), 2w), so
cal_max_gap(0) = 1, (10) = 5, (30) = 25, (40) = 35."""
import numpy as np

from ferrousalign_amd import _native as N
from ferrousalign_amd.pipeline import ChainBatch, MemOpt, collect_extension_jobs, merge_scores_to_regions
from tests import pipeline_oracle as PO

L_PAC = 100000


def _batch(reads):
    """reads: list of (qlen, chains); a chain is a list of (ref_pos,
    query_pos, len) seeds, all kept."""
    qlens, rcs, rcc, kept, frac, css, csc, cs, seeds = [], [], [], [], [], [], [], [], []
    for qlen, chains in reads:
        qlens.append(qlen)
        rcs.append(len(kept))
        rcc.append(len(chains))
        for ch in chains:
            css.append(len(cs))
            csc.append(len(ch))
            kept.append(1)
            frac.append(0.25)
            for s in ch:
                cs.append(len(seeds))
                seeds.append(s)
    sd = np.zeros(len(seeds), N.SEED_DTYPE)
    for i, x in enumerate(seeds):
        sd[i] = x
    return ChainBatch(np.array(qlens, np.int32), np.array(rcs, np.int64), np.array(rcc, np.int32),
                      np.array(kept, np.uint8), np.array(frac, np.float32), np.array(css, np.int64),
                      np.array(csc, np.int32), np.array(cs, np.int64), sd)


def _jobs_both(cb, opt=MemOpt()):
    """Jobs, per-chain (seed, left, right) maps and rmax, from the oracle and
    from the library, as plain tuples."""
    ol, orr, omaps, ormax = PO.collect_extension_jobs(opt, L_PAC, cb)
    f = ("qbeg", "qlen", "h0", "seg_beg", "seg_len", "seg_rev")
    o = ([tuple(j[k] for k in f) for j in ol], [tuple(j[k] for k in f) for j in orr], omaps, ormax)
    left, right, m = collect_extension_jobs(opt, L_PAC, cb)
    lib_l = [tuple(int(left[k][i]) for k in f) for i in range(len(left["qlen"]))]
    lib_r = [tuple(int(right[k][i]) for k in f) for i in range(len(right["qlen"]))]
    lib_maps = [[(int(m.seed[i]), int(m.left[i]), int(m.right[i]))
                 for i in range(int(m.start[c]), int(m.start[c]) + int(m.count[c]))] for c in range(cb.nchains)]
    lib_rmax = [None if int(m.rmax[c][0]) == -1 else (int(m.rmax[c][0]), int(m.rmax[c][1]))
                for c in range(cb.nchains)]
    return [o, (lib_l, lib_r, lib_maps, lib_rmax)], m


def test_interior_seed_windows():
    # qlen 100, seed query 40..60 at ref 5000: left margin 40 + 35, right
    # margin 40 + 35, window [4925, 5095); left job: the 40 query bases
    # before the seed against the 75 window bases before it (reversed), h0 =
    # seed score 20; right job: query 60..100 against window 5020..5095, h0 0
    cb = _batch([(100, [[(5000, 40, 20)]])])
    for left, right, maps, rmax in _jobs_both(cb)[0]:
        assert rmax == [(4925, 5095)]
        assert left == [(0, 40, 20, 4925, 75, 1)]
        assert right == [(60, 40, 0, 5020, 75, 0)]
        assert maps == [[(0, 0, 0)]]


def test_window_clipped_at_reference_start_and_query_ends():
    # seed at ref 30 with query 40 before it: the left edge saturates at 0
    # (30 - 75), so the left target is the 30 bases before the seed; a seed
    # at query 0 has no left job (its left margin cal_max_gap(0) = 1), a seed
    # ending at the query end no right job
    cb = _batch([(100, [[(30, 40, 20)]]), (100, [[(7000, 0, 30)]]), (100, [[(9000, 70, 30)]])])
    for left, right, maps, rmax in _jobs_both(cb)[0]:
        assert rmax[0] == (0, 30 + 20 + 40 + 35)
        assert rmax[1] == (6999, 7000 + 30 + 70 + 65)    # remaining 70: cal_max_gap(70) = 65
        assert rmax[2] == (9000 - 70 - 65, 9030 + 1)     # remaining 0: right margin 0 + 1
        assert left == [(0, 40, 20, 0, 30, 1), (0, 70, 30, 9000 - 135, 135, 1)]
        assert right == [(60, 40, 0, 50, 75, 0), (30, 70, 0, 7030, 135, 0)]
        assert maps == [[(0, 0, 0)], [(1, -1, 1)], [(2, 1, -1)]]


def test_chain_across_the_strand_boundary():
    # two seeds around l_pac: seed 0 (query 10, ref l_pac - 50) has margins
    # 10 + 5 and 70 + 65, seed 1 (query 50, ref l_pac + 10) 50 + 45 and
    # 30 + 25, so the window [l_pac - 85, l_pac + 105) crosses l_pac.  With
    # the chain's first seed on the forward strand it is cut to end at
    # l_pac; seeds are mapped in reverse chain order, and seed 1 lies outside
    # the cut window (left distance 95 > 85, right start 115 >= 85)
    cb = _batch([(100, [[(L_PAC - 50, 10, 20), (L_PAC + 10, 50, 20)]])])
    for left, right, maps, rmax in _jobs_both(cb)[0]:
        assert rmax == [(L_PAC - 85, L_PAC)]
        assert maps == [[(1, -1, -1), (0, 0, 0)]]
        assert left == [(0, 10, 20, L_PAC - 85, 35, 1)]
        assert right == [(30, 70, 0, L_PAC - 30, 30, 0)]
    # the same seeds with the reverse-strand one first: the window starts at
    # l_pac, and the forward-strand seed (now seed 1) lies left of it
    cb = _batch([(100, [[(L_PAC + 10, 50, 20), (L_PAC - 50, 10, 20)]])])
    for left, right, maps, rmax in _jobs_both(cb)[0]:
        assert rmax == [(L_PAC, L_PAC + 105)]
        assert maps == [[(1, -1, -1), (0, 0, 0)]]
        assert left == [(0, 50, 20, L_PAC, 10, 1)]
        assert right == [(70, 30, 0, L_PAC + 30, 75, 0)]


def _score(sc, tle, gtle, qle, gsc):
    return np.array([sc, tle, gtle, qle, gsc, 0], np.int32)


def _merge_both(cb, m, maps, rmax, ls, rs, rfl):
    opt = MemOpt()
    o = PO.merge_scores_to_regions(opt, cb, maps, rmax, ls, rs, rfl)
    g = merge_scores_to_regions(opt, cb, m, ls, rs, right_from_left=rfl)
    fields = ("qb", "qe", "rb", "re", "score", "truesc", "seedcov", "seed_idx")
    out = []
    for c in range(cb.nchains):
        out.append((None if o[c] is None else tuple(int(o[c][f]) for f in fields),
                    None if int(g[c]["seed_idx"]) == -1 else tuple(int(g[c][f]) for f in fields)))
    return out


def test_merge_clipping_and_scores():
    cb = _batch([(100, [[(5000, 40, 20)]])])
    both, m = _jobs_both(cb)
    _, _, maps, rmax = both[0]
    # left local (gscore 28 <= 35 - 5): qb = 40 - 15, rb = 5000 - 18, truesc = 35;
    # right global (gscore 48 > 50 - 5): qe = qlen, re = 5020 + gtle 38
    ls = np.stack([_score(35, 18, 20, 15, 28)])
    rs = np.stack([_score(50, 30, 38, 28, 48)])
    for rfl, total in ((False, 35 + 50), (True, 50)):
        for o, g in _merge_both(cb, m, maps, rmax, ls, rs, rfl):
            assert o == g == (25, 100, 4982, 5058, total, 35, 20, 0), (rfl, o, g)
    # left global (gscore 31 > 30): qb = 0, rb = 5000 - gtle 20, truesc = gscore;
    # right local with gscore <= 0: qe = 60 + qle 28, re = 5020 + tle 30
    ls = np.stack([_score(35, 18, 20, 15, 31)])
    rs = np.stack([_score(50, 30, 38, 28, 0)])
    for o, g in _merge_both(cb, m, maps, rmax, ls, rs, False):
        assert o == g == (0, 88, 4980, 5050, 85, 31, 20, 0), (o, g)
    # on the boundary, gscore == score - pen_clip (30 = 35 - 5, 45 = 50 - 5),
    # both sides clip locally (the tests are <=)
    ls = np.stack([_score(35, 18, 20, 15, 30)])
    rs = np.stack([_score(50, 30, 38, 28, 45)])
    for o, g in _merge_both(cb, m, maps, rmax, ls, rs, False):
        assert o == g == (25, 88, 4982, 5050, 85, 35, 20, 0), (o, g)


def test_merge_ties_keep_the_first_seed_in_mapping_order():
    # two seeds, equal totals: the mapping runs in reverse chain order and a
    # region replaces the best only when strictly better, so the chain's last
    # seed is kept; a seed without a left job (query 0) takes truesc = its
    # seed score, and seedcov counts the chain's seeds inside the region
    cb = _batch([(100, [[(5000, 0, 30), (5040, 40, 30)]])])
    both, m = _jobs_both(cb)
    left, right, maps, rmax = both[0]
    assert maps == [[(1, 0, 0), (0, -1, 1)]]
    ls = np.stack([_score(40, 40, 40, 40, 40)])                 # seed 1: global clip, qb 0
    rs = np.stack([_score(30, 30, 30, 30, 30), _score(70, 70, 70, 70, 70)])
    for o, g in _merge_both(cb, m, maps, rmax, ls, rs, False):
        # seed 1: 40 + 30 = 70; seed 0: 30 + 70 = 100 wins
        assert o == g == (0, 100, 5000, 5100, 100, 30, 60, 0), (o, g)
    rs = np.stack([_score(60, 30, 30, 30, 30), _score(70, 70, 70, 70, 70)])
    for o, g in _merge_both(cb, m, maps, rmax, ls, rs, False):
        # both 100: seed 1 (first in mapping order) stays: qb 0, rb 5040 - 40,
        # qe 70 + 30 = 100 (gscore 30 <= 55 clips locally), re 5070 + 30
        assert o == g == (0, 100, 5000, 5100, 100, 40, 60, 1), (o, g)


def test_seed_at_reference_start_has_no_left_job_and_takes_the_total_as_truesc():
    # seed query 20..50 at ref 0: the window starts at 0 (0 - 35 saturates),
    # so its left target is empty and there is no left job although query
    # bases precede the seed; with no left score truesc stays 0 and becomes
    # the region's total (region.rs: if region.truesc == 0)
    cb = _batch([(100, [[(0, 20, 30)]])])
    both, m = _jobs_both(cb)
    for left, right, maps, rmax in both:
        assert rmax == [(0, 125)] and left == [] and right == [(50, 50, 0, 30, 95, 0)]
        assert maps == [[(0, -1, 0)]]
    _, _, maps, rmax = both[0]
    ls = np.zeros((0, 6), np.int32)
    rs = np.stack([_score(40, 25, 30, 20, 10)])   # local: qe 50 + 20, re 30 + 25
    for rfl, total in ((False, 70), (True, 40)):
        for o, g in _merge_both(cb, m, maps, rmax, ls, rs, rfl):
            assert o == g == (20, 70, 0, 55, total, total, 30, 0), (rfl, o, g)
