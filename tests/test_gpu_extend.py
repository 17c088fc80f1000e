"""GPU parity: every output field and the cell count of the HIP path, through
the C ABI, against the CPU oracle (ksw_extend2 restatement) on seeded pairs."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _run_both(ps, sw, flags=0, lanes=None):
    from ferrousalign_amd import banded_swa as B
    flat = ps.flat()
    ora, ocells, orc = O.extend_batch(sw, *flat)
    assert (orc == 0).all()
    if lanes is None:
        gpu, gcells = B.extend_arrays(sw, *flat, flags=flags, want_cells=True)
    else:
        db = B.DeviceBatch(sw, *flat)
        db.set_launch(lanes)
        db.run(flags)
        gpu, gcells = db.download()
        db.free()
    return O.to_outscore_order(ora), ocells, gpu, gcells


def _assert_same(ora, ocells, gpu, gcells, label):
    bad = np.nonzero((ora != gpu).any(axis=1) | (ocells.astype(np.uint64) != gcells.astype(np.uint64)))[0]
    if bad.size:
        k = bad[0]
        raise AssertionError("%s: %d/%d pairs differ; first pair %d oracle=%s cells=%d gpu=%s cells=%d" % (
            label, bad.size, len(gpu), k, ora[k].tolist(), int(ocells[k]), gpu[k].tolist(), int(gcells[k])))


@pytest.mark.parametrize("name,n", [("cpu_ref_151bp_4k", 4096), ("illumina_150bp_1m", 8192),
                                    ("uneven_mix_4m", 8192)])
def test_parity_workloads(name, n):
    from ferrousalign_amd import synth
    ps = synth.workload(name, n=n)
    sw = synth.default_scoring()
    print("parity", name, n, flush=True)
    ora, oc, gpu, gc = _run_both(ps, sw)
    _assert_same(ora, oc, gpu, gc, name)


@pytest.mark.parametrize("lanes", [0, 4, 8, 16])
def test_parity_lane_groups(lanes):
    from ferrousalign_amd import synth
    ps = synth.workload("uneven_mix_4m", n=8192)
    # a lane group of L lanes holds at most 32 columns per lane (0: narrowest group per pair)
    if lanes:
        ps = ps.subset(np.nonzero(ps.qlen + 1 <= lanes * 32)[0])
    sw = synth.default_scoring()
    print("lanes", lanes, ps.n, flush=True)
    ora, oc, gpu, gc = _run_both(ps, sw, lanes=lanes)
    _assert_same(ora, oc, gpu, gc, "lanes=%d" % lanes)


def test_parity_wide_kernel():
    from ferrousalign_amd import synth, _native as N
    ps = synth.workload("uneven_mix_4m", n=8192)
    sw = synth.default_scoring()
    print("wide kernel", flush=True)
    ora, oc, gpu, gc = _run_both(ps, sw, flags=N.FA_SCORE_I32)
    _assert_same(ora, oc, gpu, gc, "32-bit kernel")


def test_overflow_fallback_mixed():
    """Pairs whose score bound does not fit the 16-bit kernel run the 32-bit
    kernel in the same call; results stay exact."""
    from ferrousalign_amd import synth
    ps = synth.workload("illumina_150bp_1m", n=6000)
    big = np.arange(0, ps.n, 7)
    ps.h0[big] = 8100 + (big % 50)  # h0 + qlen exceeds the 16-bit bound
    sw = synth.default_scoring()
    print("overflow fallback", flush=True)
    ora, oc, gpu, gc = _run_both(ps, sw)
    _assert_same(ora, oc, gpu, gc, "mixed 16/32-bit")


def _random_case_batch(rng, n, qmax=60, tmax=90):
    from ferrousalign_amd.synth import PairSet
    q = np.zeros((n, 256), np.uint8)
    t = np.zeros((n, 352), np.uint8)
    ql = rng.integers(1, qmax + 1, n).astype(np.int32)
    tl = rng.integers(0, tmax + 1, n).astype(np.int32)
    for k in range(n):
        alpha = 5 if rng.random() < 0.3 else 4
        q[k, :ql[k]] = rng.integers(0, alpha, ql[k])
        if rng.random() < 0.6 and tl[k] > 0:  # mutated copy of the query
            m = min(ql[k], tl[k])
            t[k, :m] = q[k, :m]
            flip = rng.random(m) < 0.1
            t[k, :m][flip] = rng.integers(0, alpha, flip.sum())
            t[k, m:tl[k]] = rng.integers(0, alpha, tl[k] - m)
        else:
            t[k, :tl[k]] = rng.integers(0, alpha, tl[k])
    h0 = rng.integers(1, 80, n).astype(np.int32)
    w = rng.choice([0, 1, 2, 3, 5, 10, 100], n).astype(np.int32)
    return PairSet(q, t, ql, tl, h0, w)


@pytest.mark.parametrize("seed", range(4))
def test_parity_random_edge_cases(seed):
    from ferrousalign_amd.banded_swa import BandedPairWiseSW, bwa_fill_scmat
    from ferrousalign_amd import _native as N
    rng = np.random.default_rng(seed)
    ps = _random_case_batch(rng, 3000)
    configs = [
        BandedPairWiseSW(6, 1, 6, 1, 100, 5, 5, 5, bwa_fill_scmat(1, 4, -1)),
        BandedPairWiseSW(5, 2, 3, 1, 0, 0, 5, 5, bwa_fill_scmat(2, 3, -1)),     # zdrop off
        BandedPairWiseSW(0, 1, 0, 1, 10, 5, 5, 5, bwa_fill_scmat(1, 1, 0)),     # zero gap open
        BandedPairWiseSW(12, 3, 9, 2, 30, 20, 5, 5, bwa_fill_scmat(3, 5, -2)),
    ]
    for ci, sw in enumerate(configs):
        for flags in (0, N.FA_SCORE_I32):
            print("random", seed, ci, flags, flush=True)
            ora, oc, gpu, gc = _run_both(ps, sw, flags=flags)
            _assert_same(ora, oc, gpu, gc, "random seed=%d cfg=%d flags=%d" % (seed, ci, flags))


def test_parity_large_band_and_zdrop():
    """Bands and z-drops beyond what 16-bit state can hold directly: a huge
    end_bonus keeps the w cap (ksw_affine_gap.rs:1001-1018) above 4096, and
    z-drop values of 5,000 and 100,000 exceed every reachable best - m_row."""
    from ferrousalign_amd.banded_swa import BandedPairWiseSW, bwa_fill_scmat
    from ferrousalign_amd import _native as N
    rng = np.random.default_rng(11)
    ps = _random_case_batch(rng, 2000)
    ps.w[:] = 50000
    for zdrop in (100, 5000, 100000):
        sw = BandedPairWiseSW(6, 1, 6, 1, zdrop, 100000, 5, 5, bwa_fill_scmat(1, 4, -1))
        for flags in (0, N.FA_SCORE_I32):
            print("large band", zdrop, flags, flush=True)
            ora, oc, gpu, gc = _run_both(ps, sw, flags=flags)
            _assert_same(ora, oc, gpu, gc, "large band zdrop=%d flags=%d" % (zdrop, flags))


def test_parity_wide_slot_boundary():
    """Pairs on both sides of the score bound the wide slot counts need
    (Hmax = h0 + qlen * max(mat) < 2048 for S = 18..24 on 4- and 8-lane
    groups) in one batch, for each lane configuration: the narrow and the wide
    kernels give the oracle's results."""
    from ferrousalign_amd import synth
    rng = np.random.default_rng(23)
    ps = synth.workload("uneven_mix_4m", n=16384)
    ps = ps.subset(np.nonzero((ps.qlen >= 64) & (ps.qlen <= 191))[0])
    ps.h0[:] = rng.integers(1850, 2150, ps.n).astype(np.int32)
    hmax = ps.h0 + ps.qlen
    assert (hmax < 2048).any() and (hmax >= 2048).any()
    sw = synth.default_scoring()
    for lanes in (0, 4, 8):
        print("wide slot boundary", lanes, ps.n, flush=True)
        ora, oc, gpu, gc = _run_both(ps, sw, lanes=lanes)
        _assert_same(ora, oc, gpu, gc, "wide slot boundary lanes=%d" % lanes)


def test_single_launch_wide_slots():
    """One launch for a whole batch (FA_LAUNCH_SINGLE) with 8-lane groups:
    when every pair is under the wide-key bound the launch takes a wide slot
    count (<8,18> for 150 bp reads); when one pair is over it, the launch
    stays on the non-wide slot counts.  Both give the oracle's results."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd import banded_swa as B
    sw = synth.default_scoring()
    ps = synth.workload("illumina_150bp_1m", n=16384)
    flat = ps.flat()
    ora, oc, orc = O.extend_batch(sw, *flat)
    db = B.DeviceBatch(sw, *flat)
    db.set_launch(8)
    print("single launch wide", ps.n, flush=True)
    db.run(N.FA_LAUNCH_SINGLE)
    kern = db.profile_launches(N.FA_LAUNCH_SINGLE)
    assert len(kern) == 1 and kern[0]["kernel"].startswith("extend16_plain_kernel<8,"), kern
    assert kern[0]["slots"] > 16
    gpu, gc = db.download()
    db.free()
    _assert_same(O.to_outscore_order(ora), oc, gpu, gc, "single launch, wide slots")
    # one pair over the bound (h0 + qlen >= 2048): no wide slots for the batch
    ps.h0[7] = 2000
    flat = ps.flat()
    ora, oc, orc = O.extend_batch(sw, *flat)
    db = B.DeviceBatch(sw, *flat)
    db.set_launch(8)
    db.run(N.FA_LAUNCH_SINGLE)
    kern = db.profile_launches(N.FA_LAUNCH_SINGLE)
    assert len(kern) == 1 and kern[0]["slots"] <= 16, kern
    gpu, gc = db.download()
    db.free()
    _assert_same(O.to_outscore_order(ora), oc, gpu, gc, "single launch, one pair over the wide bound")


def test_full_illumina_1m_single_launch_matches_golden():
    """The whole 1,048,576-pair 150 bp Illumina-like workload in its plain
    configuration -- one launch, length-sorted, 8-lane groups -- against the
    oracle's committed digest of all records and cells."""
    import json
    import os
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd import banded_swa as B
    from tests.golden.make_golden import output_digest
    golden = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "digests.json")))
    ps = synth.workload("illumina_150bp_1m")
    assert ps.n == 1048576
    db = B.DeviceBatch(synth.default_scoring(), *ps.flat())
    db.set_launch(8)
    print("illumina 1M single launch", flush=True)
    db.run(N.FA_LAUNCH_SINGLE | N.FA_PACK_LENGTH)
    assert len(db.profile_launches(N.FA_LAUNCH_SINGLE | N.FA_PACK_LENGTH)) == 1
    out, cells = db.download()
    db.free()
    assert int(cells.astype(np.uint64).sum()) == golden["illumina_150bp_1m:cells"]
    assert output_digest(out, cells) == golden["illumina_150bp_1m"]


def test_overflow_count_is_the_32bit_fallbacks():
    """fa_batch_overflow_count: the pairs whose score bound leaves the packed
    16-bit range (h0 + qlen * max(mat) >= 4096) and so ran the 32-bit kernel;
    0 before the first run.  Their records still equal the oracle's."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd import banded_swa as B
    sw = synth.default_scoring()
    ps = synth.workload("uneven_mix_4m", n=4096)
    rng = np.random.default_rng(11)
    big = rng.choice(ps.n, 37, replace=False)
    ps.h0[big] = 4096 - ps.qlen[big] + rng.integers(0, 50, big.size)   # bound >= 4096
    near = np.setdiff1d(rng.choice(ps.n, 20, replace=False), big)
    ps.h0[near] = 4095 - ps.qlen[near]                                 # bound 4095: still 16-bit
    flat = ps.flat()
    db = B.DeviceBatch(sw, *flat)
    assert db.overflow_count() == 0
    db.run()
    assert db.overflow_count() == big.size
    gpu, gc = db.download()
    db.free()
    ora, oc, orc = O.extend_batch(sw, *flat)
    _assert_same(O.to_outscore_order(ora), oc, gpu, gc, "overflow fallbacks")


def test_single_n_bases_among_n_free_pairs():
    """Pairs without any N read no N-mask words (a per-pair flag set at
    upload); pairs with a single N at a sequence's first or last base, or at
    a 16/32-base word edge, sit among them in the same lane groups and must
    still score their N."""
    from ferrousalign_amd import synth
    ps = synth.workload("uneven_mix_4m", n=6000)
    assert not (ps.q == 4).any() and not (ps.t == 4).any()
    rng = np.random.default_rng(11)
    q, t = ps.q.copy(), ps.t.copy()
    picks = rng.choice(ps.n, 1200, replace=False)
    for j, k in enumerate(picks):
        side, where = j % 2, (j // 2) % 6
        seq, L = (q, int(ps.qlen[k])) if side == 0 else (t, int(ps.tlen[k]))
        pos = [0, L - 1, 15, 16, 31, 32][where]
        if pos < L:
            seq[k, pos] = 4
    nps = synth.PairSet(q, t, ps.qlen, ps.tlen, ps.h0, ps.w)
    sw = synth.default_scoring()
    print("single N bases", flush=True)
    ora, oc, gpu, gc = _run_both(nps, sw, lanes=0)
    _assert_same(ora, oc, gpu, gc, "single N")
