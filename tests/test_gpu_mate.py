"""GPU mate-rescue local SW (ksw_align2, 16-bit and 8-bit modes): every Kswr
field and the cell count bit-exact against the CPU oracle."""
import json
import os
import random

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _check(out, cells, ora, ocells, label):
    bad = np.nonzero((out != ora).any(axis=1) | (cells.astype(np.uint64) != ocells))[0]
    assert bad.size == 0, "%s: %d/%d differ; first %d gpu=%s oracle=%s" % (
        label, bad.size, len(out), bad[0], out[bad[0]].tolist(), ora[bad[0]].tolist())


def test_mate_workload_subset():
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    ps = synth.mate_pairs(1024)
    flat = ps.flat()
    print("mate subset", flush=True)
    out, cells = ksw_align2_batch(*flat, synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA)
    ora, oc, rc = O.ksw_align2_batch(*flat, synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA)
    _check(out, cells, ora, oc, "mate workload")


def test_mate_golden_fixture():
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    z = np.load(os.path.join(HERE, "golden", "subset_mate.npz"))
    n = len(z["q"])
    out, cells = ksw_align2_batch(z["q"].reshape(-1), np.arange(n) * 151, np.full(n, 151), z["t"].reshape(-1),
                                  np.arange(n) * 1500, np.full(n, 1500), synth.MATE_MATRIX, 6, 1, 6, 1,
                                  synth.MATE_XTRA)
    _check(out, cells, z["out"], z["cells"].astype(np.uint64), "golden fixture")


def _random_cases(rng, n=200, qmax=255, copy_rate=0.92):
    cases = []
    for _ in range(n):
        ql = rng.randrange(1, qmax + 1)
        tl = rng.randrange(0, 700)
        q = [rng.randrange(5) for _ in range(ql)]
        t = [rng.randrange(5) for _ in range(tl)]
        if tl > ql and rng.random() < 0.7:
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                t[s + k] = q[k] if rng.random() < copy_rate else rng.randrange(4)
        cases.append((q, t))
    return cases


def _run_cases(cases, mat, gaps, xtra, label):
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    q = np.concatenate([np.asarray(c[0], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    t = np.concatenate([np.asarray(c[1], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    ql = np.array([len(c[0]) for c in cases], np.int32)
    tl = np.array([len(c[1]) for c in cases], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    print(label, hex(xtra), gaps, flush=True)
    out, cells = ksw_align2_batch(q, qo, ql, t, to, tl, mat, *gaps, xtra)
    ora, oc, rc = O.ksw_align2_batch(q, qo, ql, t, to, tl, mat, *gaps, xtra)
    assert (rc == 0).all()
    _check(out, cells, ora, oc, "%s xtra=%#x" % (label, xtra))
    return ora


@pytest.mark.parametrize("seed", [0, 1])
def test_mate_random_shapes_and_flags(seed):
    from ferrousalign_amd import synth
    cases = _random_cases(random.Random(seed))
    for xtra, gaps in ((synth.MATE_XTRA, (6, 1, 6, 1)), (O.KSW_XSTART, (5, 2, 3, 1)),
                       (O.KSW_XSUBO | O.KSW_XSTART | 40, (6, 1, 6, 1)), (0, (0, 1, 0, 1))):
        _run_cases(cases, synth.MATE_MATRIX, gaps, xtra, "mate random %d" % seed)


@pytest.mark.parametrize("byte_mode", [False, True])
def test_mate_reverse_launches_match_fused(byte_mode):
    # random query lengths spread the XSTART passes over many reverse stripe
    # lengths; both run forms against the oracle
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    cases = _random_cases(random.Random(7), n=600)
    q = np.concatenate([np.asarray(c[0], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    t = np.concatenate([np.asarray(c[1], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    ql = np.array([len(c[0]) for c in cases], np.int32)
    tl = np.array([len(c[1]) for c in cases], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    n = len(cases)
    xtra = synth.MATE_XTRA | (O.KSW_XBYTE if byte_mode else 0)
    print("mate reverse launches", byte_mode, flush=True)
    db = DeviceBatch(synth.default_scoring(), q, qo, ql, t, to, tl, np.ones(n, np.int32), np.zeros(n, np.int32))
    db.set_align2(synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
    ora, oc, rc = O.ksw_align2_batch(q, qo, ql, t, to, tl, synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
    assert (rc == 0).all()
    for flags in (N.FA_RUN_ALIGN2, N.FA_RUN_ALIGN2 | N.FA_ALIGN2_FUSED, N.FA_RUN_ALIGN2):
        db.run(flags)
        out, cells = db.download_align2()
        _check(out, cells, ora, oc, "flags %#x" % flags)
    # the reverse launches are timed apart from the forward ones
    launches = db.profile_launches(N.FA_RUN_ALIGN2)
    assert any(x["kernel"].startswith("mate16_rev") for x in launches)
    db.free()


# ---- 8-bit mode (KSW_XBYTE) ------------------------------------------------------

def test_mate_u8_workload_subset():
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    ps = synth.mate_pairs(2048)
    flat = ps.flat()
    xtra = synth.MATE_XTRA | O.KSW_XBYTE
    print("mate u8 subset", flush=True)
    out, cells = ksw_align2_batch(*flat, synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
    ora, oc, rc = O.ksw_align2_batch(*flat, synth.MATE_MATRIX, 6, 1, 6, 1, xtra)
    assert (rc == 0).all()
    _check(out, cells, ora, oc, "mate u8 workload")


@pytest.mark.parametrize("seed", [0, 1])
def test_mate_u8_random_shapes_and_flags(seed):
    from ferrousalign_amd import synth
    cases = _random_cases(random.Random(50 + seed))
    for xtra, gaps in ((synth.MATE_XTRA, (6, 1, 6, 1)), (O.KSW_XSTART, (5, 2, 3, 1)),
                       (O.KSW_XSUBO | O.KSW_XSTART | 40, (6, 1, 6, 1)), (0, (0, 1, 0, 1)),
                       (O.KSW_XSTOP | O.KSW_XSTART | 30, (6, 1, 6, 1))):
        _run_cases(cases, synth.MATE_MATRIX, gaps, xtra | O.KSW_XBYTE, "mate u8 random %d" % seed)


def test_mate_forward_stop_score():
    # KSW_XSTOP in the caller's xtra: the forward pass stops at the first best
    # score >= the stop score, in both modes
    from ferrousalign_amd import synth
    cases = _random_cases(random.Random(90), n=160)
    for xtra in (O.KSW_XSTOP | 30, O.KSW_XSTOP | O.KSW_XSTART | 60):
        for mode in (0, O.KSW_XBYTE):
            _run_cases(cases, synth.MATE_MATRIX, (6, 1, 6, 1), xtra | mode, "mate xstop")


def test_mate_u8_saturation():
    # best scores crossing 255 - shift: long near-exact copies under the mate matrix
    # (shift 4, saturating at 251) and under match 3 / mismatch -2 (shift 2, at 253);
    # a saturated pair reports score 255, qe -1 and no reverse pass
    from ferrousalign_amd import synth
    rng = random.Random(7)
    cases = _random_cases(rng, n=160, copy_rate=0.995)
    for k in range(40):  # 30 exact copies of 251..255 bases (saturating), 10 of 240..250
        ql = 251 + k % 5 if k < 30 else 240 + k % 11
        tl = rng.randrange(ql, 700)
        t = [rng.randrange(4) for _ in range(tl)]
        s = rng.randrange(0, tl - ql + 1)
        cases.append((t[s:s + ql], t))
    ora = _run_cases(cases, synth.MATE_MATRIX, (6, 1, 6, 1), synth.MATE_XTRA | O.KSW_XBYTE, "mate u8 sat")
    assert (ora[:, 0] == 255).sum() >= 30 and ((ora[:, 0] == 255) <= (ora[:, 2] == -1)).all()
    m3 = [3 if a == b and a < 4 else (0 if a == 4 or b == 4 else -2) for a in range(5) for b in range(5)]
    ora = _run_cases(cases, m3, (5, 2, 5, 2), O.KSW_XSUBO | O.KSW_XSTART | 30 | O.KSW_XBYTE, "mate u8 sat m3")
    assert (ora[:, 0] == 255).sum() >= 40


def test_mate_u8_full_size_digest():
    from ferrousalign_amd import synth
    from tests.golden.make_golden_mate import digest
    d = json.load(open(os.path.join(HERE, "golden", "mate_digests.json")))
    key = "mate_u8_%d" % synth.MATE_PAIRS
    n = synth.MATE_PAIRS
    from ferrousalign_amd.banded_swa import DeviceBatch
    from ferrousalign_amd import _native as N
    ps = synth.mate_pairs(n)
    print("mate u8 full size", n, flush=True)
    db = DeviceBatch(synth.default_scoring(), *ps.flat(), np.ones(n, np.int32), np.zeros(n, np.int32))
    db.set_align2(synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA | O.KSW_XBYTE)
    # the separate reverse launches (default) and the reverse pass inside the
    # forward kernel give the same records
    for flags in (N.FA_RUN_ALIGN2, N.FA_RUN_ALIGN2 | N.FA_ALIGN2_FUSED):
        db.run(flags)
        out, cells = db.download_align2()
        assert digest(out, cells) == d[key], hex(flags)
    db.free()


def test_mate_full_size_digest():
    from ferrousalign_amd import synth
    from tests.golden.make_golden_mate import digest
    path = os.path.join(HERE, "golden", "mate_digests.json")
    d = json.load(open(path))
    key = "mate_%d" % synth.MATE_PAIRS if ("mate_%d" % synth.MATE_PAIRS) in d else "mate_262144"
    n = int(key.split("_")[1])
    from ferrousalign_amd.banded_swa import DeviceBatch
    from ferrousalign_amd import _native as N
    ps = synth.mate_pairs(n)
    print("mate full size", n, flush=True)
    db = DeviceBatch(synth.default_scoring(), *ps.flat(), np.ones(n, np.int32), np.zeros(n, np.int32))
    db.set_align2(synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA)
    # the separate reverse launches (default) and the reverse pass inside the
    # forward kernel give the same records
    for flags in (N.FA_RUN_ALIGN2, N.FA_RUN_ALIGN2 | N.FA_ALIGN2_FUSED):
        db.run(flags)
        out, cells = db.download_align2()
        assert digest(out, cells) == d[key], hex(flags)
    db.free()


@pytest.mark.parametrize("byte_mode", [False, True])
def test_mate_score2_two_copies(byte_mode):
    # exact copy + a copy with 4 spaced mismatches far away: score2 = qlen - 20 at
    # the second copy's last row, on the GPU as in the oracle's closed form
    from ferrousalign_amd import synth
    rng = random.Random(31)
    qlen, cases, want = 151, [], []
    for s1, s2, first_exact in ((100, 700, True), (120, 900, False), (1200, 300, True)):
        t = [rng.randrange(4) for _ in range(1500)]
        q = [rng.randrange(4) for _ in range(qlen)]
        mut = list(q)
        for k in (30, 60, 90, 120):
            mut[k] = (q[k] + 1) % 4
        a, b = (q, mut) if first_exact else (mut, q)
        t[s1:s1 + qlen] = a
        t[s2:s2 + qlen] = b
        se, so = (s1, s2) if first_exact else (s2, s1)
        cases.append((q, t))
        want.append((qlen, se + qlen - 1, qlen - 1, qlen - 20, so + qlen - 1, se, 0))
    xtra = synth.MATE_XTRA | (O.KSW_XBYTE if byte_mode else 0)
    ora = _run_cases(cases, synth.MATE_MATRIX, (6, 1, 6, 1), xtra, "mate two copies")
    assert [tuple(int(x) for x in r) for r in ora] == want


def test_extension_and_mate_runs_keep_their_own_results():
    """One batch, extension and mate-rescue runs in turn: each download gives
    the results of its own last run (the two keep separate cell counts)."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps = synth.mate_pairs(512)
    q, qo, ql, t, to, tl = ps.flat()
    n = len(ql)
    h0 = np.full(n, 40, np.int32)
    w = np.full(n, 60, np.int32)
    sw = synth.default_scoring()
    eo, ec, erc = O.extend_batch(sw, q, qo, ql, t, to, tl, h0, w)
    assert (erc == 0).all()
    eo = O.to_outscore_order(eo)
    mo, mc, mrc = O.ksw_align2_batch(q, qo, ql, t, to, tl, synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA)
    assert (mrc == 0).all()
    print("extension + mate on one batch", flush=True)
    db = DeviceBatch(sw, q, qo, ql, t, to, tl, h0, w)
    db.set_align2(synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA)
    db.run(0)
    db.run(N.FA_RUN_ALIGN2)
    for _ in range(2):
        out, cells = db.download()
        assert (out == eo).all() and (cells.astype(np.uint64) == ec).all()
        mout, mcells = db.download_align2()
        _check(mout, mcells, mo, mc, "mate after extension")
        db.run(0)
    db.free()
