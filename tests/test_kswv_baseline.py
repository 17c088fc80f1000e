"""CPU timing baseline of mate rescue: the reference's horizontal kswv path
(batch_ksw_align_avx2, kswv_avx2.rs:93-635, driven as
paired/mate_rescue.rs:704-783 does) restated in oracle/kswv_baseline.c.  Not
an oracle of the GPU's results (its tb / qb are approximations and its
second-best search differs from ksw_align2's); pinned here by its plain-C
twin, bit for bit, and by closed forms of full-matrix local alignment."""
import random

import numpy as np
import pytest

from tests import _oracle as O


def _avx2():
    try:
        return "avx2" in open("/proc/cpuinfo").read()
    except OSError:
        return False


def _flat(cases):
    q = np.concatenate([np.asarray(c[0], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    t = np.concatenate([np.asarray(c[1], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
    ql = np.array([len(c[0]) for c in cases], np.int32)
    tl = np.array([len(c[1]) for c in cases], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    return q, qo, ql, t, to, tl


def _cases(rng, n, qmax=200, tmax=400, alpha=5):
    cases = []
    for _ in range(n):
        ql = rng.randrange(1, qmax + 1)
        tl = rng.randrange(1, tmax + 1)
        q = [rng.randrange(alpha) for _ in range(ql)]
        t = [rng.randrange(alpha) for _ in range(tl)]
        if tl > ql and rng.random() < 0.7:
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                if rng.random() < 0.93:
                    t[s + k] = q[k]
            if rng.random() < 0.4 and ql > 2:
                c = rng.randrange(1, ql)
                del t[s + c:s + c + rng.randrange(1, 4)]
        cases.append((q, t))
    return cases


def _score(a, b, match, mis, ambig):
    x = a ^ b  # the kernel's shuffle index: 0 match, 1-3 mismatch, 4-12 ambig
    return match if x == 0 else (mis if x <= 3 else ambig)


def _textbook(q, t, match, mis, ambig, o_del, e_del, o_ins, e_ins):
    """Full-matrix Gotoh local alignment: (best, first row reaching it, first
    column of that row holding it); deletions run along target rows, insertions
    along query columns."""
    ql = len(q)
    M = np.array([[_score(a, b, match, mis, ambig) for b in range(5)] for a in range(5)], np.int64)
    qa = np.asarray(q)
    Hp = np.zeros(ql, np.int64)
    E = np.full(ql, -(1 << 30), np.int64)
    ks = np.arange(ql)
    best, te, qe = 0, -1, 0
    for i, tb in enumerate(t):
        d = np.concatenate([[0], Hp[:-1]]) + M[tb][qa]
        E = np.maximum(E - e_del, Hp - (o_del + e_del))
        Hx = np.maximum(np.maximum(d, E), 0)
        acc = np.maximum.accumulate(Hx + e_ins * ks)
        F = np.concatenate([[-(1 << 30)], acc[:-1] - (o_ins + e_ins) - e_ins * (ks[1:] - 1)])
        H = np.maximum(Hx, F)
        if H.max() > best:
            best, te, qe = int(H.max()), i, int(np.argmax(H))
        Hp = H
    return best, te, qe


@pytest.mark.skipif(not _avx2(), reason="AVX2 absent")
@pytest.mark.parametrize("seed", range(3))
def test_avx2_matches_plain_c_twin(seed):
    rng = random.Random(seed)
    cases = _cases(rng, 150)
    # mate-rescue scoring, then others with ambiguity scores and asymmetric gaps
    for match, mis, ambig, gaps in ((1, -4, 0, (6, 1, 6, 1)), (2, -3, -1, (5, 2, 3, 1)), (3, -2, 1, (4, 1, 6, 2))):
        flat = _flat(cases)
        a = O.kswv_batch(*flat, match=match, mis=mis, ambig=ambig, gaps=gaps, simd=True, nthreads=3)
        b = O.kswv_batch(*flat, match=match, mis=mis, ambig=ambig, gaps=gaps, simd=False, nthreads=2)
        assert a is not None and b is not None
        bad = np.nonzero((a != b).any(axis=1))[0]
        assert bad.size == 0, (match, mis, bad[:5], a[bad[:3]], b[bad[:3]])


@pytest.mark.parametrize("seed", range(3))
def test_score_te_qe_are_full_matrix_local_alignment(seed):
    # each lane runs the full Smith-Waterman matrix of its job (padding scores 0
    # and stays below every real cell), so below saturation score / te / qe are
    # the textbook best, the first row reaching it and its first column there
    rng = random.Random(20 + seed)
    cases = _cases(rng, 64, qmax=120, tmax=250)
    for match, mis, ambig, gaps in ((1, -4, 0, (6, 1, 6, 1)), (2, -3, -1, (5, 2, 3, 1))):
        out = O.kswv_batch(*_flat(cases), match=match, mis=mis, ambig=ambig, gaps=gaps, simd=False)
        for k, (q, t) in enumerate(cases):
            best, te, qe = _textbook(q, t, match, mis, ambig, *gaps)
            shift = (256 - min(match, mis, ambig)) & 0xFF
            if best + shift >= 255:
                assert out[k, 0] == 255
                continue
            assert tuple(out[k, :3]) == (best, te, qe), (k, out[k], best, te, qe)
            if best > 0:  # the approximate start (kswv_avx2.rs:612-620)
                assert (out[k, 5], out[k, 6]) == (max(te - qe, 0), 0)


def test_saturation_and_mate_workload_shape():
    from ferrousalign_amd import synth
    ps = synth.mate_pairs(96)
    out = O.kswv_batch(*ps.flat(), simd=False)
    ora, _, _ = O.ksw_align2_batch(*ps.flat(), synth.MATE_MATRIX, 6, 1, 6, 1, synth.MATE_XTRA | O.KSW_XBYTE)
    # mate rescue's N scores 0 in both (ambig 0 here, the zero N row there); the
    # full matrix reaches at least what the striped pass reports, and the two agree
    # on almost every pair of this workload
    assert (out[:, 0] >= ora[:, 0]).all()
    assert (out[:, 0] == ora[:, 0]).mean() > 0.9
    # a 251-base exact copy saturates (shift 4): score 255
    rng = random.Random(5)
    t = [rng.randrange(4) for _ in range(600)]
    out = O.kswv_batch(*_flat([(t[100:351], t), (t[100:300], t)]), simd=False)
    assert out[0, 0] == 255 and out[1, 0] == 200 and tuple(out[1, 1:3]) == (299, 199)
