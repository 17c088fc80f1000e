"""The CPU timing baseline that restates the reference's SIMD batch path
(oracle/refsimd_baseline.c: dispatch.rs:95-268, soa.rs:11-71, kernel.rs
i8 :29-284, kernel_i16.rs :44-272): the vector kernels (AVX-512BW, AVX2)
against the lane-by-lane restatement of the same chunk semantics, the
reference's own kernel tests (tests/kernel_optimization_test.rs) and closed
forms.  CPU only."""
import numpy as np
import pytest

from ferrousalign_amd.banded_swa import BandedPairWiseSW, bwa_fill_scmat
from tests import _oracle as O


def _sw(zdrop=100):
    return BandedPairWiseSW(6, 1, 6, 1, zdrop, 5, 5, 5, bwa_fill_scmat(1, 4, -1))


def _batch(rng, n, qmax, tmax, nfrac=0.02, homol=True):
    ql = rng.integers(1, qmax + 1, n).astype(np.int32)
    tl = rng.integers(1, tmax + 1, n).astype(np.int32)
    qs, ts = [], []
    for k in range(n):
        t = rng.integers(0, 4, tl[k]).astype(np.uint8)
        if homol and rng.random() < 0.8:
            q = t[:ql[k]].copy() if ql[k] <= tl[k] else np.concatenate([t, rng.integers(0, 4, ql[k] - tl[k])])
            m = rng.random(ql[k]) < 0.05
            q[m] = rng.integers(0, 4, int(m.sum()))
        else:
            q = rng.integers(0, 4, ql[k])
        q = q.astype(np.uint8)
        q[rng.random(ql[k]) < nfrac] = 4
        t[rng.random(tl[k]) < nfrac] = 4
        qs.append(q)
        ts.append(t)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    h0 = rng.integers(1, 40, n).astype(np.int32)
    w = rng.choice([2, 10, 50, 100, 200], n).astype(np.int32)  # 200 wraps to -56 as the reference's i8
    return np.concatenate(qs), qo, ql, np.concatenate(ts), to, tl, h0, w


@pytest.mark.parametrize("simd,scalar", [("avx512", "scalar512"), ("avx2", "scalar2")])
@pytest.mark.parametrize("qmax,tmax,zdrop", [(100, 128, 100), (100, 128, 3), (150, 250, 100), (150, 250, 5),
                                            (20, 30, 0)])
def test_vector_kernels_match_lane_restatement(simd, scalar, qmax, tmax, zdrop):
    rng = np.random.default_rng(qmax * 7 + zdrop)
    args = _batch(rng, 700, qmax, tmax)
    got = O.refsimd_batch(_sw(zdrop), *args, isa=simd, nthreads=4)
    if got is None:
        pytest.skip("%s not available on this CPU" % simd)
    want = O.refsimd_batch(_sw(zdrop), *args, isa=scalar, nthreads=3)
    assert got[2] == want[2] == (max(qmax, tmax) > 128)
    assert np.array_equal(got[0], want[0])


def _uniform(n, pattern_q, pattern_t, h0=0, w=10):
    q = np.tile(np.asarray(pattern_q, np.uint8), n)
    t = np.tile(np.asarray(pattern_t, np.uint8), n)
    ql = np.full(n, len(pattern_q), np.int32)
    tl = np.full(n, len(pattern_t), np.int32)
    return (q, np.arange(n, dtype=np.int64) * len(pattern_q), ql, t, np.arange(n, dtype=np.int64) * len(pattern_t),
            tl, np.full(n, h0, np.int32), np.full(n, w, np.int32))


@pytest.mark.parametrize("isa", ["auto", "scalar512", "scalar2"])
def test_reference_kernel_tests(isa):
    sw = _sw()
    # identical ACGT... sequences: 0-based ends at the last position, score = length (h0 = 0)
    for n in (10, 100, 127):
        pat = [p % 4 for p in range(n)]
        out, _, wide = O.refsimd_batch(sw, *_uniform(16, pat, pat), isa=isa)
        assert not wide
        assert (out[:, 0] == n).all() and (out[:, 3] == n - 1).all() and (out[:, 1] == n - 1).all()
        assert (out[:, [2, 4, 5]] == 0).all()  # gtle, gscore, max_off are not computed
    # the same on the i16 path (a 200-base job in the batch); there the band end
    # is written back each row, so it never grows past w + 1 (kernel_i16.rs:150-160)
    pat = [p % 4 for p in range(100)]
    for w, want in ((100, 100), (10, 11)):
        args = list(_uniform(16, pat, pat, w=w))
        args[5][0] = 200
        args[3] = np.concatenate([args[3], np.zeros(200, np.uint8)])
        args[4][0] = len(args[3]) - 200
        out, _, wide = O.refsimd_batch(sw, *args, isa=isa)
        assert wide and (out[1:, 0] == want).all() and (out[1:, 3] == want - 1).all()
    # complete mismatch (all A against all T): score < 4
    out, _, _ = O.refsimd_batch(sw, *_uniform(16, [0] * 8, [3] * 8), isa=isa)
    assert (out[:, 0] < 4).all()
    # AAA then TTT against all A: saturating arithmetic keeps everything >= 0
    out, _, _ = O.refsimd_batch(sw, *_uniform(16, [0] * 6, [0, 0, 0, 3, 3, 3]), isa=isa)
    assert (out[:, 0] >= 0).all() and (out[:, 1] >= 0).all() and (out[:, 3] >= 0).all()
    # a narrow band still aligns
    pat = [p % 4 for p in range(10)]
    out, _, _ = O.refsimd_batch(sw, *_uniform(16, pat, pat, w=2), isa=isa)
    assert (out[:, 0] >= 0).all()
    # varying lengths 4..11: positive score, query end inside the query
    pats = [[p % 4 for p in range(4 + k // 2)] for k in range(16)]
    q = np.concatenate([np.asarray(p, np.uint8) for p in pats])
    ql = np.array([len(p) for p in pats], np.int32)
    off = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    out, _, _ = O.refsimd_batch(sw, q, off, ql, q, off, ql, np.zeros(16, np.int32), np.full(16, 10, np.int32),
                                isa=isa)
    assert (out[:, 0] > 0).all() and (out[:, 3] <= ql - 1).all()


def test_every_chunk_is_scored():
    """All chunks get records (the reference dispatch scores only the first W)."""
    sw = _sw()
    pat = [p % 4 for p in range(30)]
    out, _, _ = O.refsimd_batch(sw, *_uniform(1000, pat, pat), isa="auto", nthreads=8)
    assert (out[:, 0] == 30).all()
