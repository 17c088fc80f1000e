"""CPU oracle of the work predictor: FP16 rounding against numpy's IEEE
binary16, analytic gradients against central finite differences, features
against a direct numpy computation."""
import numpy as np
import pytest

from tests import _oracle as O


def test_round16_matches_ieee_binary16():
    rng = np.random.default_rng(0)
    vals = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-5, 500), rng.normal(0, 3e4, 500),
                           [0.0, -0.0, 1.0, 65504.0, 65519.0, 65520.0, 2 ** -24, 2 ** -25, 3 * 2 ** -26,
                            1 + 2 ** -11, 1 + 3 * 2 ** -11]]).astype(np.float32)
    ref = vals.astype(np.float16).astype(np.float32)
    got = np.array([O.round16(v) for v in vals], np.float32)
    assert np.array_equal(got, ref, equal_nan=True)


def _random_problem(seed, n=40):
    from ferrousalign_amd.predictor import init_params
    rng = np.random.default_rng(seed)
    p = init_params(seed)
    p[:] += rng.normal(0, 0.05, p.size).astype(np.float32)   # non-zero biases
    feat = np.stack([rng.integers(1, 230, n), rng.integers(100, 330, n), rng.integers(19, 61, n),
                     rng.choice([100, 200], n), rng.integers(0, 20, n), rng.random(n)], 1).astype(np.float32)
    norm = np.concatenate([feat.mean(0), 1.0 / (feat.std(0) + 1e-3)]).astype(np.float32)
    target = np.log1p(rng.integers(100, 20000, n)).astype(np.float32)
    return p, norm, feat, target


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_finite_differences(seed):
    p, norm, feat, target = _random_problem(seed)
    loss, g = O.mlp_grad(p, norm, feat, target, fp16=False)
    rng = np.random.default_rng(seed + 10)
    # every bias, and a sample of each weight matrix
    idx = list(range(192, 224)) + list(range(1248, 1313)) + list(rng.choice(192, 20, replace=False)) \
        + list(224 + rng.choice(1024, 40, replace=False))
    for k in idx:
        h = 1e-3 * max(1.0, abs(float(p[k])))
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        lp, _ = O.mlp_grad(pp, norm, feat, target, fp16=False)
        lm, _ = O.mlp_grad(pm, norm, feat, target, fp16=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[k]) <= 2e-3 * max(1.0, abs(g[k])) + 1e-4, (k, fd, g[k])


def test_forward_closed_form_linear_regime():
    # with all hidden units active and identity-like weights the net is affine
    from ferrousalign_amd.predictor import NPARAM
    p = np.zeros(NPARAM, np.float32)
    p[0:192].reshape(32, 6)[:6, :6] = np.eye(6)
    p[192:224] = 10.0                      # b1 keeps every unit positive
    W2 = p[224:1248].reshape(32, 32)
    W2[:, :] = np.eye(32)
    p[1280:1312] = 0.5                     # W3
    p[1312] = -3.0
    feat = np.array([[1, 2, 3, 4, 5, 0.25]], np.float32)
    norm = np.concatenate([np.zeros(6), np.ones(6)]).astype(np.float32)
    y = O.mlp_forward(p, norm, feat, fp16=False)[0]
    # h1 = x + 10 (first 6), 10 elsewhere; h2 = h1; y = 0.5*sum(h2) - 3
    expect = 0.5 * (feat.sum() + 32 * 10) - 3.0
    assert abs(y - expect) < 1e-4


def test_features_direct():
    rng = np.random.default_rng(3)
    for _ in range(50):
        ql, tl = rng.integers(1, 120), rng.integers(0, 150)
        q = rng.integers(0, 5, ql).astype(np.uint8)
        t = rng.integers(0, 5, tl).astype(np.uint8)
        f = O.pair_features(q, t, 33, 100)
        lim = min(32, ql, tl)
        assert f[4] <= np.count_nonzero(q[:lim] != t[:lim])  # the main diagonal is a candidate
        assert f[5] == np.float32(np.count_nonzero((q == 1) | (q == 2))) / np.float32(ql)
        assert tuple(f[:4]) == (ql, tl, 33, 100)


def test_pair_features_best_diagonal():
    """Feature 4: mismatches in the first min(32, qlen, tlen) bases on the
    best diagonal within three of the main one.  An exact copy gives 0; a
    read with two bases deleted after base 3 mismatches on the main diagonal
    (a random-looking count) and matches from base 3 on the diagonal two to
    the right; a random target stays high on every diagonal."""
    rng = np.random.default_rng(9)
    t = rng.integers(0, 4, 200).astype(np.uint8)
    f = O.pair_features(t[:120], t, 30, 100)
    assert f[4] == 0 and f[0] == 120 and f[1] == 200
    q = np.concatenate([t[:3], t[5:125]])          # bases 3, 4 of the target deleted in the read
    main = int((q[:32] != t[:32]).sum())
    f = O.pair_features(q, t, 30, 100)
    best = int((q[3:35] != t[5:37]).sum())          # diagonal +2 from base 3 on (q[j] vs t[j + 2])
    assert main >= 12 and f[4] == int((q[:32] != t[2:34]).sum()) <= 3 + best
    r = rng.integers(0, 4, 200).astype(np.uint8)
    assert O.pair_features(r[:120], t, 30, 100)[4] >= 12
    # brute force over random short pairs, N bases included
    for _ in range(300):
        ql, tl = int(rng.integers(0, 45)), int(rng.integers(0, 45))
        qq = rng.integers(0, 5, ql).astype(np.uint8)
        tt = rng.integers(0, 5, tl).astype(np.uint8)
        lim = min(32, ql, tl)
        want = min(int((qq[max(0, -d):max(0, -d) + lim] != tt[max(0, d):max(0, d) + lim]).sum())
                   for d in range(-3, 4) if max(0, -d) + lim <= ql and max(0, d) + lim <= tl)
        assert O.pair_features(qq, tt, 1, 1)[4] == want
