"""GPU work predictor: device features equal the oracle's, the FP16 forward
and backward kernels agree with the CPU oracle within the stated tolerance,
training lowers the loss, and learned packing keeps every output bit-exact."""
import numpy as np
import pytest

from tests import _oracle as O
from tests import rowpar

pytestmark = pytest.mark.gpu

# Tolerances: network output y = log(1 + cells) within 2e-2 absolute (FP16
# activations may round one ulp apart when the GPU contracts a*b+c into one
# FMA); gradient within 2% relative L2 error.
Y_TOL = 2e-2
G_TOL = 2e-2


@pytest.fixture(scope="module")
def setup():
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import DeviceBatch, Context
    from ferrousalign_amd.predictor import WorkPredictor
    ctx = Context(0)
    ps = synth.workload("uneven_mix_4m", n=30000)
    sw = synth.default_scoring()
    db = DeviceBatch(sw, *ps.flat(), ctx=ctx)
    db.run(0)
    out, cells = db.download()
    pred = WorkPredictor(ctx, seed=1)
    return ctx, ps, sw, db, cells, pred


def _oracle_features(ps, sw):
    f = np.zeros((ps.n, 6), np.float32)
    for k in range(ps.n):
        w = rowpar.capped_w(int(ps.qlen[k]), list(sw.mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins, int(ps.w[k]),
                            sw.end_bonus)
        f[k] = O.pair_features(ps.q[k, :ps.qlen[k]], ps.t[k, :ps.tlen[k]], int(ps.h0[k]), w)
    return f


def test_features_exact(setup):
    ctx, ps, sw, db, cells, pred = setup
    print("features", flush=True)
    g = pred.features(db)[:3000]
    o = _oracle_features(ps.subset(np.arange(3000)), sw)
    bad = np.nonzero((g[:, :5] != o[:, :5]).any(axis=1))[0]
    assert bad.size == 0, "pair %d: gpu %s oracle %s" % (bad[0], g[bad[0]].tolist(), o[bad[0]].tolist())
    # GC fraction: one float division on each side, equal to within one ulp
    assert np.max(np.abs(g[:, 5] - o[:, 5])) <= 1.2e-7


def test_forward_and_gradient_vs_oracle(setup):
    ctx, ps, sw, db, cells, pred = setup
    from ferrousalign_amd.predictor import norm_from_features, init_params
    feat = pred.features(db)
    rng = np.random.default_rng(5)
    pred.params = init_params(3) + rng.normal(0, 0.05, 1313).astype(np.float32)
    pred.norm = norm_from_features(feat)
    pred.upload()
    print("forward", flush=True)
    cp = pred.predict(db)
    y_gpu = np.log1p(cp[:4000].astype(np.float64))
    y_ora = O.mlp_forward(pred.params, pred.norm, feat[:4000], fp16=True).astype(np.float64)
    ok = y_ora > 0  # expm1 clamps negative outputs to 0 cells
    assert np.max(np.abs(y_gpu[ok] - y_ora[ok])) < Y_TOL
    print("gradient", flush=True)
    pairs = np.arange(1000, 5000, dtype=np.uint32)
    g, loss = pred.gradient(db, pairs)
    lo, go = O.mlp_grad(pred.params, pred.norm, feat[pairs], np.log1p(cells[pairs].astype(np.float32)), fp16=True)
    assert abs(loss - lo) <= 1e-3 * max(1.0, abs(lo))
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    assert rel < G_TOL, rel


def test_training_and_learned_packing(setup):
    ctx, ps, sw, db, cells, pred = setup
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.predictor import WorkPredictor, mae
    p2 = WorkPredictor(ctx, seed=2)
    train = np.arange(20000, dtype=np.uint32)
    print("train", flush=True)
    losses = p2.fit(db, epochs=4, minibatch=2048, lr=3e-3)
    assert losses[-1] < losses[0]
    cp = p2.predict(db, set_cost=True)
    held = np.arange(20000, ps.n)
    m_learned = mae(cp[held], cells[held])
    m_const = mae(np.full(held.size, np.median(cells[train])), cells[held])
    print("held-out MAE learned %.1f vs constant %.1f" % (m_learned, m_const), flush=True)
    assert m_learned < 0.8 * m_const
    ref_out, ref_cells = db.download()
    db.run(N.FA_PACK_COST | N.FA_LAUNCH_PERSISTENT | N.FA_LAUNCH_GRAPH)
    out, c2 = db.download()
    assert np.array_equal(out, ref_out) and np.array_equal(c2, ref_cells)
