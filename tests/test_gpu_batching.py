"""GPU: persistent queue, hipGraph replay, chunk sizes, lane-group widths and
pair order never change any output field (bit-exact to the oracle)."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix():
    from ferrousalign_amd import synth
    ps = synth.workload("uneven_mix_4m", n=20000)
    sw = synth.default_scoring()
    ora, oc, rc = O.extend_batch(sw, *ps.flat())
    return ps, sw, O.to_outscore_order(ora), oc


def _check(db, ora, oc, label):
    out, cells = db.download()
    bad = (out != ora).any(axis=1) | (cells != oc)
    assert not bad.any(), "%s: %d pairs differ (first %d)" % (label, bad.sum(), np.nonzero(bad)[0][0])


@pytest.mark.parametrize("chunk_tasks", [1, 3, 16])
def test_persistent_queue_chunks(mix, chunk_tasks):
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps, sw, ora, oc = mix
    db = DeviceBatch(sw, *ps.flat())
    db.set_launch(16, chunk_tasks * 4, 0)
    print("persistent chunk", chunk_tasks, flush=True)
    db.run(N.FA_LAUNCH_PERSISTENT)
    _check(db, ora, oc, "persistent chunk=%d" % chunk_tasks)
    db.free()


def test_graph_replays(mix):
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps, sw, ora, oc = mix
    db = DeviceBatch(sw, *ps.flat())
    db.set_launch(8, 16, 0)
    flags = N.FA_LAUNCH_PERSISTENT | N.FA_LAUNCH_GRAPH
    for rep in range(4):
        print("graph replay", rep, flush=True)
        db.run(flags)
        _check(db, ora, oc, "graph replay %d" % rep)
        # eager work between replays
        db.run(N.FA_LAUNCH_PLAIN)
        _check(db, ora, oc, "eager after replay %d" % rep)
    db.free()


def test_plain_graph(mix):
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps, sw, ora, oc = mix
    db = DeviceBatch(sw, *ps.flat())
    for rep in range(3):
        db.run(N.FA_LAUNCH_GRAPH)
        _check(db, ora, oc, "plain graph %d" % rep)
    db.free()


def test_order_invariance(mix):
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps, sw, ora, oc = mix
    perm = np.random.default_rng(7).permutation(ps.n)
    sub = ps.subset(perm)
    db = DeviceBatch(sw, *sub.flat())
    for flags in (N.FA_PACK_NONE, N.FA_PACK_LENGTH, N.FA_PACK_LENGTH | N.FA_LAUNCH_PERSISTENT):
        db.run(flags)
        _check(db, ora[perm], oc[perm], "shuffled flags=%#x" % flags)
    cost = np.random.default_rng(3).random(ps.n).astype(np.float32)
    db.set_cost(cost)
    db.run(N.FA_PACK_COST | N.FA_LAUNCH_PERSISTENT)
    _check(db, ora[perm], oc[perm], "random cost packing")
    for flags in (N.FA_PACK_COST_ORDER, N.FA_PACK_COST_ORDER | N.FA_LAUNCH_GRAPH):
        db.run(flags)
        _check(db, ora[perm], oc[perm], "random cost-order packing flags=%#x" % flags)
    db.free()


def test_autotuner_picks_a_valid_config(mix, tmp_path):
    from ferrousalign_amd import autotune
    from ferrousalign_amd.banded_swa import DeviceBatch
    ps, sw, ora, oc = mix
    db = DeviceBatch(sw, *ps.flat())
    best, flags, table = autotune.tune(db, ps.qlen, ps.tlen, ps.w, cache_path=str(tmp_path / "c.json"), reps=1)
    assert table and best["lanes"] in (0, 4, 8, 16)
    db.run(flags)
    _check(db, ora, oc, "autotuned")
    # second call hits the cache
    best2, flags2, table2 = autotune.tune(db, ps.qlen, ps.tlen, ps.w, cache_path=str(tmp_path / "c.json"))
    assert not table2 and best2["lanes"] == best["lanes"]
    db.free()


def test_concurrent_contexts_from_threads():
    """The reference calls the batch path from every rayon worker at once
    (single_end.rs:92-110): one context per host thread, four threads calling
    at the same time (ctypes releases the GIL during each call), each batch
    checked against the oracle."""
    import threading
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import Context, extend_arrays
    from tests import _oracle as O
    sw = synth.default_scoring()
    batches = [synth.workload("uneven_mix_4m", n=4096 + 512 * k) for k in range(4)]
    want = []
    for ps in batches:
        ora, oc, rc = O.extend_batch(sw, *ps.flat())
        assert (rc == 0).all()
        want.append((O.to_outscore_order(ora), oc.astype(np.uint32)))
    got = [None] * 4
    errors = []

    def work(k):
        try:
            ctx = Context(0)
            res = []
            for _ in range(3):
                out, cells = extend_arrays(sw, *batches[k].flat(), ctx=ctx, want_cells=True)[:2]
                res.append((out.copy(), cells.copy()))
            got[k] = res
        except Exception as e:  # reported below
            errors.append(repr(e))

    th = [threading.Thread(target=work, args=(k,)) for k in range(4)]
    for t in th:
        t.start()
    for t in th:
        t.join(120)
    assert not errors, errors
    for k in range(4):
        assert got[k] is not None
        for out, cells in got[k]:
            assert np.array_equal(out, want[k][0]) and np.array_equal(cells.astype(np.uint32), want[k][1])
