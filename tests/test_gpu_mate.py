"""GPU mate-rescue local SW (ksw_align2, 16-bit mode): every Kswr field and
the cell count bit-exact against the CPU oracle."""
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


@pytest.mark.parametrize("seed", [0, 1])
def test_mate_random_shapes_and_flags(seed):
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import ksw_align2_batch
    rng = random.Random(seed)
    cases = []
    for _ in range(200):
        ql = rng.randrange(1, 257)
        tl = rng.randrange(0, 700)
        q = [rng.randrange(5) for _ in range(ql)]
        t = [rng.randrange(5) for _ in range(tl)]
        if tl > ql and rng.random() < 0.7:
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                t[s + k] = q[k] if rng.random() < 0.92 else rng.randrange(4)
        cases.append((q, t))
    for xtra, gaps in ((synth.MATE_XTRA, (6, 1, 6, 1)), (O.KSW_XSTART, (5, 2, 3, 1)),
                       (O.KSW_XSUBO | O.KSW_XSTART | 40, (6, 1, 6, 1)), (0, (0, 1, 0, 1))):
        q = np.concatenate([np.asarray(c[0], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
        t = np.concatenate([np.asarray(c[1], np.uint8) for c in cases] + [np.zeros(1, np.uint8)])
        ql = np.array([len(c[0]) for c in cases], np.int32)
        tl = np.array([len(c[1]) for c in cases], np.int32)
        qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
        to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
        print("mate random", seed, hex(xtra), gaps, flush=True)
        out, cells = ksw_align2_batch(q, qo, ql, t, to, tl, synth.MATE_MATRIX, *gaps, xtra)
        ora, oc, rc = O.ksw_align2_batch(q, qo, ql, t, to, tl, synth.MATE_MATRIX, *gaps, xtra)
        assert (rc == 0).all()
        _check(out, cells, ora, oc, "random xtra=%#x" % xtra)


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
    db.run(N.FA_RUN_ALIGN2)
    out, cells = db.download_align2()
    assert digest(out, cells) == d[key]
    db.free()
