"""Seeded workload generator: determinism, thread independence, the golden
input digests, and the shape of each workload."""
import json
import os

import numpy as np
import pytest

from ferrousalign_amd import synth
from tests.golden.make_golden import input_digest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_deterministic_and_thread_independent():
    a = synth.extension_pairs(2, 5000, 123, threads=1)
    b = synth.extension_pairs(2, 5000, 123, threads=8)
    for x, y in ((a.q, b.q), (a.t, b.t), (a.qlen, b.qlen), (a.tlen, b.tlen), (a.h0, b.h0), (a.w, b.w)):
        assert (x == y).all()
    c = synth.extension_pairs(2, 5000, 124, threads=8)
    assert not (a.qlen == c.qlen).all()


def test_small_workload_matches_golden_digest():
    d = json.load(open(os.path.join(HERE, "golden", "digests.json")))
    ps = synth.workload("cpu_ref_151bp_4k")
    assert input_digest(ps) == d["cpu_ref_151bp_4k:inputs"]


@pytest.mark.parametrize("name", ["cpu_ref_151bp_4k", "illumina_150bp_1m", "uneven_mix_4m"])
def test_workload_shape(name):
    ps = synth.workload(name, n=20000)
    assert (ps.qlen >= 1).all()
    assert (ps.tlen == ps.qlen + 100).all()
    assert ((ps.h0 >= 19) & (ps.h0 <= 60)).all()
    if name == "uneven_mix_4m":
        assert ps.qlen.max() <= 231
        frac_wide = (ps.w == 200).mean()
        assert 0.08 < frac_wide < 0.12
    else:
        assert (ps.w == 100).all()
        assert ps.qlen.max() <= 132
    assert ps.q.max() <= 3 and ps.t.max() <= 3


def test_golden_subset_fixture_is_consistent():
    from tests import _oracle as O
    z = np.load(os.path.join(HERE, "golden", "subset_uneven_mix_4m.npz"))
    ps = synth.PairSet(z["q"], z["t"], z["qlen"], z["tlen"], z["h0"], z["w"])
    out, cells, rc = O.extend_batch(synth.default_scoring(), *ps.flat())
    assert (O.to_outscore_order(out) == z["out"]).all() and (cells == z["cells"]).all()
