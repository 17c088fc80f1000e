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


@pytest.mark.parametrize("lanes", [4, 8, 16])
def test_parity_lane_groups(lanes):
    from ferrousalign_amd import synth
    ps = synth.workload("uneven_mix_4m", n=8192)
    # a lane group of L lanes holds at most 32 columns per lane
    ps = ps.subset(np.nonzero(ps.qlen + 1 <= lanes * 32)[0])
    sw = synth.default_scoring()
    print("lanes", lanes, ps.n, flush=True)
    ora, oc, gpu, gc = _run_both(ps, sw, lanes=lanes)
    _assert_same(ora, oc, gpu, gc, "lanes=%d" % lanes)
