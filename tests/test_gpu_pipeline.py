"""GPU parity of the chain-extension path (fa_extend_chains): job producer,
left batch and right batch on the GPU with targets cut from the resident .pac,
right h0 chained from the left score, region merge -- against the same
pipeline with the extensions run by the ksw_extend2 oracle on targets cut by
the .pac oracle.  Bit-exact on every region field; exact reads extended from
one interior seed give the full-length region (closed form)."""
import numpy as np
import pytest

from tests import chains_synth
from tests.test_pipeline_host import REGION_FIELDS, _cpu_pipeline, _edge_batch

pytestmark = pytest.mark.gpu


def _gpu(d, sw, opt):
    from ferrousalign_amd.banded_swa import DeviceReference
    from ferrousalign_amd.pipeline import extend_chains
    ref = DeviceReference(d["pac"], d["l_pac"])
    try:
        return extend_chains(ref, sw, opt, d["qseq"], d["qoff"], d["cb"])
    finally:
        ref.free()


def _compare(got, want):
    for f in REGION_FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert np.array_equal(got["frac_rep"], want["frac_rep"])


@pytest.mark.parametrize("seed,w,zdrop", [(0, 100, 100), (1, 20, 40), (2, 5, 100)])
def test_extend_chains_matches_cpu_pipeline(seed, w, zdrop):
    from ferrousalign_amd import synth
    from ferrousalign_amd.pipeline import MemOpt
    sw = synth.default_scoring()
    sw.zdrop = zdrop
    opt = MemOpt(w=w, zdrop=zdrop)
    d = chains_synth.make(seed=seed, nreads=400, sub=0.03)
    got, nl, nr = _gpu(d, sw, opt)
    want, wl, wr = _cpu_pipeline(d, sw, opt)
    assert (nl, nr) == (wl, wr) and nl > 100 and nr > 100
    _compare(got, want)


def test_extend_chains_edge_windows():
    """Chains at the strand boundary and the ends of [0, 2 l_pac), seeds at
    the query ends, unkept and empty chains."""
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.pipeline import MemOpt
    rng = np.random.default_rng(7)
    cb, l_pac = _edge_batch(rng)
    ref = rng.integers(0, 4, l_pac).astype(np.uint8)
    qlen = cb.query_len
    qoff = np.zeros(cb.nreads, np.int64)
    qoff[1:] = np.cumsum(qlen)[:-1]
    d = dict(pac=pack_pac(ref), l_pac=l_pac, cb=cb, qoff=qoff,
             qseq=rng.integers(0, 5, int(qlen.sum())).astype(np.uint8))
    sw = synth.default_scoring()
    opt = MemOpt(w=30)
    got, nl, nr = _gpu(d, sw, opt)
    want, wl, wr = _cpu_pipeline(d, sw, opt)
    assert (nl, nr) == (wl, wr)
    _compare(got, want)


def test_exact_reads_full_length_on_gpu():
    from ferrousalign_amd import synth
    from ferrousalign_amd.pipeline import MemOpt
    d = chains_synth.make(seed=5, nreads=300, exact=True, single_seed=True)
    got, nl, nr = _gpu(d, synth.default_scoring(), MemOpt())
    assert nl > 0 and nr > 0
    for r, (P, R, rev, ins) in enumerate(d["truth"]):
        g = got[int(d["cb"].read_chain_start[r])]
        assert (int(g["qb"]), int(g["qe"]), int(g["rb"]), int(g["re"]), int(g["score"])) == (0, R, P, P + R, R), r
