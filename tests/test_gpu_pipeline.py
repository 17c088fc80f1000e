"""GPU parity of the chain-extension path: the producer, the packing of
queries and .pac targets, both extension batches and the merge on the device
(fa_extend_chains over host buffers, fa_chains_* with inputs resident in
HBM) against the same pipeline on the CPU -- the library's host producer and
merge (checked against the Python restatement of collect_soa.rs and
region.rs in tests/test_pipeline_host.py) around extensions by the
ksw_extend2 oracle on targets cut by the .pac oracle.  Bit-exact on every
region field; exact reads extended from one interior seed give the
full-length region (closed form); invalid inputs give the error of the first
invalid job in the reference's job order."""
import numpy as np
import pytest

from tests import chains_synth
from tests import pipeline_oracle as PO
from tests.test_pipeline_host import REGION_FIELDS, _cpu_pipeline, _edge_batch

pytestmark = pytest.mark.gpu


def _gpu(d, sw, opt, resident=False):
    from ferrousalign_amd.banded_swa import DeviceReference
    from ferrousalign_amd.pipeline import DeviceChains, extend_chains
    ref = DeviceReference(d["pac"], d["l_pac"])
    try:
        if not resident:
            return extend_chains(ref, sw, opt, d["qseq"], d["qoff"], d["cb"])
        dc = DeviceChains(d["qseq"], d["qoff"], d["cb"])
        try:
            nl, nr = dc.extend(ref, sw, opt)
            first = dc.download()
            assert dc.extend(ref, sw, opt) == (nl, nr)  # a second run on the same inputs
            assert np.array_equal(dc.download().view(np.uint8), first.view(np.uint8))
            return first, nl, nr
        finally:
            dc.free()
    finally:
        ref.free()


def _compare(got, want):
    for f in REGION_FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert np.array_equal(got["frac_rep"], want["frac_rep"])


@pytest.mark.parametrize("seed,w,zdrop,resident", [(0, 100, 100, False), (1, 20, 40, True), (2, 5, 100, False)])
def test_extend_chains_matches_cpu_pipeline(seed, w, zdrop, resident):
    from ferrousalign_amd import synth
    from ferrousalign_amd.pipeline import MemOpt
    sw = synth.default_scoring()
    sw.zdrop = zdrop
    opt = MemOpt(w=w, zdrop=zdrop)
    d = chains_synth.make(seed=seed, nreads=400, sub=0.03)
    got, nl, nr = _gpu(d, sw, opt, resident)
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
    for resident in (False, True):
        got, nl, nr = _gpu(d, sw, opt, resident)
        want, wl, wr = _cpu_pipeline(d, sw, opt)
        assert (nl, nr) == (wl, wr)
        _compare(got, want)


def test_chain_reads_workload_subset():
    """20,000 reads of the 150 bp chain workload (both strands, decoy
    chains), resident path against the CPU pipeline."""
    from ferrousalign_amd import synth, pack_pac
    from ferrousalign_amd.pipeline import MemOpt
    cb, qseq, qoff, _ = synth.chain_reads(n=20000)
    ref = synth.reference()
    d = dict(pac=pack_pac(ref), l_pac=ref.size, cb=cb, qseq=qseq, qoff=qoff)
    sw = synth.default_scoring()
    got, nl, nr = _gpu(d, sw, MemOpt(), resident=True)
    want, wl, wr = PO.extend_chains_cpu(cb, d["pac"], d["l_pac"], qseq, qoff, sw, MemOpt())
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


def _one_read_batch(reads):
    """reads: list of (query codes, [(ref_pos, query_pos, len), ...]) -> one
    kept chain per read."""
    from ferrousalign_amd import _native as N
    from ferrousalign_amd.pipeline import ChainBatch
    n = len(reads)
    ql = np.array([len(q) for q, _ in reads], np.int32)
    qoff = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    seeds = [s for _, ss in reads for s in ss]
    sd = np.zeros(len(seeds), N.SEED_DTYPE)
    for i, x in enumerate(seeds):
        sd[i] = x
    cnt = np.array([len(ss) for _, ss in reads], np.int32)
    cs = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    cb = ChainBatch(ql, np.arange(n, dtype=np.int64), np.ones(n, np.int32), np.ones(n, np.uint8),
                    np.zeros(n, np.float32), cs, cnt, np.arange(len(seeds), dtype=np.int64), sd)
    return cb, np.concatenate([q for q, _ in reads]).astype(np.uint8), qoff


def _run_err(reads, l_pac=20000):
    from ferrousalign_amd import synth, pack_pac, _native as N
    from ferrousalign_amd.banded_swa import DeviceReference
    from ferrousalign_amd.pipeline import MemOpt, extend_chains
    rng = np.random.default_rng(3)
    ref = DeviceReference(pack_pac(rng.integers(0, 4, l_pac).astype(np.uint8)), l_pac)
    cb, q, qo = _one_read_batch(reads)
    try:
        extend_chains(ref, synth.default_scoring(), MemOpt(), q, qo, cb)
        return N.FA_OK
    except N.FaError as e:
        return e.code
    finally:
        ref.free()


def test_errors_follow_job_order():
    from ferrousalign_amd import _native as N
    good = np.zeros(150, np.uint8)
    bad_left = good.copy()
    bad_left[3] = 7                    # inside the left job's query [0, 60)
    long_read = np.zeros(300, np.uint8)
    # read 0 valid; read 1: bad code in its left job; read 2: left job of 280 bases (QLEN)
    reads = [(good, [(5000, 60, 20)]), (bad_left, [(6000, 60, 20)]), (long_read, [(7000, 280, 19)])]
    assert _run_err(reads) == N.FA_ERR_CODE
    assert _run_err([reads[0], reads[2], reads[1]]) == N.FA_ERR_QLEN
    # a bad code outside every job's query is no error
    seedpart = good.copy()
    seedpart[70] = 7                   # inside the seed [60, 80)
    assert _run_err([(seedpart, [(5000, 60, 20)])]) == N.FA_OK
    # right job of 281 bases: the left batch passes, the right one fails
    assert _run_err([(good, [(5000, 60, 20)]), (long_read, [(9000, 0, 19)])]) == N.FA_ERR_QLEN
    # structural errors
    assert _run_err([(good, [(5000, 140, 20)])]) == N.FA_ERR_ARG   # seed past the read's end
    from ferrousalign_amd import pack_pac, synth
    from ferrousalign_amd.banded_swa import DeviceReference
    from ferrousalign_amd.pipeline import MemOpt, extend_chains
    cb, q, qo = _one_read_batch([(good, [(5000, 60, 20)]), (good, [(6000, 60, 20)])])
    cb.read_chain_start[1] = 0         # chain 0 in both reads' ranges
    ref = DeviceReference(pack_pac(np.zeros(20000, np.uint8)), 20000)
    try:
        with pytest.raises(N.FaError) as e:
            extend_chains(ref, synth.default_scoring(), MemOpt(), q, qo, cb)
        assert e.value.code == N.FA_ERR_ARG
    finally:
        ref.free()
