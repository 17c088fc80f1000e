"""The extension-job producer and the region merge of the library
(fa_collect_extension_jobs, fa_merge_scores_to_regions; host code, no GPU)
against the Python restatement of the reference's collect_soa.rs:28-228 and
region.rs:737-890 (tests/pipeline_oracle.py), plus closed forms: an exact
read extended from an interior seed gives the full-length region."""
import numpy as np
import pytest

from ferrousalign_amd import _native as N
from ferrousalign_amd.pipeline import ChainBatch, MemOpt, collect_extension_jobs, merge_scores_to_regions
from tests import _oracle as O
from tests import chains_synth
from tests import pipeline_oracle as PO

JOB_FIELDS = ("read_idx", "chain_idx", "seed_idx", "qbeg", "qlen", "h0", "w", "seg_beg", "seg_len", "seg_rev")
REGION_FIELDS = ("rb", "re", "qb", "qe", "score", "truesc", "seedcov", "seedlen0", "chain_idx", "seed_idx", "w")


def _edge_batch(rng, l_pac=5000, nreads=300):
    """Chains at the strand boundary, near coordinate 0 and 2 l_pac, seeds at
    the query ends, empty and unkept chains, seeds left of the window."""
    qlens, rcs, rcc, kept, frac, css, csc, cs, seeds = [], [], [], [], [], [], [], [], []
    for r in range(nreads):
        R = int(rng.integers(20, 255))
        qlens.append(R)
        rcs.append(len(kept))
        nch = int(rng.integers(0, 4))
        rcc.append(nch)
        for _ in range(nch):
            css.append(len(cs))
            ns = int(rng.integers(0, 4))
            csc.append(ns)
            kept.append(int(rng.random() < 0.85))
            frac.append(float(rng.random()))
            where = rng.integers(0, 4)
            base = [int(rng.integers(0, 300)), l_pac - int(rng.integers(0, 200)), 2 * l_pac - 600,
                    int(rng.integers(0, 2 * l_pac - 600))][where]
            for _ in range(ns):
                ln = int(rng.integers(1, min(R, 60) + 1))
                qp = int(rng.choice([0, R - ln, int(rng.integers(0, R - ln + 1))]))
                rp = max(0, min(2 * l_pac - ln, base + int(rng.integers(-150, 150))))
                cs.append(len(seeds))
                seeds.append((rp, qp, ln))
    sd = np.zeros(len(seeds), N.SEED_DTYPE)
    for i, x in enumerate(seeds):
        sd[i] = x
    return ChainBatch(np.array(qlens, np.int32), np.array(rcs, np.int64), np.array(rcc, np.int32),
                      np.array(kept, np.uint8), np.array(frac, np.float32), np.array(css, np.int64),
                      np.array(csc, np.int32), np.array(cs, np.int64), sd), l_pac


def _check_collect(opt, l_pac, cb):
    left, right, m = collect_extension_jobs(opt, l_pac, cb)
    ol, orr, omaps, ormax = PO.collect_extension_jobs(opt, l_pac, cb)
    assert len(left) == len(ol) and len(right) == len(orr)
    for got, want in ((left, ol), (right, orr)):
        for f in JOB_FIELDS:
            assert [int(x) for x in got[f]] == [j[f] for j in want], f
    for c in range(cb.nchains):
        got = [(int(m.seed[i]), int(m.left[i]), int(m.right[i]))
               for i in range(int(m.start[c]), int(m.start[c]) + int(m.count[c]))]
        assert got == omaps[c], c
        want = (-1, -1) if ormax[c] is None else ormax[c]
        assert tuple(int(x) for x in m.rmax[c]) == tuple(want), c
    return left, right, m, (ol, orr, omaps, ormax)


@pytest.mark.parametrize("seed", range(3))
def test_collect_matches_oracle_on_simulated_reads(seed):
    d = chains_synth.make(seed=seed, nreads=300)
    for opt in (MemOpt(), MemOpt(a=2, b=3, o_del=5, e_del=2, o_ins=4, e_ins=1, w=30)):
        _check_collect(opt, d["l_pac"], d["cb"])


@pytest.mark.parametrize("seed", range(4))
def test_collect_matches_oracle_on_edge_chains(seed):
    rng = np.random.default_rng(100 + seed)
    cb, l_pac = _edge_batch(rng)
    left, right, m, _ = _check_collect(MemOpt(w=int(rng.integers(5, 120))), l_pac, cb)
    # every job's target lies on one strand inside [0, 2 l_pac)
    for jobs in (left, right):
        b, e = jobs["seg_beg"], jobs["seg_beg"] + jobs["seg_len"]
        assert (b >= 0).all() and (e <= 2 * l_pac).all() and not ((b < l_pac) & (e > l_pac)).any()


@pytest.mark.parametrize("seed", range(4))
def test_merge_matches_oracle(seed):
    rng = np.random.default_rng(200 + seed)
    cb, l_pac = _edge_batch(rng)
    opt = MemOpt(pen_clip5=int(rng.integers(0, 10)), pen_clip3=int(rng.integers(0, 10)))
    left, right, m, (_, _, omaps, ormax) = _check_collect(opt, l_pac, cb)

    def scores(n):
        s = np.zeros((max(n, 1), 6), np.int32)
        s[:, 0] = rng.integers(1, 300, len(s))                        # score
        s[:, 1] = rng.integers(0, 150, len(s))                        # tle
        s[:, 2] = rng.integers(0, 150, len(s))                        # gtle
        s[:, 3] = rng.integers(0, 150, len(s))                        # qle
        s[:, 4] = s[:, 0] - rng.integers(-3, 12, len(s))              # gscore around score
        s[rng.random(len(s)) < 0.2, 4] = -1
        return s[:n] if n else s[:0]
    ls, rs = scores(len(left)), scores(len(right))
    for rfl in (True, False):
        got = merge_scores_to_regions(opt, cb, m, ls, rs, right_from_left=rfl)
        want = PO.merge_scores_to_regions(opt, cb, omaps, ormax, ls, rs, rfl)
        for c in range(cb.nchains):
            if want[c] is None:
                assert int(got[c]["seed_idx"]) == -1, c
                continue
            for f in REGION_FIELDS:
                assert int(got[c][f]) == int(want[c][f]), (c, f)
            assert np.float32(got[c]["frac_rep"]) == np.float32(want[c]["frac_rep"])


def _cpu_pipeline(d, sw, opt):
    return PO.extend_chains_cpu(d["cb"], d["pac"], d["l_pac"], d["qseq"], d["qoff"], sw, opt)


def test_exact_reads_give_full_length_regions():
    from ferrousalign_amd import synth
    sw = synth.default_scoring()
    opt = MemOpt()
    d = chains_synth.make(seed=5, nreads=120, exact=True, single_seed=True)
    regions, nl, nr = _cpu_pipeline(d, sw, opt)
    assert nl > 0 and nr > 0
    for r, (P, R, rev, ins) in enumerate(d["truth"]):
        g = regions[int(d["cb"].read_chain_start[r])]
        assert (int(g["qb"]), int(g["qe"]), int(g["rb"]), int(g["re"]), int(g["score"])) == (0, R, P, P + R, R), r
