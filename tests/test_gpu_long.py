"""GPU parity of long extension pairs: queries over 255 bases and targets
over 2048 bases run in the long-pair kernel (one wavefront per pair, eh[] in
HBM).  The reference's ksw_extend2 (ksw_affine_gap.rs:952-1169) has no length
limit; every field and the cell count are compared with the CPU oracle, alone
and mixed into a batch of lane-group pairs, under every packing and launch
form, and at the length limits of the extension entry points."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _mutate(rng, src, sub=0.02, indel=0.005):
    out = []
    for c in src:
        r = rng.random()
        if r < indel:
            continue                                   # deletion
        if r < 2 * indel:
            out.append(int(rng.integers(0, 4)))        # insertion
        out.append(int(rng.integers(0, 4)) if rng.random() < sub else int(c))
    return np.array(out, np.uint8)


def _long_batch(rng, n, qrange, trange, related=0.7, n_frac=0.002):
    """Flat (qseq, qoff, qlen, tseq, toff, tlen, h0, w): a target is a
    mutated copy of its query (cut or padded with random bases to its
    length) with probability `related`, else random; a few N bases."""
    qs, ts = [], []
    for _ in range(n):
        ql = int(rng.integers(qrange[0], qrange[1] + 1))
        tl = int(rng.integers(trange[0], trange[1] + 1))
        q = rng.integers(0, 4, ql).astype(np.uint8)
        if rng.random() < related and tl > 0:
            t = _mutate(rng, q)
            if len(t) >= tl:
                t = t[:tl]
            else:
                t = np.concatenate([t, rng.integers(0, 4, tl - len(t)).astype(np.uint8)])
        else:
            t = rng.integers(0, 4, tl).astype(np.uint8)
        q[rng.random(ql) < n_frac] = 4
        t[rng.random(tl) < n_frac] = 4
        qs.append(q)
        ts.append(t)
    ql = np.array([len(x) for x in qs], np.int32)
    tl = np.array([len(x) for x in ts], np.int32)
    qoff = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    toff = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    qseq = np.concatenate(qs + [np.zeros(1, np.uint8)])
    tseq = np.concatenate(ts + [np.zeros(1, np.uint8)])
    h0 = rng.integers(1, 300, n).astype(np.int32)
    w = rng.choice([10, 50, 100, 500, 5000], n).astype(np.int32)
    return qseq, qoff, ql, tseq, toff, tl, h0, w


def _check(sw, flat, label, flags=0):
    from ferrousalign_amd import banded_swa as B
    ora, ocells, orc = O.extend_batch(sw, *flat)
    assert (orc == 0).all()
    ora = O.to_outscore_order(ora)
    gpu, gcells = B.extend_arrays(sw, *flat, flags=flags, want_cells=True)
    bad = np.nonzero((ora != gpu).any(axis=1) | (ocells.astype(np.uint64) != gcells.astype(np.uint64)))[0]
    if bad.size:
        k = bad[0]
        raise AssertionError("%s: %d/%d pairs differ; first pair %d (qlen %d tlen %d) oracle=%s cells=%d "
                             "gpu=%s cells=%d" % (label, bad.size, len(gpu), k, flat[2][k], flat[5][k],
                                                  ora[k].tolist(), int(ocells[k]), gpu[k].tolist(),
                                                  int(gcells[k])))
    return gpu, gcells


def _scorings():
    from ferrousalign_amd.banded_swa import BandedPairWiseSW, bwa_fill_scmat
    return [BandedPairWiseSW(6, 1, 6, 1, 100, 5, 5, 5, bwa_fill_scmat(1, 4, -1)),
            BandedPairWiseSW(5, 2, 3, 1, 0, 0, 5, 5, bwa_fill_scmat(2, 3, -1)),      # z-drop off
            BandedPairWiseSW(12, 3, 9, 2, 30, 20, 5, 5, bwa_fill_scmat(3, 5, -2))]


@pytest.mark.parametrize("seed", range(2))
def test_long_queries_match_oracle(seed):
    from ferrousalign_amd import _native as N
    rng = np.random.default_rng(100 + seed)
    flat = _long_batch(rng, 240, (256, 1500), (0, 2500))
    for ci, sw in enumerate(_scorings()):
        for flags in (0, N.FA_SCORE_I32):
            print("long queries", seed, ci, flags, flush=True)
            _check(sw, flat, "long queries seed=%d cfg=%d flags=%d" % (seed, ci, flags), flags)


def test_long_targets_match_oracle():
    rng = np.random.default_rng(7)
    flat = _long_batch(rng, 240, (20, 255), (2049, 6000))
    for ci, sw in enumerate(_scorings()):
        print("long targets", ci, flush=True)
        _check(sw, flat, "long targets cfg=%d" % ci)


def test_long_pairs_mixed_into_a_workload_batch():
    """Long pairs interleaved with lane-group pairs: one batch, every
    packing and launch form, identical exact results."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd import banded_swa as B
    sw = synth.default_scoring()
    ps = synth.workload("uneven_mix_4m", n=6000)
    rng = np.random.default_rng(3)
    lq = _long_batch(rng, 150, (200, 900), (100, 3000))
    # concatenate: the workload pairs then the long ones, then shuffle the job order
    fq, fqo, fql, ft, fto, ftl, fh0, fw = ps.flat()
    qseq = np.concatenate([fq, lq[0]])
    tseq = np.concatenate([ft, lq[3]])
    qoff = np.concatenate([fqo, lq[1] + len(fq)])
    toff = np.concatenate([fto, lq[4] + len(ft)])
    perm = rng.permutation(ps.n + len(lq[2]))
    flat = (qseq, qoff[perm], np.concatenate([fql, lq[2]])[perm], tseq, toff[perm],
            np.concatenate([ftl, lq[5]])[perm], np.concatenate([fh0, lq[6]])[perm],
            np.concatenate([fw, lq[7]])[perm])
    print("mixed", len(perm), flush=True)
    ref, rcells = _check(sw, flat, "mixed batch")
    db = B.DeviceBatch(sw, *flat)
    cost = rng.random(len(perm)).astype(np.float32) * 5000
    db.set_cost(cost)
    for lanes in (0, 8):
        db.set_launch(lanes)
        for flags in (N.FA_PACK_LENGTH, N.FA_PACK_NONE, N.FA_PACK_COST, N.FA_PACK_LENGTH | N.FA_LAUNCH_SINGLE,
                      N.FA_PACK_LENGTH | N.FA_LAUNCH_GRAPH, N.FA_PACK_COST | N.FA_LAUNCH_PERSISTENT):
            print("mixed form", lanes, hex(flags), flush=True)
            db.run(flags)
            out, cells = db.download()
            assert (out == ref).all() and (cells == rcells).all(), (lanes, hex(flags))
    names = [x["kernel"] for x in db.profile_launches(N.FA_PACK_LENGTH)]
    assert names[-1] == "extend_long_kernel" and names.count("extend_long_kernel") == 1
    db.free()


def test_length_limits():
    """The longest query and target the extension entry points take run
    exactly; one base more is a per-job error while the other jobs are
    scored."""
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd import banded_swa as B
    sw = synth.default_scoring()
    rng = np.random.default_rng(5)
    qa = rng.integers(0, 4, N.FA_MAX_QLEN).astype(np.uint8)          # longest query, short target
    ta = _mutate(rng, qa[:400], sub=0.01, indel=0.0)
    qb = rng.integers(0, 4, 40).astype(np.uint8)                     # short query, longest target
    tb = np.concatenate([_mutate(rng, qb, sub=0.0, indel=0.0),
                         rng.integers(0, 4, N.FA_MAX_TLEN - 40).astype(np.uint8)])
    qc = rng.integers(0, 4, N.FA_MAX_QLEN + 1).astype(np.uint8)      # one base too many
    tc = rng.integers(0, 4, 100).astype(np.uint8)
    qd = rng.integers(0, 4, 50).astype(np.uint8)
    td = rng.integers(0, 4, N.FA_MAX_TLEN + 1).astype(np.uint8)
    qs, ts = [qa, qb, qc, qd], [ta, tb, tc, td]
    ql = np.array([len(x) for x in qs], np.int32)
    tl = np.array([len(x) for x in ts], np.int32)
    qoff = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    toff = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    qseq, tseq = np.concatenate(qs), np.concatenate(ts)
    h0 = np.array([60, 30, 30, 30], np.int32)
    w = np.array([60, 20, 20, 20], np.int32)
    out, cells, rc, code = B.extend_arrays(sw, qseq, qoff, ql, tseq, toff, tl, h0, w, want_cells=True,
                                           per_job_rc=True)
    assert rc.tolist() == [0, 0, N.FA_ERR_QLEN, N.FA_ERR_TLEN] and code == N.FA_ERR_QLEN
    assert (out[2:] == 0).all()
    ora, ocells, orc = O.extend_batch(sw, qseq, qoff[:2], ql[:2], tseq, toff[:2], tl[:2], h0[:2], w[:2])
    assert (orc == 0).all()
    assert (O.to_outscore_order(ora) == out[:2]).all(), (O.to_outscore_order(ora), out[:2])
    assert (ocells == cells[:2]).all()
