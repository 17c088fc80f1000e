"""The reference's drop-in surface on the GPU, called the way its callers
call it, every field against the ksw_extend2 oracle:

* the kernel-level entry points simd_banded_swa_batch{16,32,64}_soa and
  simd_banded_swa_batch{8,16,32}_int16_soa (banded_swa/shared.rs:282-389) on
  pad_batch + soa_transform buffers (0xFF padding, ragged batches with
  num_jobs < lanes, an empty batch, N bases);
* the pipeline-level execute_batch_simd_scoring over an ExtensionJobBatch
  (batch_extension/dispatch.rs:31-86, types.rs:20-197): one result per job,
  in job order, with the job's ids and direction;
* the reference's right-extension jobs: collect_extension_jobs_batch_soa
  builds them with h0 = 0 (collect_soa.rs:206) and orchestration_soa.rs:170
  passes them to execute_batch_simd_scoring; ksw_extend2 asserts h0 > 0
  (ksw_affine_gap.rs:968), so each such job gets FA_ERR_H0 and a zeroed
  record while the valid jobs of the same call are scored, and when the
  caller gives each right job the seed's left-extended score as h0 (or
  seed_len * a without a left job), every job is scored."""
import numpy as np
import pytest

from ferrousalign_amd import _native as N
from ferrousalign_amd import banded_swa as B
from ferrousalign_amd import synth
from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _job_tuples(n, seed):
    """(qlen, query, tlen, target, w, h0) from the config-0 and uneven-mix
    workloads, a few bases turned into N."""
    rng = np.random.default_rng(seed)
    a = synth.workload("cpu_ref_151bp_4k", n=n)
    b = synth.workload("uneven_mix_4m", n=4 * n)
    jobs = []
    for ps, k in ((a, rng.choice(a.n, n // 2, replace=False)), (b, rng.choice(b.n, n - n // 2, replace=False))):
        for i in k:
            ql, tl = int(ps.qlen[i]), int(ps.tlen[i])
            q, t = ps.q[i, :ql].copy(), ps.t[i, :tl].copy()
            if rng.random() < 0.2:
                q[rng.integers(0, ql)] = 4
                if tl:
                    t[rng.integers(0, tl)] = 4
            jobs.append((ql, q, tl, t, int(ps.w[i]), int(ps.h0[i])))
    rng.shuffle(jobs)
    return jobs


def _oracle(sw, jobs):
    if not jobs:
        return np.zeros((0, 6), np.int32)
    ql = np.array([j[0] for j in jobs], np.int32)
    tl = np.array([j[2] for j in jobs], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    q = np.concatenate([j[1] for j in jobs] + [np.zeros(1, np.uint8)])
    t = np.concatenate([j[3] for j in jobs] + [np.zeros(1, np.uint8)])
    out, _, rc = O.extend_batch(sw, q, qo, ql, t, to, tl, [j[5] for j in jobs], [j[4] for j in jobs])
    assert (rc == 0).all()
    return O.to_outscore_order(out)


def _as_rows(scores):
    return np.array([[s.score, s.target_end_pos, s.gtarget_end_pos, s.query_end_pos, s.global_score,
                      s.max_offset] for s in scores], np.int32).reshape(-1, 6)


@pytest.mark.parametrize("int16,lanes", [(False, 16), (False, 32), (False, 64), (True, 8), (True, 16), (True, 32)])
def test_soa_entry_points(int16, lanes):
    sw = synth.default_scoring()
    entry = getattr(B, "simd_banded_swa_batch%d%s_soa" % (lanes, "_int16" if int16 else ""))
    jobs = _job_tuples(3 * lanes, seed=lanes + 100 * int16)
    for num in (lanes, lanes - 3, 1, 0):
        batch = jobs[:num]
        ql, tl, h0, w, mq, mt, padded = B.pad_batch(batch, lanes)
        qs, ts = B.soa_transform(padded, lanes, 512)
        if int16:
            inputs = B.SoAInputs16(qs.astype(np.int16), ts.astype(np.int16), ql, tl, h0, w, lanes, mq, mt)
        else:
            inputs = B.SoAInputs(qs, ts, ql, tl, w, h0, lanes, mq, mt)
        got = entry(inputs, num, sw.o_del, sw.e_del, sw.o_ins, sw.e_ins, sw.zdrop, sw.mat, 5, sw.end_bonus)
        assert len(got) == num
        want = _oracle(sw, batch)
        assert (_as_rows(got) == want).all(), (lanes, num)


def test_execute_batch_simd_scoring_round_trip():
    sw = synth.default_scoring()
    jobs = _job_tuples(600, seed=7)
    batch = B.ExtensionJobBatch()
    dirs = (B.ExtensionDirection.Left, B.ExtensionDirection.Right)
    for k, (ql, q, tl, t, w, h0) in enumerate(jobs):
        batch.add_job(k // 3, k % 3, 1000 + k, dirs[k % 2], q, t, h0, w)
    assert len(batch) == len(jobs) and bytes(batch.get_query_seq(5)) == bytes(jobs[5][1])
    res = B.execute_batch_simd_scoring(sw, batch)
    want = _oracle(sw, jobs)
    assert len(res) == len(jobs)
    for k, r in enumerate(res):
        assert (r.read_idx, r.chain_idx, r.seed_idx, r.direction) == (k // 3, k % 3, 1000 + k, dirs[k % 2])
        # OutScore order: score, tle, gtle, qle, gscore, max_off
        assert (r.score, r.ref_end, r.gref_end, r.query_end, r.gscore, r.max_off) == tuple(int(x) for x in want[k])
    assert B.execute_batch_simd_scoring(sw, B.ExtensionJobBatch()) == []


def _pipeline_jobs(seed):
    """Left and right jobs of the reference's producer on synthetic chains,
    materialised as (query, target) codes."""
    from ferrousalign_amd.pipeline import MemOpt, collect_extension_jobs
    from tests import chains_synth
    d = chains_synth.make(seed=seed, nreads=300, sub=0.03)
    opt = MemOpt()
    left, right, m = collect_extension_jobs(opt, d["l_pac"], d["cb"])

    def mat(jobs, rev):
        out = []
        for j in jobs:
            b = int(d["qoff"][j["read_idx"]]) + int(j["qbeg"])
            q = d["qseq"][b:b + int(j["qlen"])]
            t = O.pac_segment(d["pac"], d["l_pac"], int(j["seg_beg"]), int(j["seg_len"]), reverse=bool(j["seg_rev"]))
            out.append((int(j["qlen"]), np.ascontiguousarray(q[::-1] if rev else q, np.uint8), len(t),
                        np.ascontiguousarray(t, np.uint8), int(j["w"]), int(j["h0"])))
        return out
    return mat(left, True), mat(right, False), left, right, m, d, opt


def test_reference_right_jobs_h0_zero():
    sw = synth.default_scoring()
    L, R, left, right, m, d, opt = _pipeline_jobs(3)
    assert len(L) > 100 and len(R) > 100
    assert all(j[5] == 0 for j in R) and all(j[5] > 0 for j in L)
    # one call with the left jobs first, then the right jobs as the producer made them
    jobs = L + R
    ql = np.array([j[0] for j in jobs], np.int32)
    tl = np.array([j[2] for j in jobs], np.int32)
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    q = np.concatenate([j[1] for j in jobs])
    t = np.concatenate([j[3] for j in jobs])
    h0 = np.array([j[5] for j in jobs], np.int32)
    w = np.array([j[4] for j in jobs], np.int32)
    out, _, rc, code = B.extend_arrays(sw, q, qo, ql, t, to, tl, h0, w, want_cells=True, per_job_rc=True)
    nl = len(L)
    assert code == N.FA_ERR_H0
    assert (rc[:nl] == 0).all() and (rc[nl:] == N.FA_ERR_H0).all()
    assert (out[nl:] == 0).all()
    assert (out[:nl] == _oracle(sw, L)).all()
    # the Python mirror of execute_batch_simd_scoring raises on such a batch
    batch = B.ExtensionJobBatch()
    for k, (a, qq, b, tt, ww, hh) in enumerate(R[:10]):
        batch.add_job(0, 0, k, B.ExtensionDirection.Right, qq, tt, hh, ww)
    with pytest.raises(N.FaError) as e:
        B.execute_batch_simd_scoring(sw, batch)
    assert e.value.code == N.FA_ERR_H0
    # h0 substitution at the call site: the left-extended score of the same
    # seed, or seed_len * a when the seed has no left job
    ls = out[:nl]
    cnt = m.count[:d["cb"].nchains].astype(np.int64)
    idx = np.repeat(m.start[:d["cb"].nchains] - (np.cumsum(cnt) - cnt), cnt) + np.arange(int(cnt.sum()))
    idx = idx[m.right[idx] >= 0]
    li = m.left[idx]
    seed_len = d["cb"].seeds["len"][m.seed[idx]].astype(np.int64)
    h0r = np.zeros(len(R), np.int64)
    h0r[m.right[idx]] = np.where(li >= 0, ls[np.maximum(li, 0), 0], seed_len * opt.a)
    R2 = [(a, qq, b, tt, ww, int(h0r[k])) for k, (a, qq, b, tt, ww, _) in enumerate(R)]
    assert all(j[5] > 0 for j in R2)
    ql2 = np.array([j[0] for j in R2], np.int32)
    tl2 = np.array([j[2] for j in R2], np.int32)
    out2, _, rc2, code2 = B.extend_arrays(
        sw, np.concatenate([j[1] for j in R2]), np.concatenate([[0], np.cumsum(ql2)[:-1]]).astype(np.int64), ql2,
        np.concatenate([j[3] for j in R2]), np.concatenate([[0], np.cumsum(tl2)[:-1]]).astype(np.int64), tl2,
        np.array([j[5] for j in R2], np.int32), np.array([j[4] for j in R2], np.int32), per_job_rc=True,
        want_cells=True)
    assert code2 == 0 and (rc2 == 0).all()
    assert (out2 == _oracle(sw, R2)).all()


def test_invalid_codes_and_scalar_errors_keep_their_order():
    """The uploads check the base codes while packing and validate in full
    only when something is wrong; the per-job codes and the batch's code
    must be those of the reference's checks in their order (h0, qlen, tlen,
    w, then the query's codes, then the target's: validate order of
    ksw_affine_gap.rs:968-985), and every valid job is still scored."""
    sw = synth.default_scoring()
    ps = synth.workload("uneven_mix_4m", n=3000)
    q, qo, ql, t, to, tl, h0, w = [np.array(x) for x in ps.flat()]
    q, t, h0 = q.copy(), t.copy(), h0.copy()
    want = np.zeros(ps.n, np.int32)
    rng = np.random.default_rng(21)
    picks = rng.choice(ps.n, 60, replace=False)
    for j, k in enumerate(picks):
        kind = j % 4
        if kind in (0, 3):        # a query code above 4 (at a random base, 5..255)
            q[qo[k] + int(rng.integers(0, ql[k]))] = int(rng.integers(5, 256))
        if kind == 1 and tl[k] > 0:  # a target code above 4
            t[to[k] + int(rng.integers(0, tl[k]))] = int(rng.integers(5, 256))
        if kind in (2, 3):        # h0 = 0: decides before any code
            h0[k] = 0
        want[k] = N.FA_ERR_H0 if kind in (2, 3) else (N.FA_ERR_CODE if kind == 0 or tl[k] > 0 else 0)
    print("codes and scalar errors", flush=True)
    out, cells, rc, code = B.extend_arrays(sw, q, qo, ql, t, to, tl, h0, w, want_cells=True, per_job_rc=True)
    assert np.array_equal(rc, want)
    first = want[np.nonzero(want)[0][0]]
    assert code == first
    bad = want != 0
    assert (out[bad] == 0).all() and (cells[bad] == 0).all()
    ora, ocells, orc = O.extend_batch(sw, q, qo, ql, t, to, tl, h0, w)
    assert np.array_equal(orc, want)  # the oracle's checks, in the same order
    ora = O.to_outscore_order(ora)
    assert (out[~bad] == ora[~bad]).all() and (cells[~bad] == ocells[~bad]).all()
    # a whole-batch upload reports the first invalid job's error
    for kinds in ((0,), (1,), (2,), (0, 2), (2, 0)):
        qq, tt, hh = ps.q.reshape(-1).copy(), ps.t.reshape(-1).copy(), np.array(ps.h0).copy()
        for pos, kind in zip((100, 200), kinds):
            if kind == 0:
                qq[qo[pos] + 1] = 9
            elif kind == 1:
                tt[to[pos] + 2] = 9
            else:
                hh[pos] = 0
        expect = {0: N.FA_ERR_CODE, 1: N.FA_ERR_CODE, 2: N.FA_ERR_H0}[kinds[0]]
        with pytest.raises(N.FaError) as e:
            B.DeviceBatch(sw, qq, qo, ql, tt, to, tl, hh, w)
        assert e.value.code == expect, kinds
