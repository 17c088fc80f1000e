"""GPU: the coalescing aggregator (fa_agg_*).  Many host threads each submit
small batches at once, as every rayon worker of the reference calls
execute_batch_simd_scoring (single_end.rs:92-110); every record must equal
the oracle's, the calls must have been run as far fewer batches, and calls
with different scorings or invalid jobs must keep their own results and
return codes."""
import threading

import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _slices(ps, rng, n_calls, lo=16, hi=64):
    """Disjoint job ranges of `ps` for n_calls calls of lo..hi jobs."""
    sizes = rng.integers(lo, hi + 1, n_calls)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    assert starts[-1] + sizes[-1] <= ps.n
    return [(int(a), int(b)) for a, b in zip(starts, sizes)]


def _flat_of(ps, a, b):
    sub = ps.subset(np.arange(a, a + b))
    return sub.flat()


def test_concurrent_small_batches_are_coalesced_and_exact():
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import Aggregator
    sw = synth.default_scoring()
    ps = synth.workload("uneven_mix_4m", n=40000)
    ora, ocells, orc = O.extend_batch(sw, *ps.flat())
    assert (orc == 0).all()
    ora = O.to_outscore_order(ora)
    rng = np.random.default_rng(0)
    threads, per_thread = 16, 24
    jobs = _slices(ps, rng, threads * per_thread)
    errors = []

    def worker(t, agg):
        try:
            for (a, b) in jobs[t * per_thread:(t + 1) * per_thread]:
                out, cells = agg.execute(sw, *_flat_of(ps, a, b), want_cells=True)
                if not ((out == ora[a:a + b]).all() and (cells == ocells[a:a + b]).all()):
                    errors.append((t, a, b))
        except Exception as e:  # noqa: BLE001 - reported below
            errors.append((t, repr(e)))

    with Aggregator(0, min_jobs=2048, max_wait_us=3000) as agg:
        th = [threading.Thread(target=worker, args=(t, agg)) for t in range(threads)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        calls, runs, njobs = agg.stats()
    print("aggregator: %d calls, %d runs, %d jobs" % (calls, runs, njobs), flush=True)
    assert not errors, errors[:5]
    assert calls == threads * per_thread
    assert njobs == sum(b for _, b in jobs)
    assert runs < calls / 4, (runs, calls)


def test_scorings_and_invalid_jobs_stay_with_their_calls():
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd.banded_swa import Aggregator, BandedPairWiseSW, bwa_fill_scmat
    ps = synth.workload("cpu_ref_151bp_4k", n=2000)
    sws = [synth.default_scoring(), BandedPairWiseSW(5, 2, 3, 1, 0, 0, 5, 5, bwa_fill_scmat(2, 3, -1))]
    expect = []
    for sw in sws:
        o, c, r = O.extend_batch(sw, *ps.flat())
        assert (r == 0).all()
        expect.append((O.to_outscore_order(o), c))
    results = {}

    def worker(t, agg):
        sw = sws[t % 2]
        a, b = 100 * t, 100
        q, qo, ql, tt, to, tl, h0, w = _flat_of(ps, a, b)
        h0 = h0.copy()
        if t % 3 == 0:
            h0[7] = 0          # the reference asserts h0 > 0 (ksw_affine_gap.rs:968)
        results[t] = agg.execute(sw, q, qo, ql, tt, to, tl, h0, w, want_cells=True, per_job_rc=True)

    with Aggregator(0, min_jobs=100000, max_wait_us=20000) as agg:
        th = [threading.Thread(target=worker, args=(t, agg)) for t in range(12)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        calls, runs, _ = agg.stats()
    assert calls == 12 and runs >= 2
    for t in range(12):
        out, cells, rc, code = results[t]
        o, c = expect[t % 2]
        a, b = 100 * t, 100
        keep = np.ones(b, bool)
        if t % 3 == 0:
            keep[7] = False
            assert code == N.FA_ERR_H0 and rc[7] == N.FA_ERR_H0 and (out[7] == 0).all() and cells[7] == 0
        else:
            assert code == N.FA_OK
        assert (rc[keep] == 0).all()
        assert (out[keep] == o[a:a + b][keep]).all() and (cells[keep] == c[a:a + b][keep]).all(), t
