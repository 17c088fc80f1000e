"""GPU parity of the batched scalar_banded_swa (CIGAR regeneration,
banded_swa/scalar/implementation.rs:5-440) through the C ABI against the CPU
oracle (oracle/swtb_oracle.c): score record, CIGAR, aligned bases and cell
count, bit for bit, on random and region-like pairs and edge cases."""
import numpy as np
import pytest

from tests import _oracle as O

pytestmark = pytest.mark.gpu


def _pairs(rng, n, max_q=255, max_t=400, n_frac=0.0):
    qs, ts, ws, hs = [], [], [], []
    for _ in range(n):
        t = rng.integers(0, 4, int(rng.integers(1, max_t))).astype(np.uint8)
        ql = int(rng.integers(1, max_q + 1))
        if rng.random() < 0.8 and len(t) > 4:
            a = int(rng.integers(0, len(t) - 1))
            q = t[a:a + ql].copy()
            for k in range(len(q)):
                if rng.random() < 0.04:
                    q[k] = (q[k] + 1 + rng.integers(0, 3)) % 4
            for _ in range(int(rng.integers(0, 3))):
                if len(q) > 3:
                    k = int(rng.integers(1, len(q) - 1))
                    L = int(rng.integers(1, 4))
                    q = np.delete(q, slice(k, k + L)) if rng.random() < 0.5 else \
                        np.insert(q, k, rng.integers(0, 4, L).astype(np.uint8))
            q = q[:max_q]
        else:
            q = rng.integers(0, 4, ql).astype(np.uint8)
        if n_frac:
            q[rng.random(len(q)) < n_frac] = 4
            t[rng.random(len(t)) < n_frac] = 4
        qs.append(q)
        ts.append(t)
        ws.append(int(rng.integers(0, 120)))
        hs.append(0 if rng.random() < 0.7 else int(rng.integers(1, 60)))
    return qs, ts, ws, hs


def _flat(seqs):
    lens = np.array([len(s) for s in seqs], np.int32)
    off = np.zeros(len(seqs), np.int64)
    off[1:] = np.cumsum(lens)[:-1]
    flat = np.concatenate(seqs) if seqs else np.zeros(0, np.uint8)
    return (flat if flat.size else np.zeros(1, np.uint8)), off, lens


def _check(qs, ts, ws, hs, label):
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import scalar_banded_swa_batch
    sw = synth.default_scoring()
    qf, qo, ql = _flat(qs)
    tf, to, tl = _flat(ts)
    out, cig, ra, qa, cells = scalar_banded_swa_batch(sw, qf, qo, ql, tf, to, tl, ws, hs, want_cells=True)
    for k in range(len(qs)):
        rec, c2, r2, q2, cl2 = O.scalar_banded_swa(qs[k], ts[k], sw, ws[k], hs[k])
        if len(qs[k]) == 0 or len(ts[k]) == 0:
            rec = np.zeros(6, np.int32)
            cl2 = 0
        ok = (np.array_equal(out[k], rec) and cig[k] == c2 and np.array_equal(ra[k], r2)
              and np.array_equal(qa[k], q2) and int(cells[k]) == cl2)
        assert ok, "%s pair %d (qlen %d tlen %d w %d h0 %d): gpu %s %s cells %d / oracle %s %s cells %d" % (
            label, k, len(qs[k]), len(ts[k]), ws[k], hs[k], out[k].tolist(), cig[k][:6], int(cells[k]),
            rec.tolist(), c2[:6], cl2)


@pytest.mark.parametrize("seed", range(3))
def test_random_pairs(seed):
    rng = np.random.default_rng(seed)
    qs, ts, ws, hs = _pairs(rng, 1500)
    print("random", seed, flush=True)
    _check(qs, ts, ws, hs, "random")


def test_n_bases_and_edges():
    rng = np.random.default_rng(7)
    qs, ts, ws, hs = _pairs(rng, 600, n_frac=0.02)
    e = np.zeros(0, np.uint8)
    one = np.array([2], np.uint8)
    long_t = rng.integers(0, 4, 2048).astype(np.uint8)
    long_q = long_t[100:355].copy()
    qs += [e, one, one, long_q, long_q]
    ts += [one, e, one, long_t, long_t]
    ws += [5, 5, 0, 300, 3]
    hs += [0, 0, 0, 0, 20]
    print("edges", flush=True)
    _check(qs, ts, ws, hs, "edges")


def test_region_like_pairs():
    # query = a read segment, reference = the window it came from (+-10 bp),
    # as region.rs builds them for CIGAR regeneration
    from ferrousalign_amd import synth
    ps = synth.workload("illumina_150bp_1m", n=2000)
    ref = synth.reference()
    rng = np.random.default_rng(3)
    qs, ts, ws, hs = [], [], [], []
    for k in range(ps.n):
        q = ps.q[k, :ps.qlen[k]]
        a = int(ps.tstart[k])
        lo = max(0, a - int(rng.integers(0, 10)))
        t = ref[lo:a + len(q) + int(rng.integers(0, 10))]
        qs.append(q.copy())
        ts.append(t.copy())
        ws.append(int(rng.integers(5, 40)))
        hs.append(0)
    print("region-like", flush=True)
    _check(qs, ts, ws, hs, "region-like")


def test_band_window_boundaries():
    """Bands at and just past each band-window configuration's limit
    2w + S + 1 <= L*S (w = 11, 23, 55 fit 4x8, 4x16, 8x16; one more goes to
    the next configuration or the full-width kernel), on 150-255 bp queries
    against long targets, so the window moves many times and the traceback
    reads codes from windows far from column 0."""
    rng = np.random.default_rng(21)
    qs, ts, ws, hs = [], [], [], []
    for w in (0, 1, 11, 12, 23, 24, 55, 56):
        for _ in range(40):
            t = rng.integers(0, 4, int(rng.integers(300, 900))).astype(np.uint8)
            ql = int(rng.integers(150, 256))
            a = int(rng.integers(0, len(t) - ql))
            q = t[a:a + ql].copy()
            q[rng.random(ql) < 0.03] = rng.integers(0, 4)
            if rng.random() < 0.5:  # an indel so the path leaves the diagonal
                k = int(rng.integers(10, ql - 10))
                q = np.insert(q, k, rng.integers(0, 4, int(rng.integers(1, 4))).astype(np.uint8))[:255]
            qs.append(q)
            ts.append(t)
            ws.append(w)
            hs.append(0 if rng.random() < 0.5 else int(rng.integers(1, 60)))
    print("band windows", flush=True)
    _check(qs, ts, ws, hs, "band windows")


def test_dense_output_matches_slots():
    # the dense entry point packs each job's CIGAR and aligned bases back to back;
    # its contents equal the slot entry point's, empty jobs included, and a batch
    # whose outputs exceed the capacities is refused
    from ferrousalign_amd import synth, _native as N
    from ferrousalign_amd.banded_swa import scalar_banded_swa_batch, scalar_banded_swa_batch_dense
    rng = np.random.default_rng(77)
    qs, ts, ws, hs = _pairs(rng, 300, n_frac=0.02)
    qs[3] = np.zeros(0, np.uint8)  # empty query
    ts[10] = np.zeros(0, np.uint8)  # empty target
    qs[-1] = np.zeros(0, np.uint8)  # empty last job
    sw = synth.default_scoring()
    qf, qo, ql = _flat(qs)
    tf, to, tl = _flat(ts)
    print("dense output", len(qs), flush=True)
    out, cells, cig, coff, nc, ra, qa, aoff, nr, nq = scalar_banded_swa_batch(sw, qf, qo, ql, tf, to, tl, ws, hs,
                                                                             raw=True)
    d = scalar_banded_swa_batch_dense(sw, qf, qo, ql, tf, to, tl, ws, hs)
    dout, dcells, dcig, dco, dra, dro, dqa, dqo = d
    assert np.array_equal(out, dout) and np.array_equal(cells, dcells)
    assert dco[0] == dro[0] == dqo[0] == 0
    assert np.array_equal(np.diff(dco), nc) and np.array_equal(np.diff(dro), nr) and np.array_equal(np.diff(dqo), nq)
    for k in range(len(qs)):
        assert np.array_equal(dcig[dco[k]:dco[k + 1]], cig[coff[k]:coff[k] + nc[k]])
        assert np.array_equal(dra[dro[k]:dro[k + 1]], ra[aoff[k]:aoff[k] + nr[k]])
        assert np.array_equal(dqa[dqo[k]:dqo[k + 1]], qa[aoff[k]:aoff[k] + nq[k]])
    # capacities one entry short of what the batch fills
    import ctypes
    n = len(qs)
    c = sw.to_c()
    o = np.empty((n, 6), np.int32)
    cl = np.empty(n, np.uint32)
    buf32 = np.empty(int(dco[n]) + 1, np.uint32)
    b1 = np.empty(int(max(dro[n], dqo[n])) + 1, np.uint8)
    b2 = np.empty_like(b1)
    offs = [np.empty(n + 1, np.int64) for _ in range(3)]
    keep = [np.ascontiguousarray(x) for x in (qf, qo, ql, tf, to, tl, np.asarray(ws, np.int32),
                                                np.asarray(hs, np.int32))]  # alive through the call
    args = [N.ptr(x) for x in keep]
    rc = N.lib().fa_scalar_banded_swa_batch_dense(
        _ctx(), ctypes.byref(c), n, *args,
        N.ptr(o), N.ptr(cl), N.ptr(buf32), int(dco[n]) - 1, N.ptr(offs[0]), N.ptr(b1), N.ptr(offs[1]), N.ptr(b2),
        N.ptr(offs[2]), int(max(dro[n], dqo[n])))
    assert rc == N.FA_ERR_ARG


def test_large_batch_with_empty_jobs():
    """131,072 region-like pairs, so the host work runs on many threads, with
    empty queries and targets scattered through every thread's range (the
    first and last job among them) and wide bands on a share of the pairs, so
    several kernels run: records, cells and CIGAR lengths of every pair equal
    the oracle's; both entry points give the same CIGARs and aligned bases;
    full contents of pairs spread over the batch equal the oracle's."""
    from ferrousalign_amd import synth
    from ferrousalign_amd.banded_swa import CIGAR_OPS, scalar_banded_swa_batch, scalar_banded_swa_batch_dense
    n = 131072
    qf, qo, ql, tf, to, tl, w = synth.region_pairs(n=n)
    ql, tl, w = ql.copy(), tl.copy(), w.copy()
    rng = np.random.default_rng(5)
    ql[::9973] = 0
    tl[5::7919] = 0
    ql[-1] = 0
    wide = rng.random(n) < 0.1
    w[wide] = rng.integers(24, 120, int(wide.sum()))
    h0 = np.where(rng.random(n) < 0.2, rng.integers(1, 60, n), 0).astype(np.int32)
    sw = synth.default_scoring()
    print("large batch", n, flush=True)
    out, cells, cig, coff, nc, ra, qa, aoff, nr, nq = scalar_banded_swa_batch(sw, qf, qo, ql, tf, to, tl, w, h0,
                                                                             raw=True)
    ora, ocells, onc = O.scalar_banded_swa_batch(sw, qf, qo, ql, tf, to, tl, w, h0)
    assert np.array_equal(out, ora)
    assert np.array_equal(cells, ocells.astype(np.uint32))
    assert np.array_equal(nc, onc)
    empty = (ql == 0) | (tl == 0)
    assert not out[empty].any() and not nc[empty].any() and not nr[empty].any()
    dout, dcells, dcig, dco, dra, dro, dqa, dqo = scalar_banded_swa_batch_dense(sw, qf, qo, ql, tf, to, tl, w, h0)
    assert np.array_equal(out, dout) and np.array_equal(cells, dcells)
    assert dco[0] == dro[0] == dqo[0] == 0
    assert np.array_equal(np.diff(dco), nc) and np.array_equal(np.diff(dro), nr) and np.array_equal(np.diff(dqo), nq)
    # the slot buffers gathered in job order equal the dense ones
    ci = np.concatenate([np.arange(coff[k], coff[k] + nc[k]) for k in range(n)])
    ai = np.concatenate([np.arange(aoff[k], aoff[k] + nr[k]) for k in range(n)])
    qi = np.concatenate([np.arange(aoff[k], aoff[k] + nq[k]) for k in range(n)])
    assert np.array_equal(cig[ci], dcig) and np.array_equal(ra[ai], dra) and np.array_equal(qa[qi], dqa)
    for k in list(range(0, n, 4099)) + [1, 9973, n - 2]:
        q = qf[qo[k]:qo[k] + ql[k]]
        t = tf[to[k]:to[k] + tl[k]]
        if len(q) == 0 or len(t) == 0:
            continue
        rec, c2, r2, q2, _ = O.scalar_banded_swa(q, t, sw, int(w[k]), int(h0[k]))
        got = [(CIGAR_OPS[int(x) & 15], int(x) >> 4) for x in dcig[dco[k]:dco[k + 1]]]
        assert got == c2, k
        assert np.array_equal(dra[dro[k]:dro[k + 1]], r2) and np.array_equal(dqa[dqo[k]:dqo[k + 1]], q2), k


def _ctx():
    from ferrousalign_amd.banded_swa import default_context
    return default_context().handle
