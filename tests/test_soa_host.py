"""Host side of the kernel-level drop-in surface: pad_batch and soa_transform
lay jobs out as shared.rs:40-111 does (lane-interleaved buf[pos * W + lane],
0xFF past each lane's length, empty lanes for missing jobs), with int32
per-lane parameters instead of the reference's i8 (no 127 clamp)."""
import numpy as np

from ferrousalign_amd.banded_swa import pad_batch, soa_transform


def _jobs(rng, n):
    out = []
    for _ in range(n):
        ql, tl = int(rng.integers(1, 240)), int(rng.integers(0, 340))
        out.append((ql, rng.integers(0, 5, ql).astype(np.uint8), tl, rng.integers(0, 5, tl).astype(np.uint8),
                    int(rng.integers(0, 300)), int(rng.integers(1, 300))))
    return out


def test_pad_batch_params_and_empty_lanes():
    rng = np.random.default_rng(0)
    jobs = _jobs(rng, 5)
    qlen, tlen, h0, w, mq, mt, padded = pad_batch(jobs, 8)
    assert len(padded) == 8 and padded[5:] == [(0, b"", 0, b"", 0, 0)] * 3
    assert list(qlen[:5]) == [j[0] for j in jobs] and list(qlen[5:]) == [0, 0, 0]
    assert list(tlen[:5]) == [j[2] for j in jobs]
    assert list(w[:5]) == [j[4] for j in jobs] and list(h0[:5]) == [j[5] for j in jobs]
    assert (mq, mt) == (max(j[0] for j in jobs), max(j[2] for j in jobs))
    assert qlen.dtype == np.int32 and (qlen > 127).any() == any(j[0] > 127 for j in jobs)  # no i8 clamp
    # more jobs than lanes: the extra ones are dropped
    q2 = pad_batch(jobs, 3)[0]
    assert list(q2) == [j[0] for j in jobs[:3]]


def test_soa_transform_layout():
    rng = np.random.default_rng(1)
    W, MAX = 16, 512
    jobs = _jobs(rng, 11)
    _, _, _, _, _, _, padded = pad_batch(jobs, W)
    q, t = soa_transform(padded, W, MAX)
    assert q.shape == (MAX * W,) and t.shape == (MAX * W,)
    for lane, (ql, qs, tl, ts, _, _) in enumerate(padded):
        col_q, col_t = q[lane::W], t[lane::W]
        qs = np.frombuffer(bytes(qs), np.uint8) if isinstance(qs, bytes) else qs
        ts = np.frombuffer(bytes(ts), np.uint8) if isinstance(ts, bytes) else ts
        assert (col_q[:ql] == qs[:ql]).all() and (col_q[ql:] == 0xFF).all()
        assert (col_t[:tl] == ts[:tl]).all() and (col_t[tl:] == 0xFF).all()
