"""The scalar_banded_swa oracle (oracle/swtb_oracle.c, restating
src/core/alignment/banded_swa/scalar/implementation.rs:5-440), pinned by
closed forms (exact match, one substitution, a single insertion or deletion
between long exact flanks, a 5' mismatching prefix, a 3' unaligned tail) and
by invariants on random pairs: the CIGAR consumes the whole query, the
aligned bases are exactly the bases its M/I/D runs consume, and a gap-free
path scores exactly the reported maximum."""
import numpy as np
import pytest

from tests import _oracle as O

A, B, O_GAP, E_GAP = 1, 4, 6, 1


def _sw():
    from ferrousalign_amd import synth
    return synth.default_scoring()  # A=1 B=4 O=6 E=1, zdrop 100, pen_clip 5


def _rand(rng, n):
    return rng.integers(0, 4, n).astype(np.uint8)


@pytest.mark.parametrize("n", [1, 7, 31, 150])
def test_exact_match(n):
    s = _rand(np.random.default_rng(n), n)
    rec, cig, ra, qa, _ = O.scalar_banded_swa(s, s, _sw(), w=n + 5)
    assert rec.tolist() == [n * A, n, n, n, n * A, 0]
    assert cig == [("M", n)]
    assert np.array_equal(ra, s) and np.array_equal(qa, s)


def test_one_substitution():
    s = _rand(np.random.default_rng(1), 41)
    t = s.copy()
    t[20] = (t[20] + 1) % 4
    rec, cig, _, _, _ = O.scalar_banded_swa(s, t, _sw(), w=50)
    assert rec[0] == 40 * A - B and cig == [("M", 41)]


@pytest.mark.parametrize("kind", ["ins", "del"])
def test_single_gap_between_flanks(kind):
    rng = np.random.default_rng(2)
    t1, t2 = _rand(rng, 30), _rand(rng, 30)
    x = np.array([(t2[0] + 2) % 4], np.uint8)  # inserted base differs from the next one
    if kind == "ins":
        q, t = np.concatenate([t1, x, t2]), np.concatenate([t1, t2])
        want = [("M", 30), ("I", 1), ("M", 30)]
    else:
        q, t = np.concatenate([t1, t2]), np.concatenate([t1, x, t2])
        want = [("M", 30), ("D", 1), ("M", 30)]
    rec, cig, ra, qa, _ = O.scalar_banded_swa(q, t, _sw(), w=20)
    assert rec[0] == 60 * A - (O_GAP + E_GAP)
    assert cig == want
    assert np.array_equal(qa, q) and np.array_equal(ra, t)


def test_five_prime_prefix_runs_into_the_insertion_row():
    # the traceback walks on until (0, 0); row 0 holds INS codes and column 0
    # DEL codes (implementation.rs:84-91), so an unaligned 5' query prefix is
    # reported as an insertion
    rng = np.random.default_rng(3)
    t = _rand(rng, 60)
    pre = (t[:8] + 2) % 4
    q = np.concatenate([pre, t])
    rec, cig, ra, qa, _ = O.scalar_banded_swa(q, t, _sw(), w=20)
    assert cig == [("I", 8), ("M", 60)] and rec[0] == 60 * A
    assert np.array_equal(qa, q) and np.array_equal(ra, t)


def test_three_prime_tail_clip_decision():
    rng = np.random.default_rng(4)
    t = _rand(rng, 60)
    tail = (_rand(rng, 12) + 0) % 4
    q = np.concatenate([t, tail])
    rec, cig, _, _, _ = O.scalar_banded_swa(q, t, _sw(), w=30)
    score, gscore = rec[0], rec[4]
    assert cig[0] == ("M", 60)
    clip = gscore <= 0 or gscore <= score - 5
    assert cig[-1] == (("S" if clip else "M"), 12)


@pytest.mark.parametrize("seed", range(4))
def test_random_pairs_invariants(seed):
    rng = np.random.default_rng(100 + seed)
    sw = _sw()
    for _ in range(150):
        n = int(rng.integers(1, 120))
        t = _rand(rng, int(rng.integers(1, 160)))
        # a mutated copy of a target window, or random
        if rng.random() < 0.8 and len(t) > 4:
            a = int(rng.integers(0, len(t) - 1))
            q = t[a:a + n].copy()
            for k in range(len(q)):
                if rng.random() < 0.05:
                    q[k] = (q[k] + 1 + rng.integers(0, 3)) % 4
            if rng.random() < 0.3 and len(q) > 3:
                k = int(rng.integers(1, len(q) - 1))
                q = np.delete(q, k) if rng.random() < 0.5 else np.insert(q, k, rng.integers(0, 4))
        else:
            q = _rand(rng, n)
        if q.size == 0:
            continue
        w = int(rng.integers(1, 200))
        rec, cig, ra, qa, _ = O.scalar_banded_swa(q, t, sw, w)
        ops = dict(M=0, I=0, D=0, S=0)
        for op, ln in cig:
            assert ln > 0
            ops[op] += ln
        assert ops["M"] + ops["I"] + ops["S"] == len(q)
        # the query after the traced end is one S or M run (3' decision)
        core = list(cig)
        ext = len(q) - int(rec[3])
        if ext > 0:
            assert core[-1][1] == ext and core[-1][0] in "SM"
            core = core[:-1]
        qs = 0
        if core and core[0][0] == "S":
            qs = core[0][1]
            core = core[1:]
        tle = int(rec[1])
        ts = tle - sum(l for o, l in core if o in "MD")
        assert ts == 0 and qs == 0  # the traceback always reaches (0, 0)
        qi, ti, rq, rt, gap_free = qs, ts, [], [], True
        path = 0
        for op, ln in core:
            if op == "M":
                for k in range(ln):
                    path += A if q[qi + k] == t[ti + k] else -B
                rq += list(q[qi:qi + ln]); rt += list(t[ti:ti + ln]); qi += ln; ti += ln
            elif op == "I":
                rq += list(q[qi:qi + ln]); qi += ln; gap_free = False
            elif op == "D":
                rt += list(t[ti:ti + ln]); ti += ln; gap_free = False
        assert ti == tle and qi == rec[3]
        assert np.array_equal(np.array(rq, np.uint8), qa) and np.array_equal(np.array(rt, np.uint8), ra)
        if gap_free and rec[0] > 0 and ts >= 0:
            assert path == rec[0]
