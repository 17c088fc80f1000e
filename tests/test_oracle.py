"""CPU oracle (oracle/ksw_oracle.c, restating ksw_affine_gap.rs:952-1169)
pinned against the reference's own known-answer test, closed forms,
hand-traced cases, invariants and an independent row-parallel restatement."""
import itertools
import random

import numpy as np
import pytest

from tests import _oracle as O
from tests import rowpar


def dna_mat_ref_tests():
    """get_dna_scoring_matrix of the reference tests (ksw_affine_gap.rs:1200-1215)."""
    mat = [0] * 25
    for i in range(5):
        for j in range(5):
            mat[i * 5 + j] = 1 if (i == j and i < 4) else -1
    for i in range(5):
        mat[20 + i] = 0
        mat[i * 5 + 4] = 0
    return mat


def bwa_mat(a=1, b=4, n=-1):
    mat = [0] * 25
    for i in range(4):
        for j in range(4):
            mat[i * 5 + j] = a if i == j else -b
        mat[i * 5 + 4] = n
    for j in range(5):
        mat[20 + j] = n
    return mat


def enc(s):
    return [{"A": 0, "C": 1, "G": 2, "T": 3}.get(c, 4) for c in s]


def ext(q, t, mat, o_del, e_del, o_ins, e_ins, w, eb, zd, h0):
    rc, out, cells = O.ksw_extend2(q, t, mat, o_del, e_del, o_ins, e_ins, w, eb, zd, h0)
    assert rc == 0
    return out, cells


# ---- the reference's own tests (ksw_affine_gap.rs:1252-1590) -------------

def test_ref_kat_simple_match():
    # :1252-1277: ACGT vs ACGT, gapo 5 gape 1, w 50, end_bonus 0, zdrop 100, h0 10
    (score, qle, tle, gtle, gscore, max_off), _ = ext(enc("ACGT"), enc("ACGT"), dna_mat_ref_tests(),
                                                       5, 1, 5, 1, 50, 0, 100, 10)
    assert score > 10
    assert qle == 4 and tle == 4


def test_ref_with_mismatch():
    (score, *_), _ = ext(enc("ACGT"), enc("ACCT"), dna_mat_ref_tests(), 5, 1, 5, 1, 50, 0, 100, 10)
    assert 0 < score <= 14


def test_ref_with_insertion():
    (score, *_), _ = ext(enc("ACGGT"), enc("ACGT"), dna_mat_ref_tests(), 5, 1, 5, 1, 50, 0, 100, 10)
    assert score > 0


def test_ref_zdrop():
    (score, qle, tle, *_), _ = ext(enc("A" * 10), enc("T" * 10), dna_mat_ref_tests(), 5, 1, 5, 1, 50, 0, 5, 10)
    assert tle < 10 or score < 10


def test_ref_asymmetric_gaps():
    (score, *_), _ = ext(enc("ACGT"), enc("ACGT"), dna_mat_ref_tests(), 5, 1, 3, 2, 50, 0, 100, 10)
    assert score >= 14


def test_ref_large_indel():
    (score, *_), _ = ext(enc("ACGTAAAAACGT"), enc("ACGTCGT"), dna_mat_ref_tests(), 5, 1, 5, 1, 100, 0, 200, 10)
    assert score > 0


def test_ref_15bp_insertion():
    q = enc("AGCTAGCTAGGGGGGGGGGGGGGGGAGCTAGCTAGCTAGCTAGCT")
    t = enc("AGCTAGCTAGAGCTAGCTAGCTAGCTAGCT")
    (score, qle, tle, gtle, gscore, max_off), _ = ext(q, t, dna_mat_ref_tests(), 5, 1, 5, 1, 50, 0, 200, 20)
    assert score > 0 or gscore > 0
    assert score >= 20


def test_ref_extend2_vs_scalar():
    (score, *_), _ = ext(enc("ACGTACGTAAAAAAAAAAAACGTACGT"), enc("ACGTACGTACGTACGT"), dna_mat_ref_tests(),
                         5, 1, 5, 1, 30, 0, 100, 10)
    assert score > 5


# ---- closed forms ------------------------------------------------------------

@pytest.mark.parametrize("qlen,extra,h0", [(1, 0, 7), (5, 3, 20), (40, 10, 30), (131, 100, 60)])
def test_closed_form_exact_match(qlen, extra, h0):
    rng = random.Random(qlen)
    q = [rng.randrange(4) for _ in range(qlen)]
    t = q + [rng.randrange(4) for _ in range(extra)]
    # exact match with w and zdrop non-binding: score climbs one per base
    (score, qle, tle, gtle, gscore, max_off), cells = ext(q, t, bwa_mat(), 6, 1, 6, 1, 100, 5, 100, h0)
    assert (score, qle, tle, gtle, gscore, max_off)[:6] == (h0 + qlen, qlen, qlen, qlen, h0 + qlen, 0)


def test_closed_form_empty_target():
    out, cells = ext(enc("ACGT"), [], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 9)
    assert out == (9, 0, 0, 0, -1, 0) and cells == 0


def test_hand_trace_single_match():
    # row 0: band [0,1) (w capped to 1), M = 10+1, H = 11, E = max(-1, 11-7) = 4,
    # eh[1] = (11, 0), gscore = 11 at i = 0
    out, cells = ext([0], [0], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 10)
    assert out == (11, 1, 1, 1, 11, 0) and cells == 1


def test_hand_trace_single_mismatch():
    # M = 10-4 = 6 < h0: no new max; z-drop 10-6-0 = 4 <= 100; gscore = 6 at i = 0
    out, cells = ext([0], [3], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 10)
    assert out == (10, 0, 0, 1, 6, 0) and cells == 1


def test_hand_trace_first_row_init():
    # h0 = 9 <= oe_ins + e_ins...: eh[1] = 2, eh[2] = 1, eh[3] = 0 ; one row, target base G
    # query A A A, target G (all mismatches): M0 = 9-4 = 5, M1 = 2-4 = -2, M2 = 1-4 = -3
    # E row: max(0-1, M-7, 0) -> 0 ; F: f1 = max(-1, max(5-7,0)) = 0, F(2) = 0
    # H = [5, 0, 0] ; row max 5 at j=0 ; 5 < 9 -> z-drop 9-5-(0)*1 = 4
    # w is capped to (3+5-6)/1+1 = 3, so end == qlen on row 0 and gscore takes
    # H(0, 2) = 0 there (gscore -1 < 0): gtle = 1, gscore = 0
    out, cells = ext([0, 0, 0], [2], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 9)
    assert out == (9, 0, 0, 1, 0, 0) and cells == 3


def test_zdrop_breaks_early():
    # random target vs query: the loop must stop long before tlen rows
    rng = random.Random(5)
    q = [rng.randrange(4) for _ in range(60)]
    t = [rng.randrange(4) for _ in range(400)]
    out, cells = ext(q, t, bwa_mat(), 6, 1, 6, 1, 100, 5, 30, 30)
    assert cells < 60 * 400 // 4


def test_w_cap_truncates_toward_zero():
    # (qlen*1 + end_bonus - o_ins)/e_ins + 1 = (2 + 0 - 9)/2 + 1 = -2.5 -> -2 -> max(.,1) = 1
    assert rowpar.capped_w(2, bwa_mat(), 9, 2, 9, 2, 100, 0) == 1
    # (10 + 5 - 6)/4 + 1 = 3.25 -> 3
    assert rowpar.capped_w(10, bwa_mat(), 6, 4, 6, 4, 100, 5) == 3


# ---- errors where the reference panics -------------------------------------

def test_error_codes():
    assert O.ksw_extend2([0, 1], [0], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 0)[0] == -2
    assert O.ksw_extend2([], [0], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 5)[0] == -3
    assert O.ksw_extend2([0, 5], [0], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 5)[0] == -4
    assert O.ksw_extend2([0, 1], [7], bwa_mat(), 6, 1, 6, 1, 100, 5, 100, 5)[0] == -4


# ---- independent row-parallel restatement -------------------------------------

def _both(q, t, mat, o_del, e_del, o_ins, e_ins, w, eb, zd, h0):
    a = ext(q, t, mat, o_del, e_del, o_ins, e_ins, w, eb, zd, h0)
    b = rowpar.extend(q, t, mat, o_del, e_del, o_ins, e_ins, w, eb, zd, h0)
    assert a == b, (q, t, w, zd, h0, a, b)


def test_rowpar_exhaustive_tiny():
    mat = bwa_mat(1, 4, -1)
    for ql in range(1, 4):
        for tl in range(0, 4):
            for q in itertools.product(range(3), repeat=ql):
                for t in itertools.product(range(3), repeat=tl):
                    for h0, w in ((1, 100), (8, 1), (12, 100)):
                        _both(list(q), list(t), mat, 6, 1, 6, 1, w, 5, 100, h0)


@pytest.mark.parametrize("seed", range(6))
def test_rowpar_random(seed):
    rng = random.Random(seed)
    for _ in range(150):
        ql = rng.randrange(1, 24)
        tl = rng.randrange(0, 30)
        alpha = rng.choice([2, 4, 5])
        q = [rng.randrange(alpha) for _ in range(ql)]
        t = [rng.randrange(alpha) for _ in range(tl)]
        a, b = rng.randrange(1, 3), rng.randrange(1, 5)
        mat = bwa_mat(a, b, rng.choice([-1, 0]))
        od, ed, oi, ei = rng.randrange(0, 7), rng.randrange(1, 3), rng.randrange(0, 7), rng.randrange(1, 3)
        w = rng.choice([0, 1, 2, 3, 5, 100])
        zd = rng.choice([0, 3, 10, 100])
        h0 = rng.randrange(1, 40)
        eb = rng.choice([0, 5])
        _both(q, t, mat, od, ed, oi, ei, w, eb, zd, h0)


def test_rowpar_homologous():
    rng = random.Random(11)
    mat = bwa_mat()
    for _ in range(40):
        ql = rng.randrange(20, 70)
        q = [rng.randrange(4) for _ in range(ql)]
        t = list(q)
        for _ in range(rng.randrange(0, 4)):
            p = rng.randrange(len(t))
            t[p] = (t[p] + 1) % 4
        if rng.random() < 0.5:
            p = rng.randrange(len(t))
            del t[p:p + rng.randrange(1, 4)]
        t += [rng.randrange(4) for _ in range(30)]
        _both(q, t, mat, 6, 1, 6, 1, rng.choice([5, 20, 100]), 5, 100, rng.randrange(19, 61))


# ---- invariants on seeded workload samples ---------------------------------------

@pytest.mark.parametrize("name", ["cpu_ref_151bp_4k", "illumina_150bp_1m", "uneven_mix_4m"])
def test_invariants(name):
    from ferrousalign_amd import synth
    ps = synth.workload(name, n=2048)
    sw = synth.default_scoring()
    out, cells, rc = O.extend_batch(sw, *ps.flat())
    assert (rc == 0).all()
    score, qle, tle, gtle, gscore, max_off = out.T
    assert (score >= ps.h0).all()
    assert ((qle >= 0) & (qle <= ps.qlen)).all() and ((tle >= 0) & (tle <= ps.tlen)).all()
    assert ((gtle == 0) == (gscore == -1)).all()
    assert (max_off >= 0).all()
    assert (cells > 0).all()
    # order independence: reversed batch gives the reversed results
    rev = ps.subset(np.arange(ps.n)[::-1])
    out2, cells2, _ = O.extend_batch(sw, *rev.flat())
    assert (out2[::-1] == out).all() and (cells2[::-1] == cells).all()
    # thread count independence
    out3, cells3, _ = O.extend_batch(sw, *ps.flat(), nthreads=1)
    assert (out3 == out).all() and (cells3 == cells).all()


def test_rows_match_rowpar_and_bounds():
    """Rows of the row loop (what a wavefront's run time follows): the C
    oracle against the independent row-parallel restatement on random pairs,
    plus closed forms -- tlen = 0 runs no row; an exact match of a query
    against itself plus a random tail runs past the query end; a random
    non-homologous target breaks early."""
    from ferrousalign_amd import synth
    sw = synth.default_scoring()
    rng = np.random.default_rng(3)
    qs, ts, ql, tl, h0, w = [], [], [], [], [], []
    for _ in range(300):
        q = rng.integers(0, 4, int(rng.integers(1, 40))).astype(np.uint8)
        t = rng.integers(0, 4, int(rng.integers(0, 60))).astype(np.uint8)
        if rng.random() < 0.5 and len(t) >= len(q):
            t[:len(q)] = q
        qs.append(q), ts.append(t), ql.append(len(q)), tl.append(len(t))
        h0.append(int(rng.integers(1, 40))), w.append(int(rng.integers(0, 60)))
    qo = np.concatenate([[0], np.cumsum(ql)[:-1]]).astype(np.int64)
    to = np.concatenate([[0], np.cumsum(tl)[:-1]]).astype(np.int64)
    rows, cells = O.extend_rows(sw, np.concatenate(qs), qo, ql, np.concatenate(ts + [np.zeros(1, np.uint8)]), to, tl,
                                h0, w)
    for k in range(len(ql)):
        _, c, r = rowpar.extend(list(qs[k]), list(ts[k]), list(sw.mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins,
                                w[k], sw.end_bonus, sw.zdrop, h0[k], with_rows=True)
        assert (int(rows[k]), int(cells[k])) == (r, c), k
        assert 0 <= rows[k] <= tl[k] and (tl[k] == 0) == (rows[k] == 0)
    # exact match, then a long random tail: the loop runs beyond the query
    q = rng.integers(0, 4, 50).astype(np.uint8)
    t = np.concatenate([q, rng.integers(0, 4, 200).astype(np.uint8)])
    r, _ = O.extend_rows(sw, q, [0], [50], t, [0], [250], [20], [100])
    assert 50 < r[0] <= 250
    # a random target against a short seed score: breaks within a few rows
    t2 = rng.integers(0, 4, 250).astype(np.uint8)
    r2, _ = O.extend_rows(sw, q, [0], [50], t2, [0], [250], [5], [100])
    assert r2[0] < 50
