"""CPU oracle of the mate-rescue local SW (ksw_align2, 16-bit mode): closed
forms, edge cases and agreement with an independent numpy emulation of the
8-lane striped execution."""
import random

import numpy as np
import pytest

from tests import _oracle as O
from tests import kswalign_np as K

MATE = [1, -4, -4, -4, 0, -4, 1, -4, -4, 0, -4, -4, 1, -4, 0, -4, -4, -4, 1, 0, 0, 0, 0, 0, 0]
XTRA = O.KSW_XSUBO | O.KSW_XSTART | 20


def test_exact_embedded_match_closed_form():
    rng = random.Random(1)
    for qlen, s in ((151, 400), (40, 0), (64, 900), (23, 5)):
        t = [rng.randrange(4) for _ in range(1500)]
        q = t[s:s + qlen]
        rc, r, cells = O.ksw_align2_i16(q, t, MATE, 6, 1, 6, 1, XTRA)
        assert rc == 0
        score, te, qe, score2, te2, tb, qb = r
        assert (score, te, qe, tb, qb) == (qlen, s + qlen - 1, qlen - 1, s, 0)
        assert score2 < 20


def test_empty_target():
    rc, r, cells = O.ksw_align2_i16([0, 1, 2], [], MATE, 6, 1, 6, 1, XTRA)
    # no rows: score 0, te -1, qe found at position 0 of the all-zero Hmax
    assert rc == 0 and r == (0, -1, 0, -1, -1, -1, -1) and cells == 0


def test_rejects_byte_mode_and_bad_codes():
    assert O.ksw_align2_i16([0, 1], [0, 1], MATE, 6, 1, 6, 1, XTRA | O.KSW_XBYTE)[0] == -1
    assert O.ksw_align2_i16([0, 9], [0, 1], MATE, 6, 1, 6, 1, XTRA)[0] == -1
    assert O.ksw_align2_i16([], [0, 1], MATE, 6, 1, 6, 1, XTRA)[0] == -1


@pytest.mark.parametrize("seed", range(4))
def test_matches_numpy_emulation(seed):
    rng = random.Random(seed)
    for _ in range(12):
        ql = rng.randrange(1, 45)
        tl = rng.randrange(0, 120)
        alpha = rng.choice([2, 4, 5])
        q = [rng.randrange(alpha) for _ in range(ql)]
        t = [rng.randrange(alpha) for _ in range(tl)]
        if tl > ql and rng.random() < 0.6:  # plant a mutated copy of the query
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                t[s + k] = q[k] if rng.random() < 0.9 else rng.randrange(4)
        xtra = rng.choice([XTRA, O.KSW_XSTART, O.KSW_XSUBO | 5, O.KSW_XSUBO | O.KSW_XSTART | 3, 0])
        gaps = rng.choice([(6, 1, 6, 1), (5, 2, 3, 1), (0, 1, 0, 1)])
        rc, r, cells = O.ksw_align2_i16(q, t, MATE, *gaps, xtra)
        assert rc == 0
        r2, cells2 = K.ksw_align2_i16(q, t, MATE, *gaps, xtra)
        assert r == r2 and cells == cells2, (q, t, xtra, gaps, r, r2)


def test_matches_numpy_on_mate_workload():
    from ferrousalign_amd import synth
    ps = synth.mate_pairs(6)
    for k in range(6):
        q = ps.q[k, :151]
        t = ps.t[k, :1500]
        rc, r, cells = O.ksw_align2_i16(q, t, MATE, 6, 1, 6, 1, XTRA)
        r2, cells2 = K.ksw_align2_i16(q, t, MATE, 6, 1, 6, 1, XTRA)
        assert r == r2 and cells == cells2


# ---- 8-bit mode (KSW_XBYTE, ksw_u8_impl :267-494) ---------------------------------

XBYTE = XTRA | O.KSW_XBYTE


def _textbook_sw(q, t, mat, o_del, e_del, o_ins, e_ins, gaps=True):
    """Gotoh local alignment score, full matrix, row by row (E along target rows,
    F along query columns, H clamped at 0); gaps=False: best ungapped run."""
    q = np.asarray(q)
    ql = len(q)
    M = np.asarray(mat, np.int64).reshape(5, 5)
    Hp = np.zeros(ql, np.int64)
    E = np.full(ql, -(1 << 30), np.int64)
    best = 0
    ks = np.arange(ql)
    for tb in t:
        d = np.concatenate([[0], Hp[:-1]]) + M[tb][q]
        if not gaps:
            H = np.maximum(d, 0)
        else:
            E = np.maximum(E - e_del, Hp - (o_del + e_del))
            Hx = np.maximum(np.maximum(d, E), 0)
            # F(j) = max_{k<j} Hx(k) - (o_ins + e_ins) - e_ins * (j - 1 - k)
            acc = np.maximum.accumulate(Hx + e_ins * ks)
            F = np.concatenate([[-(1 << 30)], acc[:-1] - (o_ins + e_ins) - e_ins * (ks[1:] - 1)])
            H = np.maximum(Hx, F)
        best = max(best, int(H.max()) if ql else 0)
        Hp = H
    return best


def test_u8_exact_embedded_match_closed_form():
    rng = random.Random(11)
    for qlen, s in ((151, 400), (40, 0), (200, 900), (23, 5), (250, 10)):
        t = [rng.randrange(4) for _ in range(1500)]
        q = t[s:s + qlen]
        rc, r, cells = O.ksw_align2(q, t, MATE, 6, 1, 6, 1, XBYTE)
        assert rc == 0
        score, te, qe, score2, te2, tb, qb = r
        assert (score, te, qe, tb, qb) == (qlen, s + qlen - 1, qlen - 1, s, 0)
        assert score2 < 20


def test_u8_overflow_closed_form():
    # shift = 256 - (-4) = 4 (as u8): a best score of 251 stops the pass and reports
    # 255, with no qe, so no reverse pass; te is the row where the score reached 251
    rng = random.Random(12)
    for qlen, s in ((251, 300), (256, 0), (400, 77)):
        t = [rng.randrange(4) for _ in range(1500)]
        q = t[s:s + qlen]
        rc, r, cells = O.ksw_align2(q, t, MATE, 6, 1, 6, 1, XBYTE)
        assert rc == 0
        assert r == (255, s + 250, -1, -1, -1, -1, -1)
        assert cells == (s + 251) * qlen
    # the 16-bit mode scores the same pair in full
    t = [rng.randrange(4) for _ in range(600)]
    rc, r, _ = O.ksw_align2(t[50:350], t, MATE, 6, 1, 6, 1, XTRA)
    assert r[:3] == (300, 349, 299)


def test_u8_bounded_by_textbook_and_equal_when_ungapped_optimal():
    # every H the striped pass computes is the score of a real path (so <= the
    # textbook optimum); when the optimum is an ungapped run the pass reaches it,
    # since each cell is at least its diagonal predecessor plus the score
    rng = random.Random(13)
    n_eq = n_all = 0
    for _ in range(60):
        ql = rng.randrange(1, 90)
        tl = rng.randrange(1, 160)
        q = [rng.randrange(5) for _ in range(ql)]
        t = [rng.randrange(5) for _ in range(tl)]
        if tl > ql and rng.random() < 0.7:
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                if rng.random() < 0.9:
                    t[s + k] = q[k]
            if rng.random() < 0.5:  # an indel inside the copy
                c = rng.randrange(1, ql) if ql > 1 else 0
                del t[s + c:s + c + rng.randrange(1, 4)]
        gaps = rng.choice([(6, 1, 6, 1), (5, 2, 3, 1), (1, 1, 1, 1)])
        rc, r, _ = O.ksw_align2(q, t, MATE, *gaps, XBYTE)
        assert rc == 0
        full = _textbook_sw(q, t, MATE, *gaps)
        ungapped = _textbook_sw(q, t, MATE, *gaps, gaps=False)
        assert r[0] <= full
        if ungapped == full:
            assert r[0] == full
        n_all += 1
        n_eq += r[0] == full
    assert n_eq >= 0.9 * n_all


def test_u8_empty_target_and_bad_input():
    rc, r, cells = O.ksw_align2([0, 1, 2], [], MATE, 6, 1, 6, 1, XBYTE)
    assert rc == 0 and r == (0, -1, 0, -1, -1, -1, -1) and cells == 0
    assert O.ksw_align2([0, 9], [0, 1], MATE, 6, 1, 6, 1, XBYTE)[0] == -1
    assert O.ksw_align2([], [0, 1], MATE, 6, 1, 6, 1, XBYTE)[0] == -1


@pytest.mark.parametrize("seed", range(4))
def test_u8_matches_numpy_emulation(seed):
    rng = random.Random(100 + seed)
    for _ in range(12):
        ql = rng.randrange(1, 70)
        tl = rng.randrange(0, 120)
        alpha = rng.choice([2, 4, 5])
        q = [rng.randrange(alpha) for _ in range(ql)]
        t = [rng.randrange(alpha) for _ in range(tl)]
        if tl > ql and rng.random() < 0.6:
            s = rng.randrange(0, tl - ql + 1)
            for k in range(ql):
                t[s + k] = q[k] if rng.random() < 0.9 else rng.randrange(4)
        xtra = rng.choice([XTRA, O.KSW_XSTART, O.KSW_XSUBO | 5, O.KSW_XSUBO | O.KSW_XSTART | 3, 0])
        gaps = rng.choice([(6, 1, 6, 1), (5, 2, 3, 1), (0, 1, 0, 1)])
        mat = MATE if rng.random() < 0.7 else [3 if a == b else -2 for a in range(5) for b in range(5)]
        rc, r, cells = O.ksw_align2(q, t, mat, *gaps, xtra | O.KSW_XBYTE)
        assert rc == 0
        r2, cells2 = K.ksw_align2(q, t, mat, *gaps, xtra | K.XBYTE)
        assert r == r2 and cells == cells2, (q, t, xtra, gaps, r, r2)


def test_u8_overflow_matches_numpy_with_a_larger_match_score():
    # match 3 against mismatch -2: shift 2, overflow at 253, reached inside a 90 bp copy
    mat = [3 if a == b else -2 for a in range(5) for b in range(5)]
    rng = random.Random(14)
    for ql in (60, 84, 85, 90):
        t = [rng.randrange(4) for _ in range(300)]
        q = t[100:100 + ql]
        rc, r, cells = O.ksw_align2(q, t, mat, 6, 1, 6, 1, XBYTE)
        r2, cells2 = K.ksw_align2(q, t, mat, 6, 1, 6, 1, XBYTE)
        assert rc == 0 and r == r2 and cells == cells2
        if 3 * ql >= 253:
            assert r[0] == 255 and r[2] == -1 and r[1] == 100 + (253 + 2) // 3 - 1
        else:
            assert r[:3] == (3 * ql, 100 + ql - 1, ql - 1)


def test_u8_matches_numpy_on_mate_workload():
    from ferrousalign_amd import synth
    ps = synth.mate_pairs(4)
    for k in range(4):
        q = ps.q[k, :151]
        t = ps.t[k, :1500]
        rc, r, cells = O.ksw_align2(q, t, MATE, 6, 1, 6, 1, XBYTE)
        r2, cells2 = K.ksw_align2(q, t, MATE, 6, 1, 6, 1, XBYTE)
        assert rc == 0 and r == r2 and cells == cells2


@pytest.mark.parametrize("byte_mode", [False, True])
def test_score2_two_copies_closed_form(byte_mode):
    # an exact copy of the query and, far away, a copy with 4 mismatches spaced 30
    # apart (score qlen - 5 x 4): the best is the exact copy, the second best is the
    # other copy's peak row, outside te +- ceil(score / max); each copy's run of
    # rows with a row maximum >= minsc is one entry of the suboptimal list
    rng = random.Random(31)
    qlen = 151
    xtra = XTRA | (O.KSW_XBYTE if byte_mode else 0)
    for s1, s2, first_exact in ((100, 700, True), (120, 900, False), (1200, 300, True)):
        t = [rng.randrange(4) for _ in range(1500)]
        q = [rng.randrange(4) for _ in range(qlen)]
        mut = list(q)
        for k in (30, 60, 90, 120):
            mut[k] = (q[k] + 1) % 4
        a, b = (q, mut) if first_exact else (mut, q)
        t[s1:s1 + qlen] = a
        t[s2:s2 + qlen] = b
        se = s1 if first_exact else s2  # the exact copy
        so = s2 if first_exact else s1
        rc, r, _ = O.ksw_align2(q, t, MATE, 6, 1, 6, 1, xtra)
        assert rc == 0
        assert (r[0], r[1], r[2]) == (qlen, se + qlen - 1, qlen - 1)
        assert (r[3], r[4]) == (qlen - 20, so + qlen - 1), r
        assert (r[5], r[6]) == (se, 0)
        r2, _ = K.ksw_align2(q, t, MATE, 6, 1, 6, 1, xtra)
        assert r2 == r
