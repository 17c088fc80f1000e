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
