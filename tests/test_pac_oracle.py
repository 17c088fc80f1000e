"""The .pac target-segment oracle (oracle/pac_oracle.c, restating
src/pipelines/linear/index/bntseq.rs:451-522, :559-565, :627-651 and the
left-job reversal of batch_extension/collect_soa.rs:156-176), pinned by what
the layout and the strand arithmetic fix: a forward segment is a slice of the
sequence that was packed, a reverse-strand segment is the reverse complement
of the matching forward slice, a left job reads its segment backwards, and a
segment across the strand boundary is empty.  Also the host packer used by the
tests and a hand-packed byte."""
import numpy as np
import pytest

from ferrousalign_amd import pack_pac
from tests import _oracle as O


def test_pack_pac_byte_layout():
    # A C G T -> 00 01 10 11, first base in the top bits (bntseq.rs:559-575)
    assert pack_pac(np.array([0, 1, 2, 3], np.uint8)).tolist() == [0b00011011]
    assert pack_pac(np.array([3, 0, 0, 0, 1], np.uint8)).tolist() == [0b11000000, 0b01000000]
    with pytest.raises(ValueError):
        pack_pac(np.array([4], np.uint8))


def test_known_answers():
    s = np.array([0, 0, 1, 2], np.uint8)  # AACG; reverse complement CGTT
    pac = pack_pac(s)
    assert O.pac_segment(pac, 4, 0, 4).tolist() == [0, 0, 1, 2]
    assert O.pac_segment(pac, 4, 4, 4).tolist() == [1, 2, 3, 3]
    assert O.pac_segment(pac, 4, 4, 4, reverse=True).tolist() == [3, 3, 2, 1]
    assert O.pac_segment(pac, 4, 5, 2).tolist() == [2, 3]


@pytest.mark.parametrize("seed", range(3))
def test_segments_are_slices_and_reverse_complements(seed):
    rng = np.random.default_rng(seed)
    l_pac = int(rng.integers(1000, 5000))
    seq = rng.integers(0, 4, l_pac).astype(np.uint8)
    pac = pack_pac(seq)
    for _ in range(300):
        length = int(rng.integers(0, 400))
        fwd = bool(rng.integers(0, 2))
        rev = bool(rng.integers(0, 2))
        if fwd:
            start = int(rng.integers(0, l_pac - length + 1))
            want = seq[start:start + length]
        else:
            start = int(rng.integers(l_pac, 2 * l_pac - length + 1))
            # coordinate p >= l_pac is the complement of forward base 2 l_pac - 1 - p
            want = 3 - seq[2 * l_pac - start - length:2 * l_pac - start][::-1]
        if rev:
            want = want[::-1]
        got = O.pac_segment(pac, l_pac, start, length, rev)
        assert got is not None and np.array_equal(got, want), (start, length, fwd, rev)


def test_bridging_and_out_of_range_are_empty():
    seq = np.random.default_rng(9).integers(0, 4, 1001).astype(np.uint8)
    pac = pack_pac(seq)
    assert O.pac_segment(pac, 1001, 990, 20) is None        # across l_pac
    assert O.pac_segment(pac, 1001, 2000, 5) is None        # past 2 l_pac
    assert O.pac_segment(pac, 1001, -1, 5) is None
    assert O.pac_segment(pac, 1001, 981, 20).size == 20     # ends exactly at l_pac
    assert O.pac_segment(pac, 1001, 1001, 20).size == 20    # starts exactly at l_pac
