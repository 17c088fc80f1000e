"""Seeded synthetic extension-pair workloads (wrapper of libfa_synth.so).

Scoring for every workload: A=1, B=4, O=6, E=1 (deletion = insertion),
w=100 (w=200 for the wide 10% of the uneven mix), zdrop=100, end_bonus=5;
matrix per bwa_fill_scmat (src/core/alignment/banded_swa/scoring.rs:2-27).
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .banded_swa import BandedPairWiseSW, bwa_fill_scmat

REF_LEN = 1 << 20
REF_SEED = 0xFA00

QSTRIDE = 256
TSTRIDE = 352

# name -> (generator kind, pair count, seed)
WORKLOADS = {
    "cpu_ref_151bp_4k": (0, 4096, 0xFA00),          # 151 bp reads, 1% subs: 4,096 pairs
    "illumina_150bp_1m": (1, 1048576, 0xFA01),      # 150 bp reads, 0.2%->2% subs: 1,048,576 pairs
    "uneven_mix_4m": (2, 4194304, 0xFA02),          # {100,151,250} bp, 5% random targets, 10% w=200
    "uneven_mix_train_16m": (2, 16777216, 0xFA03),  # predictor training pairs, same model, other seed
}

MATE_SEED = 0xFA04
MATE_PAIRS = 2097152
MATE_QLEN = 151
MATE_TLEN = 1500
# Mate-rescue scoring (src/pipelines/linear/paired/mate_rescue.rs:116-122): match 1,
# mismatch -4, N scores 0; gaps 6/1 (:642-645); xtra = KSW_XSUBO|KSW_XSTART|20.
MATE_MATRIX = np.array([1, -4, -4, -4, 0, -4, 1, -4, -4, 0, -4, -4, 1, -4, 0, -4, -4, -4, 1, 0,
                        0, 0, 0, 0, 0], np.int8)
MATE_XTRA = 0x40000 | 0x80000 | 20


def default_scoring():
    return BandedPairWiseSW(6, 1, 6, 1, 100, 5, 5, 5, bwa_fill_scmat(1, 4, -1))


_ref_cache = {}


def reference(seed=REF_SEED, length=REF_LEN):
    key = (seed, length)
    if key not in _ref_cache:
        ref = np.zeros(length, np.uint8)
        N.synth().fa_synth_reference(ctypes.c_uint64(seed), length, N.ptr(ref))
        _ref_cache[key] = ref
    return _ref_cache[key]


@dataclass
class PairSet:
    q: np.ndarray      # [n, QSTRIDE] u8 codes
    t: np.ndarray      # [n, TSTRIDE]
    qlen: np.ndarray
    tlen: np.ndarray
    h0: np.ndarray
    w: np.ndarray
    tstart: np.ndarray = None  # target k = reference()[tstart[k] : tstart[k] + tlen[k]]

    @property
    def n(self):
        return len(self.qlen)

    def flat(self):
        """(qseq, qoff, qlen, tseq, toff, tlen, h0, w) in ExtensionJobBatch layout."""
        n = self.n
        qoff = np.arange(n, dtype=np.int64) * self.q.shape[1]
        toff = np.arange(n, dtype=np.int64) * self.t.shape[1]
        return (self.q.reshape(-1), qoff, self.qlen, self.t.reshape(-1), toff, self.tlen, self.h0, self.w)

    def subset(self, idx):
        return PairSet(self.q[idx], self.t[idx], self.qlen[idx], self.tlen[idx], self.h0[idx], self.w[idx],
                       None if self.tstart is None else self.tstart[idx])


def extension_pairs(kind, n, seed, w1=100, w2=200, threads=0):
    ref = reference()
    q = np.zeros((n, QSTRIDE), np.uint8)
    t = np.zeros((n, TSTRIDE), np.uint8)
    ql = np.zeros(n, np.int32)
    tl = np.zeros(n, np.int32)
    h0 = np.zeros(n, np.int32)
    w = np.zeros(n, np.int32)
    ts = np.zeros(n, np.int64)
    rc = N.synth().fa_synth_extension_pairs(kind, ctypes.c_uint64(seed), N.ptr(ref), ref.size, n, N.ptr(q),
                                            QSTRIDE, N.ptr(ql), N.ptr(t), TSTRIDE, N.ptr(tl), N.ptr(h0), N.ptr(w),
                                            1, w1, w2, threads, N.ptr(ts))
    if rc != 0:
        raise ValueError("fa_synth_extension_pairs failed")
    return PairSet(q, t, ql, tl, h0, w, ts)


def workload(name, n=None, threads=0):
    kind, count, seed = WORKLOADS[name]
    return extension_pairs(kind, count if n is None else n, seed, threads=threads)


@dataclass
class MatePairSet:
    q: np.ndarray          # [n, 151]
    t: np.ndarray          # [n, 1500]
    has_mate: np.ndarray

    @property
    def n(self):
        return len(self.has_mate)

    def flat(self):
        n = self.n
        return (self.q.reshape(-1), np.arange(n, dtype=np.int64) * MATE_QLEN, np.full(n, MATE_QLEN, np.int32),
                self.t.reshape(-1), np.arange(n, dtype=np.int64) * MATE_TLEN, np.full(n, MATE_TLEN, np.int32))

    def subset(self, idx):
        return MatePairSet(self.q[idx], self.t[idx], self.has_mate[idx])


def mate_pairs(n=MATE_PAIRS, seed=MATE_SEED, threads=0):
    """Mate-rescue workload: 151 bp queries against 1,500 bp windows."""
    ref = reference()
    q = np.zeros((n, MATE_QLEN), np.uint8)
    t = np.zeros((n, MATE_TLEN), np.uint8)
    hm = np.zeros(n, np.uint8)
    rc = N.synth().fa_synth_mate_pairs(ctypes.c_uint64(seed), N.ptr(ref), ref.size, n, N.ptr(q), N.ptr(t),
                                       N.ptr(hm), threads)
    if rc != 0:
        raise ValueError("fa_synth_mate_pairs failed")
    return MatePairSet(q, t, hm)


REGION_SEED = 0xFA05
REGION_PAIRS = 1048576
REGION_QSTRIDE = 150
REGION_TSTRIDE = 180


def region_pairs(n=REGION_PAIRS, seed=REGION_SEED, threads=0):
    """CIGAR-regeneration workload: 150 bp reads against the reference span of
    their alignment region, band |tlen - qlen| + 10.  Returns the flat arrays
    (qseq, qoff, qlen, tseq, toff, tlen, w)."""
    ref = reference()
    q = np.zeros((n, REGION_QSTRIDE), np.uint8)
    t = np.zeros((n, REGION_TSTRIDE), np.uint8)
    ql = np.zeros(n, np.int32)
    tl = np.zeros(n, np.int32)
    w = np.zeros(n, np.int32)
    rc = N.synth().fa_synth_region_pairs(ctypes.c_uint64(seed), N.ptr(ref), ref.size, n, N.ptr(q), REGION_QSTRIDE,
                                         N.ptr(ql), N.ptr(t), REGION_TSTRIDE, N.ptr(tl), N.ptr(w), threads)
    if rc != 0:
        raise ValueError("fa_synth_region_pairs failed")
    return (q.reshape(-1), np.arange(n, dtype=np.int64) * REGION_QSTRIDE, ql,
            t.reshape(-1), np.arange(n, dtype=np.int64) * REGION_TSTRIDE, tl, w)
