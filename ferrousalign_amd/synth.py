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


CHAIN_SEED = 0xFA06
CHAIN_READS = 1048576
CHAIN_READ_LEN = 150
CHAIN_KMIN = 19
CHAIN_MAX_SEEDS = 8


def chain_reads(n=CHAIN_READS, seed=CHAIN_SEED, threads=0):
    """Chain-extension workload: n reads of 150 bases drawn from both strands
    of the synthetic reference (substitutions rising 0.2% -> 2%, 0.1% indels).
    Chain 0 of a read holds its seeds (maximal exact runs of >= 19 bases on
    the true alignment, at most 8); half the reads carry a second chain with
    one decoy seed at a uniform position, kept with probability 0.7.
    Returns (ChainBatch, qseq, qoff, true bidirectional position)."""
    from .pipeline import ChainBatch
    ref = reference()
    l_pac = ref.size
    R, S = CHAIN_READ_LEN, CHAIN_MAX_SEEDS
    q = np.zeros(n * R, np.uint8)
    srp = np.zeros(n * S, np.int64)
    sqp = np.zeros(n * S, np.int32)
    sln = np.zeros(n * S, np.int32)
    ns = np.zeros(n, np.int32)
    drp = np.zeros(n, np.int64)
    dqp = np.zeros(n, np.int32)
    dln = np.zeros(n, np.int32)
    tp = np.zeros(n, np.int64)
    rc = N.synth().fa_synth_chain_reads(ctypes.c_uint64(seed), N.ptr(ref), l_pac, n, R, CHAIN_KMIN, S, N.ptr(q),
                                        N.ptr(srp), N.ptr(sqp), N.ptr(sln), N.ptr(ns), N.ptr(drp), N.ptr(dqp),
                                        N.ptr(dln), N.ptr(tp), threads)
    if rc != 0:
        raise ValueError("fa_synth_chain_reads failed")
    rng = np.random.default_rng(seed)
    has0 = ns > 0
    has1 = dln > 0
    nch = has0.astype(np.int32) + has1
    rcs = np.zeros(n, np.int64)
    rcs[1:] = np.cumsum(nch)[:-1]
    nchains = int(nch.sum())
    ccount = np.zeros(nchains, np.int32)
    ccount[rcs[has0]] = ns[has0]
    ccount[rcs[has1] + has0[has1]] = 1
    cstart = np.zeros(nchains, np.int64)
    cstart[1:] = np.cumsum(ccount)[:-1]
    kept = np.ones(nchains, np.uint8)
    kept[rcs[has1] + has0[has1]] = (rng.random(int(has1.sum())) < 0.7).astype(np.uint8)
    seeds = np.zeros(int(ccount.sum()), N.SEED_DTYPE)
    # chain 0 seeds: per read, slots [0, ns) of its fixed-stride block
    slot = np.arange(n * S) % S
    m0 = slot < np.repeat(ns, S)
    first = np.repeat(cstart[rcs], S) + slot
    seeds["ref_pos"][first[m0]] = srp[m0]
    seeds["query_pos"][first[m0]] = sqp[m0]
    seeds["len"][first[m0]] = sln[m0]
    dc = cstart[rcs[has1] + has0[has1]]
    seeds["ref_pos"][dc] = drp[has1]
    seeds["query_pos"][dc] = dqp[has1]
    seeds["len"][dc] = dln[has1]
    cb = ChainBatch(np.full(n, R, np.int32), rcs, nch, kept, rng.random(nchains).astype(np.float32), cstart,
                    ccount, np.arange(len(seeds), dtype=np.int64), seeds)
    return cb, q, np.arange(n, dtype=np.int64) * R, tp
