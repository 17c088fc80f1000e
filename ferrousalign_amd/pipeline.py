"""The callers either side of the extension batch, over the C ABI.

* ``collect_extension_jobs`` -- the job producer
  (``collect_extension_jobs_batch_soa``, src/pipelines/linear/batch_extension/
  collect_soa.rs:28-228): per kept chain a reference window from the seeds'
  ``cal_max_gap`` margins (chaining.rs:1108-1116), then per seed (reverse
  order) a left job (query prefix reversed, window prefix read backwards,
  h0 = seed_len * a) and a right job (query suffix, window suffix).
* ``merge_scores_to_regions`` -- the consumer (region.rs:737-890): one region
  per seed from its left/right scores and the pen_clip5 / pen_clip3 clipping
  decisions, seed coverage, and the chain's first highest-scoring region.
* ``extend_chains`` / ``DeviceChains`` -- the whole path on the GPU: the
  producer, the packing of queries and of targets cut from the resident .pac,
  the left and the right extension batch and the merge (host buffers in and
  out, or inputs and regions resident in HBM).  The right jobs start from the
  seed's left-extended score, as BWA-MEM2's mem_chain2aln chains them: the
  reference gives them h0 = 0 (collect_soa.rs:206), outside ksw_extend2's
  h0 > 0 domain (ksw_affine_gap.rs:968), and adds the two scores instead.

Chains and seeds are SoA arrays (``ChainBatch``), as the reference's
SoAChainBatch / SoASeedBatch: read r owns chains
[read_chain_start[r], +read_chain_count[r]), chain c owns the seed indices
chain_seeds[chain_seed_start[c] .. +chain_seed_count[c]) into ``seeds``.
"""
import ctypes
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _native as N
from .banded_swa import BandedPairWiseSW, DeviceReference, default_context


@dataclass
class MemOpt:
    """The MemOpt fields the path reads (src/pipelines/linear/mem_opt.rs)."""
    a: int = 1
    b: int = 4
    o_del: int = 6
    e_del: int = 1
    o_ins: int = 6
    e_ins: int = 1
    w: int = 100
    zdrop: int = 100
    pen_clip5: int = 5
    pen_clip3: int = 5

    def to_c(self):
        x = np.zeros(1, N.MEM_OPT_DTYPE)
        for f in N.MEM_OPT_DTYPE.names:
            x[f] = getattr(self, f)
        return x


@dataclass
class ChainBatch:
    query_len: np.ndarray         # [nreads] int32
    read_chain_start: np.ndarray  # [nreads] int64
    read_chain_count: np.ndarray  # [nreads] int32
    chain_kept: np.ndarray        # [nchains] uint8
    chain_frac_rep: np.ndarray    # [nchains] float32
    chain_seed_start: np.ndarray  # [nchains] int64
    chain_seed_count: np.ndarray  # [nchains] int32
    chain_seeds: np.ndarray       # [sum of seed counts] int64, indices into seeds
    seeds: np.ndarray             # [nseeds] SEED_DTYPE

    @property
    def nreads(self):
        return len(self.query_len)

    @property
    def nchains(self):
        return len(self.chain_kept)

    def normalised(self):
        return ChainBatch(np.ascontiguousarray(self.query_len, np.int32),
                          np.ascontiguousarray(self.read_chain_start, np.int64),
                          np.ascontiguousarray(self.read_chain_count, np.int32),
                          np.ascontiguousarray(self.chain_kept, np.uint8),
                          np.ascontiguousarray(self.chain_frac_rep, np.float32),
                          np.ascontiguousarray(self.chain_seed_start, np.int64),
                          np.ascontiguousarray(self.chain_seed_count, np.int32),
                          np.ascontiguousarray(self.chain_seeds, np.int64),
                          np.ascontiguousarray(self.seeds, N.SEED_DTYPE))


@dataclass
class Mappings:
    """Per chain: seed mappings [start[c], +count[c]) (seed, left job, right
    job; -1 = none) and the reference window rmax[c] = (rmax0, rmax1)."""
    start: np.ndarray
    count: np.ndarray
    seed: np.ndarray
    left: np.ndarray
    right: np.ndarray
    rmax: np.ndarray


def collect_extension_jobs(opt: MemOpt, l_pac: int, cb: ChainBatch):
    """Left jobs, right jobs (JOB_DTYPE records) and the mappings."""
    cb = cb.normalised()
    cap = int(cb.chain_seed_count.astype(np.int64).sum())
    left = np.zeros(cap + 1, N.JOB_DTYPE)
    right = np.zeros(cap + 1, N.JOB_DTYPE)
    nc = cb.nchains
    m = Mappings(np.zeros(nc + 1, np.int64), np.zeros(nc + 1, np.int32), np.zeros(cap + 1, np.int64),
                 np.zeros(cap + 1, np.int64), np.zeros(cap + 1, np.int64), np.zeros((nc + 1, 2), np.int64))
    nl, nr = ctypes.c_int64(), ctypes.c_int64()
    o = opt.to_c()
    N.check(N.lib().fa_collect_extension_jobs(
        N.ptr(o), int(l_pac), cb.nreads, N.ptr(cb.query_len), N.ptr(cb.read_chain_start),
        N.ptr(cb.read_chain_count), N.ptr(cb.chain_kept), N.ptr(cb.chain_seed_start), N.ptr(cb.chain_seed_count),
        N.ptr(cb.chain_seeds), N.ptr(cb.seeds), cap, N.ptr(left), ctypes.byref(nl), N.ptr(right),
        ctypes.byref(nr), N.ptr(m.start), N.ptr(m.count), N.ptr(m.seed), N.ptr(m.left), N.ptr(m.right),
        N.ptr(m.rmax)), "fa_collect_extension_jobs")
    return left[:nl.value], right[:nr.value], m


def merge_scores_to_regions(opt: MemOpt, cb: ChainBatch, m: Mappings, left_scores, right_scores,
                            right_from_left=True):
    """One REGION_DTYPE record per chain (seed_idx -1: no region).  Scores
    are [n, 6] int32 in OutScore field order."""
    cb = cb.normalised()
    ls = np.ascontiguousarray(left_scores, np.int32).reshape(-1, 6)
    rs = np.ascontiguousarray(right_scores, np.int32).reshape(-1, 6)
    ls = ls if ls.size else np.zeros((1, 6), np.int32)
    rs = rs if rs.size else np.zeros((1, 6), np.int32)
    regions = np.zeros(cb.nchains + 1, N.REGION_DTYPE)
    o = opt.to_c()
    N.check(N.lib().fa_merge_scores_to_regions(
        N.ptr(o), cb.nreads, N.ptr(cb.query_len), N.ptr(cb.read_chain_start), N.ptr(cb.read_chain_count),
        N.ptr(cb.chain_frac_rep), N.ptr(cb.chain_seed_start), N.ptr(cb.chain_seed_count), N.ptr(cb.chain_seeds),
        N.ptr(cb.seeds), N.ptr(m.start), N.ptr(m.count), N.ptr(m.seed), N.ptr(m.left), N.ptr(m.right),
        N.ptr(m.rmax), N.ptr(ls), N.ptr(rs), 1 if right_from_left else 0, N.ptr(regions)),
        "fa_merge_scores_to_regions")
    return regions[:cb.nchains]


CHAIN_STAGES = ("upload", "producer", "left_prep", "left_run", "right_prep", "right_run", "merge", "download")


def _chain_args(query_seqs, query_offset, cb: ChainBatch):
    q = np.ascontiguousarray(query_seqs, np.uint8)
    qo = np.ascontiguousarray(query_offset, np.int64)
    keep = (q, qo, cb)
    args = [cb.nreads, N.ptr(q), q.size, N.ptr(qo), N.ptr(cb.query_len), N.ptr(cb.read_chain_start),
            N.ptr(cb.read_chain_count), cb.nchains, N.ptr(cb.chain_kept), N.ptr(cb.chain_frac_rep),
            N.ptr(cb.chain_seed_start), N.ptr(cb.chain_seed_count), len(cb.chain_seeds), N.ptr(cb.chain_seeds),
            len(cb.seeds), N.ptr(cb.seeds)]
    return args, keep


def extend_chains(ref: DeviceReference, sw: BandedPairWiseSW, opt: MemOpt, query_seqs, query_offset,
                  cb: ChainBatch, ctx=None, stage_ms=None):
    """Producer, both extension batches and the merge on the GPU over host
    buffers (fa_extend_chains): (regions, number of left jobs, number of
    right jobs).  stage_ms: an optional dict that receives the time of each
    stage (CHAIN_STAGES) in ms."""
    ctx = ctx or ref.ctx or default_context()
    cb = cb.normalised()
    args, keep = _chain_args(query_seqs, query_offset, cb)
    regions = np.zeros(cb.nchains + 1, N.REGION_DTYPE)
    nl, nr = ctypes.c_int64(), ctypes.c_int64()
    o = opt.to_c()
    c = sw.to_c()
    st = np.zeros(len(CHAIN_STAGES), np.float64)
    N.check(N.lib().fa_extend_chains(ctx.handle, ref.handle, ctypes.byref(c), N.ptr(o), *args, N.ptr(regions),
                                     ctypes.byref(nl), ctypes.byref(nr), N.ptr(st)), "fa_extend_chains")
    if stage_ms is not None:
        stage_ms.update(zip(CHAIN_STAGES, map(float, st)))
    return regions[:cb.nchains], nl.value, nr.value


class DeviceChains:
    """Reads, chains and seeds resident in HBM (fa_chains_upload); extend()
    runs the whole path on the device, download() returns the regions."""

    def __init__(self, query_seqs, query_offset, cb: ChainBatch, ctx=None):
        self.ctx = ctx or default_context()
        self.cb = cb.normalised()
        args, keep = _chain_args(query_seqs, query_offset, self.cb)
        self._h = ctypes.c_void_p()
        N.check(N.lib().fa_chains_upload(self.ctx.handle, *args, ctypes.byref(self._h)), "fa_chains_upload")

    def extend(self, ref: DeviceReference, sw: BandedPairWiseSW, opt: MemOpt, stage_ms=None):
        """(left jobs, right jobs); stage_ms as in extend_chains (device stages)."""
        nl, nr = ctypes.c_int64(), ctypes.c_int64()
        st = np.zeros(len(CHAIN_STAGES), np.float64)
        c = sw.to_c()
        o = opt.to_c()
        N.check(N.lib().fa_chains_extend(self.ctx.handle, self._h, ref.handle, ctypes.byref(c), N.ptr(o),
                                         ctypes.byref(nl), ctypes.byref(nr), N.ptr(st)), "fa_chains_extend")
        if stage_ms is not None:
            stage_ms.update(zip(CHAIN_STAGES, map(float, st)))
        return nl.value, nr.value

    def download(self):
        regions = np.zeros(self.cb.nchains + 1, N.REGION_DTYPE)
        N.check(N.lib().fa_chains_download(self.ctx.handle, self._h, N.ptr(regions)), "fa_chains_download")
        return regions[:self.cb.nchains]

    def free(self):
        if self._h:
            N.lib().fa_chains_free(self.ctx.handle, self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
