"""Host-side mirror of the reference's banded seed-extension interface.

Names, argument meaning and result records follow the Rust crate:

* ``BandedPairWiseSW``            src/core/alignment/banded_swa/types.rs:106-199
* ``OutScore``                    src/core/alignment/banded_swa/types.rs:77-84
* ``bwa_fill_scmat``              src/core/alignment/banded_swa/scoring.rs:2-27
* ``ExtensionJobBatch`` / ``BatchedExtensionJob`` / ``BatchExtensionResult``
                                  src/pipelines/linear/batch_extension/types.rs:20-197
* ``execute_batch_simd_scoring``  src/pipelines/linear/batch_extension/dispatch.rs:31-86
* ``simd_banded_swa_batch{16,32,64}_soa`` and ``..._int16_soa``
                                  src/core/alignment/banded_swa/shared.rs:161-389

Every entry point computes the reference's ``ksw_extend2``
(src/core/alignment/ksw_affine_gap.rs:952-1169) on the GPU for every job.
Unlike the reference's SIMD dispatch, every job of a batch is scored (no
first-W truncation, dispatch.rs:126-146) and lengths are not clamped to 127
(shared.rs:69-72).  Where the reference would panic (h0 <= 0, qlen == 0,
base code >= 5) the call raises :class:`FaError` with the matching code.
"""
import ctypes
from dataclasses import dataclass, field
from enum import Enum
from typing import List, Optional

import numpy as np

from . import _native as N
from ._native import FaError

Left = "Left"
Right = "Right"


class ExtensionDirection(Enum):
    Left = 0
    Right = 1


def bwa_fill_scmat(match_score, mismatch_penalty, ambig_penalty):
    """5x5 A,C,G,T,N matrix, scoring.rs:2-27 (mat[i*5+j])."""
    mat = np.zeros(25, np.int8)
    for i in range(4):
        for j in range(4):
            mat[i * 5 + j] = match_score if i == j else -mismatch_penalty
        mat[i * 5 + 4] = ambig_penalty
    mat[20:25] = ambig_penalty
    return mat


class BandedPairWiseSW:
    """Parameter bundle, types.rs:106-199 (``new`` at :122-148)."""

    def __init__(self, o_del, e_del, o_ins, e_ins, zdrop, end_bonus, pen_clip5, pen_clip3, mat,
                 w_match=1, w_mismatch=4):
        self.m = 5
        self.o_del, self.e_del, self.o_ins, self.e_ins = int(o_del), int(e_del), int(o_ins), int(e_ins)
        self.zdrop, self.end_bonus = int(zdrop), int(end_bonus)
        self.pen_clip5, self.pen_clip3 = int(pen_clip5), int(pen_clip3)
        self.mat = np.asarray(mat, np.int8).reshape(25).copy()

    def alphabet_size(self):
        return self.m

    def scoring_matrix(self):
        return self.mat

    def to_c(self):
        c = N.BandedPairwiseSW()
        c.m = self.m
        c.end_bonus, c.zdrop = self.end_bonus, self.zdrop
        c.pen_clip5, c.pen_clip3 = self.pen_clip5, self.pen_clip3
        c.o_del, c.o_ins, c.e_del, c.e_ins = self.o_del, self.o_ins, self.e_del, self.e_ins
        for k in range(25):
            c.mat[k] = int(self.mat[k])
        return c


@dataclass
class OutScore:
    score: int
    target_end_pos: int
    gtarget_end_pos: int
    query_end_pos: int
    global_score: int
    max_offset: int


@dataclass
class BatchedExtensionJob:
    read_idx: int
    chain_idx: int
    seed_idx: int
    direction: ExtensionDirection
    query_len: int
    ref_len: int
    h0: int
    band_width: int
    query_offset: int
    ref_offset: int


@dataclass
class BatchExtensionResult:
    read_idx: int
    chain_idx: int
    seed_idx: int
    direction: ExtensionDirection
    score: int
    query_end: int
    ref_end: int
    gscore: int
    gref_end: int
    max_off: int


class ExtensionJobBatch:
    """Jobs with sequences in contiguous u8-code buffers (types.rs:67-197)."""

    def __init__(self):
        self.jobs: List[BatchedExtensionJob] = []
        self._q = bytearray()
        self._r = bytearray()

    def add_job(self, read_idx, chain_idx, seed_idx, direction, query_seq, ref_seq, h0, band_width):
        q = bytes(bytearray(query_seq))
        r = bytes(bytearray(ref_seq))
        self.jobs.append(BatchedExtensionJob(read_idx, chain_idx, seed_idx, direction, len(q), len(r),
                                             int(h0), int(band_width), len(self._q), len(self._r)))
        self._q += q
        self._r += r

    def __len__(self):
        return len(self.jobs)

    def is_empty(self):
        return not self.jobs

    def clear(self):
        self.jobs.clear()
        self._q = bytearray()
        self._r = bytearray()

    def get_query_seq(self, k):
        j = self.jobs[k]
        return bytes(self._q[j.query_offset:j.query_offset + j.query_len])

    def get_ref_seq(self, k):
        j = self.jobs[k]
        return bytes(self._r[j.ref_offset:j.ref_offset + j.ref_len])

    def arrays(self):
        n = len(self.jobs)
        q = np.frombuffer(bytes(self._q) + b"\0", np.uint8)
        r = np.frombuffer(bytes(self._r) + b"\0", np.uint8)
        qo = np.fromiter((j.query_offset for j in self.jobs), np.int64, n)
        ro = np.fromiter((j.ref_offset for j in self.jobs), np.int64, n)
        ql = np.fromiter((j.query_len for j in self.jobs), np.int32, n)
        rl = np.fromiter((j.ref_len for j in self.jobs), np.int32, n)
        h0 = np.fromiter((j.h0 for j in self.jobs), np.int32, n)
        w = np.fromiter((j.band_width for j in self.jobs), np.int32, n)
        return q, qo, ql, r, ro, rl, h0, w


class Context:
    """One GPU context (fa_ctx): device, stream, events.  Not shared across threads."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        N.check(N.lib().fa_ctx_create(int(device), ctypes.byref(self._h)), "fa_ctx_create")

    @property
    def handle(self):
        return self._h

    def describe(self):
        buf = ctypes.create_string_buffer(256)
        N.check(N.lib().fa_ctx_describe(self._h, buf, 256), "fa_ctx_describe")
        return buf.value.decode()

    def valu_peak(self, packed16=True):
        """Measured integer add/max VALU rate of this device, 10^12 lane-ops/s
        (packed 16-bit counts two ops per lane per instruction)."""
        v = ctypes.c_double()
        N.check(N.lib().fa_valu_peak(self._h, 1 if packed16 else 0, ctypes.byref(v)), "fa_valu_peak")
        return v.value

    def hbm_copy(self, nbytes):
        """Read+write GB/s of one device copy of `nbytes` (multiple of 16)."""
        v = ctypes.c_double()
        N.check(N.lib().fa_hbm_copy(self._h, int(nbytes), ctypes.byref(v)), "fa_hbm_copy")
        return v.value

    def close(self):
        if self._h:
            N.lib().fa_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def extend_arrays(sw: BandedPairWiseSW, qseq, qoff, qlen, tseq, toff, tlen, h0, w, flags=0,
                  ctx: Optional[Context] = None, want_cells=False, per_job_rc=False):
    """Score every pair of a batch given as flat arrays; returns (out[n,6], cells, rc)."""
    ctx = ctx or default_context()
    n = len(qlen)
    qseq = np.ascontiguousarray(qseq, np.uint8)
    tseq = np.ascontiguousarray(tseq, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.int64)
    toff = np.ascontiguousarray(toff, np.int64)
    qlen = np.ascontiguousarray(qlen, np.int32)
    tlen = np.ascontiguousarray(tlen, np.int32)
    h0 = np.ascontiguousarray(h0, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    out = np.zeros((n, 6), np.int32)
    cells = np.zeros(n, np.uint32) if want_cells else None
    rc = np.zeros(n, np.int32) if per_job_rc else None
    c = sw.to_c()
    code = N.lib().fa_execute_batch_simd_scoring(
        ctx.handle, ctypes.byref(c), n, N.ptr(qseq), N.ptr(qoff), N.ptr(qlen), N.ptr(tseq), N.ptr(toff),
        N.ptr(tlen), N.ptr(h0), N.ptr(w), flags, N.ptr(out), N.ptr(cells), N.ptr(rc))
    if per_job_rc:
        return out, cells, rc, code
    N.check(code, "fa_execute_batch_simd_scoring")
    return out, cells


class Aggregator:
    """Coalesces the small batches of concurrent host threads (fa_agg_*): the
    reference calls execute_batch_simd_scoring from every rayon worker at once
    (single_end.rs:92-110).  `execute` takes the arguments of extend_arrays
    and returns the same records; calls from several threads at once are run
    as one GPU batch per (scoring, flags) once `min_jobs` jobs are pending or
    `max_wait_us` after the oldest call arrived."""

    def __init__(self, device=0, min_jobs=4096, max_wait_us=2000):
        self._h = ctypes.c_void_p()
        N.check(N.lib().fa_agg_create(int(device), int(min_jobs), int(max_wait_us), ctypes.byref(self._h)),
                "fa_agg_create")

    def execute(self, sw: BandedPairWiseSW, qseq, qoff, qlen, tseq, toff, tlen, h0, w, flags=0,
                want_cells=False, per_job_rc=False):
        n = len(qlen)
        qseq = np.ascontiguousarray(qseq, np.uint8)
        tseq = np.ascontiguousarray(tseq, np.uint8)
        qoff = np.ascontiguousarray(qoff, np.int64)
        toff = np.ascontiguousarray(toff, np.int64)
        qlen = np.ascontiguousarray(qlen, np.int32)
        tlen = np.ascontiguousarray(tlen, np.int32)
        h0 = np.ascontiguousarray(h0, np.int32)
        w = np.ascontiguousarray(w, np.int32)
        out = np.zeros((n, 6), np.int32)
        cells = np.zeros(n, np.uint32) if want_cells else None
        rc = np.zeros(n, np.int32) if per_job_rc else None
        c = sw.to_c()
        code = N.lib().fa_agg_execute(self._h, ctypes.byref(c), n, N.ptr(qseq), N.ptr(qoff), N.ptr(qlen),
                                      N.ptr(tseq), N.ptr(toff), N.ptr(tlen), N.ptr(h0), N.ptr(w), flags,
                                      N.ptr(out), N.ptr(cells), N.ptr(rc))
        if per_job_rc:
            return out, cells, rc, code
        N.check(code, "fa_agg_execute")
        return out, cells

    def stats(self):
        """(calls served, batches run, jobs scored)."""
        v = [ctypes.c_int64() for _ in range(3)]
        N.check(N.lib().fa_agg_stats(self._h, *[ctypes.byref(x) for x in v]), "fa_agg_stats")
        return tuple(x.value for x in v)

    def close(self):
        if self._h:
            N.lib().fa_agg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


CIGAR_OPS = "MIDNSHP=X"


def scalar_banded_swa_batch(sw: BandedPairWiseSW, qseq, qoff, qlen, tseq, toff, tlen, w, h0=None,
                            ctx: Optional[Context] = None, want_cells=False, raw=False):
    """Batched scalar_banded_swa (banded_swa/scalar/implementation.rs:5-440) on
    the GPU: banded local SW with traceback, as region.rs:1058 runs it to
    regenerate CIGARs.  Returns (out[n,6] in OutScore order, CIGARs as lists
    of (op char, len), aligned reference bases, aligned query bases, cells)."""
    ctx = ctx or default_context()
    n = len(qlen)
    qseq = np.ascontiguousarray(qseq, np.uint8)
    tseq = np.ascontiguousarray(tseq, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.int64)
    toff = np.ascontiguousarray(toff, np.int64)
    qlen = np.ascontiguousarray(qlen, np.int32)
    tlen = np.ascontiguousarray(tlen, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    h0 = np.ascontiguousarray(np.zeros(n, np.int32) if h0 is None else h0, np.int32)
    cap_c = qlen.astype(np.int64) + tlen + 4
    cap_a = qlen.astype(np.int64) + tlen + 2
    coff = np.zeros(n, np.int64)
    aoff = np.zeros(n, np.int64)
    if n:
        coff[1:] = np.cumsum(cap_c)[:-1]
        aoff[1:] = np.cumsum(cap_a)[:-1]
    out = np.zeros((n, 6), np.int32)
    cells = np.zeros(n, np.uint32)
    cig = np.zeros(int(cap_c.sum()) if n else 1, np.uint32)
    ra = np.zeros(int(cap_a.sum()) if n else 1, np.uint8)
    qa = np.zeros_like(ra)
    nc = np.zeros(n, np.int32)
    nr = np.zeros(n, np.int32)
    nq = np.zeros(n, np.int32)
    c = sw.to_c()
    N.check(N.lib().fa_scalar_banded_swa_batch(
        ctx.handle, ctypes.byref(c), n, N.ptr(qseq), N.ptr(qoff), N.ptr(qlen), N.ptr(tseq), N.ptr(toff),
        N.ptr(tlen), N.ptr(w), N.ptr(h0), N.ptr(out), N.ptr(cells), N.ptr(cig), N.ptr(coff), N.ptr(nc),
        N.ptr(ra), N.ptr(qa), N.ptr(aoff), N.ptr(nr), N.ptr(nq)), "fa_scalar_banded_swa_batch")
    if raw:  # flat buffers: records, cells, CIGAR words + offsets + counts, aligned bases + offsets + counts
        return out, cells, cig, coff, nc, ra, qa, aoff, nr, nq
    cigars = [[(CIGAR_OPS[int(x) & 15], int(x) >> 4) for x in cig[coff[k]:coff[k] + nc[k]]] for k in range(n)]
    ref_al = [ra[aoff[k]:aoff[k] + nr[k]].copy() for k in range(n)]
    qry_al = [qa[aoff[k]:aoff[k] + nq[k]].copy() for k in range(n)]
    return out, cigars, ref_al, qry_al, (cells if want_cells else None)


def scalar_banded_swa_batch_dense(sw: BandedPairWiseSW, qseq, qoff, qlen, tseq, toff, tlen, w, h0=None,
                                  ctx: Optional[Context] = None):
    """scalar_banded_swa_batch with the CIGARs and aligned bases packed back to
    back in job order (fa_scalar_banded_swa_batch_dense): returns (out[n,6],
    cells[n], cigar words, cigar offsets[n+1], ref-aligned bases, their
    offsets[n+1], query-aligned bases, their offsets[n+1]); job k's entries are
    [off[k], off[k+1]).  The buffers are sized for the worst case but only
    their filled part is written, so only those pages are touched."""
    ctx = ctx or default_context()
    n = len(qlen)
    qseq = np.ascontiguousarray(qseq, np.uint8)
    tseq = np.ascontiguousarray(tseq, np.uint8)
    qoff = np.ascontiguousarray(qoff, np.int64)
    toff = np.ascontiguousarray(toff, np.int64)
    qlen = np.ascontiguousarray(qlen, np.int32)
    tlen = np.ascontiguousarray(tlen, np.int32)
    w = np.ascontiguousarray(w, np.int32)
    h0 = np.ascontiguousarray(np.zeros(n, np.int32) if h0 is None else h0, np.int32)
    tot = int(qlen.astype(np.int64).sum() + tlen.astype(np.int64).sum())
    ccap, acap = tot + 4 * n, tot + 2 * n
    out = np.empty((n, 6), np.int32)
    cells = np.empty(n, np.uint32)
    cig = np.empty(max(ccap, 1), np.uint32)
    ra = np.empty(max(acap, 1), np.uint8)
    qa = np.empty(max(acap, 1), np.uint8)
    co = np.empty(n + 1, np.int64)
    ro = np.empty(n + 1, np.int64)
    qo = np.empty(n + 1, np.int64)
    c = sw.to_c()
    N.check(N.lib().fa_scalar_banded_swa_batch_dense(
        ctx.handle, ctypes.byref(c), n, N.ptr(qseq), N.ptr(qoff), N.ptr(qlen), N.ptr(tseq), N.ptr(toff),
        N.ptr(tlen), N.ptr(w), N.ptr(h0), N.ptr(out), N.ptr(cells), N.ptr(cig), ccap, N.ptr(co), N.ptr(ra),
        N.ptr(ro), N.ptr(qa), N.ptr(qo), acap), "fa_scalar_banded_swa_batch_dense")
    return out, cells, cig[:co[n]], co, ra[:ro[n]], ro, qa[:qo[n]], qo


def scalar_banded_swa_last_ms(ctx: Optional[Context] = None):
    """(DP ms, traceback ms) of the last scalar_banded_swa_batch call (HIP events)."""
    ctx = ctx or default_context()
    a, b = ctypes.c_double(), ctypes.c_double()
    N.check(N.lib().fa_scalar_banded_swa_last_ms(ctx.handle, ctypes.byref(a), ctypes.byref(b)),
            "fa_scalar_banded_swa_last_ms")
    return a.value, b.value


def scalar_banded_swa(sw_params: BandedPairWiseSW, qlen, query, tlen, target, w, h0,
                      ctx: Optional[Context] = None):
    """implementation.rs:5-13 signature: (OutScore, CIGAR [(op char, len)],
    ref_aligned, query_aligned) for one pair, computed on the GPU."""
    q = np.ascontiguousarray(query[:qlen], np.uint8)
    t = np.ascontiguousarray(target[:tlen], np.uint8)
    out, cig, ra, qa, _ = scalar_banded_swa_batch(sw_params, q, [0], [len(q)], t, [0], [len(t)], [w], [h0],
                                                  ctx=ctx)
    o = out[0]
    return (OutScore(int(o[0]), int(o[1]), int(o[2]), int(o[3]), int(o[4]), int(o[5])), cig[0],
            list(ra[0]), list(qa[0]))


def execute_batch_simd_scoring(sw_params: BandedPairWiseSW, batch: ExtensionJobBatch, engine=None,
                               ctx: Optional[Context] = None) -> List[BatchExtensionResult]:
    """dispatch.rs:31-86: score every job of the batch; one result per job, in job order."""
    if batch.is_empty():
        return []
    q, qo, ql, r, ro, rl, h0, w = batch.arrays()
    out, _ = extend_arrays(sw_params, q, qo, ql, r, ro, rl, h0, w, N.FA_SCORE_AUTO, ctx)
    res = []
    for j, o in zip(batch.jobs, out):
        res.append(BatchExtensionResult(j.read_idx, j.chain_idx, j.seed_idx, j.direction,
                                        score=int(o[0]), query_end=int(o[3]), ref_end=int(o[1]),
                                        gscore=int(o[4]), gref_end=int(o[2]), max_off=int(o[5])))
    return res


@dataclass
class SoAInputs:
    """shared.rs:12-22 with int32 per-lane parameters (the i8 lanes clamp at 127)."""
    query_soa: np.ndarray
    target_soa: np.ndarray
    qlen: np.ndarray
    tlen: np.ndarray
    w: np.ndarray
    h0: np.ndarray
    lanes: int
    max_qlen: int = 0
    max_tlen: int = 0


@dataclass
class SoAInputs16:
    """shared.rs:24-35: the 16-bit kernels' carrier (i16 sequences), int32
    per-lane parameters."""
    query_soa: np.ndarray
    target_soa: np.ndarray
    qlen: np.ndarray
    tlen: np.ndarray
    h0: np.ndarray
    w: np.ndarray
    lanes: int
    max_qlen: int = 0
    max_tlen: int = 0


def pad_batch(batch, lanes):
    """shared.rs:40-82: pad a list of (qlen, query, tlen, target, w, h0) jobs
    to `lanes` entries (extra jobs are dropped, missing lanes get empty jobs)
    and take the per-lane parameters.  Returns (qlen, tlen, h0, w, max_qlen,
    max_tlen, padded); the parameters are int32 here, where the reference
    casts them to i8 (lengths clamped at 127)."""
    padded = list(batch[:lanes]) + [(0, b"", 0, b"", 0, 0)] * max(0, lanes - len(batch))
    qlen = np.array([int(x[0]) for x in padded], np.int32)
    tlen = np.array([int(x[2]) for x in padded], np.int32)
    w = np.array([int(x[4]) for x in padded], np.int32)
    h0 = np.array([int(x[5]) for x in padded], np.int32)
    return qlen, tlen, h0, w, int(max(0, qlen.max(initial=0))), int(max(0, tlen.max(initial=0))), padded


def soa_transform(batch, lanes, max_len):
    """shared.rs:87-111: lane-interleaved buffers padded with 0xFF.

    batch: list of (qlen, query, tlen, target, w, h0) tuples (pad_batch order,
    shared.rs:40-82)."""
    def codes(x, n):
        if isinstance(x, (bytes, bytearray)):
            return np.frombuffer(bytes(x[:n]), np.uint8)
        return np.asarray(x[:n], np.uint8)

    q = np.full(max_len * lanes, 0xFF, np.uint8)
    t = np.full(max_len * lanes, 0xFF, np.uint8)
    for i, (ql, qs, tl, ts, _w, _h) in enumerate(batch[:lanes]):
        qn, tn = min(ql, max_len), min(tl, max_len)
        q[np.arange(qn) * lanes + i] = codes(qs, qn)
        t[np.arange(tn) * lanes + i] = codes(ts, tn)
    return q, t


def _soa_call(fn, inputs: SoAInputs, num_jobs, o_del, e_del, o_ins, e_ins, zdrop, mat, m, end_bonus, dtype,
              ctx=None):
    ctx = ctx or default_context()
    out = np.zeros((max(num_jobs, 1), 6), np.int32)
    qsoa = np.ascontiguousarray(inputs.query_soa, dtype)
    tsoa = np.ascontiguousarray(inputs.target_soa, dtype)
    ql = np.ascontiguousarray(inputs.qlen, np.int32)
    tl = np.ascontiguousarray(inputs.tlen, np.int32)
    h0 = np.ascontiguousarray(inputs.h0, np.int32)
    w = np.ascontiguousarray(inputs.w, np.int32)
    mat = np.ascontiguousarray(mat, np.int8)
    N.check(fn(ctx.handle, int(inputs.lanes), N.ptr(qsoa), N.ptr(tsoa), N.ptr(ql), N.ptr(tl), N.ptr(h0), N.ptr(w),
               int(num_jobs), o_del, e_del, o_ins, e_ins, zdrop, end_bonus, N.ptr(mat), m, N.ptr(out)),
            fn.__name__)
    return [OutScore(*map(int, o)) for o in out[:num_jobs]]


def _make_soa_entry(lanes, int16):
    def entry(inputs: SoAInputs, num_jobs, o_del, e_del, o_ins, e_ins, zdrop, mat, m, end_bonus=5, ctx=None):
        if inputs.lanes != lanes:
            raise FaError(N.FA_ERR_ARG, "SoA lanes %d != %d" % (inputs.lanes, lanes))
        fn = N.lib().fa_simd_banded_swa_batch_int16_soa if int16 else N.lib().fa_simd_banded_swa_batch_soa
        return _soa_call(fn, inputs, num_jobs, o_del, e_del, o_ins, e_ins, zdrop, mat, m, end_bonus,
                         np.int16 if int16 else np.uint8, ctx)
    entry.__name__ = "simd_banded_swa_batch%d%s_soa" % (lanes, "_int16" if int16 else "")
    entry.__doc__ = ("GPU replacement of %s (shared.rs:161-389); returns one OutScore per job "
                     "with ksw_extend2 semantics." % entry.__name__)
    return entry


simd_banded_swa_batch16_soa = _make_soa_entry(16, False)
simd_banded_swa_batch32_soa = _make_soa_entry(32, False)
simd_banded_swa_batch64_soa = _make_soa_entry(64, False)
simd_banded_swa_batch8_int16_soa = _make_soa_entry(8, True)
simd_banded_swa_batch16_int16_soa = _make_soa_entry(16, True)
simd_banded_swa_batch32_int16_soa = _make_soa_entry(32, True)


def pack_pac(codes):
    """2-bit .pac bytes of a base-code sequence (codes 0-3; 4 bases per byte,
    first base in the two most significant bits, src/pipelines/linear/index/
    bntseq.rs:559-575)."""
    codes = np.ascontiguousarray(codes, np.uint8)
    if codes.size and codes.max() > 3:
        raise ValueError("a .pac holds codes 0-3 only")
    n = codes.size
    pad = np.zeros((n + 3) // 4 * 4, np.uint8)
    pad[:n] = codes
    q = pad.reshape(-1, 4).astype(np.uint8)
    return ((q[:, 0] << 6) | (q[:, 1] << 4) | (q[:, 2] << 2) | q[:, 3]).astype(np.uint8)


class DeviceReference:
    """A 2-bit reference (.pac bytes, l_pac bases) resident in HBM (fa_ref)."""

    def __init__(self, pac, l_pac, ctx=None):
        self.ctx = ctx or default_context()
        self.l_pac = int(l_pac)
        pac = np.ascontiguousarray(pac, np.uint8)
        if pac.size < (self.l_pac + 3) // 4:
            raise ValueError("pac holds fewer than l_pac bases")
        self._h = ctypes.c_void_p()
        N.check(N.lib().fa_ref_upload(self.ctx.handle, N.ptr(pac), self.l_pac, ctypes.byref(self._h)),
                "fa_ref_upload")

    @property
    def handle(self):
        return self._h

    def free(self):
        if self._h:
            N.lib().fa_ref_free(self.ctx.handle, self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceBatch:
    """A batch resident in HBM (fa_batch): upload once, run many times."""

    @classmethod
    def from_reference(cls, sw, qseq, qoff, qlen, ref, seg_beg, seg_len, seg_rev, h0, w, ctx=None):
        """Batch whose targets are segments of a DeviceReference, read on the
        device: target j = [seg_beg[j], seg_beg[j] + seg_len[j]) of the
        bidirectional coordinates [0, 2 l_pac), backwards when seg_rev[j]."""
        self = cls.__new__(cls)
        self.ctx = ctx or ref.ctx
        self.n = len(qlen)
        k = [np.ascontiguousarray(x, dt) for x, dt in
             ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32))]
        seg = [np.ascontiguousarray(x, dt) for x, dt in
               ((seg_beg, np.int64), (seg_len, np.int32), (seg_rev, np.uint8), (h0, np.int32), (w, np.int32))]
        self._h = ctypes.c_void_p()
        c = sw.to_c()
        N.check(N.lib().fa_batch_upload_ref(self.ctx.handle, ctypes.byref(c), self.n, *[N.ptr(x) for x in k],
                                            ref.handle, *[N.ptr(x) for x in seg], ctypes.byref(self._h)),
                "fa_batch_upload_ref")
        self._tlen = seg[1]
        return self

    def download_targets(self, tlen=None):
        """The batch's targets as one array of codes (target j at the offset
        of the running sum of lengths) and those offsets."""
        tlen = np.asarray(tlen if tlen is not None else self._tlen, np.int64)
        off = np.zeros(len(tlen), np.int64)
        off[1:] = np.cumsum(tlen)[:-1]
        codes = np.zeros(int(tlen.sum()), np.uint8)
        N.check(N.lib().fa_batch_download_targets(self.ctx.handle, self._h, N.ptr(codes), N.ptr(off)),
                "fa_batch_download_targets")
        return codes, off

    def __init__(self, sw: BandedPairWiseSW, qseq, qoff, qlen, tseq, toff, tlen, h0, w, ctx=None):
        self.ctx = ctx or default_context()
        self.n = len(qlen)
        self._keep = [np.ascontiguousarray(x, dt) for x, dt in
                      ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32), (tseq, np.uint8),
                       (toff, np.int64), (tlen, np.int32), (h0, np.int32), (w, np.int32))]
        self._h = ctypes.c_void_p()
        c = sw.to_c()
        k = self._keep
        N.check(N.lib().fa_batch_upload(self.ctx.handle, ctypes.byref(c), self.n, *[N.ptr(x) for x in k],
                                        ctypes.byref(self._h)), "fa_batch_upload")
        self._tlen = k[5]
        self._keep = None

    def set_cost(self, cost):
        cost = np.ascontiguousarray(cost, np.float32)
        self._cost = cost
        N.check(N.lib().fa_batch_set_cost(self.ctx.handle, self._h, N.ptr(cost)), "fa_batch_set_cost")

    def set_launch(self, lanes_per_pair=8, chunk_pairs=0, waves_per_cu=0):
        N.check(N.lib().fa_batch_set_launch(self.ctx.handle, self._h, lanes_per_pair, chunk_pairs, waves_per_cu),
                "fa_batch_set_launch")

    def set_streams(self, streams):
        N.check(N.lib().fa_batch_set_streams(self.ctx.handle, self._h, int(streams)), "fa_batch_set_streams")

    def run(self, flags=0):
        N.check(N.lib().fa_batch_run(self.ctx.handle, self._h, flags), "fa_batch_run")

    def time(self, flags=0, reps=10, dominant=False):
        """(ms per run over `reps` back-to-back runs from HIP events, longest
        single launch in ms or None, launches per run)."""
        ms = ctypes.c_double()
        dom = ctypes.c_double()
        nl = ctypes.c_int32()
        N.check(N.lib().fa_batch_time(self.ctx.handle, self._h, flags, reps, ctypes.byref(ms),
                                      ctypes.byref(dom) if dominant else None, ctypes.byref(nl)), "fa_batch_time")
        return ms.value, (dom.value if dominant else None), nl.value

    def set_align2(self, mat, o_del, e_del, o_ins, e_ins, xtra, m=5):
        """Attach mate-rescue local SW parameters (run with FA_RUN_ALIGN2)."""
        mat = np.ascontiguousarray(mat, np.int8)
        N.check(N.lib().fa_batch_set_align2(self.ctx.handle, self._h, N.ptr(mat), m, o_del, e_del, o_ins, e_ins,
                                            xtra), "fa_batch_set_align2")

    def download_align2(self):
        out = np.zeros((self.n, 7), np.int32)
        cells = np.zeros(self.n, np.uint32)
        N.check(N.lib().fa_batch_download_align2(self.ctx.handle, self._h, N.ptr(out), N.ptr(cells)),
                "fa_batch_download_align2")
        return out, cells

    def profile_launches(self, flags=0, cap=64):
        """Each launch of one run, alone and timed: list of dicts with the
        kernel name, slots, lanes, packed range and ms."""
        nl = ctypes.c_int32()
        slots = np.zeros(cap, np.int32)
        kern = np.zeros(cap, np.int32)
        first = np.zeros(cap, np.int64)
        count = np.zeros(cap, np.int64)
        ms = np.zeros(cap, np.float64)
        N.check(N.lib().fa_batch_profile_launches(self.ctx.handle, self._h, flags, cap, ctypes.byref(nl),
                                                  N.ptr(slots), N.ptr(kern), N.ptr(first), N.ptr(count), N.ptr(ms)),
                "fa_batch_profile_launches")
        names = {0: "extend_plain_kernel<%d,%d>", 1: "extend16_plain_kernel<%d,%d>"}
        # kind 4: the long-pair kernel (one wavefront per pair, no lane groups)
        out = []
        for i in range(min(nl.value, cap)):
            kind, lanes = int(kern[i]) >> 8, int(kern[i]) & 0xFF
            if kind in (2, 3):
                # mate rescue: forward (2) or separate XSTART reverse (3) launch;
                # `first` of a reverse launch indexes the reverse order
                base = "mate16_kernel" if kind == 2 else "mate16_rev_kernel"
                name = "%s<%d%s>" % (base, int(slots[i]), "" if lanes == 8 else ",8-bit")
            elif kind == 4:
                name = "extend_long_kernel"
            else:
                name = names[kind] % (lanes, int(slots[i]))
                if flags & N.FA_LAUNCH_PERSISTENT:
                    name = name.replace("_plain_", "_persistent_")
            out.append(dict(kernel=name, slots=int(slots[i]), lanes=lanes, first=int(first[i]),
                            count=int(count[i]), ms=float(ms[i])))
        return out

    def order(self, flags=0):
        o = np.zeros(self.n, np.uint32)
        N.check(N.lib().fa_batch_get_order(self.ctx.handle, self._h, flags, N.ptr(o)), "fa_batch_get_order")
        return o

    def download(self):
        out = np.zeros((self.n, 6), np.int32)
        cells = np.zeros(self.n, np.uint32)
        N.check(N.lib().fa_batch_download(self.ctx.handle, self._h, N.ptr(out), N.ptr(cells)), "fa_batch_download")
        return out, cells

    def overflow_count(self):
        c = ctypes.c_int64()
        N.check(N.lib().fa_batch_overflow_count(self.ctx.handle, self._h, ctypes.byref(c)), "overflow_count")
        return c.value

    def free(self):
        if self._h:
            N.lib().fa_batch_free(self.ctx.handle, self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def ksw_align2_batch(qseq, qoff, qlen, tseq, toff, tlen, mat, o_del, e_del, o_ins, e_ins, xtra, m=5, ctx=None):
    """Mate-rescue local SW for a batch: one Kswr record (score, te, qe,
    score2, te2, tb, qb) per job, exactly what ksw_align2
    (ksw_affine_gap.rs:773-861) returns in the mode xtra selects: 16-bit, or
    8-bit with KSW_XBYTE (0x10000; mate_rescue.rs:314-321 sets it when
    qlen * match < 250).  Returns (out[n,7], cells[n])."""
    ctx = ctx or default_context()
    n = len(qlen)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (tseq, np.uint8), (toff, np.int64), (tlen, np.int32))]
    mat = np.ascontiguousarray(mat, np.int8)
    out = np.zeros((n, 7), np.int32)
    cells = np.zeros(n, np.uint32)
    N.check(N.lib().fa_ksw_align2_batch(ctx.handle, n, *[N.ptr(a) for a in arrs], m, N.ptr(mat), o_del, e_del, o_ins,
                                        e_ins, xtra, N.ptr(out), N.ptr(cells)), "fa_ksw_align2_batch")
    return out, cells
