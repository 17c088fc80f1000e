"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure."""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "liboracle.so")
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
        l = ctypes.CDLL(LIB)
        l.ora_ksw_extend2.argtypes = [_I32, _P, _I32, _P, _I32, _P] + [_I32] * 8 + [_P, _P]
        l.ora_ksw_extend2.restype = ctypes.c_int
        l.ora_extend_batch.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P] + [_I32] * 6 + [_P, _P, _P,
                                                                                                        ctypes.c_int]
        l.ora_extend_batch.restype = _I64
        l.ora_extend_batch_rows.argtypes = ([_I64, _P, _P, _P, _P, _P, _P, _P, _P, _I32, _P] + [_I32] * 6 +
                                            [_P, _P, _P, _P, ctypes.c_int])
        l.ora_extend_batch_rows.restype = _I64
        _lib = l
    return _lib


def P(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def ksw_extend2(query, target, mat, o_del, e_del, o_ins, e_ins, w, end_bonus, zdrop, h0, m=5):
    """Single pair; returns (rc, (score, qle, tle, gtle, gscore, max_off), cells)."""
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(target, np.uint8)
    if q.size == 0:
        q = np.zeros(1, np.uint8)[:0]
    mat = np.ascontiguousarray(mat, np.int8)
    out = np.zeros(6, np.int32)
    cells = np.zeros(1, np.uint64)
    qb = np.concatenate([q, np.zeros(1, np.uint8)])
    tb = np.concatenate([t, np.zeros(1, np.uint8)])
    rc = lib().ora_ksw_extend2(len(query), P(qb), len(target), P(tb), m, P(mat), o_del, e_del, o_ins, e_ins, w,
                               end_bonus, zdrop, h0, P(out), P(cells))
    return rc, tuple(int(x) for x in out), int(cells[0])


def extend_batch(sw, qseq, qoff, qlen, tseq, toff, tlen, h0, w, nthreads=0):
    """Oracle over a batch; returns (out[n,6] in KswExtendResult order, cells, rc)."""
    n = len(qlen)
    out = np.zeros((n, 6), np.int32)
    cells = np.zeros(n, np.uint64)
    rc = np.zeros(n, np.int32)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (tseq, np.uint8), (toff, np.int64), (tlen, np.int32),
                                                      (h0, np.int32), (w, np.int32))]
    mat = np.ascontiguousarray(sw.mat, np.int8)
    lib().ora_extend_batch(n, *[P(a) for a in arrs], 5, P(mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins,
                           sw.end_bonus, sw.zdrop, P(out), P(cells), P(rc), nthreads)
    return out, cells, rc


def extend_rows(sw, qseq, qoff, qlen, tseq, toff, tlen, h0, w, nthreads=0):
    """Target rows each job's row loop ran (the break row included), with the
    cells: (rows int32[n], cells uint64[n])."""
    n = len(qlen)
    out = np.zeros((n, 6), np.int32)
    cells = np.zeros(n, np.uint64)
    rc = np.zeros(n, np.int32)
    rows = np.zeros(n, np.int32)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (tseq, np.uint8), (toff, np.int64), (tlen, np.int32),
                                                      (h0, np.int32), (w, np.int32))]
    mat = np.ascontiguousarray(sw.mat, np.int8)
    lib().ora_extend_batch_rows(n, *[P(a) for a in arrs], 5, P(mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins,
                                sw.end_bonus, sw.zdrop, P(out), P(cells), P(rc), P(rows), nthreads)
    return rows, cells


def to_outscore_order(ora_out):
    """KswExtendResult (score,qle,tle,gtle,gscore,max_off) -> OutScore field order
    (score, target_end_pos=tle, gtarget_end_pos=gtle, query_end_pos=qle, global_score, max_offset)."""
    o = np.asarray(ora_out)
    return np.stack([o[:, 0], o[:, 2], o[:, 3], o[:, 1], o[:, 4], o[:, 5]], axis=1)


def _mlp_lib():
    l = lib()
    if not getattr(l, "_mlp", False):
        l.ora_round16.argtypes = [ctypes.c_float]
        l.ora_round16.restype = ctypes.c_float
        l.ora_mlp_forward.argtypes = [_P, _P, _P, _I64, ctypes.c_int, _P]
        l.ora_mlp_forward.restype = None
        l.ora_mlp_grad.argtypes = [_P, _P, _P, _P, _I64, ctypes.c_int, _P]
        l.ora_mlp_grad.restype = ctypes.c_double
        l.ora_pair_features.argtypes = [_P, _I32, _P, _I32, _I32, _I32, _P]
        l.ora_pair_features.restype = None
        l._mlp = True
    return l


# Every array handed to C is bound to a local name first, so that it stays
# alive for the duration of the call.

def mlp_forward(params, norm, feat, fp16=True):
    feat = np.ascontiguousarray(feat, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    y = np.zeros(len(feat), np.float32)
    _mlp_lib().ora_mlp_forward(P(params), P(norm), P(feat), len(feat), int(fp16), P(y))
    return y


def mlp_grad(params, norm, feat, target, fp16=True):
    feat = np.ascontiguousarray(feat, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    norm = np.ascontiguousarray(norm, np.float32)
    target = np.ascontiguousarray(target, np.float32)
    g = np.zeros(1313, np.float64)
    loss = _mlp_lib().ora_mlp_grad(P(params), P(norm), P(feat), P(target), len(feat), int(fp16), P(g))
    return loss, g


def pair_features(q, t, h0, w):
    qb = np.concatenate([np.asarray(q, np.uint8), np.zeros(1, np.uint8)])
    tb = np.concatenate([np.asarray(t, np.uint8), np.zeros(1, np.uint8)])
    out = np.zeros(6, np.float32)
    _mlp_lib().ora_pair_features(P(qb), len(qb) - 1, P(tb), len(tb) - 1, h0, w, P(out))
    return out


def round16(x):
    return _mlp_lib().ora_round16(float(x))


def _align_lib():
    l = lib()
    if not getattr(l, "_align", False):
        l.ora_ksw_align2_i16.argtypes = [_I32, _P, _I32, _P, _I32, _P, _I32, _I32, _I32, _I32, ctypes.c_uint32, _P, _P]
        l.ora_ksw_align2_i16.restype = ctypes.c_int
        l.ora_ksw_align2.argtypes = l.ora_ksw_align2_i16.argtypes
        l.ora_ksw_align2.restype = ctypes.c_int
        l.ora_ksw_align2_batch.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32,
                                           ctypes.c_uint32, _P, _P, _P, ctypes.c_int]
        l.ora_ksw_align2_batch.restype = None
        l._align = True
    return l


KSW_XSTOP, KSW_XSUBO, KSW_XSTART, KSW_XBYTE = 0x20000, 0x40000, 0x80000, 0x10000


def _align2_call(fn, query, target, mat, o_del, e_del, o_ins, e_ins, xtra):
    qb = np.concatenate([np.asarray(query, np.uint8), np.zeros(1, np.uint8)])
    tb = np.concatenate([np.asarray(target, np.uint8), np.zeros(1, np.uint8)])
    mat = np.ascontiguousarray(mat, np.int8)
    out = np.zeros(7, np.int32)
    cells = np.zeros(1, np.uint64)
    rc = fn(len(qb) - 1, P(qb), len(tb) - 1, P(tb), 5, P(mat), o_del, e_del, o_ins, e_ins, xtra, P(out), P(cells))
    return rc, tuple(int(x) for x in out), int(cells[0])


def ksw_align2_i16(query, target, mat, o_del, e_del, o_ins, e_ins, xtra):
    """16-bit mode only (rc -1 when KSW_XBYTE is set)."""
    return _align2_call(_align_lib().ora_ksw_align2_i16, query, target, mat, o_del, e_del, o_ins, e_ins, xtra)


def ksw_align2(query, target, mat, o_del, e_del, o_ins, e_ins, xtra):
    """The mode xtra selects: 8-bit with KSW_XBYTE, else 16-bit."""
    return _align2_call(_align_lib().ora_ksw_align2, query, target, mat, o_del, e_del, o_ins, e_ins, xtra)


def ksw_align2_batch(q, qoff, qlen, t, toff, tlen, mat, o_del, e_del, o_ins, e_ins, xtra, nthreads=0):
    n = len(qlen)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((q, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (t, np.uint8), (toff, np.int64), (tlen, np.int32))]
    mat = np.ascontiguousarray(mat, np.int8)
    out = np.zeros((n, 7), np.int32)
    cells = np.zeros(n, np.uint64)
    rc = np.zeros(n, np.int32)
    _align_lib().ora_ksw_align2_batch(n, *[P(a) for a in arrs], 5, P(mat), o_del, e_del, o_ins, e_ins, xtra,
                                      P(out), P(cells), P(rc), nthreads)
    return out, cells, rc


def pac_segment(pac, l_pac, start, length, reverse=False):
    """Target segment of a .pac reference (oracle/pac_oracle.c), None when the
    reference would return an empty segment (outside [0, 2 l_pac) or across
    l_pac)."""
    l = lib()
    l.ora_pac_segment.restype = ctypes.c_int64
    l.ora_pac_segment.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int,
                                  ctypes.c_void_p]
    pac = np.ascontiguousarray(pac, np.uint8)
    out = np.zeros(max(int(length), 1), np.uint8)
    n = l.ora_pac_segment(P(pac), int(l_pac), int(start), int(length), int(bool(reverse)), P(out))
    return None if n < 0 else out[:n]


CIGAR_OPS = "MIDNSHP=X"


def scalar_banded_swa(query, target, sw, w, h0=0):
    """Oracle scalar_banded_swa (oracle/swtb_oracle.c): (OutScore-order record,
    CIGAR [(op char, len)], ref_aligned, query_aligned, cells)."""
    l = lib()
    f = l.ora_scalar_banded_swa
    f.restype = ctypes.c_int
    f.argtypes = [_I32, _P, _I32, _P, _P] + [_I32] * 8 + [_P, _P, _P, _P, _P, _P, _P, _P]
    q = np.ascontiguousarray(query, np.uint8)
    t = np.ascontiguousarray(target, np.uint8)
    mat = np.ascontiguousarray(sw.mat, np.int8)
    cap = len(q) + len(t) + 4
    out = np.zeros(6, np.int32)
    cig = np.zeros(cap, np.uint32)
    ra = np.zeros(cap, np.uint8)
    qa = np.zeros(cap, np.uint8)
    nc, nr, nq = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
    cells = ctypes.c_uint64()
    rc = f(len(q), P(q), len(t), P(t), P(mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins, sw.zdrop, sw.pen_clip3,
           int(w), int(h0), P(out), P(cig), ctypes.byref(nc), P(ra), ctypes.byref(nr), P(qa), ctypes.byref(nq),
           ctypes.byref(cells))
    assert rc == 0
    score, qle, tle, gtle, gscore, max_off = out.tolist()
    rec = np.array([score, tle, gtle, qle, gscore, max_off], np.int32)  # OutScore order
    cigar = [(CIGAR_OPS[int(c) & 15], int(c) >> 4) for c in cig[:nc.value]]
    return rec, cigar, ra[:nr.value].copy(), qa[:nq.value].copy(), int(cells.value)


def scalar_banded_swa_batch(sw, qseq, qoff, qlen, tseq, toff, tlen, w, h0, nthreads=0):
    """Oracle scalar_banded_swa over a batch (records in OutScore order, cells)."""
    l = lib()
    f = l.ora_scalar_banded_swa_batch
    f.restype = _I64
    f.argtypes = [_I64] + [_P] * 9 + [_I32] * 6 + [_P, _P, _P, ctypes.c_int]
    arrs = [np.ascontiguousarray(a, dt) for a, dt in ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (tseq, np.uint8), (toff, np.int64), (tlen, np.int32),
                                                      (w, np.int32), (h0, np.int32))]
    n = len(arrs[2])
    mat = np.ascontiguousarray(sw.mat, np.int8)
    out = np.zeros((n, 6), np.int32)
    cells = np.zeros(n, np.uint64)
    nc = np.zeros(n, np.int32)
    f(n, *[P(a) for a in arrs], P(mat), sw.o_del, sw.e_del, sw.o_ins, sw.e_ins, sw.zdrop, sw.pen_clip3,
      P(out), P(cells), P(nc), int(nthreads))
    return to_outscore_order(out), cells, nc


REFSIMD_ISA = {"auto": 0, "avx2": 1, "avx512": 2, "scalar512": 3, "scalar2": 4}


def refsimd_batch(sw, qseq, qoff, qlen, tseq, toff, tlen, h0, w, isa="auto", nthreads=1):
    """The restatement of the reference's SIMD batch path (oracle/refsimd_baseline.c):
    (out[n,6] in OutScore order, isa run, whether the i16 kernel ran); None when
    the requested instruction set is missing."""
    l = lib()
    if not hasattr(l, "_rs"):
        l.ora_refsimd_batch.argtypes = [_I64] + [_P] * 8 + [_I32] * 5 + [_P, _I32, _I32, _P, _P]
        l.ora_refsimd_batch.restype = ctypes.c_int
        l._rs = True
    n = len(qlen)
    out = np.zeros((max(n, 1), 6), np.int32)
    info = np.zeros(2, np.int32)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((qseq, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (tseq, np.uint8), (toff, np.int64), (tlen, np.int32),
                                                      (h0, np.int32), (w, np.int32))]
    mat = np.ascontiguousarray(sw.mat, np.int8)
    rc = l.ora_refsimd_batch(n, *[P(a) for a in arrs], sw.o_del, sw.e_del, sw.o_ins, sw.e_ins, sw.zdrop, P(mat),
                             REFSIMD_ISA[isa], int(nthreads), P(out), P(info))
    if rc != 0:
        return None
    names = {v: k for k, v in REFSIMD_ISA.items()}
    return out[:n], names[int(info[0])], bool(info[1])


def _kswv_lib():
    l = lib()
    if not getattr(l, "_kswv", False):
        l.kv_batch.argtypes = [_I64, _P, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32,
                               ctypes.c_int, _P, ctypes.c_int]
        l.kv_batch.restype = ctypes.c_int
        l._kswv = True
    return l


def kswv_batch(q, qoff, qlen, t, toff, tlen, match=1, mis=-4, ambig=0, gaps=(6, 1, 6, 1), simd=True,
               nthreads=0, xtra=None):
    """The reference's horizontal kswv mate-rescue path restated
    (oracle/kswv_baseline.c): chunks of 32 jobs through the AVX2 kernel
    (simd=True) or its plain-C twin.  Returns out[n, 7] (score, te, qe,
    score2, te2, tb, qb), or None when AVX2 is absent or the input is refused."""
    n = len(qlen)
    arrs = [np.ascontiguousarray(x, dt) for x, dt in ((q, np.uint8), (qoff, np.int64), (qlen, np.int32),
                                                      (t, np.uint8), (toff, np.int64), (tlen, np.int32))]
    xt = None if xtra is None else np.ascontiguousarray(xtra, np.int32)
    out = np.zeros((n, 7), np.int32)
    rc = _kswv_lib().kv_batch(n, *[P(a) for a in arrs], P(xt) if xt is not None else None, match, mis, ambig,
                              *gaps, 1 if simd else 0, P(out), nthreads)
    return None if rc else out
