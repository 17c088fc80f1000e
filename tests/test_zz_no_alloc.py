"""fa_execute_batch_device_noalloc (the allocation-free entry point) against
fa_execute_batch_simd_scoring on the same jobs: records, cells, per-job codes
and the return code bit for bit, on two calls in a row (the second smaller,
so stale scratch contents from the first would show), plus the scratch-size
rejection."""
import ctypes

import numpy as np
import pytest

from ferrousalign_amd import _native as N
from ferrousalign_amd import synth
from ferrousalign_amd.banded_swa import extend_arrays, Context

pytestmark = pytest.mark.gpu


def _jobs(name, n, seed):
    """Workload pairs plus invalid jobs and pairs outside the 16-bit bound."""
    ps = synth.workload(name, n=n)
    q, qo, ql, t, to, tl, h0, w = [np.array(x) for x in ps.flat()]
    rng = np.random.default_rng(seed)
    k = rng.choice(n, 64, replace=False)
    h0[k[:8]] = 0                      # FA_ERR_H0
    ql[k[8:12]] = 0                    # FA_ERR_QLEN
    w[k[12:16]] = -1                   # FA_ERR_BAND
    q[qo[k[16]] + 3] = 9               # FA_ERR_CODE (query)
    t[to[k[17]] + min(5, int(tl[k[17]]) - 1)] = 7  # FA_ERR_CODE (target)
    h0[k[20:40]] = rng.integers(3000, 5000, 20)    # scores beyond the packed 16-bit kernel
    tl[k[40:44]] = 0                   # empty targets
    return q, qo, ql, t, to, tl, h0, w


class _Hip:
    """Device buffers through the HIP runtime the library itself links
    (torch bundles a runtime of its own, which cannot share the process)."""

    def __init__(self):
        N.lib()  # loads libamdhip64.so.7 with the library
        self.rt = ctypes.CDLL("libamdhip64.so.7")
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 16)) == 0
        self.live.append(p)
        assert self.rt.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p

    def download(self, p, like):
        a = np.empty_like(like)
        assert self.rt.hipMemcpy(a.ctypes.data, p, a.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return a

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


def _noalloc(ctx, sw, jobs, max_ref_len):
    hip = _Hip()
    q, qo, ql, t, to, tl, h0, w = jobs
    n = len(ql)
    dev = [hip.upload(x) for x in (q, qo, ql, t, to, tl, h0, w)]
    out0 = np.full((n, 6), -7, np.int32)
    cells0 = np.full(n, 12345, np.uint32)
    rc0 = np.full(n, 99, np.int32)
    d_out, d_cells, d_rc = hip.upload(out0), hip.upload(cells0), hip.upload(rc0)
    c = sw.to_c()
    code = N.lib().fa_execute_batch_device_noalloc(ctx.handle, ctypes.byref(c), n, *dev, int(max_ref_len),
                                                   d_out, d_cells, d_rc)
    res = hip.download(d_out, out0), hip.download(d_cells, cells0), hip.download(d_rc, rc0), code
    hip.free()
    return res


def test_noalloc_matches_execute_batch_twice():
    ctx = Context(0)
    sw = synth.default_scoring()
    first = _jobs("uneven_mix_4m", 20000, 1)
    second = _jobs("illumina_150bp_1m", 3000, 2)
    for jobs in (first, second):
        ref_out, ref_cells, ref_rc, ref_code = extend_arrays(sw, *jobs, ctx=ctx, want_cells=True, per_job_rc=True)
        assert (ref_rc != 0).sum() >= 17 and ref_code != 0
        out, cells, rc, code = _noalloc(ctx, sw, jobs, int(jobs[5].max()))
        print("noalloc jobs", len(jobs[2]), "code", code, flush=True)
        assert code == ref_code
        assert (rc == ref_rc).all()
        assert (out == ref_out).all(), int((out != ref_out).any(axis=1).sum())
        assert (cells == ref_cells).all()
    ctx.close()


def test_noalloc_rejects_oversize_batch():
    ctx = Context(0)
    sw = synth.default_scoring()
    c = sw.to_c()
    hip = _Hip()
    p = hip.upload(np.zeros(64, np.int64))
    # 1,000,000 jobs with 2048-base targets need ~0.9 GB of scratch: refused on
    # the host before any launch
    code = N.lib().fa_execute_batch_device_noalloc(ctx.handle, ctypes.byref(c), 1000000, p, p, p, p, p, p, p, p,
                                                   2048, p, None, None)
    assert code == N.FA_ERR_SCRATCH
    assert N.lib().fa_execute_batch_device_noalloc(ctx.handle, ctypes.byref(c), 4, p, p, p, p, p, p, p, p,
                                                   4096, p, None, None) == N.FA_ERR_ARG
    hip.free()
    ctx.close()
