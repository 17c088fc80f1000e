"""The C-ABI library loads, exports every symbol include/ferrousalign_amd.h
declares, validates its arguments, and its host-side 2-bit packing round-trips
(no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "ferrousalign_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(fa_[a-z0-9_]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    from ferrousalign_amd import _native as N
    lib = ctypes.CDLL(N.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_device_is_an_error_not_a_crash():
    from ferrousalign_amd import _native as N
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    h = ctypes.c_void_p()
    assert N.lib().fa_ctx_create(0, ctypes.byref(h)) == N.FA_ERR_NODEVICE
    assert N.lib().fa_strerror(N.FA_ERR_H0) == b"h0 must be positive"
    assert N.lib().fa_ctx_destroy(None) == N.FA_ERR_ARG


@pytest.mark.parametrize("seed", range(4))
def test_pack_2bit_roundtrip(seed):
    from ferrousalign_amd import _native as N
    lib = N.lib()
    lib.fa_pack_2bit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    lib.fa_unpack_2bit.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    rng = np.random.default_rng(seed)
    for n in (0, 1, 15, 16, 17, 31, 32, 33, 151, 331, 2048):
        codes = rng.integers(0, 5, n).astype(np.uint8)
        nw = ((n + 31) // 32) * 2
        w2 = np.zeros(max(nw, 1), np.uint32)
        wn = np.zeros(max(nw // 2, 1), np.uint32)
        assert lib.fa_pack_2bit(N.ptr(codes), n, N.ptr(w2), N.ptr(wn)) == 0
        back = np.zeros(max(n, 1), np.uint8)
        assert lib.fa_unpack_2bit(N.ptr(w2), N.ptr(wn), n, N.ptr(back)) == 0
        assert (back[:n] == codes).all()
        # closed form of the layout: base p lives in bits 2*(p%16) of word p//16
        if n:
            p = n - 1
            c = codes[p]
            if c != 4:
                assert (int(w2[p // 16]) >> (2 * (p % 16))) & 3 == c
    bad = np.array([0, 5], np.uint8)
    assert lib.fa_pack_2bit(N.ptr(bad), 2, N.ptr(np.zeros(2, np.uint32)), N.ptr(np.zeros(1, np.uint32))) \
        == N.FA_ERR_CODE
