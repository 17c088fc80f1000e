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
    # the aggregator opens its own context: no device, no aggregator
    assert N.lib().fa_agg_create(0, 1, 0, ctypes.byref(h)) == N.FA_ERR_NODEVICE
    assert N.lib().fa_agg_create(0, 0, 0, ctypes.byref(h)) == N.FA_ERR_ARG
    assert N.lib().fa_agg_destroy(None) == N.FA_ERR_ARG
    assert N.lib().fa_agg_stats(None, None, None, None) == N.FA_ERR_ARG


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


def _np_pack(codes):
    """Plain restatement of the packed layout: base p at bits 2*(p%16) of word
    p//16 (N as code 0), N bit p%32 of mask word p//32, padding zero."""
    n = len(codes)
    nw = ((n + 31) // 32) * 2
    w2 = np.zeros(max(nw, 1), np.uint64)
    wn = np.zeros(max(nw // 2, 1), np.uint64)
    for p, c in enumerate(codes):
        if c == 4:
            wn[p // 32] |= np.uint64(1 << (p % 32))
        else:
            w2[p // 16] |= np.uint64(int(c) << (2 * (p % 16)))
    return w2.astype(np.uint32), wn.astype(np.uint32)


@pytest.mark.parametrize("seed", range(3))
def test_pack_2bit_words_and_code_check(seed):
    """Every packed word (padding included) equals the plain restatement, and
    a code above 4 anywhere (both sides of 8-byte boundaries, 5..255) is
    reported."""
    from ferrousalign_amd import _native as N
    lib = N.lib()
    lib.fa_pack_2bit.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(100 + seed)
    for n in list(range(0, 70)) + [127, 128, 129, 250, 1000]:
        codes = rng.integers(0, 5, n).astype(np.uint8)
        nw = ((n + 31) // 32) * 2
        w2 = np.full(max(nw, 1), 0xDEADBEEF, np.uint32)
        wn = np.full(max(nw // 2, 1), 0xDEADBEEF, np.uint32)
        assert lib.fa_pack_2bit(N.ptr(codes), n, N.ptr(w2), N.ptr(wn)) == 0
        e2, en = _np_pack(codes)
        assert np.array_equal(w2[:nw], e2[:nw]) and np.array_equal(wn[:nw // 2], en[:nw // 2]), n
        if n:
            for bad in (5, 6, 127, 128, 200, 255):
                c2 = codes.copy()
                c2[int(rng.integers(0, n))] = bad
                rc = lib.fa_pack_2bit(N.ptr(c2), n, N.ptr(w2), N.ptr(wn))
                assert rc == N.FA_ERR_CODE, (n, bad)
