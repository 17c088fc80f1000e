"""ctypes bindings of the C ABI declared in include/ferrousalign_amd.h.

The product path has no CPU fallback: if the HIP library is missing, or no
gfx950 device can be opened, the calls raise.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
# FA_LIBRARY selects another build of the same library (the host-sanitized
# one that `python -m ferrousalign_amd.build --sanitize` tests with)
LIB_PATH = os.environ.get("FA_LIBRARY") or os.path.join(_PKG, "libferrousalign_amd.so")
SYNTH_PATH = os.path.join(_PKG, "libfa_synth.so")

FA_OK = 0
FA_ERR_ARG = -1
FA_ERR_H0 = -2
FA_ERR_QLEN = -3
FA_ERR_CODE = -4
FA_ERR_TLEN = -5
FA_ERR_BAND = -6
FA_ERR_SCORING = -7
FA_ERR_DEVICE = -8
FA_ERR_NODEVICE = -9
FA_ERR_OVERFLOW = -10
FA_ERR_SEGMENT = -11
FA_ERR_SCRATCH = -12

FA_MAX_QLEN = 32767
FA_MAX_TLEN = 65536
FA_MAX_QLEN_PACKED = 255
FA_MAX_TLEN_PACKED = 2048

FA_SCORE_I16 = 0x0
FA_SCORE_U8 = 0x1
FA_SCORE_AUTO = 0x2
FA_SCORE_I32 = 0x4
FA_PACK_LENGTH = 0x00
FA_PACK_NONE = 0x10
FA_PACK_COST = 0x20
FA_PACK_COST_ORDER = 0x30
FA_LAUNCH_PLAIN = 0x000
FA_LAUNCH_PERSISTENT = 0x100
FA_LAUNCH_GRAPH = 0x200
FA_LAUNCH_SINGLE = 0x400
FA_RUN_ALIGN2 = 0x1000
FA_PLAN_FRESH = 0x2000
FA_ALIGN2_FUSED = 0x4000


class FaError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = _strerror(code) if _lib is not None else str(code)
        super().__init__("%s: %s (code %d)" % (what, msg, code) if what else "%s (code %d)" % (msg, code))


class BandedPairwiseSW(ctypes.Structure):
    """fa_banded_pairwise_sw == BandedPairWiseSW (banded_swa/types.rs:106-199)."""
    _fields_ = [
        ("m", ctypes.c_int32),
        ("end_bonus", ctypes.c_int32),
        ("zdrop", ctypes.c_int32),
        ("pen_clip5", ctypes.c_int32),
        ("pen_clip3", ctypes.c_int32),
        ("o_del", ctypes.c_int32),
        ("o_ins", ctypes.c_int32),
        ("e_del", ctypes.c_int32),
        ("e_ins", ctypes.c_int32),
        ("mat", ctypes.c_int8 * 25),
    ]


# Record layouts of include/ferrousalign_amd.h (numpy structured dtypes)
MEM_OPT_DTYPE = np.dtype([(f, np.int32) for f in ("a", "b", "o_del", "e_del", "o_ins", "e_ins", "w", "zdrop",
                                                    "pen_clip5", "pen_clip3")])
SEED_DTYPE = np.dtype([("ref_pos", np.int64), ("query_pos", np.int32), ("len", np.int32)])
JOB_DTYPE = np.dtype([("read_idx", np.int32), ("chain_idx", np.int32), ("seed_idx", np.int64), ("qbeg", np.int32),
                      ("qlen", np.int32), ("h0", np.int32), ("w", np.int32), ("seg_beg", np.int64),
                      ("seg_len", np.int32), ("seg_rev", np.int32)])
REGION_DTYPE = np.dtype([("rb", np.int64), ("re", np.int64), ("qb", np.int32), ("qe", np.int32), ("score", np.int32),
                         ("truesc", np.int32), ("seedcov", np.int32), ("seedlen0", np.int32), ("chain_idx", np.int32),
                         ("w", np.int32), ("seed_idx", np.int64), ("frac_rep", np.float32), ("pad", np.int32)])
assert MEM_OPT_DTYPE.itemsize == 40 and SEED_DTYPE.itemsize == 16
assert JOB_DTYPE.itemsize == 48 and REGION_DTYPE.itemsize == 64

# OutScore field order (banded_swa/types.rs:77-84)
OUT_FIELDS = ("score", "target_end_pos", "gtarget_end_pos", "query_end_pos",
              "global_score", "max_offset")

_lib = None
_synth = None

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U32 = ctypes.c_uint32


def _declare(lib):
    lib.fa_strerror.restype = ctypes.c_char_p
    lib.fa_strerror.argtypes = [ctypes.c_int]
    lib.fa_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_P)]
    lib.fa_ctx_destroy.argtypes = [_P]
    lib.fa_ctx_describe.argtypes = [_P, ctypes.c_char_p, ctypes.c_size_t]
    lib.fa_valu_peak.argtypes = [_P, _I32, ctypes.POINTER(ctypes.c_double)]
    lib.fa_hbm_copy.argtypes = [_P, _I64, ctypes.POINTER(ctypes.c_double)]
    lib.fa_ref_upload.argtypes = [_P, _P, _I64, ctypes.POINTER(_P)]
    lib.fa_ref_free.argtypes = [_P, _P]
    lib.fa_ref_length.argtypes = [_P, ctypes.POINTER(_I64)]
    lib.fa_collect_extension_jobs.argtypes = [_P, _I64, _I64] + [_P] * 8 + [_I64, _P, ctypes.POINTER(_I64), _P,
                                                                         ctypes.POINTER(_I64)] + [_P] * 6
    lib.fa_merge_scores_to_regions.argtypes = [_P, _I64] + [_P] * 16 + [_I32, _P]
    chain_in = [_I64, _P, _I64, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _I64, _P, _I64, _P]
    lib.fa_extend_chains.argtypes = ([_P, _P, ctypes.POINTER(BandedPairwiseSW), _P] + chain_in +
                                     [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), _P])
    lib.fa_chains_upload.argtypes = [_P] + chain_in + [ctypes.POINTER(_P)]
    lib.fa_chains_extend.argtypes = [_P, _P, _P, ctypes.POINTER(BandedPairwiseSW), _P, ctypes.POINTER(_I64),
                                     ctypes.POINTER(_I64), _P]
    lib.fa_chains_download.argtypes = [_P, _P, _P]
    lib.fa_chains_free.argtypes = [_P, _P]
    lib.fa_batch_upload_ref.argtypes = [_P, ctypes.POINTER(BandedPairwiseSW), _I64, _P, _P, _P, _P, _P, _P, _P,
                                        _P, _P, ctypes.POINTER(_P)]
    lib.fa_batch_download_targets.argtypes = [_P, _P, _P, _P]
    lib.fa_scalar_banded_swa_batch.argtypes = [_P, ctypes.POINTER(BandedPairwiseSW), _I64] + [_P] * 18
    lib.fa_scalar_banded_swa_batch_dense.argtypes = ([_P, ctypes.POINTER(BandedPairwiseSW), _I64] + [_P] * 11
                                                     + [_I64, _P, _P, _P, _P, _P, _I64])
    lib.fa_scalar_banded_swa_last_ms.argtypes = [_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    seqargs = [_I64, _P, _P, _P, _P, _P, _P, _P, _P]
    lib.fa_execute_batch_simd_scoring.argtypes = [_P, ctypes.POINTER(BandedPairwiseSW)] + seqargs + [_U32, _P, _P, _P]
    lib.fa_batch_upload.argtypes = [_P, ctypes.POINTER(BandedPairwiseSW)] + seqargs + [ctypes.POINTER(_P)]
    lib.fa_execute_batch_device_noalloc.argtypes = ([_P, ctypes.POINTER(BandedPairwiseSW)] + seqargs +
                                                    [_I32, _P, _P, _P])
    lib.fa_batch_set_cost.argtypes = [_P, _P, _P]
    lib.fa_batch_set_launch.argtypes = [_P, _P, _I32, _I32, _I32]
    lib.fa_batch_run.argtypes = [_P, _P, _U32]
    lib.fa_batch_set_streams.argtypes = [_P, _P, _I32]
    lib.fa_batch_set_align2.argtypes = [_P, _P, _P, _I32, _I32, _I32, _I32, _I32, _U32]
    lib.fa_batch_download_align2.argtypes = [_P, _P, _P, _P]
    lib.fa_ksw_align2_batch.argtypes = [_P, _I64, _P, _P, _P, _P, _P, _P, _I32, _P, _I32, _I32, _I32, _I32, _U32,
                                        _P, _P]
    lib.fa_batch_time.argtypes = [_P, _P, _U32, _I32, ctypes.POINTER(ctypes.c_double),
                                  ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I32)]
    lib.fa_batch_download.argtypes = [_P, _P, _P, _P]
    lib.fa_batch_profile_launches.argtypes = [_P, _P, _U32, _I32, ctypes.POINTER(_I32), _P, _P, _P, _P, _P]
    lib.fa_batch_get_order.argtypes = [_P, _P, _U32, _P]
    lib.fa_batch_overflow_count.argtypes = [_P, _P, ctypes.POINTER(_I64)]
    lib.fa_batch_free.argtypes = [_P, _P]
    lib.fa_agg_create.argtypes = [ctypes.c_int, _I32, _I32, ctypes.POINTER(_P)]
    lib.fa_agg_destroy.argtypes = [_P]
    lib.fa_agg_execute.argtypes = [_P, ctypes.POINTER(BandedPairwiseSW)] + seqargs + [_U32, _P, _P, _P]
    lib.fa_agg_stats.argtypes = [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(_I64)]
    soa = [_P, _I32, _P, _P, _P, _P, _P, _P, _I32, _I32, _I32, _I32, _I32, _I32, _I32, _P, _I32, _P]
    lib.fa_simd_banded_swa_batch_soa.argtypes = soa
    lib.fa_simd_banded_swa_batch_int16_soa.argtypes = soa
    for name in ("fa_ctx_create", "fa_ctx_destroy", "fa_ctx_describe", "fa_valu_peak", "fa_hbm_copy", "fa_ref_upload", "fa_ref_free",
                 "fa_ref_length", "fa_collect_extension_jobs", "fa_merge_scores_to_regions", "fa_extend_chains",
                 "fa_chains_upload", "fa_chains_extend", "fa_chains_download", "fa_chains_free",
                 "fa_batch_upload_ref", "fa_batch_download_targets", "fa_scalar_banded_swa_batch",
                 "fa_scalar_banded_swa_batch_dense", "fa_scalar_banded_swa_last_ms", "fa_execute_batch_simd_scoring",
                 "fa_execute_batch_device_noalloc",
                 "fa_batch_upload", "fa_batch_set_cost", "fa_batch_set_launch", "fa_batch_set_streams", "fa_batch_run", "fa_batch_set_align2",
                 "fa_batch_download_align2", "fa_ksw_align2_batch",
                 "fa_batch_time", "fa_batch_download", "fa_batch_profile_launches", "fa_batch_get_order", "fa_batch_overflow_count", "fa_batch_free",
                 "fa_simd_banded_swa_batch_soa", "fa_simd_banded_swa_batch_int16_soa",
                 "fa_agg_create", "fa_agg_destroy", "fa_agg_execute", "fa_agg_stats"):
        getattr(lib, name).restype = ctypes.c_int


def lib():
    """The loaded HIP library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FaError(FA_ERR_NODEVICE, "libferrousalign_amd.so is not built (run __graft_entry__.build())")
        l = ctypes.CDLL(LIB_PATH)
        _declare(l)
        _lib = l
    return _lib


def synth():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise FaError(FA_ERR_ARG, "libfa_synth.so is not built")
        s = ctypes.CDLL(SYNTH_PATH)
        s.fa_synth_reference.argtypes = [ctypes.c_uint64, _I64, _P]
        s.fa_synth_reference.restype = None
        s.fa_synth_extension_pairs.argtypes = [ctypes.c_int, ctypes.c_uint64, _P, _I64, _I64, _P, _I64, _P,
                                               _P, _I64, _P, _P, _P, _I32, _I32, _I32, ctypes.c_int, _P]
        s.fa_synth_extension_pairs.restype = ctypes.c_int
        s.fa_synth_mate_pairs.argtypes = [ctypes.c_uint64, _P, _I64, _I64, _P, _P, _P, ctypes.c_int]
        s.fa_synth_region_pairs.argtypes = [ctypes.c_uint64, _P, _I64, _I64, _P, _I64, _P, _P, _I64, _P, _P,
                                            ctypes.c_int]
        s.fa_synth_region_pairs.restype = ctypes.c_int
        s.fa_synth_mate_pairs.restype = ctypes.c_int
        s.fa_synth_chain_reads.argtypes = [ctypes.c_uint64, _P, _I64, _I64, _I32, _I32, _I32, _P, _P, _P, _P, _P,
                                           _P, _P, _P, _P, ctypes.c_int]
        s.fa_synth_chain_reads.restype = ctypes.c_int
        _synth = s
    return _synth


def _strerror(code):
    return lib().fa_strerror(code).decode()


def check(rc, what=""):
    if rc != FA_OK:
        raise FaError(rc, what)
    return rc


def ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)
