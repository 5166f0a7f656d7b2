"""ctypes binding of libpolfed_hip (the C ABI declared in include/polfed_hip.h).

The library is built in-tree (``polfed.jl_amd/lib/libpolfed_hip.so``) by
``polfed_amd.build.build_library``.  Loading fails loudly when it is missing:
there is no CPU fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import pathlib

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libpolfed_hip.so"

PF_OK = 0
PF_ERR_INVALID = -1
PF_ERR_NOMEM = -2
PF_ERR_HIP = -3
PF_ERR_UNSUPPORTED = -4
PF_ERR_NUMERIC = -5

PF_F64 = 0
PF_F32 = 1

PF_CHEBYSHEV = 0
PF_LEGENDRE = 1
PF_HERMITE = 2
PF_TAYLOR = 3

PF_TUNE_XCD_REMAP = 1
PF_TUNE_VEC = 2
PF_TUNE_RESET = 3
PF_TUNE_NONTEMPORAL = 4
PF_TUNE_PBLOCKS = 5
PF_TUNE_WIDE = 6
PF_TUNE_GRAIN_SHUFFLE = 7
PF_TUNE_OCCUPANCY = 8
PF_TUNE_ROW_PARTITION = 9

PF_BATCH_GRAPH = 1
PF_BATCH_GRAPH_STEPS = 2
PF_BATCH_VEC = 3
PF_BATCH_ROWS_PER_LANE = 4
PF_BATCH_XCD = 5
PF_BATCH_GRAIN_SHUFFLE = 6
PF_BATCH_LDS_ROWS = 7
PF_BATCH_ROW_VECTORS = 8

PF_PROBE_GAUSSIAN = 0
PF_PROBE_PHASE = 1

POLYTYPES = {"Chebyshev": PF_CHEBYSHEV, "Legendre": PF_LEGENDRE, "Hermite": PF_HERMITE,
             "Taylor": PF_TAYLOR}


class PolfedHIPError(RuntimeError):
    """A libpolfed_hip call failed (status code + library message)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"libpolfed_hip error {code}: {msg}")
        self.code = code


class NumericBreakdown(PolfedHIPError):
    """PF_ERR_NUMERIC: e.g. Cholesky-QR of a rank-deficient block."""


_c = ctypes
_vp = _c.c_void_p
_i32 = _c.c_int32
_i64 = _c.c_int64
_u64 = _c.c_uint64
_dbl = _c.c_double
_pdbl = _c.POINTER(_c.c_double)
_pi32 = _c.POINTER(_c.c_int32)

# name -> (restype, argtypes); every function returns an int status unless noted
SIGNATURES = {
    "pf_version": (_c.c_int, []),
    "pf_last_error": (_c.c_char_p, []),
    "pf_create": (_c.c_int, [_c.POINTER(_vp), _c.c_int]),
    "pf_destroy": (_c.c_int, [_vp]),
    "pf_op_create_spin_chain": (_c.c_int, [_vp, _i32, _i32, _i32, _pi32, _pi32, _pdbl, _pdbl,
                                           _pdbl, _c.POINTER(_vp)]),
    "pf_op_create_qsun": (_c.c_int, [_vp, _i32, _i32, _dbl, _pdbl, _pi32, _pdbl, _pdbl,
                                     _c.POINTER(_vp)]),
    "pf_op_create_packed": (_c.c_int, [_vp, _i64, _pdbl, _i32, _pdbl, _c.POINTER(_i64), _pi32,
                                       _i64, _c.POINTER(_vp)]),
    "pf_op_destroy": (_c.c_int, [_vp]),
    "pf_op_dim": (_i64, [_vp]),
    "pf_op_set_rescale": (_c.c_int, [_vp, _dbl, _dbl]),
    "pf_op_connectivity": (_c.c_int, [_vp, _i64, _i64, _i32, _vp, _vp, _vp, _vp, _vp]),
    "pf_op_connectivity_order": (_c.c_int, [_vp, _i32, _i64, _i64, _i32, _vp, _vp, _vp, _vp,
                                            _vp]),
    "pf_apply": (_c.c_int, [_vp, _i32, _i32, _vp, _i64, _vp, _i64, _i32, _vp]),
    "pf_clenshaw_recurrence": (_c.c_int, [_vp, _i32, _vp, _vp, _vp, _dbl, _dbl, _dbl, _vp, _i64,
                                          _i32, _vp]),
    "pf_clenshaw_finalsum": (_c.c_int, [_vp, _i32, _vp, _vp, _dbl, _dbl, _dbl, _dbl, _vp, _vp,
                                        _i64, _i32, _vp]),
    "pf_filter": (_c.c_int, [_vp, _i32, _i32, _pdbl, _i32, _vp, _i64, _vp, _i64, _i32, _vp]),
    "pf_filter_noalloc": (_c.c_int, [_vp, _i32, _i32, _pdbl, _i32, _vp, _i64, _vp, _i64, _i32,
                                     _vp]),
    "pf_op_set_tuning": (_c.c_int, [_vp, _i32, _i32]),
    "pf_autotune": (_c.c_int, [_vp, _i32, _i32, _i32, _pdbl, _pi32, _pi32, _vp]),
    "pf_op_get_tuning": (_c.c_int, [_vp, _i32, _i32, _pi32]),
    "pf_fill_probes": (_c.c_int, [_i32, _u64, _i64, _i32, _vp, _i64, _vp]),
    "pf_kpm_moments": (_c.c_int, [_vp, _i32, _i32, _i32, _vp, _i64, _pdbl, _vp]),
    "pf_gram": (_c.c_int, [_vp, _vp, _i64, _i32, _vp, _i64, _i32, _i64, _vp, _i32, _vp]),
    "pf_gemm_update": (_c.c_int, [_vp, _vp, _i64, _i32, _vp, _i32, _vp, _i64, _i32, _i64, _dbl,
                                  _dbl, _vp]),
    "pf_reorth": (_c.c_int, [_vp, _vp, _i64, _i32, _vp, _i64, _i32, _i64, _i32, _vp]),
    "pf_block_qr": (_c.c_int, [_vp, _vp, _i64, _i64, _i32, _pdbl, _vp]),
    "pf_rotate": (_c.c_int, [_vp, _vp, _i64, _i64, _i32, _pdbl, _vp]),
    "pf_residuals": (_c.c_int, [_vp, _vp, _vp, _i64, _i32, _pdbl, _pdbl, _vp]),
    "pf_coldots": (_c.c_int, [_vp, _vp, _i64, _vp, _i64, _i64, _i32, _pdbl, _vp]),
    "pf_batch_create": (_c.c_int, [_vp, _i32, _c.POINTER(_vp), _i32, _c.POINTER(_vp)]),
    "pf_batch_destroy": (_c.c_int, [_vp]),
    "pf_batch_set_filter": (_c.c_int, [_vp, _i32, _pi32, _pdbl]),
    "pf_batch_filter": (_c.c_int, [_vp, _pi32, _pi32, _c.POINTER(_vp), _c.POINTER(_i64),
                                   _c.POINTER(_vp), _c.POINTER(_i64), _vp]),
    "pf_batch_set_tuning": (_c.c_int, [_vp, _i32, _i32]),
    "pf_batch_get_tuning": (_c.c_int, [_vp, _pi32]),
    "pf_batch_autotune": (_c.c_int, [_vp, _i32, _i32, _pdbl, _vp]),
}

_LIB = None


def load(path: pathlib.Path | str | None = None):
    """Load (once) and return the ctypes library; raise if it is not built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = pathlib.Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise PolfedHIPError(PF_ERR_UNSUPPORTED,
                             f"{p} is missing: build it with polfed_amd.build.build_library() "
                             "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ctypes.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(status: int) -> None:
    """Raise PolfedHIPError for a non-zero status from the library."""
    if status != PF_OK:
        msg = load().pf_last_error().decode(errors="replace")
        if status == PF_ERR_NUMERIC:
            raise NumericBreakdown(status, msg)
        raise PolfedHIPError(status, msg)


def dptr(arr) -> ctypes.POINTER(ctypes.c_double):
    """Host float64 numpy array -> double*."""
    return arr.ctypes.data_as(_pdbl)


def iptr(arr) -> ctypes.POINTER(ctypes.c_int32):
    """Host int32 numpy array -> int32_t*."""
    return arr.ctypes.data_as(_pi32)
