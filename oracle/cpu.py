"""ctypes wrapper of the oracle's C restatement (oracle/cpu_ref.c) -- test
infrastructure and the bench's CPU-baseline leg only."""
from __future__ import annotations

import ctypes
import pathlib

import numpy as np

LIB = pathlib.Path(__file__).resolve().parent / "lib" / "liborc.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            from .build import build
            build()
        _lib = ctypes.CDLL(str(LIB))
        P = ctypes.POINTER
        args = [ctypes.c_int64, P(ctypes.c_double), ctypes.c_int, P(ctypes.c_double),
                P(P(ctypes.c_int64)), P(P(ctypes.c_int64)), P(ctypes.c_int64), P(ctypes.c_double),
                ctypes.c_int, P(ctypes.c_double), P(ctypes.c_double), ctypes.c_int, ctypes.c_int]
        _lib.cpu_clenshaw_filter.argtypes = args
        _lib.cpu_clenshaw_filter.restype = ctypes.c_int
        _lib.cpu_clenshaw_filter_rows.argtypes = args[:10] + [P(ctypes.c_double), ctypes.c_int]
        _lib.cpu_clenshaw_filter_rows.restype = ctypes.c_int
    return _lib


class PackedCPU:
    """Packed diag + value buckets (0-based, from oracle.packing) on the CPU."""

    def __init__(self, diagonals: np.ndarray, buckets):
        self.D = int(diagonals.size)
        self.diag = np.ascontiguousarray(diagonals, dtype=np.float64)
        self.vals = np.ascontiguousarray([v for v, _, _ in buckets], dtype=np.float64)
        self.flat = [np.ascontiguousarray(f, dtype=np.int64) for _, f, _ in buckets]
        self.starts = [np.ascontiguousarray(s, dtype=np.int64) for _, _, s in buckets]
        self.flat_len = np.ascontiguousarray([f.size for f in self.flat], dtype=np.int64)
        PI64 = ctypes.POINTER(ctypes.c_int64)
        self._flat_p = (PI64 * len(self.flat))(*[f.ctypes.data_as(PI64) for f in self.flat])
        self._starts_p = (PI64 * len(self.starts))(*[s.ctypes.data_as(PI64) for s in self.starts])

    def _common(self):
        P = ctypes.POINTER(ctypes.c_double)
        return (self.D, self.diag.ctypes.data_as(P), len(self.flat), self.vals.ctypes.data_as(P),
                self._flat_p, self._starts_p,
                self.flat_len.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)))

    def filter(self, coeffs: np.ndarray, X: np.ndarray, nthreads: int = 0) -> np.ndarray:
        """Y = p(H) X, column-parallel (one thread per column)."""
        lib = _load()
        P = ctypes.POINTER(ctypes.c_double)
        X = np.asfortranarray(X, dtype=np.float64)
        if X.ndim == 1:
            X = X.reshape(-1, 1, order="F")
        Y = np.zeros_like(X, order="F")
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        rc = lib.cpu_clenshaw_filter(*self._common(), c.ctypes.data_as(P), c.size - 1,
                                     X.ctypes.data_as(P), Y.ctypes.data_as(P), X.shape[1],
                                     nthreads)
        if rc:
            raise MemoryError("cpu_clenshaw_filter failed")
        return Y

    def filter_rows(self, coeffs: np.ndarray, x: np.ndarray, nthreads: int = 0) -> np.ndarray:
        """y = p(H) x for one vector, row-chunk threaded."""
        lib = _load()
        P = ctypes.POINTER(ctypes.c_double)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        rc = lib.cpu_clenshaw_filter_rows(*self._common(), c.ctypes.data_as(P), c.size - 1,
                                          x.ctypes.data_as(P), y.ctypes.data_as(P), nthreads)
        if rc:
            raise MemoryError("cpu_clenshaw_filter_rows failed")
        return y
