"""Batched filters of many Quantum-Sun instances (wrapper of ``pf_batch``).

The reference solves one problem per ``polfed`` call and applies its filter
once per Lanczos iteration (``clenshaw(transform, W, V, b, ::GPU,
::NoParallel)``, src/Polfed/Algorithm/SpectralTransformation/ParallelizationClenshaw.jl:41-51).
A ``QSunBatch`` applies the filters of many instances -- disorder
realizations of possibly different sizes, each with its own polynomial and
rescaling -- as one multi-instance kernel launch per Clenshaw step, replayed
from a captured HIP graph.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check
from .device import colmajor
from .operators import QSunOperator, _dtype

_TUNE = {"graph": _lib.PF_BATCH_GRAPH, "graph_steps": _lib.PF_BATCH_GRAPH_STEPS,
         "vec": _lib.PF_BATCH_VEC, "rows_per_lane": _lib.PF_BATCH_ROWS_PER_LANE,
         "xcd": _lib.PF_BATCH_XCD, "grain_shuffle": _lib.PF_BATCH_GRAIN_SHUFFLE,
         "lds_rows": _lib.PF_BATCH_LDS_ROWS, "row_vectors": _lib.PF_BATCH_ROW_VECTORS}


class QSunBatch:
    """n Quantum-Sun operators (same device) filtered together, ``ncols``
    columns each.  The operators must stay alive while the batch is used."""

    def __init__(self, ops: Sequence[QSunOperator], ncols: int = 1):
        if not ops:
            raise ValueError("QSunBatch needs at least one operator")
        for op in ops:
            if not isinstance(op, QSunOperator):
                raise TypeError("QSunBatch batches Quantum-Sun operators only")
        dev = ops[0].device
        if any(op.device is not dev for op in ops):
            raise ValueError("QSunBatch: all operators must live on one Device")
        self.ops = list(ops)
        self.device = dev
        self.lib = dev.lib
        self.n = len(ops)
        self.ncols = int(ncols)
        arr = (ctypes.c_void_p * self.n)(*[op.op for op in ops])
        b = ctypes.c_void_p()
        with torch.cuda.device(dev.index):
            check(self.lib.pf_batch_create(dev.h, self.n, arr, self.ncols, ctypes.byref(b)))
        self.b = b
        self.orders = None

    def close(self):
        if getattr(self, "b", None):
            self.lib.pf_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_filter(self, coeffs: Sequence[np.ndarray], polytype: str = "Chebyshev") -> None:
        """Instance i's filter: ``coeffs[i]`` (order + 1 normalized values).
        Reads every operator's current rescaling."""
        if len(coeffs) != self.n:
            raise ValueError("set_filter: one coefficient array per instance")
        orders = np.array([len(c) - 1 for c in coeffs], dtype=np.int32)
        if np.any(orders < 0):
            raise ValueError("set_filter: empty coefficient array")
        flat = np.ascontiguousarray(np.concatenate([np.asarray(c, np.float64) for c in coeffs]))
        check(self.lib.pf_batch_set_filter(
            self.b, _lib.POLYTYPES[polytype],
            orders.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            flat.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        self.orders = orders

    def filter(self, requests: Sequence) -> None:
        """One batched application.  ``requests[i]`` is None (instance idle)
        or ``(X, Y, dtype)``: Y_i = p_i(H~_i) X_i, with (ncols, D_i) float64
        blocks and dtype "f64" or "f32" for the arithmetic."""
        if len(requests) != self.n:
            raise ValueError("filter: one request (or None) per instance")
        n = self.n
        active = np.zeros(n, np.int32)
        dtypes = np.zeros(n, np.int32)
        xs = (ctypes.c_void_p * n)()
        ys = (ctypes.c_void_p * n)()
        ldx = np.zeros(n, np.int64)
        ldy = np.zeros(n, np.int64)
        for i, req in enumerate(requests):
            if req is None:
                continue
            X, Y, dt = req
            px, lx, D, nc = colmajor(X)
            py, ly, D2, nc2 = colmajor(Y)
            if D != self.ops[i].D or D2 != D or nc != self.ncols or nc2 != self.ncols:
                raise ValueError(f"filter: instance {i} expects ({self.ncols}, {self.ops[i].D}) blocks")
            active[i] = 1
            dtypes[i] = _dtype(dt)[0]
            xs[i], ys[i] = px.value, py.value
            ldx[i], ldy[i] = lx, ly
        i64 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        i32 = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        check(self.lib.pf_batch_filter(self.b, i32(active), i32(dtypes), xs, i64(ldx), ys,
                                       i64(ldy), self.device.stream))

    def set_tuning(self, **kw) -> None:
        """graph (0/1), graph_steps, vec, rows_per_lane, xcd, grain_shuffle (0/1),
        lds_rows (0/1), row_vectors (0/1/2)."""
        for k, v in kw.items():
            if k not in _TUNE:
                raise ValueError(f"unknown batch tuning key {k!r}")
            check(self.lib.pf_batch_set_tuning(self.b, _TUNE[k], int(v)))

    def tuning(self) -> dict:
        """The launch settings in force (pf_batch_get_tuning); lds_used and
        row_vectors_used say whether those forms run (requested and allowed by
        the tiling; row vectors take precedence)."""
        cfg = (ctypes.c_int32 * 10)()
        check(self.lib.pf_batch_get_tuning(self.b, cfg))
        keys = ("vec", "rows_per_lane", "xcd", "grain_shuffle", "lds_rows", "lds_used",
                "graph", "graph_steps", "row_vectors", "row_vectors_used")
        return dict(zip(keys, list(cfg)))

    def autotune(self, dtype="f32", nsteps: int = 40) -> float:
        """Pick the fastest tiling and graph/stream mode; returns the time per
        batched step in microseconds.  Overwrites the batch work buffers."""
        us = ctypes.c_double()
        check(self.lib.pf_batch_autotune(self.b, _dtype(dtype)[0], int(nsteps), ctypes.byref(us),
                                         self.device.stream))
        return us.value
