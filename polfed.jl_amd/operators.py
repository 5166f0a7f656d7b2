"""Matrix-free Hamiltonians on the GPU (wrappers of ``pf_op``).

These replace the matrices the reference builds in ``src/Models`` and
``src/QSun`` and the packed operator of ``src/Optimization``: the HIP kernels
compute every matrix element from the bit pattern of the row's state, so no
sparse matrix is ever stored.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, dptr, iptr
from .device import Device, colmajor, default_device

_DT = {"f64": (_lib.PF_F64, torch.float64), "f32": (_lib.PF_F32, torch.float32),
       torch.float64: (_lib.PF_F64, torch.float64), torch.float32: (_lib.PF_F32, torch.float32)}


def _dtype(dtype):
    if dtype not in _DT:
        raise ValueError(f"dtype must be 'f64' or 'f32', got {dtype!r}")
    return _DT[dtype]


class Operator:
    """Base wrapper: owns a pf_op on one Device."""

    def __init__(self, device: Device, handle: ctypes.c_void_p):
        self.device = device
        self.lib = device.lib
        self.op = handle
        self.D = int(self.lib.pf_op_dim(handle))
        self.Emin, self.Emax = -1.0, 1.0

    def close(self):
        if getattr(self, "op", None):
            self.lib.pf_op_destroy(self.op)
            self.op = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------- scaling
    def set_rescale(self, Emin: float, Emax: float) -> None:
        """H~ = (H - b)/a with a = (Emax-Emin)/2, b = (Emax+Emin)/2."""
        check(self.lib.pf_op_set_rescale(self.op, float(Emin), float(Emax)))
        self.Emin, self.Emax = float(Emin), float(Emax)

    @property
    def a(self) -> float:
        return (self.Emax - self.Emin) / 2

    @property
    def b(self) -> float:
        return (self.Emax + self.Emin) / 2

    # ------------------------------------------------------------ mappings
    def apply(self, X: torch.Tensor, Y: torch.Tensor | None = None, rescaled: bool = False,
              dtype="f64") -> torch.Tensor:
        """Y = H X (or H~ X); X, Y blocks of the given dtype."""
        code, tdt = _dtype(dtype)
        if Y is None:
            Y = torch.empty_like(X)
        px, ldx, D, n = colmajor(X, tdt)
        py, ldy, _, _ = colmajor(Y, tdt)
        self._check_rows(D)
        check(self.lib.pf_apply(self.op, code, int(rescaled), px, ldx, py, ldy, n,
                                self.device.stream))
        return Y

    def __call__(self, Y: torch.Tensor, X: torch.Tensor) -> None:
        """``f!(Y, X)`` calling convention of the reference (Y = H X)."""
        self.apply(X, Y)

    def recurrence(self, b1, b2, b3, c: float, X, alpha: float = 2.0, beta: float = -1.0,
                   dtype="f64") -> None:
        """b1 = c X + alpha H~ b2 + beta b3 (the clenshaw_recurrence hook)."""
        code, tdt = _dtype(dtype)
        p1, ld, D, n = colmajor(b1, tdt)
        p2, ld2, _, _ = colmajor(b2, tdt)
        px, ldx, _, _ = colmajor(X, tdt)
        p3 = colmajor(b3, tdt)[0] if b3 is not None else None
        if not (ld == ld2 == ldx):
            raise ValueError("recurrence blocks must share one leading dimension")
        self._check_rows(D)
        check(self.lib.pf_clenshaw_recurrence(self.op, code, p1, p2, p3, float(c), float(alpha),
                                              float(beta), px, ld, n, self.device.stream))

    def finalsum(self, b2, b3, c0: float, Y, X, beta1: float = -1.0, phi0: float = 1.0,
                 phi1: float = 1.0, dtype="f64") -> None:
        """Y = c0 phi0 X + phi1 H~ b2 + beta1 phi0 b3 (the clenshaw_finalsum hook)."""
        code, tdt = _dtype(dtype)
        py, ld, D, n = colmajor(Y, tdt)
        p2 = colmajor(b2, tdt)[0]
        px = colmajor(X, tdt)[0]
        p3 = colmajor(b3, tdt)[0] if b3 is not None else None
        self._check_rows(D)
        check(self.lib.pf_clenshaw_finalsum(self.op, code, p2, p3, float(c0), float(beta1),
                                            float(phi0), float(phi1), py, px, ld, n,
                                            self.device.stream))

    def filter(self, coeffs, X: torch.Tensor, Y: torch.Tensor | None = None, dtype="f64",
               polytype: str = "Chebyshev") -> torch.Tensor:
        """Y = p(H~) X with p = sum_n coeffs[n] P_n; X, Y FP64, arithmetic ``dtype``."""
        code, _ = _dtype(dtype)
        c = np.ascontiguousarray(coeffs, dtype=np.float64)
        if Y is None:
            Y = torch.empty_like(X)
        px, ldx, D, n = colmajor(X)
        py, ldy, _, _ = colmajor(Y)
        self._check_rows(D)
        check(self.lib.pf_filter(self.op, code, _lib.POLYTYPES[polytype], dptr(c), c.size - 1,
                                 px, ldx, py, ldy, n, self.device.stream))
        return Y

    def kpm_moments(self, N: int, probes: torch.Tensor, dtype="f64") -> np.ndarray:
        """mu_n = (D/R) sum_r <r|T_n(H~)|r> with each probe column normalized."""
        code, _ = _dtype(dtype)
        pp, ldp, D, R = colmajor(probes)
        self._check_rows(D)
        mu = np.zeros(N)
        check(self.lib.pf_kpm_moments(self.op, code, int(N), R, pp, ldp, dptr(mu),
                                      self.device.stream))
        return mu

    def set_tuning(self, xcd: int | None = None, vec: int | None = None, nt: int | None = None,
                   pblocks: int | None = None, wide: int | None = None,
                   grain_shuffle: int | None = None, occupancy: int | None = None,
                   row_partition: int | None = None, reset: bool = False):
        """Force launch knobs of the row-interleaved kernels (see pf_op_set_tuning)."""
        if reset:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_RESET, 0))
            self.__dict__.pop("_tuned", None)
        if pblocks is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_PBLOCKS, int(pblocks)))
        if wide is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_WIDE, int(wide)))
        if grain_shuffle is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_GRAIN_SHUFFLE, int(grain_shuffle)))
        if occupancy is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_OCCUPANCY, int(occupancy)))
        if row_partition is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_ROW_PARTITION, int(row_partition)))
        if xcd is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_XCD_REMAP, int(xcd)))
        if vec is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_VEC, int(vec)))
        if nt is not None:
            check(self.lib.pf_op_set_tuning(self.op, _lib.PF_TUNE_NONTEMPORAL, int(nt)))

    def autotune(self, ncols: int, dtype="f64", nsteps: int = 20):
        """Pick the fastest step-kernel launch configuration for (dtype, ncols)
        on this operator; returns (us per step, vector width, xcd mapping)."""
        code, _ = _dtype(dtype)
        us = ctypes.c_double()
        vec = ctypes.c_int32()
        xcd = ctypes.c_int32()
        check(self.lib.pf_autotune(self.op, code, int(ncols), int(nsteps), ctypes.byref(us),
                                   ctypes.byref(vec), ctypes.byref(xcd), self.device.stream))
        return us.value, vec.value, xcd.value

    def tuning(self, ncols: int, dtype="f64") -> dict:
        """The launch configuration the drivers use for (dtype, ncols): the
        autotuned one with any set_tuning overrides (pf_op_get_tuning)."""
        code, _ = _dtype(dtype)
        cfg = (ctypes.c_int32 * 8)()
        check(self.lib.pf_op_get_tuning(self.op, code, int(ncols), cfg))
        keys = ("vec", "xcd", "nt", "pblocks", "wide", "grain_shuffle", "occupancy",
                "row_partition")
        return dict(zip(keys, (int(v) for v in cfg)))

    def ensure_tuned(self, ncols: int, dtype="f64", nsteps: int = 30):
        """Autotune (dtype, ncols) once per operator."""
        done = self.__dict__.setdefault("_tuned", {})
        if (dtype, ncols) not in done:
            done[(dtype, ncols)] = self.autotune(ncols, dtype, nsteps)
        return done[(dtype, ncols)]

    def connectivity(self, row0: int = 0, nrows: int | None = None, max_per_row: int = 64,
                     order: int = 0):
        """Per-row off-diagonal (cols, vals, counts) and the unscaled diagonal,
        as computed by the device kernels (host numpy arrays): the step
        kernels' rank function and diagonal table.  ``order`` 1 takes the
        spin-chain XCD partition order's tables (rows and columns mapped back
        to the reference order)."""
        if nrows is None:
            nrows = self.D - row0
        dev = self.device.torch_device
        cols = torch.empty((nrows, max_per_row), dtype=torch.int64, device=dev)
        vals = torch.empty((nrows, max_per_row), dtype=torch.float64, device=dev)
        counts = torch.empty(nrows, dtype=torch.int32, device=dev)
        diag = torch.empty(nrows, dtype=torch.float64, device=dev)
        vp = lambda t: ctypes.c_void_p(t.data_ptr())
        check(self.lib.pf_op_connectivity_order(self.op, int(order), row0, nrows, max_per_row,
                                                vp(cols), vp(vals), vp(counts), vp(diag),
                                                self.device.stream))
        return (cols.cpu().numpy(), vals.cpu().numpy(), counts.cpu().numpy(),
                diag.cpu().numpy())

    def _check_rows(self, D: int) -> None:
        if D != self.D:
            raise ValueError(f"block has {D} rows, operator dimension is {self.D}")


class SpinChainOperator(Operator):
    """Spin-1/2 chain in a fixed-Lup sector with arbitrary two-site XXZ bonds
    (``_spin_half_hamiltonian``, src/Models/SpinHalfChains.jl:267-315)."""

    def __init__(self, L: int, Lup: int, bonds, couplings, anisotropies, fields,
                 device: Device | None = None):
        device = device or default_device()
        bi = np.ascontiguousarray([b[0] for b in bonds], dtype=np.int32)
        bj = np.ascontiguousarray([b[1] for b in bonds], dtype=np.int32)
        J = np.ascontiguousarray(couplings, dtype=np.float64)
        Dl = np.ascontiguousarray(anisotropies, dtype=np.float64)
        h = np.ascontiguousarray(fields, dtype=np.float64)
        if h.shape != (L,):
            raise ValueError(f"fields must have length L={L}; got length {h.size}.")
        op = ctypes.c_void_p()
        with torch.cuda.device(device.index):
            check(device.lib.pf_op_create_spin_chain(device.h, L, Lup, len(bonds), iptr(bi),
                                                     iptr(bj), dptr(J), dptr(Dl), dptr(h),
                                                     ctypes.byref(op)))
        super().__init__(device, op)
        self.L, self.Lup = L, Lup
        self.bonds = [tuple(b) for b in bonds]
        self.couplings, self.anisotropies, self.fields = J, Dl, h


class QSunOperator(Operator):
    """Quantum Sun Hamiltonian in the full space for spin 1/2
    (``build_hamiltonian``, src/QSun/QuantumSun.jl:1-102)."""

    def __init__(self, L_loc: int, L_grain: int, g0: float, disorder, neighbors, couplings,
                 H_grain, device: Device | None = None):
        device = device or default_device()
        G = np.asfortranarray(H_grain, dtype=np.float64)
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        cp = np.ascontiguousarray(couplings, dtype=np.float64)
        hz = np.ascontiguousarray(disorder, dtype=np.float64)
        op = ctypes.c_void_p()
        Gc = np.ascontiguousarray(G.reshape(-1, order="F"))
        with torch.cuda.device(device.index):
            check(device.lib.pf_op_create_qsun(device.h, L_grain, L_loc, float(g0), dptr(Gc),
                                               iptr(nb), dptr(cp), dptr(hz), ctypes.byref(op)))
        super().__init__(device, op)
        self.L_loc, self.L_grain, self.g0 = L_loc, L_grain, g0
        self.disorder, self.neighbors, self.couplings, self.H_grain = hz, nb, cp, G


class PackedOperator(Operator):
    """Any real symmetric matrix in the reference's value-bucket form
    (``optimize_spectral_transform``, src/Optimization/optimization.jl:75-119,
    and its packed row kernel, src/Optimization/cpu/single_row_mappings.jl:66-79)
    -- the operator ``polfed(mat, ...)`` runs on.  ``diag``, ``vals``,
    ``starts`` and ``flat`` as returned by :func:`packing.pack_by_value`."""

    def __init__(self, diag, vals, starts, flat, device: Device | None = None):
        device = device or default_device()
        dg = np.ascontiguousarray(diag, dtype=np.float64)
        D = dg.size
        vv = np.ascontiguousarray(vals, dtype=np.float64).reshape(-1)
        st = np.ascontiguousarray(starts, dtype=np.int64)
        fl = np.ascontiguousarray(flat, dtype=np.int32).reshape(-1)
        if st.shape != (vv.size, D + 1):
            raise ValueError(f"starts must have shape ({vv.size}, {D + 1}), got {st.shape}")
        if vv.size and st[-1, -1] > fl.size:
            raise ValueError("starts point past the end of flat")
        op = ctypes.c_void_p()
        i64p = st.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        with torch.cuda.device(device.index):
            check(device.lib.pf_op_create_packed(device.h, D, dptr(dg), vv.size, dptr(vv), i64p,
                                                 iptr(fl), fl.size, ctypes.byref(op)))
        super().__init__(device, op)
        self.diag, self.vals, self.starts, self.flat = dg, vv, st, fl

    @classmethod
    def from_matrix(cls, H, device: Device | None = None, tol: float = 1e-13,
                    round_digits: int = 15) -> "PackedOperator":
        """Pack a scipy sparse or dense matrix (``get_diags_and_offdiagonals_by_value``,
        src/Optimization/optimization.jl:134-182) and build the operator."""
        from .packing import pack_by_value
        return cls(*pack_by_value(H, tol, round_digits), device=device)

