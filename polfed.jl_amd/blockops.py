"""FP64 block operations of the block-Lanczos step (wrappers of the dense C ABI)."""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, dptr
from .device import Device, colmajor

MAX_N = 64  # widest block of one pf_gram / pf_gemm_update call


def gram(dev: Device, A: torch.Tensor, B: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
    """G = A^T B (m x n) on the device; returned as an (n, m) tensor = column-major m x n."""
    pa, lda, D, m = colmajor(A)
    _, _, D2, n = colmajor(B)
    if D != D2:
        raise ValueError("gram: row counts differ")
    if out is None:
        out = torch.empty((n, m), dtype=torch.float64, device=A.device)
    for j0 in range(0, n, MAX_N):           # B columns are independent
        Bj = B[j0:j0 + MAX_N] if B.dim() == 2 else B
        pb, ldb, _, nj = colmajor(Bj)
        pg, ldg, _, _ = colmajor(out[j0:j0 + nj])
        check(dev.lib.pf_gram(dev.h, pa, lda, m, pb, ldb, nj, D, pg, ldg, dev.stream))
    return out


def update(dev: Device, A: torch.Tensor, G: torch.Tensor, B: torch.Tensor, alpha: float = 1.0,
           beta: float = 1.0) -> None:
    """B = beta B + alpha A G with G given as an (n, m) tensor (column-major m x n)."""
    pa, lda, D, m = colmajor(A)
    _, _, m2, n = colmajor(G)
    _, _, D2, n2 = colmajor(B)
    if m != m2 or n != n2 or D != D2:
        raise ValueError("update: shape mismatch")
    for j0 in range(0, n, MAX_N):           # output columns are independent
        Gj = G[j0:j0 + MAX_N] if G.dim() == 2 else G
        Bj = B[j0:j0 + MAX_N] if B.dim() == 2 else B
        pg, ldg, _, nj = colmajor(Gj)
        pb, ldb, _, _ = colmajor(Bj)
        check(dev.lib.pf_gemm_update(dev.h, pa, lda, m, pg, ldg, pb, ldb, nj, D, float(alpha),
                                     float(beta), dev.stream))


def reorth(dev: Device, V: torch.Tensor, W: torch.Tensor, passes: int = 2) -> None:
    """W -= V (V^T W), ``passes`` times (classical Gram-Schmidt, 2 = CGS2)."""
    pv, ldv, D, k = colmajor(V)
    pw, ldw, D2, s = colmajor(W)
    if D != D2:
        raise ValueError("reorth: row counts differ")
    check(dev.lib.pf_reorth(dev.h, pv, ldv, k, pw, ldw, s, D, passes, dev.stream))


def block_qr(dev: Device, W: torch.Tensor) -> np.ndarray:
    """In-place Cholesky-QR2 of W; returns R (s x s, upper).  Raises
    ``NumericBreakdown`` (W unchanged) for numerically rank-deficient blocks."""
    pw, ldw, D, s = colmajor(W)
    R = np.zeros((s, s), order="F")
    Rf = np.zeros(s * s)
    check(dev.lib.pf_block_qr(dev.h, pw, ldw, D, s, dptr(Rf), dev.stream))
    R[:, :] = Rf.reshape((s, s), order="F")
    return R


def rotate(dev: Device, W: torch.Tensor, M: np.ndarray) -> None:
    """W <- W M in place for a small n x n host matrix M."""
    pw, ldw, D, n = colmajor(W)
    # column-major M: element (i, j) at i + j*n
    Mf = np.ascontiguousarray(np.asarray(M, dtype=np.float64).reshape(-1, order="F"))
    check(dev.lib.pf_rotate(dev.h, pw, ldw, D, n, dptr(Mf), dev.stream))


def coldots(dev: Device, A: torch.Tensor, B: torch.Tensor) -> np.ndarray:
    """<A_j, B_j> for every column j (FP64, deterministic order)."""
    pa, lda, D, n = colmajor(A)
    pb, ldb, D2, n2 = colmajor(B)
    if D != D2 or n != n2:
        raise ValueError("coldots: shape mismatch")
    out = np.zeros(n)
    check(dev.lib.pf_coldots(dev.h, pa, lda, pb, ldb, D, n, dptr(out), dev.stream))
    return out


def residuals(dev: Device, op, psi: torch.Tensor):
    """(E, res): E_i = <psi_i|H|psi_i>, res_i = ||H psi_i - E_i psi_i|| (unscaled H)."""
    pp, ld, D, n = colmajor(psi)
    E = np.zeros(n)
    res = np.zeros(n)
    check(dev.lib.pf_residuals(dev.h, op.op, pp, ld, n, dptr(E), dptr(res), dev.stream))
    return E, res


def fill_probes(dev: Device, D: int, R: int, seed: int, kind: str = "gaussian") -> torch.Tensor:
    """R probe columns from the counter-based generator (not normalized)."""
    k = {"gaussian": _lib.PF_PROBE_GAUSSIAN, "phase": _lib.PF_PROBE_PHASE}[kind]
    P = dev.block(D, R)
    pp, ld, _, _ = colmajor(P)
    check(dev.lib.pf_fill_probes(k, ctypes.c_uint64(seed & (2**64 - 1)), D, R, pp, ld, dev.stream))
    return P
