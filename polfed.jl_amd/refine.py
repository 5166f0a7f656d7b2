"""FP64 polish of Ritz pairs from a Krylov basis built with FP32 filters.

Not in the reference: it serves the FP32-filter form (``TransformConfig``
dtype "mixed"/"f32"), whose Krylov vectors carry the FP32 rounding of the
filter.  That rounding is white across the spectrum: a Ritz vector's error
along eigenvectors far from the target dominates its residual
||H psi - E psi|| (the far components are weighted by |lambda - E|), while the
components near the target are what the filtered Lanczos resolves.  The
polish removes the far part with one FP64 application of a low-degree
Jackson-damped Chebyshev window q centred on the target to the wanted Ritz
vectors, then takes Rayleigh-Ritz pairs of the unscaled H
on span(q(H~) Y) (the extraction of getvecs.jl / getvals.jl,
src/Lanczos/Convergence/getvecs.jl:10-53, getvals.jl:10-36, on a different
search space).  Everything here runs in FP64 on the device except the small
projected eigenproblem.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import blockops as bo
from . import spectral as spc


def window_coefficients(target: float, degree: int) -> np.ndarray:
    """Jackson-damped Chebyshev expansion of a delta at ``target``:
    g_n (2 - [n = 0]) cos(n acos t), normalized to 1 at the target."""
    n = np.arange(degree + 1)
    g = np.array([spc.jackson(int(k), degree) for k in n])
    c = g * np.where(n == 0, 1.0, 2.0) * np.cos(n * math.acos(target))
    return c / float(spc.series_value(c, target))


def polish_degree(order: int, fraction: float = 1.0 / 32.0, minimum: int = 32) -> int:
    """Degree of the polish window: a fraction of the main filter's order."""
    return max(minimum, int(order * fraction))


def _orthonormalize(dev, Z: torch.Tensor) -> torch.Tensor:
    """Cholesky-QR2 of the (n, D) block Z on the device (Gram on the FP64
    matrix cores, the n x n Cholesky factor on the host); Householder QR on
    the host if the Gram matrix is numerically singular."""
    for _ in range(2):
        G = bo.gram(dev, Z, Z).cpu().numpy().T
        G = (G + G.T) / 2
        try:
            R = np.linalg.cholesky(G).T
        except np.linalg.LinAlgError:
            Q, _ = np.linalg.qr(Z.detach().cpu().numpy().T)
            return torch.from_numpy(np.ascontiguousarray(Q.T)).to(Z.device)
        Rinv = np.linalg.solve(R, np.eye(R.shape[0]))
        W = torch.zeros_like(Z)
        bo.update(dev, Z, torch.from_numpy(np.ascontiguousarray(Rinv.T)).to(Z.device), W,
                  alpha=1.0, beta=0.0)
        Z = W
    return Z


def polish(dev, op, Y: torch.Tensor, keep: int, *, target: float, order: int, coeffs: np.ndarray,
           polytype: str, which: str, fraction: float = 1.0 / 32.0):
    """Polished pairs from the Ritz block Y ((n, D) device tensor, n >= keep:
    converged Ritz vectors only, see lanczos.lanczos).

    Returns (vals, vecs, res, pvals): the ``keep`` Rayleigh-Ritz pairs of the
    unscaled H on span(q(H~) Y) that rank first under ``which`` by the value of
    the main filter p (``coeffs``) at their rescaled energy -- the order in
    which the Lanczos run on p(H~) found them -- their residual norms and
    their filter values."""
    from .lanczos import sortvals
    m = polish_degree(order, fraction)
    q = window_coefficients(target, m)
    Z = op.filter(q, Y, dtype="f64")
    Z = _orthonormalize(dev, Z)
    HZ = op.apply(Z, rescaled=False)
    A = bo.gram(dev, Z, HZ).cpu().numpy().T
    A = (A + A.T) / 2
    E, U = np.linalg.eigh(A)
    a = (op.Emax - op.Emin) / 2
    b = (op.Emax + op.Emin) / 2
    x = np.clip((E - b) / a, -1.0, 1.0)
    pv = spc.series_value(coeffs, x, polytype)
    sel = sortvals(pv, which)[:keep]
    Usel = torch.from_numpy(np.ascontiguousarray(U[:, sel].T)).to(Z.device)
    psi = torch.zeros((len(sel), Z.shape[1]), dtype=torch.float64, device=Z.device)
    bo.update(dev, Z, Usel, psi, alpha=1.0, beta=0.0)
    vals, res = bo.residuals(dev, op, psi)
    return vals, psi, res, pv[sel]
