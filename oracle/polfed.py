"""Oracle restatement of the POLFED pipeline on the CPU in FP64 (test infrastructure).

Restates ``polfed(f!, x0, howmany, target; ...)``
(src/Polfed/PolfedCore.jl:194-263), ``build_mapping_plan`` (Plan.jl:128-191),
``polfed_algorithm`` (src/Polfed/Algorithm/polfed_algorithm.jl:15-37) and
``_prepare_polfed_initial_state`` (PolfedCore.jl:38-95).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .clenshaw import clenshaw_apply, filter_coefficients
from .dos import dos_function, dos_moments
from .lanczos import LanczosResult, lanczos
from .spectral import TransformPlan, build_transform_plan, expectedkrylovdim, getspectraltransform


def initial_state_tolerance(dtype=np.float64) -> float:
    """``_initial_state_tolerance`` (PolfedCore.jl:38-41)."""
    return max(float(np.asarray(1e-10, dtype=dtype)), 10 * float(np.finfo(dtype).eps))


def prepare_initial_state(x0: np.ndarray) -> np.ndarray:
    """``_prepare_polfed_initial_state`` (PolfedCore.jl:43-95)."""
    x0 = np.asarray(x0)
    tol = initial_state_tolerance(x0.dtype if x0.dtype.kind == "f" else np.float64)
    if x0.ndim == 1:
        if x0.size == 0:
            raise ValueError("POLFED initial vector must not be empty.")
        nrm = np.linalg.norm(x0)
        if not np.isfinite(nrm):
            raise ValueError("POLFED initial vector must contain finite values.")
        if nrm == 0:
            raise ValueError("POLFED initial vector must be nonzero.")
        if abs(nrm - 1) <= max(tol, tol * max(nrm, 1.0)):
            return x0
        return x0.astype(np.float64) / np.linalg.norm(x0.astype(np.float64))
    if x0.shape[0] == 0:
        raise ValueError("POLFED initial matrix must have at least one row.")
    if x0.shape[1] == 0:
        raise ValueError("POLFED initial matrix must have at least one column.")
    if x0.shape[1] > x0.shape[0]:
        raise ValueError("POLFED initial matrix cannot have orthonormal columns when it has more columns than rows.")
    gram = x0.T @ x0
    if not np.all(np.isfinite(gram)):
        raise ValueError("POLFED initial matrix must contain finite values.")
    eye = np.eye(gram.shape[0])
    diff = np.linalg.norm(gram - eye)
    if diff <= max(tol, tol * max(np.linalg.norm(gram), np.linalg.norm(eye))):
        return x0
    q, _ = np.linalg.qr(x0)
    return q[:, :x0.shape[1]]


@dataclass
class PolfedOracleResult:
    vals: np.ndarray
    vecs: np.ndarray
    lanczos: LanczosResult
    plan: TransformPlan
    Emin: float
    Emax: float
    moments: np.ndarray
    coeffs: np.ndarray


def polfed(f, x0, howmany: int, target, *, Emin=None, Emax=None, order=None, left=None,
           right=None, cutoff=0.17, normalization=1.0, order_safety_factor=0.97,
           tol=1e-14, eigentol=1e-9, overestimate_iters=1.25, which="LR",
           N=250, R=300, kernel="Jackson", probes=None, rng=None) -> PolfedOracleResult:
    """CPU FP64 POLFED with explicit probes (or ``rng`` standard normals)."""
    x0 = prepare_initial_state(x0)
    D = x0.shape[0]
    s = x0.shape[1] if x0.ndim == 2 else 1
    v0 = x0 if x0.ndim == 1 else x0[:, 0]
    if Emin is None:
        Emin = lanczos(f, v0, 1, which="SR", maxdim=1000).vals[0]
    if Emax is None:
        Emax = lanczos(f, v0, 1, which="LR", maxdim=1000).vals[-1]
    a = (Emax - Emin) / 2
    b = (Emax + Emin) / 2

    def f_rescaled(X):
        Y = np.asarray(f(X)) * (1 / a)
        return Y - (b / a) * X

    plan = build_transform_plan(a, b, howmany, target, left=left, right=right, order=order,
                                cutoff=cutoff, normalization=normalization,
                                order_safety_factor=order_safety_factor)
    maxdim = expectedkrylovdim(howmany, s, overestimate_iters)
    if probes is None:
        if rng is None:
            rng = np.random.default_rng()
        probes = rng.standard_normal((D, R))
    moments = dos_moments(f_rescaled, N, probes)
    rho = dos_function(moments, kernel)
    getspectraltransform(plan, a, b, rho)
    c = filter_coefficients(plan.target, plan.order, normalization)
    f_transformed = lambda X: clenshaw_apply(f_rescaled, c, X)
    res = lanczos(f_transformed, x0, howmany, maxdim=maxdim, which=which, tol=tol,
                  eigentol=eigentol, mapvals=f)
    return PolfedOracleResult(res.vals, res.vecs, res, plan, Emin, Emax, moments, c)
