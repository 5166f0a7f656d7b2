"""Oracle restatement of the (block) Lanczos eigensolver (test infrastructure).

Restates ``src/Lanczos/{Lanczos,LanczosMethod,LanczosAlgorithm}.jl``,
``src/Lanczos/Factorization/**`` (FullRO), ``src/Lanczos/Convergence/**`` and
``src/Lanczos/OrthonormalBasis/MatrixBasis.jl`` in FP64 numpy.  Both the
scalar (vector x0) and block (matrix x0) factorizations are restated,
including the reference's residual estimate, which multiplies the Ritz
vectors' last block by the *previous* coupling block (``getnorm``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla


@dataclass
class LanczosResult:
    vals: np.ndarray
    vecs: np.ndarray
    converged: int
    eigenconverged: int
    numiter: int
    maxiter: int
    krylovdim: int
    residual: float
    eigenresidual: float
    converged_history: list = field(default_factory=list)
    krylovdim_history: list = field(default_factory=list)
    maxresidual_history: list = field(default_factory=list)


def sortvals(vals: np.ndarray, which: str) -> np.ndarray:
    """``sortvals`` (src/Lanczos/Convergence/EigSorter.jl:41-57), stable sorts."""
    if which == "SR":
        return np.argsort(vals, kind="stable")
    if which == "LR":
        return np.argsort(-vals, kind="stable")
    if which == "LM":
        return np.argsort(-np.abs(vals), kind="stable")
    if which == "SM":
        return np.argsort(np.abs(vals), kind="stable")
    raise ValueError(f"Unknown sorting mode: {which}")


def next_check(k: int, residual: float) -> int:
    """``next_checking!`` (src/Lanczos/Convergence/NextChecking.jl:23-42)."""
    if residual > 1e-2:
        return k + 20
    if residual > 1e-3:
        return k + 10
    return k + 5


def _sym_upper(a: np.ndarray) -> np.ndarray:
    """``Hermitian(A)`` with the default upper triangle (BlockLanczosFactorization.jl:108)."""
    return np.triu(a) + np.triu(a, 1).T


def lanczos(f, x0: np.ndarray, howmany: int, maxdim: int | None = None, which: str = "SR",
            tol: float = 1e-14, eigentol: float = 1e-8, mapvals=None) -> LanczosResult:
    """``lanczos(f!, x0, howmany; ...)`` (src/Lanczos/Lanczos.jl:119-143).

    ``f(X)`` returns the operator applied to a vector or D x s block.
    Returns the converged pairs sorted ascending (LanczosMethod.jl:55-83).
    """
    if mapvals is None:
        mapvals = f
    if maxdim is None:
        maxdim = 500 * howmany
    block = x0.ndim == 2
    s = x0.shape[1] if block else 1
    maxiter = math.ceil(maxdim / s)
    maxdim = maxiter * s
    D = x0.shape[0]
    X0 = x0 if block else x0.reshape(D, 1)

    basis = np.empty((D, maxdim))
    nvecs = 0
    mat = np.zeros((maxdim, maxdim))         # block case (BlockLanczosFactorization.jl:115-123)
    alphas = np.zeros(maxdim)                # scalar case (LanczosFactorization.jl:24-25)
    betas = np.zeros(max(maxdim - 1, 1))
    v_last = np.zeros((D, s))
    v_second = np.zeros((D, s))
    k = 0

    def update(v):                           # Factorization.jl:26-38, MatrixBasis.jl:24-45
        nonlocal nvecs, v_last, v_second, k
        basis[:, nvecs:nvecs + s] = v
        nvecs += s
        v_second = v_last
        v_last = v.copy()
        k += s

    def apply(v):
        y = f(v if block else v[:, 0])
        return np.asarray(y).reshape(D, s)

    # ---- lanczositer (Factorization.jl:50-67)
    update(X0)
    r = apply(X0)
    if block:
        a = _sym_upper(v_last.T @ r)
        mat[k - s:k, k - s:k] = a
        r = r - X0 @ a
    else:
        alphas[k - 1] = float(v_last[:, 0] @ r[:, 0])
        r = r - X0 * alphas[k - 1]
    numiter = 1
    nextcheck = max(howmany // s + 1, 2)     # ConvergenceInfo.jl:43
    converged = eigenconverged = 0
    residual = eigenresidual = 1.0
    hist_c, hist_k, hist_r = [], [], []
    vals = np.zeros(howmany)
    vecs = np.zeros((D, howmany))

    while True:
        # [1] tridiagonalization! (BlockLanczosFactorization.jl:115-131 / LanczosFactorization.jl:55-63)
        if block:
            Q, R = np.linalg.qr(r)
            mat[k:k + s, k - s:k] = R
            mat[k - s:k, k:k + s] = R.T
            update(Q)
        else:
            beta = float(np.linalg.norm(r))
            r = r / beta
            betas[k - 1] = beta
            update(r)
        # [2] mapping!
        r = apply(v_last)
        # [3] calcoverlap!
        if block:
            a = _sym_upper(v_last.T @ r)
            mat[k - s:k, k - s:k] = a
            b = mat[k - s:k, k - 2 * s:k - s]
            # [4] orthogonalization! (orthogonalization.jl:31-33)
            r = r - (v_last @ a + v_second @ b.T)
        else:
            alphas[k - 1] = float(v_last[:, 0] @ r[:, 0])
            r = r - (alphas[k - 1] * v_last + betas[k - 2] * v_second)
        # [5] reorthogonalization!(FullRO): one CGS pass (reorthogonalization.jl:12-28, 77-84)
        nb = nvecs - 2 * s
        if nb > 0:
            B = basis[:, :nb]
            r = r - B @ (B.T @ r)
        # [6] getnumberofconvergedvecs! (ConvergenceInfo.jl:63-130)
        numiter += 1
        if nextcheck == numiter or numiter >= maxiter:
            if block:
                T = mat[:k, :k]
                lam, phi = sla.eigh(T, lower=False)
                bnorm = mat[k - s:k, k - 2 * s:k - s]
                res_of = lambda i: float(np.linalg.norm(bnorm @ phi[k - s:k, i]))
            else:
                lam, phi = sla.eigh_tridiagonal(alphas[:k], betas[:k - 1])
                bnorm = betas[k - 2]
                res_of = lambda i: abs(bnorm * phi[k - 1, i])
            idxs = sortvals(lam, which)
            converged = 0
            maxres = 0.0
            for i in range(howmany):
                rr = res_of(idxs[i])
                if rr > maxres:
                    maxres = rr
                if maxres < tol:
                    converged += 1
            residual = maxres
            hist_c.append(converged)
            hist_k.append(k)
            hist_r.append(maxres)
            if converged == howmany or numiter >= maxiter:
                # calculate_eigenvectors! (getvecs.jl:10-53)
                V = basis[:, :nvecs]
                vecs[:, :converged] = V @ phi[:, idxs[:converged]]
                # calculate_eigenvalues! with the original operator (getvals.jl:10-36)
                for i in range(converged):
                    hv = np.asarray(mapvals(vecs[:, i])).reshape(D)
                    vals[i] = float(vecs[:, i] @ hv)
                # calculate_eigenresiduals! (ConvergenceInfo.jl:139-173)
                eigenconverged = 0
                counting = True
                maxres_e = 0.0
                for i in range(converged):
                    hv = np.asarray(mapvals(vecs[:, i])).reshape(D)
                    re = float(np.linalg.norm(hv - vals[i] * vecs[:, i]))
                    if counting and re < eigentol:
                        eigenconverged += 1
                    else:
                        counting = False
                    maxres_e = max(maxres_e, re)
                eigenresidual = maxres_e
            nextcheck = next_check(numiter, residual)
        if converged == howmany or numiter >= maxiter:
            break

    nconv = min(converged, eigenconverged)
    order = np.argsort(vals[:nconv], kind="stable")
    return LanczosResult(vals[:nconv][order], vecs[:, :nconv][:, order], converged,
                         eigenconverged, numiter, maxiter, k, residual, eigenresidual,
                         hist_c, hist_k, hist_r)
