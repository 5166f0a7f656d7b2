"""Block Lanczos on the GPU (FP64), mirroring the reference's factorization.

Restates ``lanczos`` (src/Lanczos/Lanczos.jl:119-143), ``lanczos_algorithm!``
(LanczosAlgorithm.jl:11-54), the step sequence of ``lanczositer!``
(Factorization/Factorization.jl:75-96), the block operations of
BlockLanczosFactorization.jl / FactorizationSteps/*.jl, the convergence checks
of Convergence/ConvergenceInfo.jl, CalculatingResiduals.jl, NextChecking.jl,
getvecs.jl, getvals.jl and the final ordering of LanczosMethod.jl:55-83.

Device work (all FP64): basis storage (one D x maxdim column-major tensor, or
device columns plus pinned host columns streamed through the device: basis.py),
QR of the residual block (Cholesky-QR2, Householder on the host only when the
block is numerically rank deficient), overlaps and three-term updates on the
FP64 matrix cores, full reorthogonalization (``passes`` CGS sweeps), Ritz
vectors and residuals, and the projected k x k eigenproblem from
k >= DEVICE_EIGH_MIN (below it, LAPACK's banded solver on the host).
A vector x0 runs the same code with block size 1 (the reference's scalar
factorization normalizes by ||r|| > 0, which is what Cholesky-QR returns).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np
import scipy.linalg as sla
import torch

from . import blockops as bo
from ._lib import NumericBreakdown
from .basis import make_basis
from .configs import FullRO, PartialRO
from .device import Device


@dataclass
class LanczosResult:
    vals: np.ndarray
    vecs: torch.Tensor           # (nconv, D) block = column-major D x nconv
    residuals: np.ndarray        # ||H psi - E psi|| of the returned pairs
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
    times: dict = field(default_factory=dict)
    all_vals: np.ndarray = None       # every Lanczos-converged pair (before the eigentol cut)
    all_residuals: np.ndarray = None
    fast_iterations: int = 0          # iterations whose filter ran in the fast (FP32) form
    switch_iteration: int = -1        # iteration at which the mixed mode switched to FP32
    fallback: bool = False            # mixed mode went back to FP64 after a failed residual check
    basis_device_vectors: int = 0     # basis columns held in device memory
    basis_host_vectors: int = 0       # basis columns held in pinned host memory (hybrid basis)


def sortvals(vals: np.ndarray, which: str) -> np.ndarray:
    """``sortvals`` (Convergence/EigSorter.jl:41-57)."""
    if which == "SR":
        return np.argsort(vals, kind="stable")
    if which == "LR":
        return np.argsort(-vals, kind="stable")
    if which == "LM":
        return np.argsort(-np.abs(vals), kind="stable")
    if which == "SM":
        return np.argsort(np.abs(vals), kind="stable")
    raise ValueError(f"Unknown sorting mode: {which}")


def _next_check(k: int, residual: float) -> int:
    """``next_checking!`` (Convergence/NextChecking.jl:23-42)."""
    if residual > 1e-2:
        return k + 20
    if residual > 1e-3:
        return k + 10
    return k + 5


def _to_dev_matrix(M: np.ndarray, like: torch.Tensor) -> torch.Tensor:
    """host m x n -> device (n, m) tensor (column-major m x n)."""
    return torch.from_numpy(np.ascontiguousarray(M.T)).to(like.device)


def _exhausted(R: np.ndarray, tmax: float) -> bool:
    """True when the QR factor of the residual block is rounding noise against
    the projected matrix (||R|| <= 1000 eps max|T|, ``tmax`` = max|T| so far): the
    Krylov space is invariant, as when it fills a space of dimension <= maxdim."""
    return float(np.linalg.norm(R)) <= 1000 * np.finfo(np.float64).eps * max(tmax, 1e-300)


def _host_qr(block: torch.Tensor) -> np.ndarray:
    """Householder QR on the host for a rank-deficient block (in place on the
    device tensor); the reference's qr(fact.r) (BlockLanczosFactorization.jl:123)."""
    A = block.detach().cpu().numpy().T
    Q, R = np.linalg.qr(A)
    block.copy_(torch.from_numpy(np.ascontiguousarray(Q.T)).to(block.device))
    return R


class _Timer:
    def __init__(self, sync: bool):
        self.sync = sync
        self.t = {}

    def __call__(self, key):
        timer = self

        class _Ctx:
            def __enter__(self):
                if timer.sync:
                    torch.cuda.synchronize()
                self.t0 = time.perf_counter()

            def __exit__(self, *a):
                if timer.sync:
                    torch.cuda.synchronize()
                timer.t[key] = timer.t.get(key, 0.0) + time.perf_counter() - self.t0

        return _Ctx()


def lanczos(dev: Device, f, x0: torch.Tensor, howmany: int, *, maxdim: int | None = None,
            which: str = "SR", tol: float = 1e-14, eigentol: float = 1e-8, mapvals=None,
            rot=None, stage_timers: bool = True, f_fast=None, fast_mode: str = "never",
            switch_tol: float = 1e-6, basistype=None, polish=None) -> LanczosResult:
    """Block Lanczos for ``howmany`` pairs of the operator ``f``.

    ``f(X, Y)`` writes the operator applied to the (s, D) block X into Y.
    ``mapvals`` (default ``f``'s operator) evaluates the final Rayleigh
    quotients and residuals: an ``Operator`` (unscaled H) or a callable
    ``mapvals(X, Y)``.

    ``f_fast`` is a cheaper, less accurate version of ``f`` (the FP32 filter).
    ``fast_mode`` "always" uses it for every iteration; "mixed" starts with
    ``f`` and switches to ``f_fast`` once the residual estimate
    ||R_{k+1} phi_last|| of the ``howmany`` wanted Ritz pairs drops below
    ``switch_tol`` (inexact-Krylov relaxation: late products may be inexact,
    early ones may not), and returns to ``f`` if the final residuals miss
    ``eigentol``.

    ``basistype`` selects the basis storage (``MatrixBasis`` default,
    ``HybridMatrixBasis``: see basis.py).

    ``polish(Y, keep) -> (vals, vecs, res)`` (refine.polish) replaces the
    plain Ritz extraction when some filter applications ran in the fast form:
    it receives the Ritz vectors of the Lanczos-converged pairs, best first,
    and returns as many pairs in the same order.  (Unconverged Ritz vectors
    are not added as a guard set: Rayleigh-Ritz on such mixtures yields
    spurious interior Ritz values that rank among the wanted ones.)
    """
    gen = lanczos_coroutine(dev, x0, howmany, maxdim=maxdim, which=which, tol=tol,
                            eigentol=eigentol, mapvals=mapvals, rot=rot,
                            stage_timers=stage_timers, has_fast=f_fast is not None,
                            fast_mode=fast_mode, switch_tol=switch_tol, basistype=basistype,
                            polish=polish)
    try:
        X, Y, fast = next(gen)
        while True:
            (f_fast if fast else f)(X, Y)
            X, Y, fast = gen.send(None)
    except StopIteration as stop:
        return stop.value


# From this projected dimension on, the dense symmetric eigensolver runs on the
# device (rocSOLVER through torch.linalg.eigh; the k x k matrix crosses PCIe
# both ways): on MI355X 19 ms against 60 ms for LAPACK's banded solver on the
# host at k=848 (block 16), 37 against 215 ms at k=1472 (block 32); below it
# the host banded solver is as fast (10 ms at k=424, block 8).
DEVICE_EIGH_MIN = 512


def _proj_eigh(T: np.ndarray, s: int, subset=None, device=None):
    """Eigenpairs (ascending) of the block-tridiagonal projected matrix T
    (symmetric, bandwidth s: the off-diagonal blocks are upper-triangular R
    factors); subset = [lo, hi] selects eigenvalues by index.  LAPACK's banded
    solver on the host (its tridiagonal solver for blocks of one), or the
    device's dense solver for k >= DEVICE_EIGH_MIN when ``device`` is a GPU."""
    n = T.shape[0]
    if device is not None and device.type == "cuda" and n >= DEVICE_EIGH_MIN:
        w, v = torch.linalg.eigh(torch.from_numpy(np.ascontiguousarray(T)).to(device))
        lam, phi = w.cpu().numpy(), v.cpu().numpy()
        if subset is None:
            return lam, phi
        return lam[subset[0]:subset[1] + 1], phi[:, subset[0]:subset[1] + 1]
    if s == 1 and n > 1:  # tridiagonal: LAPACK's MRRR solver (stemr)
        d, e = np.diagonal(T).copy(), np.diagonal(T, offset=1).copy()
        if subset is None:
            return sla.eigh_tridiagonal(d, e, check_finite=False)
        return sla.eigh_tridiagonal(d, e, select="i", select_range=(subset[0], subset[1]),
                                    check_finite=False)
    w = min(s, n - 1)
    band = np.zeros((w + 1, n))
    for d in range(w + 1):
        band[w - d, d:] = np.diagonal(T, offset=d)
    if subset is None:
        return sla.eig_banded(band, lower=False, check_finite=False)
    return sla.eig_banded(band, lower=False, check_finite=False, select="i",
                          select_range=(subset[0], subset[1]))


def lanczos_coroutine(dev: Device, x0: torch.Tensor, howmany: int, *,
                      maxdim: int | None = None, which: str = "SR", tol: float = 1e-14,
                      eigentol: float = 1e-8, mapvals=None, rot=None, stage_timers: bool = True,
                      has_fast: bool = False, fast_mode: str = "never",
                      switch_tol: float = 1e-6, basistype=None, polish=None):
    """The block Lanczos of ``lanczos`` as a generator: it yields every
    operator application it needs as ``(X, Y, fast)`` -- write f(X) (or
    f_fast(X) when ``fast``) into Y, then resume it with ``send(None)`` -- and
    returns the ``LanczosResult``.  A batch driver can so advance many
    independent solves in lockstep and serve their requests with one
    multi-instance filter launch per step."""
    rot = rot if rot is not None else FullRO()
    if maxdim is None:
        maxdim = 500 * howmany
    X0 = x0 if x0.dim() == 2 else x0.reshape(1, -1)
    s, D = X0.shape
    maxiter = math.ceil(maxdim / s)
    maxdim = maxiter * s
    tm = _Timer(stage_timers)
    basis = make_basis(dev, basistype, maxdim, D, s, X0.device)
    mat = np.zeros((maxdim, maxdim))
    r = torch.empty((s, D), dtype=torch.float64, device=X0.device)

    def sym_upper(a):
        return np.triu(a) + np.triu(a, 1).T

    def overlap(V, W):
        return bo.gram(dev, V, W).cpu().numpy().T

    # ---- lanczositer (Factorization.jl:50-67)
    V0 = basis.block(0)
    V0.copy_(X0)
    basis.commit(0)
    k = s
    with tm("mapping"):
        yield V0, r, False
    a = sym_upper(overlap(V0, r))
    mat[0:s, 0:s] = a
    tmax = float(np.max(np.abs(a)))  # max|mat| so far
    bo.update(dev, V0, _to_dev_matrix(a, r), r, alpha=-1.0, beta=1.0)
    numiter = 1
    nextcheck = max(howmany // s + 1, 2)
    converged = eigenconverged = 0
    residual = eigenresidual = 1.0
    hist_c, hist_k, hist_r = [], [], []
    vals = np.zeros(0)
    vecs = torch.empty((0, D), dtype=torch.float64, device=X0.device)
    res_out = np.zeros(0)

    def check(k, bnorm, final):
        """Residual estimates ||bnorm phi_last|| of the wanted Ritz pairs of
        mat[:k, :k] (counted while the running maximum stays below tol), and the
        Ritz vectors with their eigen-residuals once all converged or final."""
        nonlocal converged, eigenconverged, residual, eigenresidual, vals, vecs, res_out
        nonlocal all_vals, all_res
        lam, phi = _proj_eigh(mat[:k, :k], s, device=X0.device)
        idxs = sortvals(lam, which)
        converged = 0
        maxres = 0.0
        for i in range(min(howmany, k)):
            rr = float(np.linalg.norm(bnorm @ phi[k - s:k, idxs[i]]))
            if rr > maxres:
                maxres = rr
            if maxres < tol:
                converged += 1
        residual = maxres
        hist_c.append(converged)
        hist_k.append(k)
        hist_r.append(maxres)
        if converged == howmany or final:
            if converged and polish is not None and fast_iters > 0:
                # FP32 products in the basis: FP64 polish of the converged
                # pairs (refine.polish)
                Y = torch.empty((converged, D), dtype=torch.float64, device=X0.device)
                basis.combine(_to_dev_matrix(phi[:, idxs[:converged]], r), Y, k)
                with tm("polish"):
                    vals, vecs, res_out = polish(Y, converged)
                del Y
            else:
                Phi = phi[:, idxs[:converged]]
                vecs = torch.empty((converged, D), dtype=torch.float64, device=X0.device)
                if converged:
                    basis.combine(_to_dev_matrix(Phi, r), vecs, k)
                    vals, res_out = _eigen_residuals(dev, mapvals, vecs)
                else:
                    vals, res_out = np.zeros(0), np.zeros(0)
            eigenconverged = 0
            for rv in res_out:
                if rv < eigentol:
                    eigenconverged += 1
                else:
                    break
            eigenresidual = float(res_out.max()) if res_out.size else 0.0
            all_vals, all_res = vals.copy(), res_out.copy()

    use_fast = fast_mode == "always" and has_fast
    switch_est, switch_tested = math.inf, -10
    fast_iters = 0
    switch_it = -1
    fallback = False
    all_vals = all_res = None
    while True:
        # [1] tridiagonalization!: QR of r into the next basis slot
        with tm("tridiagonalization"):
            basis.prepare(k)
            slot = basis.block(k)
            slot.copy_(r)
            try:
                R = bo.block_qr(dev, slot)
            except NumericBreakdown:
                R = _host_qr(slot)
            invariant = _exhausted(R, tmax)
            if not invariant:
                tmax = max(tmax, float(np.max(np.abs(R))))
                basis.commit(k)
                mat[k:k + s, k - s:k] = R
                mat[k - s:k, k:k + s] = R.T
                k += s
        if invariant:
            # the residual block vanished: the basis spans an invariant
            # subspace and mat[:k, :k] holds its exact eigenpairs (residual
            # estimate zero; the eigen-residuals still decide); the next block
            # would be rounding noise, so the run ends here
            with tm("convergence"):
                check(k, np.zeros_like(R), True)
            break
        # the switch test costs a projected eigensolve: while the last
        # estimate is far above switch_tol it runs every fourth iteration only
        due = switch_est <= 100 * switch_tol or numiter - switch_tested >= 4
        if fast_mode == "mixed" and has_fast and not use_fast and not fallback \
                and k - s >= howmany and due:
            switch_tested = numiter
            with tm("switch_test"):
                kp = k - s
                if which == "LR":
                    sub = [kp - howmany, kp - 1]
                elif which == "SR":
                    sub = [0, howmany - 1]
                else:
                    sub = None
                lam_p, ph_p = _proj_eigh(mat[:kp, :kp], s, sub, device=X0.device)
                if sub is None:
                    ph_p = ph_p[:, sortvals(lam_p, which)[:howmany]]
                est = float(np.max(np.linalg.norm(R @ ph_p[kp - s:kp, :], axis=0)))
                switch_est = est
                if est < switch_tol:
                    use_fast = True
                    switch_it = numiter + 1
        V_last, V_prev = basis.block(k - s), basis.block(k - 2 * s)
        # [2] mapping!
        with tm("mapping"):
            yield V_last, r, use_fast
            if use_fast:
                fast_iters += 1
        # [3] calculate_overlap!  [4] orthogonalization!
        with tm("orthogonalization"):
            a = sym_upper(overlap(V_last, r))
            mat[k - s:k, k - s:k] = a
            tmax = max(tmax, float(np.max(np.abs(a))))
            b = mat[k - s:k, k - 2 * s:k - s]
            bo.update(dev, V_last, _to_dev_matrix(a, r), r, alpha=-1.0, beta=1.0)
            bo.update(dev, V_prev, _to_dev_matrix(b.T, r), r, alpha=-1.0, beta=1.0)
        # [5] reorthogonalization! against all but the last two blocks
        with tm("reorthogonalization"):
            nb = k - 2 * s
            do_ro = nb > 0
            if isinstance(rot, PartialRO):
                it = k // s
                do_ro = do_ro and (it % (2 + rot.skip) == 0 or (it - 1) % (2 + rot.skip) == 0)
            if do_ro:
                basis.reorth(r, nb, rot.passes)
        # [6] getnumberofconvergedvecs!
        numiter += 1
        if nextcheck == numiter or numiter >= maxiter:
            with tm("convergence"):
                check(k, mat[k - s:k, k - 2 * s:k - s], numiter >= maxiter)
                nextcheck = _next_check(numiter, residual)
        if converged == howmany and fast_mode == "mixed" and use_fast \
                and eigenconverged < converged and numiter < maxiter:
            # the fast products left residuals above eigentol: continue in FP64
            use_fast = False
            fallback = True
            converged = 0
            continue
        if converged == howmany or numiter >= maxiter:
            break

    basis.close()
    nconv = min(converged, eigenconverged)
    order = np.argsort(vals[:nconv], kind="stable")
    idx = torch.as_tensor(order, device=X0.device, dtype=torch.long)
    vecs_sorted = vecs[:nconv].index_select(0, idx) if nconv else vecs[:0]
    return LanczosResult(vals[:nconv][order], vecs_sorted, res_out[:nconv][order], converged,
                         eigenconverged, numiter, maxiter, k, residual, eigenresidual, hist_c,
                         hist_k, hist_r, dict(tm.t), all_vals, all_res, fast_iters, switch_it,
                         fallback, basis.device_vectors, basis.host_vectors)


def _eigen_residuals(dev: Device, mapvals, vecs: torch.Tensor):
    """Rayleigh quotients and residual norms with the original operator
    (getvals.jl:10-36, ConvergenceInfo.jl:139-173)."""
    from .operators import Operator
    if isinstance(mapvals, Operator):
        return bo.residuals(dev, mapvals, vecs)
    Hv = torch.empty_like(vecs)
    mapvals(vecs, Hv)
    E = bo.coldots(dev, vecs, Hv)
    R = Hv - torch.from_numpy(E).to(vecs.device)[:, None] * vecs
    res = np.sqrt(bo.coldots(dev, R, R))
    return E, res
