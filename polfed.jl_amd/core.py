"""``polfed(...)``: the public solver entry of the reference, on the GPU.

Restates ``polfed(f!, x0, howmany, target; mapping, transform, fact, dos)``
(src/Polfed/PolfedCore.jl:153-263): initial-state checks (:38-95), the mapping
plan with Emin/Emax Lanczos probes (src/Polfed/Plan/Plan.jl:128-191), the
transform plan (:203-236), ``polfed_algorithm`` (src/Polfed/Algorithm/polfed_algorithm.jl:15-37):
KPM density of states -> interval and order -> block Lanczos on the filtered
operator, with eigenvalues and residuals from the unscaled operator.

The operator is a matrix-free ``Operator`` (``models``), so the filter runs as
one fused HIP call per Lanczos iteration (``pf_filter``).  The
``MappingConfig`` hooks (f_rescaled / clenshaw_recurrence / clenshaw_finalsum)
replace the corresponding pieces with user callables, as in the reference
(DefineTransformation.jl:92-168).
"""
from __future__ import annotations

import dataclasses
import time
import warnings
from dataclasses import dataclass, field
from typing import Optional

import numpy as np
import torch

from . import blockops as bo
from . import refine
from . import spectral as spc
from .configs import (DoSConfig, FactorizationConfig, MappingConfig, PolfedDefaults,
                      TransformConfig)
from .device import Device, default_device
from .lanczos import LanczosResult, lanczos
from .operators import Operator, PackedOperator


# widest Lanczos block of the device path (pf_block_qr, pf_reorth, pf_rotate)
MAX_BLOCK = bo.MAX_N


@dataclass
class Report:
    """Spectral-transform, factorization and timing summary of a run
    (the reference's Report, src/Polfed/Structs/Report/*.jl)."""
    Emin: float
    Emax: float
    target: float
    left: float
    right: float
    order: int
    blocksize: int
    maxdim: int
    iterations: int
    converged: int
    eigenconverged: int
    residual: float
    eigenresidual: float
    num_vecmuls: int
    filter_dtype: str
    moments: np.ndarray
    times: dict = field(default_factory=dict)
    lanczos: Optional[LanczosResult] = None


def initial_state_tolerance(dtype=torch.float64) -> float:
    """``_initial_state_tolerance`` (PolfedCore.jl:38-41)."""
    eps = torch.finfo(dtype).eps
    return max(1e-10 if dtype == torch.float64 else float(torch.tensor(1e-10, dtype=dtype)),
               10 * eps)


def prepare_initial_state(dev: Device, x0) -> torch.Tensor:
    """``_prepare_polfed_initial_state`` (PolfedCore.jl:43-95), on the device.
    Returns a (D,) vector or an (s, D) block; the caller's array is not mutated."""
    if not isinstance(x0, torch.Tensor):
        x0 = dev.to_device(np.asarray(x0, dtype=np.float64))
    if x0.dim() == 1:
        if x0.numel() == 0:
            raise ValueError("POLFED initial vector must not be empty.")
        nrm = float(np.sqrt(bo.coldots(dev, x0, x0)[0]))
        if not np.isfinite(nrm):
            raise ValueError("POLFED initial vector must contain finite values.")
        if nrm == 0:
            raise ValueError("POLFED initial vector must be nonzero.")
        tol = initial_state_tolerance()
        if abs(nrm - 1.0) <= max(tol, tol * max(nrm, 1.0)):
            return x0
        warnings.warn(f"POLFED initial vector is not normalized; using a normalized copy "
                      f"(norm {nrm:.6g}, tolerance {tol:g}).", stacklevel=3)
        return x0 * (1.0 / nrm)
    s, D = x0.shape
    if D == 0:
        raise ValueError("POLFED initial matrix must have at least one row.")
    if s == 0:
        raise ValueError("POLFED initial matrix must have at least one column.")
    if s > D:
        raise ValueError("POLFED initial matrix cannot have orthonormal columns when it has "
                         "more columns than rows.")
    G = bo.gram(dev, x0, x0).cpu().numpy().T
    if not np.all(np.isfinite(G)):
        raise ValueError("POLFED initial matrix must contain finite values.")
    tol = initial_state_tolerance()
    I = np.eye(s)
    defect = np.linalg.norm(G - I)
    if defect <= max(tol, tol * max(np.linalg.norm(G), np.linalg.norm(I))):
        return x0
    warnings.warn(f"POLFED initial matrix is not orthonormal; using an orthonormalized copy "
                  f"(Gram defect {defect:.6g}, tolerance {tol:g}).", stacklevel=3)
    q = x0.clone()
    try:
        bo.block_qr(dev, q)
    except Exception:
        Q, _ = np.linalg.qr(x0.cpu().numpy().T)
        q = torch.from_numpy(np.ascontiguousarray(Q.T)).to(x0.device)
    return q


def _extremal(dev: Device, op, v0: torch.Tensor, which: str, tol: float = 1e-14,
              eigentol: float = 1e-8, maxdim: int = 1000) -> float:
    """``lanczos(f!, v0, 1; which, maxdim=1000)`` probe of build_mapping_plan
    (Plan.jl:136-168) with the unscaled operator (tolerances of
    optimize_spectral_transform's probe: optimization.jl:40-52).  ``op`` is an
    Operator or a callable f(X, Y) on device blocks."""
    f = (lambda X, Y: op.apply(X, Y)) if isinstance(op, Operator) else op
    res = lanczos(dev, f, v0, 1, maxdim=maxdim, which=which, tol=tol, eigentol=eigentol,
                  mapvals=op, stage_timers=False)
    if res.vals.size == 0:
        raise RuntimeError(f"Failed to estimate {which} extremal eigenvalue with Lanczos. "
                           "Set MappingConfig(Emin=..., Emax=...) explicitly.")
    return float(res.vals[0] if which == "SR" else res.vals[-1])


def _matrix_operator(H, mapping: MappingConfig, dev: Device):
    """``polfed(mat, ...)`` (PolfedCore.jl:153-184): the matrix becomes the
    packed value-bucket operator of optimize_spectral_transform
    (optimization.jl:75-119); with ``mapping.optimize_mapping`` its extremal
    eigenvalues come from that function's probes -- a random normalized start
    vector, tol = eigentol = sqrt(eps) (optimization.jl:40-52,90-97) -- unless
    Emin/Emax are given.  The packed operator serves both settings of
    ``optimize_mapping`` (the reference's other branch multiplies by the
    matrix itself, ``mul!``, the same operator)."""
    op = PackedOperator.from_matrix(H, device=dev)
    if mapping.optimize_mapping and (mapping.Emin is None or mapping.Emax is None):
        v = np.random.default_rng(0).standard_normal(op.D)
        v0 = dev.to_device(v / np.linalg.norm(v))
        tol = float(np.sqrt(np.finfo(np.float64).eps))
        Emin = mapping.Emin if mapping.Emin is not None else _extremal(dev, op, v0, "SR", tol, tol)
        Emax = mapping.Emax if mapping.Emax is not None else _extremal(dev, op, v0, "LR", tol, tol)
        mapping = dataclasses.replace(mapping, Emin=Emin, Emax=Emax)
    return op, mapping


def _filter_callable(op: Operator, mapping: MappingConfig, coeffs: np.ndarray, polytype: str,
                     dtype: str):
    """The transformed operator f!_transformed (DefineTransformation.jl:11-48)."""
    if mapping.clenshaw_recurrence is None and mapping.clenshaw_finalsum is None \
            and mapping.f_rescaled is None:
        return lambda X, Y: op.filter(coeffs, X, Y, dtype=dtype, polytype=polytype)
    # user hooks: Clenshaw loop on the host, each piece on the device
    alpha, beta, phi0, phi1 = spc._poly_alpha_beta(polytype)
    K = len(coeffs) - 1

    def rec(b1, b2, b3, c, k, X):
        if mapping.clenshaw_recurrence is not None:
            fn = mapping.clenshaw_recurrence
            try:
                fn(b1, b2, b3, c, k, X)
            except TypeError:
                fn(b1, b2, b3, c, X)
        elif mapping.f_rescaled is not None:
            mapping.f_rescaled(b1, b2)
            b1.mul_(alpha(k)).add_(X, alpha=c).add_(b3, alpha=beta(k + 1))
        else:
            op.recurrence(b1, b2, b3, c, X, alpha=alpha(k), beta=beta(k + 1))

    def fin(b2, b3, c0, Y, X):
        if mapping.clenshaw_finalsum is not None:
            mapping.clenshaw_finalsum(b2, b3, c0, Y, X)
        elif mapping.f_rescaled is not None:
            mapping.f_rescaled(Y, b2)
            Y.mul_(phi1).add_(X, alpha=c0 * phi0).add_(b3, alpha=beta(1) * phi0)
        else:
            op.finalsum(b2, b3, c0, Y, X, beta1=beta(1), phi0=phi0, phi1=phi1)

    def apply(X, Y):
        b = [torch.zeros_like(X) for _ in range(3)]
        for k in range(K, 0, -1):
            rec(b[0], b[1], b[2], float(coeffs[k]), k, X)
            b = [b[2], b[0], b[1]]
        fin(b[1], b[2], float(coeffs[0]), Y, X)

    return apply


def _polish_callable(dev: Device, op: Operator, mapping: MappingConfig,
                     transform: TransformConfig, prep: "_Prepared", which: str):
    """The FP64 polish of a run with FP32 filters (refine.polish), or None
    (FP64 filters, polish off, or user mapping hooks: the library's own
    operator is then not the filtered one)."""
    if transform.dtype == "f64" or not transform.polish or not isinstance(op, Operator):
        return None
    if mapping.clenshaw_recurrence is not None or mapping.clenshaw_finalsum is not None \
            or mapping.f_rescaled is not None:
        return None
    plan = prep.plan

    def polish(Y, keep):
        vals, vecs, res, _ = refine.polish(dev, op, Y, keep, target=plan.target,
                                           order=plan.order, coeffs=prep.coeffs,
                                           polytype=plan.polynomialtype, which=which,
                                           fraction=transform.polish_fraction)
        return vals, vecs, res

    return polish


@dataclass
class _Prepared:
    """Per-problem state between the set-up stages and the Lanczos stage."""
    x0: torch.Tensor
    s: int
    Emin: float
    Emax: float
    plan: object
    coeffs: np.ndarray
    moments: np.ndarray
    maxdim: int
    times: dict


def _prepare(op: Operator, x0, howmany: int, target, mapping: MappingConfig,
             transform: TransformConfig, fact: FactorizationConfig, dos: DoSConfig,
             probes: Optional[torch.Tensor], dev: Device) -> _Prepared:
    """Everything before the Lanczos stage of polfed_algorithm
    (PolfedCore.jl:153-263, polfed_algorithm.jl:15-24): initial state, the
    mapping plan with the Emin/Emax probes, KPM DoS, interval and order, and
    the normalized filter coefficients."""
    times = {}
    if getattr(x0, "ndim", 1) == 2 and min(x0.shape) > MAX_BLOCK:
        raise ValueError(f"POLFED on the GPU supports blocks of at most {MAX_BLOCK} columns "
                         f"(the block QR and reorthogonalization kernels); got "
                         f"{min(x0.shape)}.")
    x0 = prepare_initial_state(dev, x0)
    s = 1 if x0.dim() == 1 else x0.shape[0]
    D = x0.shape[-1]
    if D != op.D:
        raise ValueError(f"x0 has {D} rows, operator dimension is {op.D}")

    # ---- mapping plan (Plan.jl:128-191)
    t0 = time.perf_counter()
    v0 = x0 if x0.dim() == 1 else x0[0]
    Emin = mapping.Emin if mapping.Emin is not None else _extremal(dev, op, v0, "SR")
    Emax = mapping.Emax if mapping.Emax is not None else _extremal(dev, op, v0, "LR")
    op.set_rescale(Emin, Emax)
    a, b = (Emax - Emin) / 2, (Emax + Emin) / 2
    torch.cuda.synchronize()
    times["bounds"] = time.perf_counter() - t0

    plan = spc.build_transform_plan(transform, a, b, howmany, target)
    maxdim = PolfedDefaults.expectedkrylovdim(howmany, s, fact.overestimate_iters)

    # ---- getdos! (DensitiesOfStates.jl:14-43): KPM moments on the device
    t0 = time.perf_counter()
    if probes is None:
        probes = bo.fill_probes(dev, D, dos.R, dos.seed, dos.probe_kind)
    moments = op.kpm_moments(dos.N, probes, dtype=dos.dtype)
    rho = spc.DensityOfStates(moments, dos.kernel)
    times["dos"] = time.perf_counter() - t0

    # ---- getspectraltransform! (SpectralTransformation.jl:19-31)
    t0 = time.perf_counter()
    spc.getbounds(plan, a, b, rho, "mean")
    spc.getorderofexpansion(plan)
    coeffs = spc.normalized_coefficients(plan.coefficients, plan.target, plan.order,
                                         plan.normalization, plan.polynomialtype)
    times["transform"] = time.perf_counter() - t0
    if transform.dtype not in ("f64", "f32", "mixed"):
        raise ValueError(f"TransformConfig.dtype must be 'f64', 'f32' or 'mixed', got {transform.dtype!r}")
    return _Prepared(x0, s, Emin, Emax, plan, coeffs, moments, maxdim, times)


def lanczos_extrema(op, x0=None, *, maxdim: int = 1000, tol: float = 1e-14,
                    eigentol: float = 1e-8, seed: Optional[int] = None,
                    device: Optional[Device] = None):
    """``lanczos_extrema`` (src/Lanczos/Lanczos.jl:146-212): ``(Emin, Emax)`` from
    two single-vector Lanczos probes (which "SR", then "LR") on the device.

    ``op`` is a GPU operator, a square real symmetric matrix (packed like the
    matrix input of ``polfed``) or a callable ``f(X, Y)`` writing the operator
    applied to the device block X into Y.  Both probes start from a normalized
    copy of ``x0`` (host array or device vector; the caller's array is not
    changed); without ``x0`` an operator or matrix gets a random normal start
    (``seed``).  A zero start, a non-square matrix or a start of the wrong
    length raise ``ValueError``."""
    dev = device or (op.device if isinstance(op, Operator) else default_device())
    if isinstance(op, Operator):
        D = op.D
    elif callable(op):
        if x0 is None:
            raise ValueError("lanczos_extrema with an operator callable needs x0.")
        D = int(x0.shape[0])
    else:
        shape = getattr(op, "shape", None)
        if shape is None or len(shape) != 2 or shape[0] != shape[1]:
            raise ValueError("lanczos_extrema requires a square matrix.")
        D = int(shape[0])
    if x0 is None:
        x0 = np.random.default_rng(seed).standard_normal(D)
    if isinstance(x0, torch.Tensor):
        v = x0.detach().to(dev.torch_device, dtype=torch.float64).reshape(-1).clone()
    else:
        v = dev.to_device(np.asarray(x0, dtype=np.float64).reshape(-1))
    if v.numel() != D:
        raise ValueError("x0 length must match the matrix dimension.")
    nrm = float(np.sqrt(bo.coldots(dev, v, v)[0]))
    if nrm == 0 or not np.isfinite(nrm):
        raise ValueError("lanczos_extrema requires a nonzero, finite starting vector.")
    v = v * (1.0 / nrm)
    if not isinstance(op, Operator) and not callable(op):
        op = PackedOperator.from_matrix(op, device=dev)
    return (_extremal(dev, op, v, "SR", tol, eigentol, maxdim),
            _extremal(dev, op, v, "LR", tol, eigentol, maxdim))


def display_report(report: "Report", file=None) -> str:
    """Text summary of a ``polfed`` report (the reference's display_report,
    src/Polfed/Structs/Report/*.jl): spectral transform, factorization and
    stage times.  Printed to ``file`` (default stdout) and returned."""
    r = report
    lines = [
        "POLFED report",
        f"  spectrum      Emin = {r.Emin:.12g}, Emax = {r.Emax:.12g}",
        f"  target        {r.target:.12g} (rescaled), interval [{r.left:.6g}, {r.right:.6g}]",
        f"  filter        order {r.order}, {r.filter_dtype}",
        f"  factorization block {r.blocksize}, maxdim {r.maxdim}, iterations {r.iterations}",
        f"  converged     {r.converged} (Lanczos), {r.eigenconverged} (eigen-residual)",
        f"  residuals     Lanczos {r.residual:.3e}, eigen {r.eigenresidual:.3e}",
        f"  operator applications {r.num_vecmuls}",
    ]
    if r.lanczos is not None and r.lanczos.basis_host_vectors:
        lines.append(f"  basis         {r.lanczos.basis_device_vectors} device + "
                     f"{r.lanczos.basis_host_vectors} host columns")
    for k, t in r.times.items():
        lines.append(f"  time {k:<28s} {t:10.4f} s")
    text = "\n".join(lines)
    print(text, file=file)
    return text


def _report(prep: _Prepared, res: LanczosResult, howmany: int, dtype: str, times: dict) -> Report:
    plan = prep.plan
    return Report(prep.Emin, prep.Emax, plan.target, plan.left, plan.right, plan.order, prep.s,
                  prep.maxdim, res.numiter, res.converged, res.eigenconverged, res.residual,
                  res.eigenresidual, plan.order * prep.s * res.numiter + howmany, dtype,
                  prep.moments, times, res)


def polfed(op: Operator, x0, howmany: int, target, *, produce_report: bool = False,
           mapping: Optional[MappingConfig] = None, transform: Optional[TransformConfig] = None,
           fact: Optional[FactorizationConfig] = None, dos: Optional[DoSConfig] = None,
           probes: Optional[torch.Tensor] = None, device: Optional[Device] = None):
    """Eigenpairs of ``op`` closest (by the polynomial filter) to ``target``.

    ``op`` is a GPU operator, or a real symmetric matrix (scipy sparse or
    dense), which is packed into a :class:`PackedOperator` as the reference's
    matrix method does.  Returns ``(vals, vecs)`` -- vals ascending (host numpy), vecs an
    (nconv, D) device block -- or ``(vals, vecs, report)`` with
    ``produce_report=True``.  ``probes`` (D x R device block) overrides the
    generated KPM probe vectors.
    """
    mapping = mapping or MappingConfig()
    transform = transform or TransformConfig()
    fact = fact or FactorizationConfig()
    dos = dos or DoSConfig()
    if not isinstance(op, Operator):  # matrix input (scipy sparse or dense)
        dev = device or default_device()
        op, mapping = _matrix_operator(op, mapping, dev)
    dev = device or op.device or default_device()
    t_all = time.perf_counter()
    prep = _prepare(op, x0, howmany, target, mapping, transform, fact, dos, probes, dev)
    times = dict(prep.times)
    plan = prep.plan
    f_t = _filter_callable(op, mapping, prep.coeffs, plan.polynomialtype, "f64")
    f_fast = None
    if transform.dtype != "f64":
        f_fast = _filter_callable(op, mapping, prep.coeffs, plan.polynomialtype, "f32")
    fast_mode = {"f64": "never", "f32": "always", "mixed": "mixed"}[transform.dtype]
    # launch autotuning of the step kernel for this block width (cached per operator)
    for dt in (("f64",) if transform.dtype == "f64" else ("f64", "f32")):
        op.ensure_tuned(min(prep.s, 32 if dt == "f64" else 64), dt)

    # ---- lanczos on the filtered operator (polfed_algorithm.jl:25-34)
    t0 = time.perf_counter()
    res = lanczos(dev, f_t, prep.x0, howmany, maxdim=prep.maxdim, which=fact.which, tol=fact.tol,
                  eigentol=fact.eigentol if fact.eigentol is not None else 1e-8, mapvals=op,
                  rot=fact.rot, f_fast=f_fast, fast_mode=fast_mode,
                  switch_tol=transform.switch_tol, basistype=fact.basistype,
                  polish=_polish_callable(dev, op, mapping, transform, prep, fact.which))
    torch.cuda.synchronize()
    times["lanczos"] = time.perf_counter() - t0
    times.update({f"lanczos.{k}": v for k, v in res.times.items()})
    times["total"] = time.perf_counter() - t_all

    if not produce_report:
        return res.vals, res.vecs
    return res.vals, res.vecs, _report(prep, res, howmany, transform.dtype, times)


@dataclass
class BatchReport:
    """Batch-level timing of ``polfed_batch``."""
    n: int
    setup_s: float
    lanczos_s: float
    filter_s: float
    total_s: float
    filter_calls: int
    batched_steps: int
    autotune_us: Optional[float]
    reports: list


def polfed_batch(ops, x0s, howmany: int, target, *, produce_report: bool = False,
                 mapping: Optional[MappingConfig] = None,
                 transform: Optional[TransformConfig] = None,
                 fact: Optional[FactorizationConfig] = None, dos: Optional[DoSConfig] = None,
                 probes=None, device: Optional[Device] = None, autotune: bool = True,
                 progress=None):
    """``polfed`` on many Quantum-Sun instances at once (disorder realizations,
    possibly of different sizes).  Every instance runs the reference algorithm
    exactly as ``polfed`` does -- its own bounds, DoS, interval, order and
    block Lanczos -- but the filter applications of all instances that need
    one are served together by ``QSunBatch`` (one multi-instance kernel launch
    per Clenshaw step, graph-replayed).  ``transform.dtype`` "f32"/"mixed"
    applies per instance as in ``polfed``.

    Returns a list of ``(vals, vecs)`` per instance, or
    ``(list of (vals, vecs, report), BatchReport)`` with ``produce_report``.
    ``probes``: None or a list of per-instance D_i x R probe blocks.
    ``progress``: optional callable(str) for one line per set-up instance and
    per ten batched filter applications.
    """
    from .batch import QSunBatch
    from .lanczos import lanczos_coroutine
    mapping = mapping or MappingConfig()
    transform = transform or TransformConfig()
    fact = fact or FactorizationConfig()
    dos = dos or DoSConfig()
    if mapping.clenshaw_recurrence is not None or mapping.clenshaw_finalsum is not None \
            or mapping.f_rescaled is not None:
        raise ValueError("polfed_batch: user Clenshaw/mapping hooks are not batched; use polfed")
    n = len(ops)
    if len(x0s) != n:
        raise ValueError("polfed_batch: one initial state per operator")
    dev = device or ops[0].device or default_device()
    t_all = time.perf_counter()
    preps = []
    for i, (op, x0) in enumerate(zip(ops, x0s)):
        preps.append(_prepare(op, x0, howmany, target, mapping, transform, fact, dos,
                              None if probes is None else probes[i], dev))
        if progress:
            progress(f"set-up {i + 1}/{n}: D={op.D} order={preps[-1].plan.order}")
    s = preps[0].s
    if any(p.s != s for p in preps):
        raise ValueError("polfed_batch: every instance must use the same block size")
    polytypes = {p.plan.polynomialtype for p in preps}
    if len(polytypes) != 1:
        raise ValueError("polfed_batch: one polynomial family for the whole batch")
    batch = QSunBatch(ops, s)
    batch.set_filter([p.coeffs for p in preps], polytypes.pop())
    tuned = None
    if autotune:
        # one tiling serves both arithmetics: tune for the one that carries
        # most of the work (FP64 in the mixed form, whose FP32 iterations come
        # last)
        tuned = batch.autotune("f32" if transform.dtype == "f32" else "f64")
    setup_s = time.perf_counter() - t_all

    fast_mode = {"f64": "never", "f32": "always", "mixed": "mixed"}[transform.dtype]
    eigentol = fact.eigentol if fact.eigentol is not None else 1e-8
    gens = [lanczos_coroutine(dev, p.x0, howmany, maxdim=p.maxdim, which=fact.which, tol=fact.tol,
                              eigentol=eigentol, mapvals=op, rot=fact.rot, stage_timers=False,
                              has_fast=transform.dtype != "f64", fast_mode=fast_mode,
                              switch_tol=transform.switch_tol, basistype=fact.basistype,
                              polish=_polish_callable(dev, op, mapping, transform, p, fact.which))
            for op, p in zip(ops, preps)]
    t0 = time.perf_counter()
    pending = {i: next(g) for i, g in enumerate(gens)}
    results = [None] * n
    events = []
    calls = 0
    steps = 0
    while pending:
        reqs = [None] * n
        for i, (X, Y, fast) in pending.items():
            reqs[i] = (X, Y, "f32" if fast else "f64")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        batch.filter(reqs)
        e1.record()
        events.append((e0, e1))
        calls += 1
        steps += int(max(batch.orders[i] for i in pending)) + 1
        if progress and calls % 10 == 0:
            progress(f"batched filter {calls}: {len(pending)} instances running")
        for i in list(pending):
            try:
                pending[i] = gens[i].send(None)
            except StopIteration as stop:
                results[i] = stop.value
                del pending[i]
    torch.cuda.synchronize()
    lanczos_s = time.perf_counter() - t0
    filter_s = sum(a.elapsed_time(b) for a, b in events) * 1e-3
    total_s = time.perf_counter() - t_all
    batch.close()
    if not produce_report:
        return [(r.vals, r.vecs) for r in results]
    out = []
    for p, r in zip(preps, results):
        out.append((r.vals, r.vecs, _report(p, r, howmany, transform.dtype, dict(p.times))))
    return out, BatchReport(n, setup_s, lanczos_s, filter_s, total_s, calls, steps, tuned,
                            [o[2] for o in out])
