"""Oracle restatement of interval / order selection (test infrastructure).

Restates ``src/Polfed/Algorithm/SpectralTransformation/{GetBounds,
GetOrderOfExpansion,Bisection}.jl``, ``src/Polfed/Plan/Plan.jl:203-281`` and
``src/Polfed/Structs/TargetSpec.jl``.  QuadGK is restated with
``scipy.integrate.quad`` at the same absolute/relative tolerance (1e-12).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
from scipy.integrate import quad

from .clenshaw import clenshaw_scalar, coefficients as default_coefficients


def bisection(f, x1: float, x2: float, tol: float = 1e-8, max_iter: int = 500) -> float:
    """``bisection`` (Bisection.jl:18-40)."""
    if f(x1) * f(x2) > 0:
        raise RuntimeError("The function must have opposite signs at a and b")
    it = 0
    while (x2 - x1) / 2 >= tol and it < max_iter:
        x = (x1 + x2) / 2
        fc = f(x)
        if fc == 0 or (x2 - x1) / 2 <= tol:
            return x
        elif f(x1) * fc < 0:
            x2 = x
        else:
            x1 = x
        it += 1
    return (x1 + x2) / 2


BOUNDS_FUN = {
    "mean": (lambda t, d: t - d, lambda t, d: t + d),
    "left": (lambda t, d: t, lambda t, d: t + 2 * d),
    "right": (lambda t, d: t - 2 * d, lambda t, d: t),
}


# ---------------------------------------------------------------- target spec
@dataclass
class TargetSpec:
    kind: str                      # maxdos | middle | offset | unrescaled | rescaled
    value: float | None = None


def normalize_target_spec(target) -> TargetSpec:
    """``normalize_target_spec`` (TargetSpec.jl:26-58)."""
    if isinstance(target, TargetSpec):
        return target
    if isinstance(target, str):
        if target == "maxdos":
            return TargetSpec("maxdos")
        if target == "middle":
            return TargetSpec("middle")
        raise ValueError(f"Unknown target symbol: {target}")
    if isinstance(target, tuple) and len(target) == 2 and isinstance(target[0], str):
        tag, val = target
        if isinstance(val, bool) or not isinstance(val, (int, float)):
            raise ValueError(f"Target tuple value must be Real, got {type(val)}")
        if tag == "absolute":
            raise ValueError("`(:absolute, value)` is no longer supported. Use `(:unrescaled, value)` instead.")
        if tag == "quantile":
            raise ValueError("`(:quantile, q)` is no longer supported.")
        if tag == "offset":
            if not -1 <= val <= 1:
                raise ValueError(f"Offset target requires eta in [-1, 1], got {val}")
            return TargetSpec("offset", float(val))
        if tag == "unrescaled":
            return TargetSpec("unrescaled", float(val))
        if tag == "rescaled":
            return TargetSpec("rescaled", float(val))
        raise ValueError(f"Unknown target tuple tag: {tag}")
    if isinstance(target, (int, float)) and not isinstance(target, bool):
        return TargetSpec("unrescaled", float(target))
    raise ValueError(f"Unsupported target specification: {target}")


@dataclass
class TransformPlan:
    """``TransformPlan`` (Plan.jl:32-44) in rescaled coordinates."""
    target_spec: TargetSpec
    target: float | None
    left: float | None
    right: float | None
    howmany: int
    order: int | None
    cutoff: float = 0.17
    normalization: float = 1.0
    order_safety_factor: float = 0.97
    polynomialtype: str = "Chebyshev"
    coefficients: object = field(default=default_coefficients)


def build_transform_plan(a: float, b: float, howmany: int, target, left=None, right=None,
                         order=None, cutoff=0.17, normalization=1.0, order_safety_factor=0.97,
                         polynomialtype="Chebyshev", coefficients=default_coefficients):
    """``build_transform_plan`` (Plan.jl:203-236); a, b are the rescaling
    half-width and center."""
    spec = normalize_target_spec(target)
    lr = None if left is None else (left - b) / a
    rr = None if right is None else (right - b) / a
    t = None
    if spec.kind == "unrescaled":
        t = (spec.value - b) / a
    elif spec.kind == "rescaled":
        t = spec.value
    elif spec.kind == "middle":
        t = 0.0
    return TransformPlan(spec, t, lr, rr, howmany, order, cutoff, normalization,
                         order_safety_factor, polynomialtype, coefficients)


def findmaximaldensity(rho) -> float:
    """``findmaximaldensity`` (Plan.jl:243-249): argmax over 5000 points in
    [-0.99, 0.99] (first maximum, like ``findmax``)."""
    x = np.linspace(-0.99, 0.99, 5000)
    vals = np.array([rho(float(xi)) for xi in x])
    return float(x[int(np.argmax(vals))])


def resolve_target(plan: TransformPlan, a: float, b: float, rho) -> None:
    """``resolve_target!`` (Plan.jl:260-281)."""
    if plan.target is not None:
        return
    s = plan.target_spec
    if s.kind == "maxdos":
        plan.target = findmaximaldensity(rho)
    elif s.kind == "offset":
        peak = findmaximaldensity(rho)
        eta = s.value
        plan.target = peak + eta * (1.0 - peak) if eta >= 0 else peak + eta * (1.0 + peak)
    elif s.kind == "middle":
        plan.target = 0.0
    elif s.kind == "rescaled":
        plan.target = s.value
    elif s.kind == "unrescaled":
        plan.target = (s.value - b) / a
    else:
        raise ValueError(f"Unsupported target specification: {s}")


def _normalized_filter(plan: TransformPlan, center: float, order: int):
    coeff = lambda n: plan.coefficients(center, n)
    norm = clenshaw_scalar(coeff, order, center, plan.polynomialtype)
    return lambda x: clenshaw_scalar(coeff, order, x, plan.polynomialtype) / norm * plan.normalization


def getbounds_from_dos(plan: TransformPlan, rho, position: str = "mean") -> None:
    """``getbounds_from_dos!`` (GetBounds.jl:57-81): smallest half-width delta
    (bisection, tol 1e-8) whose interval holds >= howmany states of rho."""
    lo_f, hi_f = BOUNDS_FUN[position]
    t = plan.target

    def f(d):
        a = max(lo_f(t, d), -1.0)
        b = min(hi_f(t, d), 1.0)
        n, _ = quad(rho, a, b, epsabs=1e-12, epsrel=1e-12, limit=500)
        return -1 if n >= plan.howmany else +1

    d = bisection(f, 0.0, 1.0, tol=1e-8)
    plan.left = max(lo_f(t, d), -1.0)
    plan.right = min(hi_f(t, d), 1.0)


def getbounds_from_K(plan: TransformPlan, position: str = "mean") -> None:
    """``getbounds_from_K!`` (GetBounds.jl:90-118): half-width where the
    normalized filter of order K reaches ``cutoff`` at an endpoint."""
    lo_f, hi_f = BOUNDS_FUN[position]
    t = plan.target
    p = _normalized_filter(plan, t, plan.order)

    def f(d):
        a = max(lo_f(t, d), -1.0)
        b = min(hi_f(t, d), 1.0)
        return -1 if max(p(a), p(b)) >= plan.cutoff else +1

    d = bisection(f, 0.0, 1.0, tol=1e-8)
    plan.left = max(lo_f(t, d), -1.0)
    plan.right = min(hi_f(t, d), 1.0)


def analytical_orderofexpansion(lam: float, delta: float) -> float:
    """``analytical_orderofexpansion`` (GetOrderOfExpansion.jl:9-11)."""
    return 2.65499 * math.sqrt(1 - lam ** 2) / delta


def getorderofexpansion_in_interval(plan: TransformPlan) -> None:
    """``getorderofexpansion_in_interval!`` (GetOrderOfExpansion.jl:36-70):
    bisection (tol 1) on K in [10, ceil(2 K_analytic)] for the smallest order
    with max p(edges) < cutoff, K <- floor(K * safety), then bounds re-derived
    from the reduced K when safety < 1.  The order bisection uses the interval
    midpoint as the target, like the reference."""
    left, right = plan.left, plan.right
    center = (left + right) / 2
    half = (right - left) / 2

    def f(order):
        order = math.floor(order)
        p = _normalized_filter(plan, center, order)
        return -1 if max(p(left), p(right)) >= plan.cutoff else +1

    Kmax = math.ceil(analytical_orderofexpansion(center, half) * 2)
    K = bisection(f, 10, Kmax, tol=1)
    plan.order = math.floor(K * plan.order_safety_factor)
    if plan.order_safety_factor < 1.0:
        getbounds_from_K(plan, "mean")


def getbounds(plan: TransformPlan, a: float, b: float, rho, position: str = "mean") -> None:
    """``getbounds!`` (GetBounds.jl:19-48)."""
    bounds_set = plan.left is not None and plan.right is not None
    if plan.target is None:
        resolve_target(plan, a, b, rho)
    if bounds_set and plan.order is not None:
        raise RuntimeError("Both interval bounds and polynomial order are specified. "
                           "Please set only one or neither.")
    elif bounds_set:
        pass
    elif plan.order is not None:
        getbounds_from_K(plan, position)
    else:
        getbounds_from_dos(plan, rho, position)


def getspectraltransform(plan: TransformPlan, a: float, b: float, rho) -> None:
    """Bounds + order part of ``getspectraltransform!``
    (SpectralTransformation.jl:19-31)."""
    getbounds(plan, a, b, rho, "mean")
    if plan.order is None:
        getorderofexpansion_in_interval(plan)


def expectedkrylovdim(howmany: int, blocksize: int, eta: float) -> int:
    """``expectedkrylovdim`` (src/Polfed/Structs/PolfedDefaults.jl:61-62)."""
    return math.ceil((20.427 * blocksize + 1.696 * howmany) * eta)
