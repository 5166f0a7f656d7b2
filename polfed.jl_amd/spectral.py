"""Host-side spectral setup: target specs, KPM density of states, interval and
polynomial-order selection, filter coefficients.

Scalar work of ``src/Polfed/Structs/TargetSpec.jl``, ``src/Polfed/Plan/Plan.jl:203-281``,
``src/DensetiesOfStates/{DensitiesOfStates,Kernels}.jl`` and
``src/Polfed/Algorithm/SpectralTransformation/{GetBounds,GetOrderOfExpansion,
Bisection,DefineTransformation}.jl``.  Chebyshev series are evaluated here as
direct cosine sums sum_n c_n cos(n acos x), vectorized over n.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np
from scipy.integrate import quad

from .configs import PolfedDefaults


# ------------------------------------------------------------- target spec
@dataclass(frozen=True)
class TargetSpec:
    kind: str                 # "maxdos" | "middle" | "offset" | "unrescaled" | "rescaled"
    value: Optional[float] = None


def normalize_target_spec(target) -> TargetSpec:
    """Accepted forms of ``normalize_target_spec`` (TargetSpec.jl:26-58):
    "maxdos", "middle", ("offset", eta in [-1, 1]), ("unrescaled", E),
    ("rescaled", x), or a plain number (unrescaled energy)."""
    if isinstance(target, TargetSpec):
        return target
    if isinstance(target, str):
        if target in ("maxdos", "middle"):
            return TargetSpec(target)
        raise ValueError(f"Unknown target symbol: {target}")
    if isinstance(target, tuple) and len(target) == 2 and isinstance(target[0], str):
        tag, val = target
        if isinstance(val, bool) or not isinstance(val, (int, float, np.floating, np.integer)):
            raise ValueError(f"Target tuple value must be Real, got {type(val).__name__}")
        val = float(val)
        if tag == "absolute":
            raise ValueError("`(:absolute, value)` is no longer supported. Use `(:unrescaled, value)` instead.")
        if tag == "quantile":
            raise ValueError("`(:quantile, q)` is no longer supported.")
        if tag == "offset":
            if not -1.0 <= val <= 1.0:
                raise ValueError(f"Offset target requires eta in [-1, 1], got {val}")
            return TargetSpec("offset", val)
        if tag in ("unrescaled", "rescaled"):
            return TargetSpec(tag, val)
        raise ValueError(f"Unknown target tuple tag: {tag}")
    if isinstance(target, (int, float, np.floating, np.integer)) and not isinstance(target, bool):
        return TargetSpec("unrescaled", float(target))
    raise ValueError(f"Unsupported target specification: {target!r}")


# ------------------------------------------------------------ KPM kernels
def jackson(n, N):
    return ((N - n + 1) * np.cos(np.pi * n / (N + 1))
            + np.sin(np.pi * n / (N + 1)) / np.tan(np.pi / (N + 1))) / (N + 1)


def lorentz(n, N, lam=1.0):
    return np.sinh(lam * (1 - n / N)) / np.sinh(lam)


def fejer(n, N):
    return 1 - n / N


def lanczos_sigma(n, N, M=3):
    n = np.asarray(n, dtype=float)
    with np.errstate(invalid="ignore", divide="ignore"):
        v = (np.sin(np.pi * n / N) / (np.pi * n / N)) ** M
    return np.where(n == 0, 1.0, v)


def wang_zunger(n, N, alpha=1.0, beta=1.0):
    return np.exp(-(alpha * n / N) ** beta)


def dirichlet(n, N):
    return np.ones_like(np.asarray(n, dtype=float))


KERNELS = {"Jackson": jackson, "Lorentz": lorentz, "Fejer": fejer, "LanczosK": lanczos_sigma,
           "WangZunger": wang_zunger, "Dirichlet": dirichlet}


def get_kernel(name: str):
    """``get_kernel`` (Kernels.jl:63-67)."""
    if name not in KERNELS:
        raise ValueError(f"Unknown kernel: {name}. Available: {list(KERNELS)}")
    return KERNELS[name]


class DensityOfStates:
    """rho(x) = sum_n g_n mu_n (2 - delta_n0) T_n(x) / (pi sqrt(1 - x^2))
    (getdos!, src/DensetiesOfStates/DensitiesOfStates.jl:14-43); integrates
    to the number of states."""

    def __init__(self, moments: np.ndarray, kernel: str = "Jackson"):
        N = len(moments)
        n = np.arange(N)
        self.moments = np.asarray(moments, dtype=float)
        self.coeffs = get_kernel(kernel)(n, N) * self.moments * np.where(n == 0, 1.0, 2.0)
        self.n = n

    def __call__(self, x):
        x = np.asarray(x, dtype=float)
        th = np.arccos(np.clip(x, -1.0, 1.0))
        val = np.cos(np.multiply.outer(th, self.n)) @ self.coeffs
        return val / (np.pi * np.sqrt(1.0 - x * x))

    def integral(self, a: float, b: float) -> float:
        """Number of states in [a, b] (-1 <= a <= b <= 1), in closed form:
        with x = cos(theta) each term integrates to c_n cos(n theta) / pi over
        theta, i.e. (1/pi) [c_0 (theta_a - theta_b)
        + sum_{n>=1} c_n (sin(n theta_a) - sin(n theta_b)) / n]."""
        ta = math.acos(min(max(a, -1.0), 1.0))
        tb = math.acos(min(max(b, -1.0), 1.0))
        n = self.n[1:]
        rest = float(self.coeffs[1:] @ ((np.sin(n * ta) - np.sin(n * tb)) / n))
        return (float(self.coeffs[0]) * (ta - tb) + rest) / math.pi


# ------------------------------------------------------- filter polynomial
def _poly_alpha_beta(polytype: str):
    table = {
        "Chebyshev": (lambda k: 2.0, lambda k: -1.0, 1.0, 1.0),
        "Legendre": (lambda k: (2 * k + 1) / (k + 1), lambda k: -k / (k + 1), 1.0, 1.0),
        "Hermite": (lambda k: 2.0, lambda k: -2.0 * k, 1.0, 2.0),
        "Taylor": (lambda k: 1.0, lambda k: 0.0, 1.0, 1.0),
    }
    if polytype not in table:
        raise ValueError(f"unknown polynomial type {polytype}")
    return table[polytype]


def raw_coefficients(coefficients: Callable, target: float, order: int) -> np.ndarray:
    """c_n = coefficients(target, n) for n = 0..order (vectorized for the default)."""
    if coefficients is PolfedDefaults.coefficients:
        n = np.arange(order + 1)
        return np.where(n == 0, 1.0, 2.0) * np.cos(n * math.acos(target))
    return np.array([float(coefficients(target, n)) for n in range(order + 1)])


def series_value(c: np.ndarray, x, polytype: str = "Chebyshev"):
    """sum_n c_n P_n(x); Chebyshev as a cosine sum, others by the three-term
    recurrence P_{n+1} = alpha(n) x P_n + beta(n) P_{n-1} evaluated forward."""
    x = np.asarray(x, dtype=float)
    if polytype == "Chebyshev" and np.all(np.abs(x) <= 1.0):
        th = np.arccos(x)
        return np.cos(np.multiply.outer(th, np.arange(len(c)))) @ c
    alpha, beta, phi0, phi1 = _poly_alpha_beta(polytype)
    # forward recurrence of the basis: P_0 = phi0, P_1 = phi1 x,
    # P_{n+1} = alpha(n) x P_n + beta(n) P_{n-1}
    p_prev = np.ones_like(x) * phi0
    total = c[0] * p_prev
    if len(c) == 1:
        return total
    p_cur = phi1 * x
    total = total + c[1] * p_cur
    for n in range(1, len(c) - 1):
        p_next = alpha(n) * x * p_cur + beta(n) * p_prev
        total = total + c[n + 1] * p_next
        p_prev, p_cur = p_cur, p_next
    return total


def normalized_coefficients(coefficients, target: float, order: int, normalization: float = 1.0,
                            polytype: str = "Chebyshev") -> np.ndarray:
    """Coefficients scaled so that p(target) = normalization
    (definetransformation!, DefineTransformation.jl:11-30)."""
    c = raw_coefficients(coefficients, target, order)
    return c * (normalization / float(series_value(c, target, polytype)))


# -------------------------------------------------------------- bisection
def bisection(f, x1: float, x2: float, tol: float = 1e-8, max_iter: int = 500) -> float:
    """Bisection of ``bisection`` (Bisection.jl:18-40); f returns a sign."""
    if f(x1) * f(x2) > 0:
        raise RuntimeError("The function must have opposite signs at a and b")
    it = 0
    while (x2 - x1) / 2 >= tol and it < max_iter:
        x = (x1 + x2) / 2
        fc = f(x)
        if fc == 0 or (x2 - x1) / 2 <= tol:
            return x
        if f(x1) * fc < 0:
            x2 = x
        else:
            x1 = x
        it += 1
    return (x1 + x2) / 2


_BOUNDS = {"mean": (lambda t, d: t - d, lambda t, d: t + d),
           "left": (lambda t, d: t, lambda t, d: t + 2 * d),
           "right": (lambda t, d: t - 2 * d, lambda t, d: t)}


@dataclass
class TransformPlan:
    """Resolved transform plan in rescaled coordinates (Plan.jl:32-44)."""
    target_spec: TargetSpec
    target: Optional[float]
    left: Optional[float]
    right: Optional[float]
    howmany: int
    order: Optional[int]
    cutoff: float = PolfedDefaults.cutoff
    normalization: float = PolfedDefaults.normalization
    order_safety_factor: float = PolfedDefaults.order_safety_factor
    polynomialtype: str = PolfedDefaults.polynomialtype
    coefficients: Callable = field(default=PolfedDefaults.coefficients)


def build_transform_plan(transform, a: float, b: float, howmany: int, target) -> TransformPlan:
    """``build_transform_plan`` (Plan.jl:203-236)."""
    spec = normalize_target_spec(target)
    lr = None if transform.left is None else (transform.left - b) / a
    rr = None if transform.right is None else (transform.right - b) / a
    t = {"unrescaled": (lambda: (spec.value - b) / a), "rescaled": (lambda: spec.value),
         "middle": (lambda: 0.0)}.get(spec.kind, lambda: None)()
    return TransformPlan(spec, t, lr, rr, howmany, transform.order, transform.cutoff,
                         transform.normalization, transform.order_safety_factor,
                         transform.polynomialtype, transform.coefficients)


def findmaximaldensity(rho) -> float:
    """``findmaximaldensity`` (Plan.jl:243-249)."""
    x = np.linspace(-0.99, 0.99, 5000)
    return float(x[int(np.argmax(rho(x)))])


def resolve_target(plan: TransformPlan, a: float, b: float, rho) -> None:
    """``resolve_target!`` (Plan.jl:260-281)."""
    if plan.target is not None:
        return
    s = plan.target_spec
    if s.kind == "maxdos":
        plan.target = findmaximaldensity(rho)
    elif s.kind == "offset":
        peak = findmaximaldensity(rho)
        plan.target = peak + s.value * (1.0 - peak) if s.value >= 0 else peak + s.value * (1.0 + peak)
    elif s.kind == "middle":
        plan.target = 0.0
    elif s.kind == "rescaled":
        plan.target = s.value
    elif s.kind == "unrescaled":
        plan.target = (s.value - b) / a
    else:
        raise ValueError(f"Unsupported target specification: {s}")


def _dirichlet(phi: float, K: int) -> float:
    """1 + 2 sum_{n=1}^K cos(n phi) = sin((K + 1/2) phi) / sin(phi / 2)."""
    s = math.sin(phi / 2)
    if abs(s) < 1e-12:
        return float(2 * K + 1)
    return math.sin((K + 0.5) * phi) / s


def _filter_fn(plan: TransformPlan, center: float, order: int):
    """x -> p(x) of the normalized filter centred at ``center`` of this order.
    With the default coefficients (c_0 = 1, c_n = 2 cos(n theta_t),
    PolfedDefaults.jl:104) on the Chebyshev basis the series sums in closed
    form: sum_n c_n cos(n theta) = [Dir_K(theta_t - theta) + Dir_K(theta_t +
    theta)] / 2 with the Dirichlet kernel Dir_K, so each value costs O(1)
    rather than O(K) (the order and bound bisections evaluate it ~100 times
    at K up to ~10^5); other coefficients or bases sum the series."""
    if (plan.coefficients is PolfedDefaults.coefficients and plan.polynomialtype == "Chebyshev"
            and -1.0 <= center <= 1.0):
        tt = math.acos(center)
        K = int(order)
        raw = lambda th: 0.5 * (_dirichlet(tt - th, K) + _dirichlet(tt + th, K))
        scale = plan.normalization / raw(tt)

        def p(x):
            x = float(x)
            if abs(x) > 1.0:  # outside [-1, 1]: the generic series
                c = normalized_coefficients(plan.coefficients, center, order,
                                            plan.normalization, plan.polynomialtype)
                return float(series_value(c, x, plan.polynomialtype))
            return scale * raw(math.acos(x))
        return p
    c = normalized_coefficients(plan.coefficients, center, order, plan.normalization,
                                plan.polynomialtype)
    return lambda x: float(series_value(c, x, plan.polynomialtype))


def getbounds_from_dos(plan: TransformPlan, rho, position: str = "mean") -> None:
    """``getbounds_from_dos!`` (GetBounds.jl:57-81).  The state count over the
    interval is the KPM density's closed-form integral
    (``DensityOfStates.integral``), the quantity the reference integrates with
    QuadGK; any other density is integrated numerically (scipy ``quad``)."""
    lo, hi = _BOUNDS[position]
    t = plan.target

    def f(d):
        a_, b_ = max(lo(t, d), -1.0), min(hi(t, d), 1.0)
        if isinstance(rho, DensityOfStates):
            n = rho.integral(a_, b_)
        else:
            n, _ = quad(lambda x: float(rho(x)), a_, b_, epsabs=1e-12, epsrel=1e-12, limit=500)
        return -1 if n >= plan.howmany else 1

    d = bisection(f, 0.0, 1.0, tol=1e-8)
    plan.left, plan.right = max(lo(t, d), -1.0), min(hi(t, d), 1.0)


def getbounds_from_K(plan: TransformPlan, position: str = "mean") -> None:
    """``getbounds_from_K!`` (GetBounds.jl:90-118)."""
    lo, hi = _BOUNDS[position]
    t = plan.target
    p = _filter_fn(plan, t, plan.order)

    def f(d):
        a_, b_ = max(lo(t, d), -1.0), min(hi(t, d), 1.0)
        return -1 if max(p(a_), p(b_)) >= plan.cutoff else 1

    d = bisection(f, 0.0, 1.0, tol=1e-8)
    plan.left, plan.right = max(lo(t, d), -1.0), min(hi(t, d), 1.0)


def analytical_orderofexpansion(lam: float, delta: float) -> float:
    """``analytical_orderofexpansion`` (GetOrderOfExpansion.jl:9-11)."""
    return 2.65499 * math.sqrt(1 - lam ** 2) / delta


def getorderofexpansion(plan: TransformPlan) -> None:
    """``getorderofexpansion!`` / ``getorderofexpansion_in_interval!``
    (GetOrderOfExpansion.jl:21-70)."""
    if plan.order is not None:
        return
    left, right = plan.left, plan.right
    center, half = (left + right) / 2, (right - left) / 2

    def f(order):
        p = _filter_fn(plan, center, math.floor(order))
        return -1 if max(p(left), p(right)) >= plan.cutoff else 1

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
    if bounds_set:
        return
    if plan.order is not None:
        getbounds_from_K(plan, position)
    else:
        getbounds_from_dos(plan, rho, position)
