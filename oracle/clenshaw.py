"""Oracle restatement of the Clenshaw polynomial filter (test infrastructure).

Restates ``src/Clenshaw/{ClenshawMapping,Clenshaw}.jl``,
``src/Clenshaw/Algorithm/{ClenshawAlgorithm,RecurrenceRelations,FinalSum}.jl``
and the default coefficients of ``src/Polfed/Structs/PolfedDefaults.jl:104``.
"""
from __future__ import annotations

import math

import numpy as np

# ``polynomial_properties`` (src/Clenshaw/ClenshawMapping.jl:7-12):
# (alpha(k), beta(k), phi0, phi1)
POLYNOMIAL_PROPERTIES = {
    "Chebyshev": (lambda k: 2.0, lambda k: -1.0, 1.0, 1.0),
    "Legendre": (lambda k: (2 * k + 1) / (k + 1), lambda k: -k / (k + 1), 1.0, 1.0),
    "Hermite": (lambda k: 2.0, lambda k: -2.0 * k, 1.0, 2.0),
    "Taylor": (lambda k: 1.0, lambda k: 0.0, 1.0, 1.0),
}


def coefficients(lam: float, n: int) -> float:
    """``PolfedDefaults.coefficients`` (src/Polfed/Structs/PolfedDefaults.jl:104):
    (2 - (n == 0)) * cos(n * acos(lam)) -- c_0 = 1, c_n = 2 cos(n theta)."""
    return float((2 - (n == 0)) * math.cos(n * math.acos(lam)))


def clenshaw_scalar(coeff, order: int, x: float, polytype: str = "Chebyshev") -> float:
    """Scalar ``clenshaw_algorithm!`` (src/Clenshaw/Algorithm/ClenshawAlgorithm.jl:34-54):
    b1 = c_k + alpha(k) x b2 + beta(k+1) b3 for k = order..1, rotate (3,1,2);
    y = c_0 phi0 + phi1 x b2 + beta(1) phi0 b3 (FinalSum.jl:26-30)."""
    alpha, beta, phi0, phi1 = POLYNOMIAL_PROPERTIES[polytype]
    b = [0.0, 0.0, 0.0]
    for k in range(order, 0, -1):
        b[0] = coeff(k) + alpha(k) * x * b[1] + beta(k + 1) * b[2]
        b = [b[2], b[0], b[1]]
    return coeff(0) * phi0 + phi1 * x * b[1] + beta(1) * phi0 * b[2]


def filter_coefficients(target: float, order: int, normalization: float = 1.0,
                        coeff=coefficients, polytype: str = "Chebyshev") -> np.ndarray:
    """Normalized coefficients of ``definetransformation!``
    (src/Polfed/Algorithm/SpectralTransformation/DefineTransformation.jl:11-30):
    c_n * normalization / p(target), n = 0..order."""
    raw = lambda n: coeff(target, n)
    p_t = clenshaw_scalar(raw, order, target, polytype)
    norm_ = 1.0 / p_t * normalization
    return np.array([coeff(target, n) * norm_ for n in range(order + 1)])


def clenshaw_apply(mapping, c: np.ndarray, X: np.ndarray, polytype: str = "Chebyshev") -> np.ndarray:
    """Block ``clenshaw_algorithm!`` (ClenshawAlgorithm.jl:12-29) with the
    mapping-derived kernels of ``build_clenshaw_kernels`` (Clenshaw.jl:141-153):
    b1 = c_k X + alpha(k) H b2 + beta(k+1) b3; Y = c_0 phi0 X + phi1 H b2 + beta(1) phi0 b3.
    ``mapping(Z)`` returns H Z; ``c`` holds c_0..c_K."""
    alpha, beta, phi0, phi1 = POLYNOMIAL_PROPERTIES[polytype]
    K = len(c) - 1
    b1 = np.zeros_like(X)
    b2 = np.zeros_like(X)
    b3 = np.zeros_like(X)
    for k in range(K, 0, -1):
        b1 = c[k] * X + alpha(k) * mapping(b2) + beta(k + 1) * b3
        b1, b2, b3 = b3, b1, b2
    return c[0] * phi0 * X + phi1 * mapping(b2) + beta(1) * phi0 * b3


def chebyshev_filter_dense(evals: np.ndarray, c: np.ndarray) -> np.ndarray:
    """p(lambda) = sum_n c_n T_n(lambda) via numpy's chebval (independent of the
    Clenshaw restatement; used to pin it)."""
    return np.polynomial.chebyshev.chebval(evals, c)
