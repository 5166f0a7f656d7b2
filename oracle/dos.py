"""Oracle restatement of the KPM density of states (test infrastructure).

Restates ``src/DensetiesOfStates/{DensitiesOfStates,Moments,Kernels}.jl``.
Probe vectors are explicit arrays (the reference draws them from Julia's
``randn``, which cannot be reproduced here).
"""
from __future__ import annotations

import math

import numpy as np

from .clenshaw import clenshaw_scalar


def jackson(n: int, N: int) -> float:
    """``Jackson`` (src/DensetiesOfStates/Kernels.jl:14-16)."""
    return ((N - n + 1) * math.cos(math.pi * n / (N + 1))
            + math.sin(math.pi * n / (N + 1)) / math.tan(math.pi / (N + 1))) / (N + 1)


def lorentz(n: int, N: int, lam: float = 1.0) -> float:
    """``Lorentz`` (Kernels.jl:18-20)."""
    return math.sinh(lam * (1 - n / N)) / math.sinh(lam)


def fejer(n: int, N: int) -> float:
    """``Fejer`` (Kernels.jl:23-25)."""
    return 1 - n / N


def lanczos_kernel(n: int, N: int, M: int = 3) -> float:
    """``LanczosK`` (Kernels.jl:28-34)."""
    if n == 0:
        return 1.0
    return (math.sin(math.pi * n / N) / (math.pi * n / N)) ** M


def wang_zunger(n: int, N: int, alpha: float = 1.0, beta: float = 1.0) -> float:
    """``WangZunger`` (Kernels.jl:37-39)."""
    return math.exp(-(alpha * n / N) ** beta)


def dirichlet(n: int, N: int) -> float:
    """``Dirichlet`` (Kernels.jl:42-44)."""
    return 1.0


KERNELS = {
    "Jackson": jackson,
    "Lorentz": lorentz,
    "Fejer": fejer,
    "LanczosK": lanczos_kernel,
    "WangZunger": wang_zunger,
    "Dirichlet": dirichlet,
}


def get_kernel(name: str):
    """``get_kernel`` (Kernels.jl:63-67)."""
    if name not in KERNELS:
        raise ValueError(f"Unknown kernel: {name}. Available: {list(KERNELS)}")
    return KERNELS[name]


def trace(mapping, alpha: np.ndarray, mus: np.ndarray) -> None:
    """``trace!`` (src/DensetiesOfStates/Moments.jl:124-140): accumulates
    mu_n += <alpha | T_n(H) alpha> for n = 0..len(mus)-1 in place, one
    operator application per moment (three-term recurrence)."""
    v1 = alpha.copy()
    mus[0] += float(alpha @ v1)
    v2 = mapping(v1)
    mus[1] += float(alpha @ v2)
    for i in range(2, len(mus)):
        v3 = mapping(v2)
        v3 = v3 * 2.0
        v3 = v3 - v1
        mus[i] += float(alpha @ v3)
        v1, v2 = v2, v3


def dos_moments(mapping, N: int, probes: np.ndarray) -> np.ndarray:
    """``dos_moments`` / ``_dos_moments_impl`` (Moments.jl:42-109) with
    explicit probes (D x R); each probe is normalized (Moments.jl:66, :101)
    and the sum is scaled by D/R (:71, :107)."""
    D, R = probes.shape
    mus = np.zeros(N)
    for r in range(R):
        a = probes[:, r] / np.linalg.norm(probes[:, r])
        trace(mapping, a, mus)
    return mus * (D / R)


def dos_coefficients(moments: np.ndarray, kernel: str = "Jackson") -> np.ndarray:
    """``getdos!`` coefficient line (src/DensetiesOfStates/DensitiesOfStates.jl:29):
    g_n * mu_n * (2 - delta_{n0})."""
    g = get_kernel(kernel)
    N = len(moments)
    return np.array([g(i, N) * moments[i] * (2 - (i == 0)) for i in range(N)])


def dos_function(moments: np.ndarray, kernel: str = "Jackson"):
    """``getdos!`` (DensitiesOfStates.jl:14-43): rho(x) = T(x)/(pi sqrt(1-x^2))
    with T the degree N-1 Chebyshev series evaluated by scalar Clenshaw."""
    coeffs = dos_coefficients(moments, kernel)
    N = len(moments)

    def rho(x: float) -> float:
        return clenshaw_scalar(lambda i: coeffs[i], N - 1, x) / (math.pi * math.sqrt(1 - x * x))

    return rho


def exact_moments(evals_rescaled: np.ndarray, N: int) -> np.ndarray:
    """Exact trace moments Tr T_n(H~) = sum_i T_n(lambda_i) (the R -> infinity
    expectation of the stochastic estimate), from a dense spectrum."""
    th = np.arccos(np.clip(evals_rescaled, -1.0, 1.0))
    n = np.arange(N)
    return np.cos(np.outer(n, th)).sum(axis=1)
