"""Model constructors mirroring ``Polfed.Models`` (src/Models/Models.jl:85-117).

They return matrix-free device operators instead of SparseMatrixCSC: the
basis, connectivity and matrix elements are computed by the HIP kernels.
Random fields/disorder use a numpy ``Generator`` (Julia's RNG streams cannot be
reproduced); explicit arrays reproduce any realization exactly.
"""
from __future__ import annotations

from math import sqrt

import numpy as np

from .operators import QSunOperator, SpinChainOperator

SUPPORTED_BOUNDARIES = ("open", "periodic")


def chain_bonds(L: int, distance: int, boundary: str):
    """``_chain_bonds`` (src/Models/SpinHalfChains.jl:227-243), 1-based sites."""
    if boundary not in SUPPORTED_BOUNDARIES:
        raise ValueError(f"boundary must be one of {SUPPORTED_BOUNDARIES}; got {boundary}.")
    if boundary == "open":
        return [(i, i + distance) for i in range(1, L - distance + 1)] if distance < L else []
    out = []
    for i in range(1, L + 1):
        j = (i + distance - 1) % L + 1
        if i != j:
            out.append((i, j))
    return out


def resolve_fields(L: int, field: float, W: float, fields=None, rng=None) -> np.ndarray:
    """``_resolve_fields`` (src/Models/SpinHalfChains.jl:245-257)."""
    if fields is not None:
        f = np.asarray(fields, dtype=np.float64)
        if f.shape != (L,):
            raise ValueError(f"fields must have length L={L}; got length {f.size}.")
        return f.copy()
    if W < 0:
        raise ValueError(f"W must be non-negative; got {float(W)}.")
    if W == 0:
        return np.full(L, float(field))
    rng = rng if rng is not None else np.random.default_rng()
    return float(field) + float(W) * (2.0 * rng.random(L) - 1.0)


def _validate(L: int, Lup: int):
    if L < 2:
        raise ValueError(f"L must be at least 2; got L={L}.")
    if not 0 <= Lup <= L:
        raise ValueError(f"Lup must satisfy 0 <= Lup <= L; got Lup={Lup} for L={L}.")


def xxz_hamiltonian(L: int, Lup: int, J: float, Delta: float, W: float, *,
                    boundary: str = "periodic", field: float = 0.0, fields=None, rng=None,
                    device=None) -> SpinChainOperator:
    """``xxz_hamiltonian`` (src/Models/SpinHalfChains.jl:74-99)."""
    _validate(L, Lup)
    bonds = chain_bonds(L, 1, boundary)
    h = resolve_fields(L, field, W, fields, rng)
    return SpinChainOperator(L, Lup, bonds, [float(J)] * len(bonds), [float(Delta)] * len(bonds),
                             h, device=device)


def j1j2_hamiltonian(L: int, Lup: int, J1: float, J2: float, Delta1: float, Delta2: float,
                     W: float, *, boundary: str = "periodic", field: float = 0.0, fields=None,
                     rng=None, device=None) -> SpinChainOperator:
    """``j1j2_hamiltonian`` (src/Models/SpinHalfChains.jl:189-220)."""
    _validate(L, Lup)
    nn = chain_bonds(L, 1, boundary)
    nnn = chain_bonds(L, 2, boundary)
    h = resolve_fields(L, field, W, fields, rng)
    return SpinChainOperator(L, Lup, nn + nnn, [float(J1)] * len(nn) + [float(J2)] * len(nnn),
                             [float(Delta1)] * len(nn) + [float(Delta2)] * len(nnn), h,
                             device=device)


def qsun_draws(L_loc: int, L_grain: int, alpha: float, *, gamma: float = 1.0, w: float = 0.5,
               hz: float = 1.0, zeta: float = 0.2, rng=None):
    """Disorder, grain neighbors, couplings and the GOE grain of
    ``build_hamiltonian`` (src/QSun/QuantumSun.jl:21-30,
    src/QSun/RandomUtils.jl:1-26), drawn in the reference's order."""
    rng = rng if rng is not None else np.random.default_rng()
    disorder = rng.random(L_loc) * (2 * w) + (hz - w)
    neighbors = rng.integers(0, L_grain, size=L_loc)
    couplings = np.zeros(L_loc)
    if alpha > 0:
        couplings[0] = 1.0
        for j in range(2, L_loc + 1):
            couplings[j - 1] = alpha ** (j - 1 + rng.random() * 2 * zeta - zeta)
    dg = 2 ** L_grain
    X = rng.standard_normal((dg, dg))
    H_grain = gamma * ((X + X.T) / sqrt(2)) / sqrt(dg + 1)
    return disorder, neighbors, couplings, H_grain


def qsun_hamiltonian(L_loc: int, L_grain: int, g0: float, alpha: float, *, gamma: float = 1.0,
                     w: float = 0.5, hz: float = 1.0, zeta: float = 0.2, S: float = 0.5,
                     rng=None, use_U1: bool = False, draws=None, device=None) -> QSunOperator:
    """``qsun_hamiltonian`` (src/QSun/QSun.jl:37-84) for the full space, S = 1/2.
    ``draws`` = (disorder, neighbors, couplings, H_grain) reproduces a given
    realization."""
    if (2 * S) != int(2 * S):
        raise ValueError(f"Wrong value of spin. Only integer spins or half-integer spins are possible. You chose S = {S}")
    if S != 0.5 or use_U1:
        raise NotImplementedError("the HIP Quantum Sun operator covers the full space of spin 1/2")
    if draws is None:
        draws = qsun_draws(L_loc, L_grain, alpha, gamma=gamma, w=w, hz=hz, zeta=zeta, rng=rng)
    disorder, neighbors, couplings, H_grain = draws
    return QSunOperator(L_loc, L_grain, g0, disorder, neighbors, couplings, H_grain,
                        device=device)
