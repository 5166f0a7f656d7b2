"""Model constructors mirroring ``Polfed.Models`` (src/Models/Models.jl:85-117)
and ``Polfed.QSun`` (src/QSun).

Spin-1/2 chains and the spin-1/2 full-space Quantum Sun return matrix-free
device operators: the basis, connectivity and matrix elements are computed by
the HIP kernels.  The other Quantum-Sun variants (spin S, the U(1) sector)
build their sparse matrix on the host, as the reference does, and run on the
packed operator.
Random fields/disorder use a numpy ``Generator`` (Julia's RNG streams cannot be
reproduced); explicit arrays reproduce any realization exactly.
"""
from __future__ import annotations

from math import sqrt

import numpy as np

import scipy.sparse as sp

from .operators import Operator, PackedOperator, QSunOperator, SpinChainOperator

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


def qsun_spin_draws(L_loc: int, L_grain: int, alpha: float, d: int, *, w: float = 0.5,
                    hz: float = 1.0, zeta: float = 0.2, rng=None):
    """Disorder, neighbors, couplings and the raw GOE grain matrix
    ``(X + X^T)/sqrt(2)`` of dimension d^L_grain, in the reference's draw order
    (QuantumSun.jl:21-28, QuantumSun_U1_conserving.jl:23-31)."""
    rng = rng if rng is not None else np.random.default_rng()
    disorder = rng.random(L_loc) * (2 * w) + (hz - w)
    neighbors = rng.integers(0, L_grain, size=L_loc)
    couplings = np.zeros(L_loc)
    if alpha > 0:
        couplings[0] = 1.0
        for j in range(2, L_loc + 1):
            couplings[j - 1] = alpha ** (j - 1 + rng.random() * 2 * zeta - zeta)
    n = d ** L_grain
    X = rng.standard_normal((n, n))
    return disorder, neighbors, couplings, (X + X.T) / sqrt(2)


def qsun_spin_matrix(L_loc: int, L_grain: int, g0: float, alpha: float, *, gamma: float = 1.0,
                     w: float = 0.5, hz: float = 1.0, zeta: float = 0.2, S: float = 0.5,
                     use_U1: bool = False, S_z: float = 0.0, rng=None, draws=None):
    """The Quantum Sun as a scipy CSR matrix for spin S: the full space
    (``build_hamiltonian``, src/QSun/QuantumSun.jl:1-102: S^x S^x outer-grain
    couplings, grain gamma*GOE/sqrt(d^Lg + 1)) or, with ``use_U1``, the sector
    of total S^z = ``S_z`` (``build_hamiltonian_cons``,
    QuantumSun_U1_conserving.jl:1-212: S^x S^x + S^y S^y couplings, grain
    entries between different grain magnetizations removed and the grain part
    normalized to gamma * sqrt(tr(R^2)/dim) = gamma).

    Basis (src/QSun/HilbertSpace.jl:11-41, SpinOperators.jl:5-31): site k
    holds the digit m_k + S in [0, d), d = 2S+1, and the states sort like the
    mixed-radix number sum_k digit_k d^k, so that is the row index (in the
    sector, the rank among the sector's states).  ``draws`` =
    (disorder, neighbors, couplings, raw GOE) as from :func:`qsun_spin_draws`.
    """
    if (2 * S) != int(2 * S) or S <= 0:
        raise ValueError(f"Wrong value of spin. Only integer spins or half-integer spins are possible. You chose S = {S}")
    d = int(round(2 * S)) + 1
    L = L_loc + L_grain
    if use_U1 and (abs(S_z) > L * S or (S_z + L * S) != int(S_z + L * S)):
        raise ValueError(f"Wrong total value of spin to given spin-S particles. You chose S = {S} and S_z = {S_z}.")
    if draws is None:
        draws = qsun_spin_draws(L_loc, L_grain, alpha, d, w=w, hz=hz, zeta=zeta, rng=rng)
    disorder, neighbors, couplings, goe = draws
    N = d ** L
    idx = np.arange(N, dtype=np.int64)
    pw = d ** np.arange(L, dtype=np.int64)
    m = (idx[:, None] // pw[None, :]) % d - S            # (N, L) spin values
    if use_U1:
        keep = m.sum(axis=1) == S_z
        rank = np.full(N, -1, dtype=np.int64)
        rank[keep] = np.arange(int(keep.sum()))
    else:
        keep = np.ones(N, dtype=bool)
        rank = idx
    dim = int(keep.sum())
    rows, cols, vals = [], [], []
    k = idx[keep]
    # fields on the outer spins: sum_j h_j S^z
    rows.append(rank[k])
    cols.append(rank[k])
    vals.append((m[k][:, L_grain:] * disorder[None, :]).sum(axis=1))

    def ladder(mm, up):  # S^+ / S^- amplitude sqrt(S(S+1) - m(m +- 1))
        return np.sqrt(np.maximum(S * (S + 1) - mm * (mm + (1 if up else -1)), 0.0))

    for j in range(L_loc):
        sl, sg = L_grain + j, int(neighbors[j])
        ml, mg = m[k, sl], m[k, sg]
        for ul in (True, False):
            okl = ml < S if ul else ml > -S
            for ug in (True, False):
                okg = mg < S if ug else mg > -S
                ok = okl & okg
                # S^x S^x: (1/2)(1/2) amplitudes; + S^y S^y in the U(1) form:
                # (1/2i)(1/2i) sign_l sign_g, i.e. -(1/4) sign_l sign_g
                coef = 0.25 if not use_U1 else 0.25 * (1 - (1 if ul == ug else -1))
                if coef == 0.0:
                    continue
                new = k + (pw[sl] if ul else -pw[sl]) + (pw[sg] if ug else -pw[sg])
                ok &= rank[np.where(ok, new, 0)] >= 0
                amp = ladder(ml, ul) * ladder(mg, ug)
                rows.append(rank[new[ok]])
                cols.append(rank[k[ok]])
                vals.append(g0 * couplings[j] * coef * amp[ok])
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(dim, dim)).tocsr()
    # grain: kron(I_loc, G) on the low L_grain sites
    n = d ** L_grain
    G = np.array(goe, dtype=np.float64, copy=True)
    if use_U1:
        gm = ((np.arange(n)[:, None] // pw[None, :L_grain]) % d - S).sum(axis=1)
        G[gm[:, None] != gm[None, :]] = 0.0
    else:
        G = gamma * G / sqrt(n + 1)
    a = idx[keep] % n
    blk = idx[keep] // n
    # pairs of sector states in the same outer configuration (blk) and any grain states
    order = np.argsort(blk, kind="stable")
    blk_s, a_s, r_s = blk[order], a[order], rank[idx[keep]][order]
    bounds = np.flatnonzero(np.diff(np.r_[-1, blk_s, -2]))
    gr, gc, gv = [], [], []
    for lo, hi in zip(bounds[:-1], bounds[1:]):
        aa, rr = a_s[lo:hi], r_s[lo:hi]
        vv = G[np.ix_(aa, aa)]
        nz = vv != 0.0
        ii, jj = np.nonzero(nz)
        gr.append(rr[ii])
        gc.append(rr[jj])
        gv.append(vv[nz])
    R = sp.coo_matrix((np.concatenate(gv), (np.concatenate(gr), np.concatenate(gc))),
                      shape=(dim, dim)).tocsr()
    if use_U1:
        R = gamma * R / sqrt((R.multiply(R)).sum() / dim)
    return (H + R).tocsr()


def qsun_hamiltonian(L_loc: int, L_grain: int, g0: float, alpha: float, *, gamma: float = 1.0,
                     w: float = 0.5, hz: float = 1.0, zeta: float = 0.2, S: float = 0.5,
                     rng=None, use_U1: bool = False, S_z: float = 0.0, draws=None,
                     device=None) -> Operator:
    """``qsun_hamiltonian`` (src/QSun/QSun.jl:37-84).  The spin-1/2 full space
    is the matrix-free Quantum-Sun operator (``draws`` = (disorder, neighbors,
    couplings, H_grain) reproduces a realization); spin S and the U(1) sector
    (``use_U1``, total S^z = ``S_z``) are built by :func:`qsun_spin_matrix`
    and run on the packed operator (``draws`` then holds the raw GOE)."""
    if (2 * S) != int(2 * S):
        raise ValueError(f"Wrong value of spin. Only integer spins or half-integer spins are possible. You chose S = {S}")
    if S != 0.5 or use_U1:
        H = qsun_spin_matrix(L_loc, L_grain, g0, alpha, gamma=gamma, w=w, hz=hz, zeta=zeta, S=S,
                             use_U1=use_U1, S_z=S_z, rng=rng, draws=draws)
        return PackedOperator.from_matrix(H, device=device)
    if draws is None:
        draws = qsun_draws(L_loc, L_grain, alpha, gamma=gamma, w=w, hz=hz, zeta=zeta, rng=rng)
    disorder, neighbors, couplings, H_grain = draws
    return QSunOperator(L_loc, L_grain, g0, disorder, neighbors, couplings, H_grain,
                        device=device)
