"""Oracle restatement of the reference model constructors (test infrastructure).

Restates ``src/Models/SpinHalfChains.jl`` and ``src/QSun/*.jl`` of the
reference with numpy/scipy.  Random draws cannot reproduce Julia's RNG streams,
so every random quantity is an explicit array produced by a numpy ``Generator``
in the *same draw order* as the reference (documented per function).
"""
from __future__ import annotations

from math import comb, sqrt

import numpy as np
import scipy.sparse as sp

SUPPORTED_BOUNDARIES = ("open", "periodic")


def validate_chain_length(L: int) -> None:
    """``_validate_chain_length`` (src/Models/SpinHalfChains.jl:222-225)."""
    if L < 2:
        raise ValueError(f"L must be at least 2; got L={L}.")


def chain_bonds(L: int, distance: int, boundary: str) -> list[tuple[int, int]]:
    """``_chain_bonds`` (src/Models/SpinHalfChains.jl:227-243); 1-based sites.

    Periodic bonds are (i, mod1(i+d, L)) for i = 1..L, skipping i == j; this
    double-counts distance-2 bonds at L = 4 exactly like the reference.
    """
    if boundary not in SUPPORTED_BOUNDARIES:
        raise ValueError(f"boundary must be one of {SUPPORTED_BOUNDARIES}; got {boundary}.")
    if boundary == "open":
        if not distance < L:
            return []
        return [(i, i + distance) for i in range(1, L - distance + 1)]
    bonds = []
    for i in range(1, L + 1):
        j = (i + distance - 1) % L + 1
        if i == j:
            continue
        bonds.append((i, j))
    return bonds


def resolve_fields(L: int, field: float, W: float, fields=None, rng=None) -> np.ndarray:
    """``_resolve_fields`` (src/Models/SpinHalfChains.jl:245-257).

    h = field + W (2u - 1) with u = rng.random(L) (one uniform per site, site
    order), or the explicit ``fields``.
    """
    if fields is not None:
        fields = np.asarray(fields, dtype=np.float64)
        if fields.shape != (L,):
            raise ValueError(f"fields must have length L={L}; got length {fields.size}.")
        return fields.copy()
    W = float(W)
    if W < 0:
        raise ValueError(f"W must be non-negative; got {W}.")
    center = float(field)
    if W == 0:
        return np.full(L, center)
    if rng is None:
        rng = np.random.default_rng()
    u = rng.random(L)
    return center + W * (2.0 * u - 1.0)


def lup_basis(L: int, Lup: int) -> np.ndarray:
    """``_lup_basis`` (src/Models/SpinHalfChains.jl:259-262): ascending states
    with popcount Lup (brute-force enumeration)."""
    if not 0 <= Lup <= L:
        raise ValueError(f"Lup must satisfy 0 <= Lup <= L; got Lup={Lup} for L={L}.")
    if L > 22:
        # sparse sectors of long chains: the same set from the Lup-subsets of
        # the sites, sorted ascending (agreement with the enumeration below is
        # checked for L <= 16 in the oracle tests)
        return lup_basis_combinations(L, Lup)
    states = np.arange(1 << L, dtype=np.int64)
    pc = np.zeros_like(states)
    x = states.copy()
    while np.any(x):
        pc += x & 1
        x >>= 1
    return states[pc == Lup]


def combinadic_rank(states: np.ndarray) -> np.ndarray:
    """Closed-form 0-based position of a fixed-popcount state in the ascending
    basis: r(s) = sum_t C(p_t, t+1) over set-bit positions p_0 < p_1 < ...

    Equals the reference's Dict index - 1 (src/Models/SpinHalfChains.jl:278);
    cross-checked against ``np.searchsorted(lup_basis(...))`` in the tests.
    """
    states = np.asarray(states, dtype=np.int64)
    rank = np.zeros_like(states)
    t = np.zeros_like(states)
    nbits = int(states.max()).bit_length() if states.size else 0
    table = np.array([[comb(p, k) for k in range(nbits + 2)] for p in range(nbits + 1)],
                     dtype=np.int64)
    for p in range(nbits):
        bit = (states >> p) & 1
        rank += bit * table[p, t + 1]
        t += bit
    return rank


def spin_z(states: np.ndarray, site: int) -> np.ndarray:
    """``_spin_z`` (src/Models/SpinHalfChains.jl:264): +-0.5 from bit site-1."""
    return np.where(((states >> (site - 1)) & 1) == 1, 0.5, -0.5)


def spin_half_parts(L, bonds, couplings, anisotropies, fields, Lup):
    """Diagonal and off-diagonal triplets of ``_spin_half_hamiltonian``
    (src/Models/SpinHalfChains.jl:267-315), vectorized over states.

    The diagonal is accumulated in the reference's order -- fields for sites
    1..L, then J*Delta*zi*zj for bonds in order -- so it is bit-identical to
    the reference's H[i, i].  Off-diagonal entries J/2 are emitted per
    antiparallel bond; the index of the flipped state uses ``searchsorted`` on
    the ascending basis (the reference uses a Dict, same result).
    Returns (basis, diag, rows, cols, vals) with 0-based indices.
    """
    basis = lup_basis(L, Lup)
    diag = np.zeros(basis.size)
    for site in range(1, L + 1):
        diag = diag + fields[site - 1] * spin_z(basis, site)
    rows, cols, vals = [], [], []
    col_idx = np.arange(basis.size, dtype=np.int64)
    for b, (i, j) in enumerate(bonds):
        J = float(couplings[b])
        D = float(anisotropies[b])
        zi = spin_z(basis, i)
        zj = spin_z(basis, j)
        diag = diag + (J * D) * zi * zj
        anti = zi != zj
        flipped = basis[anti] ^ (1 << (i - 1)) ^ (1 << (j - 1))
        rows.append(np.searchsorted(basis, flipped))
        cols.append(col_idx[anti])
        vals.append(np.full(int(anti.sum()), J / 2))
    return basis, diag, rows, cols, vals


def lup_basis_combinations(L: int, Lup: int) -> np.ndarray:
    """Ascending L-bit states with popcount Lup built from the Lup-subsets of
    the sites (for sectors whose 2^L enumeration is too large)."""
    import itertools
    out = np.fromiter((sum(1 << p for p in c) for c in itertools.combinations(range(L), Lup)),
                      dtype=np.int64)
    return np.sort(out)


def spin_half_hamiltonian(L, bonds, couplings, anisotropies, fields, Lup, use_sparse=True):
    """``_spin_half_hamiltonian`` (src/Models/SpinHalfChains.jl:267-315) as a
    scipy CSC matrix (duplicates summed like ``sparse`` at :313)."""
    basis, diag, rows, cols, vals = spin_half_parts(L, bonds, couplings, anisotropies, fields, Lup)
    dim = basis.size
    idx = np.arange(dim, dtype=np.int64)
    r = np.concatenate(rows + [idx])
    c = np.concatenate(cols + [idx])
    v = np.concatenate(vals + [diag])
    H = sp.coo_matrix((v, (r, c)), shape=(dim, dim)).tocsc()
    H.sum_duplicates()
    return H if use_sparse else H.toarray()


def xxz_bonds(L, J, Delta, boundary="periodic"):
    bonds = chain_bonds(L, 1, boundary)
    return bonds, [float(J)] * len(bonds), [float(Delta)] * len(bonds)


def j1j2_bonds(L, J1, J2, D1, D2, boundary="periodic"):
    """Bond list of ``j1j2_hamiltonian`` (src/Models/SpinHalfChains.jl:203-208):
    nearest bonds first, then next-nearest."""
    nn = chain_bonds(L, 1, boundary)
    nnn = chain_bonds(L, 2, boundary)
    return (nn + nnn, [float(J1)] * len(nn) + [float(J2)] * len(nnn),
            [float(D1)] * len(nn) + [float(D2)] * len(nnn))


def xxz_hamiltonian(L, Lup, J, Delta, W, boundary="periodic", field=0.0, fields=None,
                    rng=None, use_sparse=True):
    """``xxz_hamiltonian`` (src/Models/SpinHalfChains.jl:74-99)."""
    validate_chain_length(L)
    bonds, Js, Ds = xxz_bonds(L, J, Delta, boundary)
    h = resolve_fields(L, field, W, fields, rng)
    return spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup, use_sparse)


def j1j2_hamiltonian(L, Lup, J1, J2, D1, D2, W, boundary="periodic", field=0.0, fields=None,
                     rng=None, use_sparse=True):
    """``j1j2_hamiltonian`` (src/Models/SpinHalfChains.jl:189-220)."""
    validate_chain_length(L)
    bonds, Js, Ds = j1j2_bonds(L, J1, J2, D1, D2, boundary)
    h = resolve_fields(L, field, W, fields, rng)
    return spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup, use_sparse)


# --------------------------------------------------------------------------
# Quantum Sun (full Hilbert space, spin-1/2)
# --------------------------------------------------------------------------

def qsun_draws(L_loc, L_grain, alpha, gamma=1.0, w=0.5, hz=1.0, zeta=0.2, rng=None):
    """Random inputs of ``build_hamiltonian`` (src/QSun/QuantumSun.jl:21-30) in
    the reference's draw order: disorder (``uniform_array``), neighbors
    (``random_neighbors``), couplings (``long_range_couplings``), then the GOE
    grain (src/QSun/RandomUtils.jl:1-26).  numpy stands in for Julia's RNG.
    """
    if rng is None:
        rng = np.random.default_rng()
    disorder = rng.random(L_loc) * ((hz + w) - (hz - w)) + (hz - w)
    neighbors = rng.integers(0, L_grain, size=L_loc)
    couplings = np.zeros(L_loc)
    if alpha > 0:
        couplings[0] = 1.0
        for j in range(2, L_loc + 1):
            u = j - 1 + rng.random() * 2 * zeta - zeta
            couplings[j - 1] = alpha ** u
    dg = 2 ** L_grain
    X = rng.standard_normal((dg, dg))
    goe = (X + X.T) / sqrt(2)
    H_grain = gamma * goe
    H_grain = H_grain / sqrt(dg + 1)
    return disorder, neighbors.astype(np.int64), couplings, H_grain


def qsun_hamiltonian_from_draws(L_loc, L_grain, g0, disorder, neighbors, couplings, H_grain,
                                use_sparse=True):
    """``build_hamiltonian`` (src/QSun/QuantumSun.jl:1-102) for S = 1/2 with
    explicit draws.  State = index (``build_hilbert`` sorts all 2^L integers,
    src/QSun/HilbertSpace.jl:11-31); grain = low L_grain bits.

    Diagonal: sum_j disorder[j] * S^z(site L_grain+j-1), plus H_grain[a, a]
    from ``kron(I, H_grain)`` (:99).  Off-diagonal per j: g0*couplings[j]*1/4
    flipping outer bit L_grain+j-1 and grain bit neighbors[j]
    (``sigma_x`` factors 1/2 * 1/2, src/QSun/SpinOperators.jl:34-53).
    """
    L = L_loc + L_grain
    dim = 1 << L
    states = np.arange(dim, dtype=np.int64)
    diag = np.zeros(dim)
    for j in range(1, L_loc + 1):
        site = L_grain + j - 1
        diag = diag + disorder[j - 1] * np.where(((states >> site) & 1) == 1, 0.5, -0.5)
    rows, cols, vals = [states], [states], [diag]
    for j in range(1, L_loc + 1):
        site_loc = L_grain + j - 1
        site_grain = int(neighbors[j - 1])
        new = states ^ (1 << site_loc) ^ (1 << site_grain)
        rows.append(new)
        cols.append(states)
        vals.append(np.full(dim, g0 * couplings[j - 1] * 0.5 * 0.5))
    H = sp.coo_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(dim, dim)).tocsc()
    H.sum_duplicates()
    dim_loc = 1 << L_loc
    H = (H + sp.kron(sp.identity(dim_loc, format="csc"), sp.csc_matrix(H_grain))).tocsc()
    return H if use_sparse else H.toarray()


def qsun_hamiltonian(L_loc, L_grain, g0, alpha, gamma=1.0, w=0.5, hz=1.0, zeta=0.2, rng=None,
                     use_sparse=True):
    """``qsun_hamiltonian`` (src/QSun/QSun.jl:37-84), full space, S = 1/2."""
    draws = qsun_draws(L_loc, L_grain, alpha, gamma, w, hz, zeta, rng)
    return qsun_hamiltonian_from_draws(L_loc, L_grain, g0, *draws, use_sparse=use_sparse)
