"""GPU parity of the operator kernels, Clenshaw hooks, filter, KPM and the
dense FP64 block kernels against the CPU oracle (called through the C ABI)."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import clenshaw as oc
from oracle import dos as odos
from oracle import models as om

pytestmark = pytest.mark.gpu

CHAINS = [
    # (name, L, Lup, bonds-builder, fields seed)
    ("xxz_L10_periodic", 10, 5, lambda L: om.xxz_bonds(L, 1.0, 1.0, "periodic"), 1),
    ("xxz_L11_open_Lup4", 11, 4, lambda L: om.xxz_bonds(L, 0.7, 1.3, "open"), 2),
    ("j1j2_L12_periodic", 12, 6, lambda L: om.j1j2_bonds(L, 1.0, 0.5, 1.0, 0.8, "periodic"), 3),
    ("j1j2_L4_periodic_dup", 4, 2, lambda L: om.j1j2_bonds(L, 0.3, 1.0, 1.0, 1.0, "periodic"), 4),
    ("j1j2_L9_open", 9, 3, lambda L: om.j1j2_bonds(L, 1.2, -0.8, 0.5, 1.5, "open"), 5),
]


def _chain(dev, L, Lup, bonds_fn, seed):
    from polfed_amd.operators import SpinChainOperator
    bonds, Js, Ds = bonds_fn(L)
    h = np.random.default_rng(seed).uniform(-2, 2, L)
    op = SpinChainOperator(L, Lup, bonds, Js, Ds, h, device=dev)
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup)
    return op, H


def _qsun(dev, L_loc, L_grain, seed):
    from polfed_amd.operators import QSunOperator
    draws = om.qsun_draws(L_loc, L_grain, 0.75, rng=np.random.default_rng(seed))
    op = QSunOperator(L_loc, L_grain, 1.0, *draws, device=dev)
    H = om.qsun_hamiltonian_from_draws(L_loc, L_grain, 1.0, *draws)
    return op, H


def _dense_rows(cols, vals, counts, D):
    """device connectivity -> scipy CSR (duplicates summed)."""
    rr, cc, vv = [], [], []
    for r in range(len(counts)):
        n = counts[r]
        rr += [r] * n
        cc += list(cols[r, :n])
        vv += list(vals[r, :n])
    return sp.coo_matrix((vv, (rr, cc)), shape=(len(counts), D)).tocsr()


def _check_chain_export(name, H, cols, vals, counts, diag):
    D = H.shape[0]
    # diagonal: bit-identical (the table the kernels stream is summed in the
    # reference's order)
    assert np.array_equal(diag, H.diagonal())
    off = H.tolil()
    off.setdiag(0)
    off = off.tocsr()
    off.eliminate_zeros()
    got = _dense_rows(cols, vals, counts, D)
    got.eliminate_zeros()
    diff = (got - off).tocsr()
    diff.eliminate_zeros()
    assert diff.nnz == 0, f"{name}: connectivity differs in {diff.nnz} entries"


@pytest.mark.parametrize("name,L,Lup,bfn,seed", CHAINS)
def test_chain_connectivity_kernel_path_bit_exact(dev, name, L, Lup, bfn, seed):
    """The export runs the step kernels' rank function (rank tables) in both
    row orders the kernels use; both equal the reference's sparse matrix."""
    op, H = _chain(dev, L, Lup, bfn, seed)
    assert op.D == H.shape[0]
    _check_chain_export(name, H, *op.connectivity(max_per_row=64))
    _check_chain_export(name + "/partition", H, *op.connectivity(max_per_row=64, order=1))


def test_qsun_connectivity_kernel_path(dev):
    """Off-diagonal entries bit-exact; the diagonal is the kernels' split table
    Dlo[r mod 2^lb] + Dhi[r >> lb] (a different summation order from the
    reference's one pass over the sites: equal to rounding)."""
    op, H = _qsun(dev, 5, 3, 11)
    cols, vals, counts, diag = op.connectivity(max_per_row=64)
    ref = H.diagonal()
    assert np.max(np.abs(diag - ref)) <= 4 * np.finfo(float).eps * np.max(np.abs(ref))
    off = H.tolil()
    off.setdiag(0)
    off = off.tocsr()
    got = _dense_rows(cols, vals, counts, H.shape[0])
    assert abs(got - off).max() == 0.0


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_long_chain_without_rank_tables(dev):
    """L = 28 > 26: no LDS rank tables, the step kernel uses the O(1)
    neighbour-rank update (with the O(Lup) periodic bond); apply and a filter
    against the oracle's sparse matrix, connectivity bit-exact."""
    from polfed_amd.operators import SpinChainOperator
    L, Lup = 28, 3
    bonds, Js, Ds = om.j1j2_bonds(L, 1.0, 0.5, 1.0, 0.7, "periodic")
    h = np.random.default_rng(12).uniform(-1, 1, L)
    op = SpinChainOperator(L, Lup, bonds, Js, Ds, h, device=dev)
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup)
    assert op.D == H.shape[0] == 3276
    cols, vals, counts, diag = op.connectivity()
    assert np.array_equal(diag, H.diagonal())
    Hd = _dense_rows(cols, vals, counts, op.D) + sp.diags(diag)
    assert abs(Hd - H).max() == 0.0
    X = np.random.default_rng(2).standard_normal((op.D, 4))
    assert _rel(dev.to_host(op.apply(dev.to_device(X))), H @ X) < 1e-13
    w = np.linalg.eigvalsh(H.toarray())
    op.set_rescale(w[0], w[-1])
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * sp.identity(op.D)) / a
    c = oc.filter_coefficients(0.3, 80)
    assert _rel(dev.to_host(op.filter(c, dev.to_device(X))), oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < 1e-12


@pytest.mark.parametrize("pblocks,wide,vec,occ", [(0, 0, 2, 0), (1, 1, 1, 0), (8, 1, 4, 0),
                                                 (8, 0, 8, 0), (8, 0, 2, 1), (0, 0, 4, 1)])
def test_filter_persistent_grid_settings_agree(dev, pblocks, wide, vec, occ):
    """The persistent-grid, gather-batch and vector-width knobs change only
    how rows and neighbours are scheduled, not the result."""
    op, H = _chain(dev, *CHAINS[2][1:])
    w = np.linalg.eigvalsh(H.toarray())
    op.set_rescale(w[0], w[-1])
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    c = oc.filter_coefficients(-0.2, 150)
    X = np.random.default_rng(7).standard_normal((H.shape[0], 8))
    op.set_tuning(pblocks=pblocks, wide=wide, vec=vec, occupancy=occ)
    t = op.tuning(8, "f64")  # the configuration the driver launches (pf_op_get_tuning)
    assert (t["pblocks"], t["wide"], t["vec"], t["occupancy"]) == (pblocks, wide, vec, occ)
    Y = dev.to_host(op.filter(c, dev.to_device(X)))
    assert _rel(Y, oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < 1e-12


PARTITION_CHAINS = CHAINS + [
    ("xxz_L16_periodic", 16, 8, lambda L: om.xxz_bonds(L, 1.0, 0.6, "periodic"), 6),
    ("j1j2_L15_open_Lup6", 15, 6, lambda L: om.j1j2_bonds(L, 1.0, 0.4, 1.0, 0.9, "open"), 7),
]


@pytest.mark.parametrize("name,L,Lup,bfn,seed", PARTITION_CHAINS)
def test_row_partition_order_changes_nothing(dev, name, L, Lup, bfn, seed):
    """The filter and KPM drivers run spin chains on the basis reordered by
    segments (XCD partition); inputs and outputs stay in the reference order
    and every row sums the same terms in the same order, so the filter output
    is bit-identical with the partition on and off (FP64 and FP32), and the
    KPM moments differ only by the order of the dot-product sums."""
    from polfed_amd.blockops import fill_probes
    op, H = _chain(dev, L, Lup, bfn, seed)
    g = abs(H).sum(axis=1).max()  # Gershgorin bound of the spectrum
    op.set_rescale(-g, g)
    a, b = (op.Emax - op.Emin) / 2, (op.Emax + op.Emin) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    c = oc.filter_coefficients(0.15, 60)
    X = np.random.default_rng(seed).standard_normal((H.shape[0], 8))
    Xd = dev.to_device(X)
    out = {}
    for part in (1, 0):
        op.set_tuning(row_partition=part)
        out[part] = [dev.to_host(op.filter(c, Xd, dtype=dt)) for dt in ("f64", "f32")]
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert _rel(out[1][0], oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < 1e-12
    P = fill_probes(dev, H.shape[0], 12, seed=seed)
    mus = {}
    for part in (1, 0):
        op.set_tuning(row_partition=part)
        mus[part] = op.kpm_moments(40, P)
    assert np.max(np.abs(mus[1] - mus[0])) / H.shape[0] < 1e-13


@pytest.mark.parametrize("L_grain,L_loc", [(1, 7), (2, 6), (4, 5)])
@pytest.mark.parametrize("s", [1, 3, 8, 16])
@pytest.mark.parametrize("shuffle", [1, 0])
def test_qsun_grain_sizes_filter(dev, L_grain, L_loc, s, shuffle):
    """Grain blocks of 2, 4 and 16 rows with 1, 3, 8 and 16 columns: the
    cross-lane shuffle path (power-of-two lanes per row) and the load path
    (3 lanes per row) against the oracle, FP64 and FP32."""
    op, H = _qsun(dev, L_loc, L_grain, 20 + L_grain)
    op.set_tuning(grain_shuffle=shuffle)
    w = np.linalg.eigvalsh(H.toarray())
    op.set_rescale(w[0] - 0.01, w[-1] + 0.01)
    a, b = (op.Emax - op.Emin) / 2, (op.Emax + op.Emin) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    c = oc.filter_coefficients(0.2, 40)
    X = np.random.default_rng(s).standard_normal((op.D, s))
    ref = oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)
    assert _rel(dev.to_host(op.filter(c, dev.to_device(X))), ref) < 1e-12
    assert _rel(dev.to_host(op.filter(c, dev.to_device(X), dtype="f32")), ref) < 5e-4
    assert _rel(dev.to_host(op.apply(dev.to_device(X))), H @ X) < 1e-13


@pytest.mark.parametrize("s", [1, 3, 8])
@pytest.mark.parametrize("kind", ["chain", "qsun"])
def test_apply_matches_sparse(dev, kind, s):
    if kind == "chain":
        op, H = _chain(dev, *CHAINS[2][1:])
    else:
        op, H = _qsun(dev, 6, 3, 7)
    D = H.shape[0]
    X = np.random.default_rng(0).standard_normal((D, s))
    Xd = dev.to_device(X)
    Y = dev.to_host(op.apply(Xd))
    assert _rel(Y, H @ X) < 1e-13
    # rescaled operator (H - b)/a
    op.set_rescale(-7.0, 9.0)
    Yr = dev.to_host(op.apply(Xd, rescaled=True))
    assert _rel(Yr, (H @ X - 1.0 * X) / 8.0) < 1e-13
    # FP32 arithmetic
    X32 = Xd.to(torch.float32)
    Y32 = op.apply(X32, rescaled=True, dtype="f32")
    assert _rel(dev.to_host(Y32).astype(np.float64), (H @ X - X) / 8.0) < 2e-6


def test_clenshaw_hooks_closed_form(dev):
    """recurrence = cX + 2 H b2 - b3 ; final = cX + H b1 - b2
    (closed forms of test/optimization.jl:81-99)."""
    op, H = _chain(dev, *CHAINS[0][1:])
    D = H.shape[0]
    rng = np.random.default_rng(3)
    X, b1, b2, b3 = (rng.standard_normal((D, 4)) for _ in range(4))
    Xd, b1d, b2d, b3d = (dev.to_device(a) for a in (X, b1, b2, b3))
    out = torch.empty_like(Xd)
    op.recurrence(out, b2d, b3d, 0.5, Xd)
    assert _rel(dev.to_host(out), 0.5 * X + 2 * (H @ b2) - b3) < 1e-13
    y = torch.empty_like(Xd)
    op.finalsum(b1d, b2d, 0.5, y, Xd)
    assert _rel(dev.to_host(y), 0.5 * X + H @ b1 - b2) < 1e-13


@pytest.mark.parametrize("s,K", [(1, 37), (8, 200), (5, 3), (16, 1), (4, 0)])
def test_filter_matches_oracle_fp64(dev, s, K):
    op, H = _chain(dev, *CHAINS[2][1:])
    ev = np.linalg.eigvalsh(H.toarray())
    Emin, Emax = ev[0], ev[-1]
    op.set_rescale(Emin, Emax)
    a, b = (Emax - Emin) / 2, (Emax + Emin) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    c = oc.filter_coefficients(0.1, K) if K > 0 else np.array([0.7])
    X = np.random.default_rng(5).standard_normal((H.shape[0], s))
    Yref = oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)
    Y = dev.to_host(op.filter(c, dev.to_device(X)))
    assert _rel(Y, Yref) < 1e-12


def test_filter_dense_spectrum_and_fp32(dev):
    """p(H~) X against V diag(p(lambda)) V^T X (numpy chebval on the exact
    spectrum), FP64 and FP32 arithmetic."""
    op, H = _qsun(dev, 6, 3, 9)
    w, V = np.linalg.eigh(H.toarray())
    op.set_rescale(w[0], w[-1])
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    c = oc.filter_coefficients(0.2, 120)
    p = oc.chebyshev_filter_dense((w - b) / a, c)
    X = np.random.default_rng(1).standard_normal((H.shape[0], 6))
    Yref = V @ (p[:, None] * (V.T @ X))
    Y64 = dev.to_host(op.filter(c, dev.to_device(X)))
    assert _rel(Y64, Yref) < 1e-11
    Y32 = dev.to_host(op.filter(c, dev.to_device(X), dtype="f32"))
    assert _rel(Y32, Yref) < 5e-4


@pytest.mark.parametrize("polytype", ["Legendre", "Hermite", "Taylor"])
def test_filter_other_polynomials(dev, polytype):
    op, H = _chain(dev, *CHAINS[0][1:])
    w = np.linalg.eigvalsh(H.toarray())
    op.set_rescale(w[0], w[-1])
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    c = np.random.default_rng(2).standard_normal(9) * 0.1
    X = np.random.default_rng(4).standard_normal((H.shape[0], 3))
    Yref = oc.clenshaw_apply(lambda Z: Ht @ Z, c, X, polytype)
    Y = dev.to_host(op.filter(c, dev.to_device(X), polytype=polytype))
    assert _rel(Y, Yref) < 1e-12


@pytest.mark.parametrize("dtype,tol", [("f64", 1e-12), ("f32", 2e-4)])
def test_kpm_moments_match_oracle(dev, dtype, tol):
    from polfed_amd.blockops import fill_probes
    op, H = _chain(dev, *CHAINS[2][1:])
    w = np.linalg.eigvalsh(H.toarray())
    op.set_rescale(w[0], w[-1])
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    R, N = 40, 64
    P = fill_probes(dev, H.shape[0], R, seed=1234)
    mu = op.kpm_moments(N, P, dtype=dtype)
    mu_ref = odos.dos_moments(lambda v: Ht @ v, N, dev.to_host(P))
    assert np.max(np.abs(mu - mu_ref)) / H.shape[0] < tol


def test_gram_update_reorth_qr(dev):
    from polfed_amd import blockops as bo
    rng = np.random.default_rng(7)
    for D, m, n in [(1000, 37, 8), (777, 130, 16), (4096, 64, 32), (50, 5, 3), (3001, 200, 64)]:
        A = rng.standard_normal((D, m))
        B = rng.standard_normal((D, n))
        Ad, Bd = dev.to_device(A), dev.to_device(B)
        G = dev.to_host(bo.gram(dev, Ad, Bd))
        assert _rel(G, A.T @ B) < 1e-13, (D, m, n)
        Gm = rng.standard_normal((m, n))
        Bd2 = dev.to_device(B)
        bo.update(dev, Ad, dev.to_device(Gm), Bd2, alpha=-0.5, beta=2.0)
        assert _rel(dev.to_host(Bd2), 2.0 * B - 0.5 * A @ Gm) < 1e-13, (D, m, n)
    # CGS2 reorth against an orthonormal basis
    D, k, s = 5000, 120, 8
    V, _ = np.linalg.qr(rng.standard_normal((D, k)))
    W = rng.standard_normal((D, s))
    Wd = dev.to_device(W)
    bo.reorth(dev, dev.to_device(V), Wd, passes=2)
    Wr = dev.to_host(Wd)
    ref = W - V @ (V.T @ W)
    assert _rel(Wr, ref) < 1e-12
    assert np.max(np.abs(V.T @ Wr)) < 1e-13 * np.max(np.abs(W))
    # Cholesky-QR2
    Wd = dev.to_device(W)
    R = bo.block_qr(dev, Wd)
    Q = dev.to_host(Wd)
    assert np.max(np.abs(Q.T @ Q - np.eye(s))) < 1e-14
    assert np.max(np.abs(Q @ R - W)) < 1e-12 * np.max(np.abs(W))
    assert np.allclose(R, np.triu(R))
    # rank-deficient block: NumericBreakdown, W unchanged
    from polfed_amd import NumericBreakdown
    Wdef = W.copy()
    Wdef[:, 3] = Wdef[:, 1]
    Wd = dev.to_device(Wdef)
    with pytest.raises(NumericBreakdown):
        bo.block_qr(dev, Wd)
    assert np.array_equal(dev.to_host(Wd), Wdef)


def test_residuals_and_coldots(dev):
    from polfed_amd import blockops as bo
    op, H = _chain(dev, *CHAINS[0][1:])
    w, V = np.linalg.eigh(H.toarray())
    psi = V[:, :5] + 1e-6 * np.random.default_rng(0).standard_normal((H.shape[0], 5))
    E, res = bo.residuals(dev, op, dev.to_device(psi))
    Eref = np.einsum("ij,ij->j", psi, H @ psi)
    rref = np.linalg.norm(H @ psi - psi * Eref, axis=0)
    assert np.max(np.abs(E - Eref)) < 1e-12
    assert np.max(np.abs(res - rref) / rref) < 1e-6
    d = bo.coldots(dev, dev.to_device(psi), dev.to_device(H @ psi))
    assert np.max(np.abs(d - Eref)) < 1e-12


@pytest.mark.parametrize("L,Lup,bc", [(8, 0, "periodic"), (8, 8, "open"), (2, 1, "periodic"),
                                      (3, 1, "open"), (26, 1, "periodic")])
def test_sector_edges(dev, L, Lup, bc):
    """Edge sectors: D = 1 (no flippable bond at all), the two-site chain (its
    periodic bonds coincide), a chain too short for next-nearest bonds, and
    the largest L with rank tables at one particle: connectivity, apply and a
    filter against the oracle."""
    bonds, Js, Ds = om.j1j2_bonds(L, 1.0, 0.4, 0.9, 1.1, bc) if L >= 4 else om.xxz_bonds(L, 1.0, 0.8, bc)
    h = np.random.default_rng(L + Lup).uniform(-1, 1, L)
    from polfed_amd.operators import SpinChainOperator
    op = SpinChainOperator(L, Lup, bonds, Js, Ds, h, device=dev)
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup)
    assert op.D == H.shape[0]
    cols, vals, counts, diag = op.connectivity(max_per_row=128)
    assert np.array_equal(diag, H.diagonal())
    X = np.random.default_rng(1).standard_normal((op.D, 3))
    assert _rel(dev.to_host(op.apply(dev.to_device(X))), H @ X) < 1e-13
    g = abs(H).sum(axis=1).max() + 1.0
    op.set_rescale(-g, g)
    Ht = H / g
    c = oc.filter_coefficients(0.1, 30)
    assert _rel(dev.to_host(op.filter(c, dev.to_device(X))), oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < 1e-12


@pytest.mark.parametrize("dtype,s,tol", [("f64", 40, 1e-12), ("f32", 70, 5e-4)])
def test_filter_and_kpm_span_several_column_chunks(dev, dtype, s, tol):
    """Blocks wider than one internal chunk (32 FP64 / 64 FP32 columns) and
    more KPM probes than one chunk: every chunk against the oracle."""
    from polfed_amd.blockops import fill_probes
    op, H = _chain(dev, *CHAINS[2][1:])
    g = abs(H).sum(axis=1).max()
    op.set_rescale(-g, g)
    Ht = H / g
    c = oc.filter_coefficients(-0.05, 50)
    X = np.random.default_rng(5).standard_normal((H.shape[0], s))
    Y = dev.to_host(op.filter(c, dev.to_device(X), dtype=dtype))
    assert _rel(Y, oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < tol
    P = fill_probes(dev, H.shape[0], s, seed=9)
    mu = op.kpm_moments(24, P, dtype=dtype)
    mu_ref = odos.dos_moments(lambda v: Ht @ v, 24, dev.to_host(P))
    assert np.max(np.abs(mu - mu_ref)) / H.shape[0] < (1e-12 if dtype == "f64" else 2e-4)
