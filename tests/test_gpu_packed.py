"""GPU parity of the packed sparse operator (pf_op_create_packed: any real
symmetric matrix in the reference's value-bucket form) and of polfed(matrix),
against scipy, the oracle and dense eigensolvers."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import clenshaw as oc
from oracle import dos as odos
from oracle import models as om

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _random_symmetric(D, nvals, density, seed, empty_row=None):
    """Sparse symmetric matrix with nvals distinct off-diagonal values."""
    rng = np.random.default_rng(seed)
    A = sp.random(D, D, density=density, random_state=seed, format="coo")
    levels = rng.uniform(-1, 1, nvals)
    A = sp.coo_matrix((rng.choice(levels, A.nnz), (A.row, A.col)), shape=(D, D))
    A = sp.triu(A, 1)
    A = (A + A.T).tolil()
    if empty_row is not None:
        A[empty_row, :] = 0
        A[:, empty_row] = 0
    A.setdiag(rng.uniform(-2, 2, D))
    return A.tocsr()


def _packed_matrix(diag, vals, starts, flat):
    """The matrix the packed arrays describe (rounded values), from the arrays."""
    D = len(diag)
    rows, cols, vv = [], [], []
    for v in range(len(vals)):
        for r in range(D):
            cs = flat[starts[v, r]:starts[v, r + 1]]
            rows += [r] * len(cs)
            cols += list(cs)
            vv += [vals[v]] * len(cs)
    return (sp.coo_matrix((vv, (rows, cols)), shape=(D, D)) + sp.diags(diag)).tocsr()


def _matrices():
    bonds, Js, Ds = om.xxz_bonds(12, 1.0, 1.0, "periodic")
    h = np.random.default_rng(1).uniform(-2, 2, 12)
    xxz = om.spin_half_hamiltonian(12, bonds, Js, Ds, h, 6)
    bonds, Js, Ds = om.j1j2_bonds(12, 1.0, 0.5, 1.0, 0.7, "periodic")
    j1j2 = om.spin_half_hamiltonian(12, bonds, Js, Ds, h, 6)
    draws = om.qsun_draws(6, 3, 0.75, rng=np.random.default_rng(3))
    qsun = om.qsun_hamiltonian_from_draws(6, 3, 1.0, *draws)
    return {"xxz_L12": xxz, "j1j2_L12": j1j2, "qsun_6_3": qsun,
            "random_200vals": _random_symmetric(1500, 200, 0.01, 4, empty_row=7)}


MATS = _matrices()


@pytest.mark.parametrize("name", list(MATS))
def test_packed_apply_filter_kpm(dev, name):
    from polfed_amd.blockops import fill_probes
    from polfed_amd.operators import PackedOperator
    from polfed_amd.packing import pack_by_value
    packed = pack_by_value(MATS[name])
    H = _packed_matrix(*packed)
    op = PackedOperator(*packed, device=dev)
    D = H.shape[0]
    assert op.D == D
    # connectivity: exactly the packed matrix
    cols, vals, counts, diag = op.connectivity(max_per_row=int(np.diff(H.indptr).max()) + 1)
    assert np.array_equal(diag, H.diagonal())
    rr = np.repeat(np.arange(D), counts)
    cc = np.concatenate([cols[r, :counts[r]] for r in range(D)])
    vv = np.concatenate([vals[r, :counts[r]] for r in range(D)])
    assert abs(sp.coo_matrix((vv, (rr, cc)), shape=(D, D)) + sp.diags(diag) - H).max() == 0.0
    for s in (1, 3, 8):
        X = np.random.default_rng(s).standard_normal((D, s))
        assert _rel(dev.to_host(op.apply(dev.to_device(X))), H @ X) < 1e-13, s
        Y32 = dev.to_host(op.apply(dev.to_device(X).float(), dtype="f32"))
        assert _rel(Y32, H @ X) < 1e-5, s
    g = abs(H).sum(axis=1).max()
    op.set_rescale(-g, g)
    Ht = H / g
    c = oc.filter_coefficients(0.2, 90)
    X = np.random.default_rng(9).standard_normal((D, 8))
    assert _rel(dev.to_host(op.filter(c, dev.to_device(X))), oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)) < 1e-12
    P = fill_probes(dev, D, 10, seed=2)
    mu = op.kpm_moments(40, P)
    mu_ref = odos.dos_moments(lambda v: Ht @ v, 40, dev.to_host(P))
    assert np.max(np.abs(mu - mu_ref)) / D < 1e-12


def test_packed_clenshaw_hooks_closed_form(dev):
    """recurrence b1 = c X + 2 H~ b2 - b3 and final sum Y = c0 X + H~ b1 - b2 on
    the caller's column-major blocks (test/optimization.jl:81-99)."""
    from polfed_amd.operators import PackedOperator
    from polfed_amd.packing import pack_by_value
    packed = pack_by_value(MATS["random_200vals"])
    H = _packed_matrix(*packed)
    op = PackedOperator(*packed, device=dev)
    op.set_rescale(-3.0, 5.0)
    Ht = (H - 1.0 * sp.identity(H.shape[0])) / 4.0
    rng = np.random.default_rng(11)
    X, b2, b3 = (rng.standard_normal((H.shape[0], 4)) for _ in range(3))
    b1 = dev.to_device(np.zeros_like(X))
    op.recurrence(b1, dev.to_device(b2), dev.to_device(b3), 0.3, dev.to_device(X))
    assert _rel(dev.to_host(b1), 0.3 * X + 2 * Ht @ b2 - b3) < 1e-13
    Y = dev.to_device(np.zeros_like(X))
    op.finalsum(dev.to_device(b2), dev.to_device(b3), 0.7, Y, dev.to_device(X))
    assert _rel(dev.to_host(Y), 0.7 * X + Ht @ b2 - b3) < 1e-13


def test_packed_equals_matrix_free_chain(dev):
    """Two independent GPU operators for one XXZ chain (the packed matrix and
    the matrix-free rank-table kernel) give the same filter output."""
    from polfed_amd.operators import PackedOperator, SpinChainOperator
    L, Lup = 14, 7
    bonds, Js, Ds = om.xxz_bonds(L, 1.0, 0.8, "periodic")
    h = np.random.default_rng(5).uniform(-1, 1, L)
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup)
    pk = PackedOperator.from_matrix(H, device=dev)
    mf = SpinChainOperator(L, Lup, bonds, Js, Ds, h, device=dev)
    for op in (pk, mf):
        op.set_rescale(-12.0, 12.0)
    c = oc.filter_coefficients(-0.3, 120)
    X = dev.to_device(np.random.default_rng(6).standard_normal((H.shape[0], 8)))
    assert _rel(dev.to_host(pk.filter(c, X)), dev.to_host(mf.filter(c, X))) < 1e-12


def test_polfed_on_a_matrix(dev):
    """polfed(mat, ...) (PolfedCore.jl:153-184): the matrix is packed, the
    bounds come from the sqrt(eps) probes of optimize_spectral_transform, and
    the returned pairs are eigenpairs of the matrix to 1e-10."""
    from polfed_amd import FactorizationConfig, polfed
    H = MATS["random_200vals"]
    w = np.linalg.eigvalsh(H.toarray())
    x0 = np.random.default_rng(1).standard_normal(H.shape[0])
    vals, vecs, rep = polfed(H, x0, 10, 0.1, produce_report=True,
                             fact=FactorizationConfig(eigentol=1e-10))
    width = w[-1] - w[0]
    assert len(vals) == 10
    assert rep.Emin >= w[0] - 1e-6 * width and rep.Emax <= w[-1] + 1e-6 * width
    assert max(np.min(np.abs(w - v)) for v in vals) <= 1e-10 * width
    psi = dev.to_host(vecs)
    from polfed_amd.packing import pack_by_value
    Hp = _packed_matrix(*pack_by_value(H))
    assert np.linalg.norm(Hp @ psi - psi * vals, axis=0).max() <= 1e-10
    # the 10 returned values are the 10 eigenvalues nearest the target's
    # position in the rescaled spectrum (a contiguous run of the spectrum)
    idx = np.searchsorted(w, vals - 1e-9 * width)
    assert np.all(np.diff(idx) == 1)


def test_packed_rejects_too_many_values(dev):
    from polfed_amd import PolfedHIPError
    from polfed_amd.operators import PackedOperator
    D = 40
    rows, cols = np.triu_indices(D, 1)
    v = np.arange(1, rows.size + 1) * 1e-3  # 780 distinct values
    H = sp.coo_matrix((np.r_[v, v], (np.r_[rows, cols], np.r_[cols, rows])), shape=(D, D))
    with pytest.raises(PolfedHIPError):
        PackedOperator.from_matrix(H, device=dev)


@pytest.mark.parametrize("kw", [dict(L_loc=10, L_grain=3, S=0.5, use_U1=True, S_z=0.5),
                                dict(L_loc=5, L_grain=2, S=1.0, use_U1=False)])
def test_polfed_on_qsun_spin_and_u1_sectors(dev, kw):
    """qsun_hamiltonian(S=..., use_U1=...) (src/QSun/QSun.jl:37-84) runs on
    the packed operator; its eigenpairs match the dense spectrum of the
    host-built matrix."""
    from polfed_amd import FactorizationConfig, models, polfed
    from polfed_amd.operators import PackedOperator
    d = int(round(2 * kw["S"])) + 1
    draws = models.qsun_spin_draws(kw["L_loc"], kw["L_grain"], 0.75, d,
                                   rng=np.random.default_rng(21))
    op = models.qsun_hamiltonian(g0=1.0, alpha=0.75, draws=draws, device=dev, **kw)
    assert isinstance(op, PackedOperator)
    H = models.qsun_spin_matrix(g0=1.0, alpha=0.75, draws=draws, **kw)
    w = np.linalg.eigvalsh(H.toarray())
    width = w[-1] - w[0]
    x0 = np.random.default_rng(2).standard_normal(op.D)
    vals, vecs = polfed(op, x0, 12, "middle", fact=FactorizationConfig(eigentol=1e-10))
    assert len(vals) == 12
    assert max(np.min(np.abs(w - v)) for v in vals) <= 1e-10 * width
    psi = dev.to_host(vecs)
    from polfed_amd.packing import pack_by_value
    Hp = _packed_matrix(*pack_by_value(H))
    assert np.linalg.norm(Hp @ psi - psi * vals, axis=0).max() <= 1e-10


def test_packed_rejects_inconsistent_arrays(dev):
    """The C ABI checks every value's start table against the flat length and
    every column against D (no host read past the caller's arrays)."""
    import ctypes
    from polfed_amd import _lib
    from polfed_amd.packing import pack_by_value
    diag, vals, starts, flat = pack_by_value(MATS["j1j2_L12"])
    lib = dev.lib
    i64p = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    out = ctypes.c_void_p()
    D = diag.size
    for fl, n in [(flat, flat.size - 1), (np.where(flat == 0, D, flat).astype(np.int32), flat.size)]:
        fl = np.ascontiguousarray(fl, dtype=np.int32)
        rc = lib.pf_op_create_packed(dev.h, D, _lib.dptr(diag), vals.size, _lib.dptr(vals),
                                     i64p(np.ascontiguousarray(starts)), _lib.iptr(fl), n,
                                     ctypes.byref(out))
        assert rc != 0 and not out.value


def test_small_polfed_problem_on_the_device(dev):
    """test/test_polfed.jl:29-61 and its CPU/GPU ordering checks (:166-197):
    diag(-3, -1.25, -0.2, 0.55, 1.8, 3.2), block 2, three pairs at 0.0, order
    12, DoS N=8 R=2, tol 1e-11, eigentol 1e-7, no Krylov overestimate.  The
    six-dimensional space is exhausted after three blocks; whatever pairs come
    back are eigenpairs of the matrix (residual <= 1e-6 max(1, |E|)), in
    ascending order, with one column per value."""
    from polfed_amd import DoSConfig, FactorizationConfig, MappingConfig, TransformConfig, polfed
    mat = np.diag([-3.0, -1.25, -0.2, 0.55, 1.8, 3.2])
    q, _ = np.linalg.qr(np.random.default_rng(1234).standard_normal((6, 2)))
    vals, vecs = polfed(mat, q, 3, 0.0, mapping=MappingConfig(Emin=-3.0, Emax=3.2),
                        transform=TransformConfig(order=12),
                        fact=FactorizationConfig(tol=1e-11, eigentol=1e-7, overestimate_iters=1.0),
                        dos=DoSConfig(N=8, R=2))
    psi = dev.to_host(vecs).reshape(6, -1)
    assert psi.shape[1] == len(vals) <= 3
    assert np.all(np.diff(vals) >= 0)
    for j, v in enumerate(vals):
        assert np.linalg.norm(mat @ psi[:, j] - v * psi[:, j]) <= 1e-6 * max(1.0, abs(v))
        assert np.min(np.abs(np.diag(mat) - v)) <= 1e-6


def test_device_initial_state_checks(dev):
    """_prepare_polfed_initial_state on device blocks (test/test_polfed.jl:138-164):
    a copy is normalized / orthonormalized with a warning, the caller's array
    is untouched, and empty, zero or over-wide starts are rejected."""
    import warnings
    from polfed_amd.core import initial_state_tolerance, prepare_initial_state
    assert initial_state_tolerance() == 1e-10
    v0 = np.array([3.0, 4.0])
    v = v0.copy()
    with pytest.warns(UserWarning, match="initial vector is not normalized"):
        vp = prepare_initial_state(dev, v)
    assert np.array_equal(v, v0)
    assert abs(np.linalg.norm(dev.to_host(vp)) - 1.0) < 1e-14
    x0 = np.array([[1.0, 1.0], [0.0, 1.0], [1.0, 0.0]])
    x = x0.copy()
    with pytest.warns(UserWarning, match="initial matrix is not orthonormal"):
        xp = dev.to_host(prepare_initial_state(dev, x))
    assert np.array_equal(x, x0)
    assert np.allclose(xp.T @ xp, np.eye(2), atol=1e-14)
    # an orthonormal start is used as given, without a warning
    q = np.linalg.qr(np.random.default_rng(0).standard_normal((5, 2)))[0]
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert np.array_equal(dev.to_host(prepare_initial_state(dev, q)), q)
    for bad in (np.zeros(3), np.ones((2, 3)), np.zeros(0)):
        with pytest.raises(ValueError):
            prepare_initial_state(dev, bad)
