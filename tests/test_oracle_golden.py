"""CPU: pin the oracle against the reference's own known answers
(tests/golden/reference_kat.json, copied from the reference's test files) and
against closed forms / brute force that a plausible restatement mistake fails."""
import json
import math
import pathlib

import numpy as np
import pytest
import scipy.sparse as sp

from oracle import clenshaw as oc
from oracle import dos as odos
from oracle import lanczos as olz
from oracle import models as om
from oracle import packing as opk
from oracle import polfed as opf
from oracle import spectral as osp

KAT = json.loads((pathlib.Path(__file__).parent / "golden" / "reference_kat.json").read_text())


def _xxz(a):
    return om.xxz_hamiltonian(a["L"], a["Lup"], a["J"], a["Delta"], a["W"],
                              boundary=a.get("boundary", "periodic"), fields=a.get("fields"),
                              use_sparse=False)


def _j1j2(a):
    return om.j1j2_hamiltonian(a["L"], a["Lup"], a["J1"], a["J2"], a["D1"], a["D2"], a["W"],
                               boundary=a.get("boundary", "periodic"), fields=a.get("fields"),
                               use_sparse=False)


# ------------------------------------------------------------------ models
def test_xxz_known_matrices():
    k = KAT["xxz_2site_open"]
    H = _xxz(k["args"])
    assert np.allclose(H, k["expected"], atol=1e-15)
    per = dict(k["args"], boundary="periodic")
    assert np.allclose(_xxz(per), KAT["xxz_2site_periodic_is_twice_open"]["factor"] * H)
    k = KAT["xxz_three_site"]
    assert np.allclose(_xxz(k["args"]), k["expected"], atol=1e-15)
    k = KAT["xxz_fields_only"]
    assert np.allclose(_xxz(k["args"]), np.diag(k["expected_diag"]))


def test_j1j2_known_matrices():
    k = KAT["j1j2_hardcoded"]
    assert np.allclose(_j1j2(k["args"]), k["expected"], atol=1e-15)
    a = KAT["j1j2_j1_only_equals_xxz"]["args"]
    xxz = dict(L=a["L"], Lup=a["Lup"], J=a["J1"], Delta=a["D1"], W=0.0, boundary="open",
               fields=a["fields"])
    assert np.allclose(_j1j2(a), _xxz(xxz))
    a = KAT["j1j2_periodic_j2_doubles_at_L4"]["args"]
    assert np.allclose(_j1j2(dict(a, boundary="periodic")), 2 * _j1j2(dict(a, boundary="open")))
    k = KAT["j1j2_triangle_equals_xxz"]
    assert np.allclose(_j1j2(k["j1j2"]), _xxz(k["xxz"]))


def test_model_sizes_and_hermiticity():
    s = KAT["model_sizes"]
    assert om.xxz_hamiltonian(4, 2, 1.0, 0.7, 0.0, fields=[0.2, -0.1, 0.3, -0.4]).shape == tuple(s["xxz_L4_Lup2"])
    H = om.j1j2_hamiltonian(6, 3, 1.0, 0.5, 1.0, 0.6, 0.4, rng=np.random.default_rng(12))
    assert H.shape == tuple(s["j1j2_L6_Lup3"])
    assert abs(H - H.T).max() == 0
    assert om.j1j2_hamiltonian(5, 2, 0.9, -0.4, 0.7, 1.2, 0.0,
                               fields=[0.4, -0.2, 0.1, -0.3, 0.5]).shape == tuple(s["j1j2_L5_Lup2"])
    Hq = om.qsun_hamiltonian(2, 2, 1.0, 0.5, rng=np.random.default_rng(1234))
    assert Hq.shape == tuple(s["qsun_Lloc2_Lgrain2"])
    assert abs(Hq - Hq.T).max() < 1e-15
    assert om.qsun_hamiltonian(1, 1, 1.0, 0.5, rng=np.random.default_rng(1)).shape == tuple(s["qsun_Lloc1_Lgrain1"])


def test_model_argument_validation():
    """test/test_models.jl:212-226."""
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(1, 0, 1.0, 1.0, 0.0)
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(4, 2, 1.0, 1.0, 0.0, boundary="twisted")
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(4, -1, 1.0, 1.0, 0.0)
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(4, 5, 1.0, 1.0, 0.0)
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(4, 2, 1.0, 1.0, 0.0, fields=[0.0, 1.0])
    with pytest.raises(ValueError):
        om.xxz_hamiltonian(4, 2, 1.0, 1.0, -1.0)
    with pytest.raises(ValueError):
        om.j1j2_hamiltonian(4, 2, 1.0, 0.5, 1.0, 0.8, 0.0, boundary="twisted")


@pytest.mark.parametrize("L,Lup", [(4, 2), (9, 0), (12, 5), (16, 8), (16, 3)])
def test_basis_from_combinations_matches_enumeration(L, Lup):
    assert np.array_equal(om.lup_basis_combinations(L, Lup), om.lup_basis(L, Lup))


@pytest.mark.parametrize("L", [4, 9, 14, 16])
def test_combinadic_rank_is_ascending_index(L):
    """closed-form rank == position in the brute-force ascending basis."""
    for Lup in range(L + 1):
        b = om.lup_basis(L, Lup)
        assert np.array_equal(om.combinadic_rank(b), np.arange(b.size))
        assert np.all(np.diff(b) > 0)


def test_spin_chain_matches_bruteforce_dense():
    """Dense construction from the definition of H (Pauli algebra over the
    full 2^L space, projected to the Lup sector) equals the restatement."""
    L, Lup = 6, 3
    h = np.random.default_rng(0).uniform(-1, 1, L)
    bonds, Js, Ds = om.j1j2_bonds(L, 1.1, 0.4, 0.7, 1.3, "periodic")
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, Lup, use_sparse=False)
    sx = np.array([[0, 0.5], [0.5, 0]])
    sy = np.array([[0, -0.5j], [0.5j, 0]])
    sz = np.array([[-0.5, 0], [0, 0.5]])   # bit 1 = up = +1/2

    def op(single, site):   # site 1-based, bit site-1 (little-endian kron)
        m = np.array([[1.0]])
        for k in range(L, 0, -1):
            m = np.kron(m, single if k == site else np.eye(2))
        return m

    Hf = np.zeros((2 ** L, 2 ** L), complex)
    for b, (i, j) in enumerate(bonds):
        Hf += Js[b] * (op(sx, i) @ op(sx, j) + op(sy, i) @ op(sy, j) + Ds[b] * op(sz, i) @ op(sz, j))
    for i in range(L):
        Hf += h[i] * op(sz, i + 1)
    basis = om.lup_basis(L, Lup)
    assert np.allclose(Hf[np.ix_(basis, basis)].real, H, atol=1e-14)
    assert np.allclose(Hf[np.ix_(basis, basis)].imag, 0)


def test_qsun_matches_kron_definition():
    """H = I (x) H_grain + sum_j h_j Sz_j + sum_j g0 alpha^u_j Sx_{n(j)} Sx_j (dense)."""
    Lloc, Lg, g0 = 3, 2, 0.8
    d = om.qsun_draws(Lloc, Lg, 0.75, rng=np.random.default_rng(5))
    H = om.qsun_hamiltonian_from_draws(Lloc, Lg, g0, *d, use_sparse=False)
    disorder, nbr, coup, G = d
    L = Lloc + Lg
    sx = np.array([[0, 0.5], [0.5, 0]])
    sz = np.array([[-0.5, 0], [0, 0.5]])

    def op(single, bit):
        m = np.array([[1.0]])
        for k in range(L - 1, -1, -1):
            m = np.kron(m, single if k == bit else np.eye(2))
        return m

    Href = np.kron(np.eye(2 ** Lloc), G)
    for j in range(Lloc):
        Href += disorder[j] * op(sz, Lg + j)
        Href += g0 * coup[j] * op(sx, int(nbr[j])) @ op(sx, Lg + j)
    assert np.allclose(H, Href, atol=1e-14)


# ----------------------------------------------------------------- packing
def test_packing_known_answer():
    k = KAT["packing_5x5"]
    H = sp.lil_matrix((5, 5))
    for i, v in enumerate(k["diag"]):
        H[i, i] = v
    for i, j, v in k["offdiag_entries"]:
        H[i - 1, j - 1] = v
        H[j - 1, i - 1] = v
    diag, buckets = opk.get_diags_and_offdiagonals_by_value(H.tocsc())
    assert list(diag) == k["diag"]
    assert len(buckets) == len(k["expected_buckets_1based"])
    for (v, flat, starts), e in zip(buckets, k["expected_buckets_1based"]):
        assert v == e["value"]
        assert list(flat + 1) == e["flat"]
        assert list(starts + 1) == e["starts"]
    dd, bd = opk.get_diags_and_offdiagonals_by_value(H.toarray())
    assert list(dd) == list(diag)
    for a, b in zip(bd, buckets):
        assert a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    x = np.arange(1.0, 6.0)
    X = np.stack([x, 2 * x], 1)
    Hd = H.toarray()
    assert np.allclose(opk.optimized_mapping(diag, buckets, X), Hd @ X)
    ci = k["clenshaw_inputs"]
    c, xx, b1, b2, b3 = (np.asarray(ci[n], float) for n in ("c", "x", "b1", "b2", "b3"))
    assert np.allclose(opk.clenshaw_recurrence(diag, buckets, b2, b3, c, xx), c * xx + 2 * Hd @ b2 - b3)
    assert np.allclose(opk.clenshaw_final_sum(diag, buckets, b1, b2, c, xx), c * xx + Hd @ b1 - b2)


def test_single_bucket():
    k = KAT["single_bucket_3x3"]
    H = np.diag(k["diag"])
    for i, j, v in k["offdiag_entries"]:
        H[i - 1, j - 1] = H[j - 1, i - 1] = v
    d, b = opk.get_diags_and_offdiagonals_by_value(H)
    assert len(b) == 1
    assert np.allclose(opk.optimized_mapping(d, b, np.array(k["x"])), H @ np.array(k["x"]))


def test_cpu_restatement_matches_numpy_oracle():
    """oracle/cpu_ref.c (the CPU baseline) == numpy Clenshaw on the packed H."""
    from oracle import cpu as ocpu
    H = om.j1j2_hamiltonian(12, 6, 1.0, 0.5, 1.0, 1.0, 0.5, rng=np.random.default_rng(2))
    w = np.linalg.eigvalsh(H.toarray())
    d, b = opk.get_diags_and_offdiagonals_by_value(H)
    dr, br = opk.rescaled_packing(d, b, w[0], w[-1])
    P = ocpu.PackedCPU(dr, br)
    c = oc.filter_coefficients(0.2, 150)
    X = np.random.default_rng(3).standard_normal((H.shape[0], 3))
    Yref = oc.clenshaw_apply(lambda Z: opk.optimized_mapping(dr, br, Z), c, X)
    assert np.max(np.abs(P.filter(c, X, nthreads=2) - Yref)) < 1e-12 * np.max(np.abs(Yref))
    y = P.filter_rows(c, X[:, 0], nthreads=2)
    assert np.max(np.abs(y - Yref[:, 0])) < 1e-12 * np.max(np.abs(Yref))


# ---------------------------------------------------------------- Clenshaw
@pytest.mark.parametrize("polytype", ["Chebyshev", "Legendre", "Hermite", "Taylor"])
def test_scalar_clenshaw_matches_numpy_series(polytype):
    c = np.random.default_rng(1).standard_normal(12)
    x = 0.37
    ref = {"Chebyshev": np.polynomial.chebyshev.chebval,
           "Legendre": np.polynomial.legendre.legval,
           "Hermite": np.polynomial.hermite.hermval,
           "Taylor": np.polynomial.polynomial.polyval}[polytype](x, c)
    assert abs(oc.clenshaw_scalar(lambda n: c[n], len(c) - 1, x, polytype) - ref) < 1e-12


def test_filter_coefficients_and_normalization():
    """c_0 = 1, c_n = 2 cos(n acos t) (PolfedDefaults.jl:104), p(t) = 1."""
    t, K = 0.3, 40
    assert oc.coefficients(t, 0) == 1.0
    assert abs(oc.coefficients(t, 5) - 2 * math.cos(5 * math.acos(t))) < 1e-15
    c = oc.filter_coefficients(t, K)
    assert abs(oc.chebyshev_filter_dense(np.array([t]), c)[0] - 1.0) < 1e-13
    # the unnormalized filter at its target is 1 + 2 sum cos^2(n theta) (closed form)
    th = math.acos(t)
    praw = 1 + 2 * sum(math.cos(n * th) ** 2 for n in range(1, K + 1))
    assert abs(1 / c[0] - praw) < 1e-10


def test_block_clenshaw_matches_dense_spectrum():
    H = om.xxz_hamiltonian(10, 5, 1.0, 1.0, 2.0, rng=np.random.default_rng(4)).toarray()
    w, V = np.linalg.eigh(H)
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * np.eye(H.shape[0])) / a
    c = oc.filter_coefficients(-0.1, 80)
    X = np.random.default_rng(0).standard_normal((H.shape[0], 3))
    Y = oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)
    p = oc.chebyshev_filter_dense((w - b) / a, c)
    assert np.max(np.abs(Y - V @ (p[:, None] * (V.T @ X)))) < 1e-11


# -------------------------------------------------------------------- DoS
def test_kernel_values():
    k = KAT["kernels"]
    funcs = {"Dirichlet": odos.dirichlet, "Fejer": odos.fejer, "Lorentz": odos.lorentz,
             "LanczosK": odos.lanczos_kernel, "WangZunger": odos.wang_zunger, "Jackson": odos.jackson}
    for name, cases in k.items():
        if name == "cite":
            continue
        for n, N, v in cases:
            assert abs(funcs[name](n, N) - v) < 1e-12, (name, n, N)
    assert odos.wang_zunger(5, 10) < odos.wang_zunger(0, 10)
    with pytest.raises(ValueError):
        odos.get_kernel("UnknownKernel")


def test_trace_and_moments_identity():
    k = KAT["trace_identity"]
    mus = np.zeros(k["nmoments"])
    odos.trace(lambda v: v.copy(), np.array(k["alpha"]), mus)
    assert np.allclose(mus, k["expected"])
    k = KAT["dos_moments_identity"]
    probes = np.random.default_rng(0).standard_normal((k["D"], k["R"]))
    assert np.allclose(odos.dos_moments(lambda v: v.copy(), k["N"], probes), k["expected"])
    k = KAT["dos_rho0_identity_dirichlet"]
    mu = odos.dos_moments(lambda v: v.copy(), k["N"], np.random.default_rng(1).standard_normal((k["D"], k["R"])))
    rho = odos.dos_function(mu, k["kernel"])
    assert abs(rho(0.0) - k["expected_rho0"]) < 1e-12


def test_moments_converge_to_exact_traces():
    """stochastic moments (R probes) -> sum_i T_n(lambda_i) within a few 1/sqrt(R) sigma."""
    H = om.xxz_hamiltonian(10, 5, 1.0, 1.0, 1.0, rng=np.random.default_rng(7)).toarray()
    w = np.linalg.eigvalsh(H)
    a, b = (w[-1] - w[0]) / 2, (w[-1] + w[0]) / 2
    Ht = (H - b * np.eye(H.shape[0])) / a
    mu = odos.dos_moments(lambda v: Ht @ v, 20, np.random.default_rng(3).standard_normal((H.shape[0], 400)))
    ex = odos.exact_moments((w - b) / a, 20)
    assert abs(mu[0] - ex[0]) < 1e-9
    assert np.max(np.abs(mu - ex)) < 0.1 * H.shape[0] ** 0.5 * 4


# ---------------------------------------------------------- order/bounds
def test_bisection_and_analytical_order():
    k = KAT["bisection"]
    assert abs(osp.bisection(lambda x: x - k["root"], 0.0, 1.0, tol=k["tol"]) - k["root"]) < k["atol"]
    with pytest.raises(RuntimeError):
        osp.bisection(lambda x: x * x + 1, -1.0, 1.0)
    k = KAT["analytical_order"]
    assert abs(osp.analytical_orderofexpansion(k["lam"], k["delta"]) - k["expected"]) < 1e-12
    assert osp.analytical_orderofexpansion(0.0, 0.25) > osp.analytical_orderofexpansion(0.0, 0.5)
    assert osp.analytical_orderofexpansion(0.8, 0.5) < osp.analytical_orderofexpansion(0.0, 0.5)


def test_bounds_and_order_selection():
    k = KAT["bounds_from_flat_dos"]
    plan = osp.build_transform_plan(1.0, 0.0, k["howmany"], "middle")
    osp.getbounds_from_dos(plan, lambda x: k["rho"], "mean")
    assert abs(plan.left - k["expected"][0]) < k["atol"]
    assert abs(plan.right - k["expected"][1]) < k["atol"]
    k = KAT["bounds_from_order16_cutoff02"]
    plan = osp.build_transform_plan(1.0, 0.0, 1, "middle", order=k["order"], cutoff=k["cutoff"])
    osp.getbounds_from_K(plan, "mean")
    assert -1 <= plan.left < plan.target < plan.right <= 1
    assert abs(plan.left + plan.right) < 1e-7
    plan = osp.build_transform_plan(1.0, 0.0, 1, "middle", order=KAT["explicit_order_kept"]["order"])
    osp.getspectraltransform(plan, 1.0, 0.0, lambda x: 1.0)
    assert plan.order == KAT["explicit_order_kept"]["order"]
    k = KAT["order_from_bounds"]
    plan = osp.build_transform_plan(1.0, 0.0, 1, "middle", left=k["left"], right=k["right"],
                                    cutoff=k["cutoff"], order_safety_factor=k["order_safety_factor"])
    osp.getorderofexpansion_in_interval(plan)
    assert isinstance(plan.order, int) and plan.order >= k["min_order"]
    # bisection returns the middle of the final bracket [K, K+1], so floor(K) is
    # the last order whose edge value still reaches the cutoff and K+1 the first
    # below it (GetOrderOfExpansion.jl:58-63, Bisection.jl:39)
    p = lambda K, x: oc.chebyshev_filter_dense(np.array([x]), oc.filter_coefficients(0.0, K))[0]
    assert max(p(plan.order, k["left"]), p(plan.order, k["right"])) >= k["cutoff"]
    assert max(p(plan.order + 1, k["left"]), p(plan.order + 1, k["right"])) < k["cutoff"]
    plan = osp.build_transform_plan(1.0, 0.0, 1, "middle", left=-0.2, right=0.2, order=12)
    with pytest.raises(RuntimeError):
        osp.getbounds(plan, 1.0, 0.0, lambda x: 1.0)


def test_target_specs():
    k = KAT["target_specs"]
    for inp, kind, val in k["valid"]:
        spec = osp.normalize_target_spec(tuple(inp) if isinstance(inp, list) else inp)
        assert spec.kind == kind and spec.value == val
    for inp in k["invalid"]:
        with pytest.raises(ValueError):
            osp.normalize_target_spec(tuple(inp) if isinstance(inp, list) else inp)


def test_expectedkrylovdim():
    for howmany, s, eta, exp in KAT["expectedkrylovdim"]["cases"]:
        assert osp.expectedkrylovdim(howmany, s, eta) == exp


def test_initial_state():
    k = KAT["initial_state"]
    assert opf.initial_state_tolerance() == k["tolerance_f64"]
    v = np.array(k["vector"])
    vp = opf.prepare_initial_state(v)
    assert np.array_equal(v, k["vector"]) and abs(np.linalg.norm(vp) - 1) < 1e-15
    X = np.array(k["matrix"])
    Xp = opf.prepare_initial_state(X)
    assert np.allclose(Xp.T @ Xp, np.eye(2))
    with pytest.raises(ValueError):
        opf.prepare_initial_state(np.zeros(3))
    with pytest.raises(ValueError):
        opf.prepare_initial_state(np.ones((2, 3)))


# ------------------------------------------------------------------ Lanczos
@pytest.mark.parametrize("s", [1, 3])
def test_lanczos_extremal_pairs_match_dense(s):
    H = om.xxz_hamiltonian(10, 5, 1.0, 1.0, 1.0, rng=np.random.default_rng(9)).toarray()
    w = np.linalg.eigvalsh(H)
    x0 = np.random.default_rng(1).standard_normal((H.shape[0], s))
    x0, _ = np.linalg.qr(x0)
    if s == 1:
        x0 = x0[:, 0]
    r = olz.lanczos(lambda X: H @ X, x0, 4, maxdim=180, which="SR", tol=1e-12, eigentol=1e-8)
    assert len(r.vals) == 4
    assert np.max(np.abs(r.vals - w[:4])) < 1e-10
    assert np.max(np.linalg.norm(H @ r.vecs - r.vecs * r.vals, axis=0)) < 1e-8


def test_small_polfed_problem_reference_behaviour():
    """test/test_polfed.jl:29-61: D=6, s=2, K=12.  The Krylov space is
    exhausted before the first convergence check after k=4, so the reference
    algorithm returns no pairs (its test only checks the returned ones)."""
    mat = np.diag([-3.0, -1.25, -0.2, 0.55, 1.8, 3.2])
    q, _ = np.linalg.qr(np.random.default_rng(1234).standard_normal((6, 2)))
    r = opf.polfed(lambda X: mat @ X, q, 3, 0.0, Emin=-3.0, Emax=3.2, order=12, tol=1e-11,
                   eigentol=1e-7, overestimate_iters=1.0, N=8, R=2, rng=np.random.default_rng(0))
    assert r.plan.order == 12
    for j, v in enumerate(r.vals):
        assert np.linalg.norm(mat @ r.vecs[:, j] - v * r.vecs[:, j]) <= 1e-6 * max(1, abs(v))
    assert np.all(np.diff(r.vals) >= 0)


def test_polfed_oracle_config1_shape():
    """L=12 stand-in of config 1: the returned set is the top-nev by p(lambda~)
    of the exact spectrum, with residuals <= 1e-10."""
    L = 12
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(0))
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    x0 = np.random.default_rng(100).standard_normal(H.shape[0])
    x0 /= np.linalg.norm(x0)
    r = opf.polfed(lambda X: H @ X, x0, 20, ("rescaled", 0.0), eigentol=1e-10, N=64, R=30,
                   rng=np.random.default_rng(2))
    w = np.linalg.eigvalsh(H.toarray())
    a, b = (r.Emax - r.Emin) / 2, (r.Emax + r.Emin) / 2
    p = oc.chebyshev_filter_dense((w - b) / a, r.coeffs)
    exp = np.sort(w[np.argsort(-p)[:len(r.vals)]])
    assert len(r.vals) == 20
    assert np.max(np.abs(r.vals - exp)) < 1e-10 * (w[-1] - w[0])
