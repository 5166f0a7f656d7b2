"""CPU: the C ABI library loads and exports every declared symbol; the host
mirror's scalar logic (target specs, KPM kernels and DoS, bounds, order,
coefficients, configs) agrees with the oracle.  No GPU calls."""
import ctypes
import json
import pathlib
import re

import numpy as np
import pytest

import polfed_amd
from oracle import clenshaw as oc
from oracle import dos as odos
from oracle import spectral as osp
from polfed_amd import _lib
from polfed_amd import spectral as spc
from polfed_amd.configs import PolfedDefaults, TransformConfig

REPO = pathlib.Path(__file__).resolve().parent.parent
KAT = json.loads((REPO / "tests" / "golden" / "reference_kat.json").read_text())


def _declared_symbols():
    names = set()
    for h in (REPO / "include").glob("*.h"):
        text = h.read_text()
        names |= set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pf_\w+)\s*\(", text, re.M))
    return names


def test_library_exports_every_declared_symbol():
    declared = _declared_symbols()
    assert len(declared) >= 20
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    for name in declared:
        assert hasattr(lib, name), name
    # every declared function has a ctypes signature in the binding
    assert declared <= set(_lib.SIGNATURES)
    lib2 = _lib.load()
    assert lib2.pf_version() >= 100
    assert lib2.pf_last_error() == b""


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(polfed_amd.PolfedHIPError):
        _lib.load(tmp_path / "nope.so")


def test_target_specs_host_mirror():
    k = KAT["target_specs"]
    for inp, kind, val in k["valid"]:
        spec = spc.normalize_target_spec(tuple(inp) if isinstance(inp, list) else inp)
        assert spec.kind == kind and spec.value == val
    for inp in k["invalid"]:
        with pytest.raises(ValueError):
            spc.normalize_target_spec(tuple(inp) if isinstance(inp, list) else inp)


def test_kernels_and_dos_host_mirror():
    for name, fn in odos.KERNELS.items():
        for n, N in [(0, 10), (3, 10), (7, 11), (249, 250)]:
            assert abs(float(spc.get_kernel(name)(np.array(n), N)) - fn(n, N)) < 1e-12, name
    mu = np.random.default_rng(0).standard_normal(64) * 3 + 50
    rho_o = odos.dos_function(mu, "Jackson")
    rho_h = spc.DensityOfStates(mu, "Jackson")
    xs = np.linspace(-0.95, 0.95, 31)
    assert np.max(np.abs(rho_h(xs) - np.array([rho_o(x) for x in xs]))) < 1e-9 * np.max(np.abs(rho_h(xs)))
    k = KAT["dos_rho0_identity_dirichlet"]
    assert abs(spc.DensityOfStates(np.full(k["N"], 4.0), "Dirichlet")(0.0) - k["expected_rho0"]) < 1e-12


@pytest.mark.parametrize("polytype", ["Chebyshev", "Legendre", "Hermite", "Taylor"])
def test_series_value_matches_oracle_clenshaw(polytype):
    c = np.random.default_rng(3).standard_normal(25)
    for x in (-0.8, 0.1, 0.65):
        ref = oc.clenshaw_scalar(lambda n: c[n], len(c) - 1, x, polytype)
        assert abs(float(spc.series_value(c, x, polytype)) - ref) < 1e-11 * max(1, abs(ref))


def test_normalized_coefficients_match_oracle():
    for t, K in [(0.0, 12), (0.37, 400), (-0.9, 3000)]:
        a = spc.normalized_coefficients(PolfedDefaults.coefficients, t, K)
        b = oc.filter_coefficients(t, K)
        assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(b))


def test_bounds_and_order_host_mirror_match_oracle():
    # flat DoS known answer
    plan = spc.build_transform_plan(TransformConfig(), 1.0, 0.0, 1, "middle")
    spc.getbounds_from_dos(plan, lambda x: 1.0)
    assert abs(plan.left + 0.5) < 1e-7 and abs(plan.right - 0.5) < 1e-7
    # a realistic DoS: compare the full interval/order resolution with the oracle
    mu = odos.exact_moments(np.cos(np.linspace(0.2, 2.9, 3000)), 250)
    rho_h = spc.DensityOfStates(mu, "Jackson")
    rho_o = odos.dos_function(mu, "Jackson")
    for target, howmany in [(("rescaled", 0.0), 50), (("rescaled", 0.4), 100), ("maxdos", 20)]:
        ph = spc.build_transform_plan(TransformConfig(), 1.0, 0.0, howmany, target)
        spc.getbounds(ph, 1.0, 0.0, rho_h)
        spc.getorderofexpansion(ph)
        po = osp.build_transform_plan(1.0, 0.0, howmany, target)
        osp.getspectraltransform(po, 1.0, 0.0, rho_o)
        assert abs(ph.target - po.target) < 1e-12
        assert abs(ph.order - po.order) <= 1, (target, ph.order, po.order)
        if ph.order == po.order:
            assert abs(ph.left - po.left) < 1e-7 and abs(ph.right - po.right) < 1e-7
    # over-constrained transform raises
    bad = spc.build_transform_plan(TransformConfig(left=-0.2, right=0.2, order=12), 1.0, 0.0, 1,
                                   "middle")
    with pytest.raises(RuntimeError):
        spc.getbounds(bad, 1.0, 0.0, rho_h)


def test_closed_form_filter_and_state_count_match_the_series():
    """The O(1) evaluations behind the order and bound bisections: the default
    Chebyshev filter through the Dirichlet kernel against its K-term series
    (numpy chebval, an independent summation), the KPM state count against
    adaptive quadrature, and whole interval/order selections against the
    series-and-quadrature path (a coefficient function that is not the
    default object, and a density given as a plain callable)."""
    from numpy.polynomial import chebyshev as npc
    for t, K in [(0.0, 12), (0.37, 401), (-0.8, 5000), (0.999, 33)]:
        plan = spc.build_transform_plan(TransformConfig(), 1.0, 0.0, 1, ("rescaled", t))
        p = spc._filter_fn(plan, t, K)
        c = spc.raw_coefficients(PolfedDefaults.coefficients, t, K)
        c = c / npc.chebval(t, c)
        for x in (-1.0, -0.5, t - 1e-3, t, 0.25, 1.0):
            x = min(max(x, -1.0), 1.0)
            assert abs(p(x) - npc.chebval(x, c)) <= 1e-12 * max(1.0, abs(npc.chebval(x, c)))
    mu = odos.exact_moments(np.cos(np.linspace(0.1, 3.0, 2000)), 250) * 300.0
    rho = spc.DensityOfStates(mu, "Jackson")
    from scipy.integrate import quad
    for a, b in [(-0.9, -0.2), (-0.01, 0.01), (0.3, 0.99), (-1.0, 1.0)]:
        q, _ = quad(lambda x: float(rho(x)), a, b, epsabs=1e-12, epsrel=1e-12, limit=500)
        assert abs(rho.integral(a, b) - q) <= 1e-8 * max(1.0, abs(q))
    generic = lambda t, n: PolfedDefaults.coefficients(t, n)
    for target, howmany in [(("rescaled", 0.0), 50), (("rescaled", -0.45), 120), ("maxdos", 20)]:
        pf = spc.build_transform_plan(TransformConfig(), 1.0, 0.0, howmany, target)
        spc.getbounds(pf, 1.0, 0.0, rho)
        spc.getorderofexpansion(pf)
        pg = spc.build_transform_plan(TransformConfig(coefficients=generic), 1.0, 0.0, howmany,
                                      target)
        spc.getbounds(pg, 1.0, 0.0, lambda x: rho(x))
        spc.getorderofexpansion(pg)
        assert pf.order == pg.order, (target, pf.order, pg.order)
        assert abs(pf.left - pg.left) < 1e-9 and abs(pf.right - pg.right) < 1e-9


def test_defaults_and_configs():
    for howmany, s, eta, exp in KAT["expectedkrylovdim"]["cases"]:
        assert PolfedDefaults.expectedkrylovdim(howmany, s, eta) == exp
    assert PolfedDefaults.coefficients(0.3, 0) == 1.0
    assert abs(PolfedDefaults.coefficients(0.3, 4) - 2 * np.cos(4 * np.arccos(0.3))) < 1e-15
    fc = polfed_amd.FactorizationConfig()
    assert fc.tol == 1e-14 and fc.eigentol == 1e-9 and fc.which == "LR"
    assert isinstance(fc.rot, polfed_amd.FullRO) and fc.rot.passes == 2
    d = polfed_amd.DoSConfig()
    assert (d.N, d.R, d.kernel) == (250, 300, "Jackson")


# ------------------------------------------------------------ packed operator
def _per_value_1based(vals, starts, flat):
    """The C ABI layout -> the reference's per-value 1-based (value, flat, starts)."""
    D = starts.shape[1] - 1
    out = []
    for v in range(len(vals)):
        f = flat[starts[v, 0]:starts[v, D]] + 1
        out.append((vals[v], list(f), list(starts[v, :D] - starts[v, 0] + 1)))
    return out


def test_pack_by_value_reference_known_answer():
    """The product's packing (host side of pf_op_create_packed) reproduces the
    flat/starts arrays of test/optimization.jl:41-55, sparse and dense input."""
    import scipy.sparse as sps
    from polfed_amd.packing import pack_by_value
    k = KAT["packing_5x5"]
    H = sps.lil_matrix((5, 5))
    for i, v in enumerate(k["diag"]):
        H[i, i] = v
    for i, j, v in k["offdiag_entries"]:
        H[i - 1, j - 1] = H[j - 1, i - 1] = v
    for M in (H.tocsc(), H.toarray()):
        diag, vals, starts, flat = pack_by_value(M)
        assert list(diag) == k["diag"]
        got = _per_value_1based(vals, starts, flat)
        assert len(got) == len(k["expected_buckets_1based"])
        for (v, f, s), e in zip(got, k["expected_buckets_1based"]):
            assert v == e["value"] and f == e["flat"] and s == e["starts"]


def test_pack_by_value_matches_oracle_on_random_matrices():
    """Same buckets as the oracle's restatement on matrices with rounding-level
    value duplicates, tiny entries below tol and empty rows."""
    import scipy.sparse as sps
    from oracle import packing as opk
    from polfed_amd.packing import pack_by_value
    rng = np.random.default_rng(5)
    D = 60
    A = sps.random(D, D, density=0.1, random_state=7, format="coo")
    vals = rng.choice([-1.0, 0.5, 0.5 + 1e-17, 2.25, 1e-14], size=A.nnz)
    A = sps.coo_matrix((vals, (A.row, A.col)), shape=(D, D)).tocsr()
    A = (A + A.T).tolil()
    A[3, :] = 0
    A[:, 3] = 0
    A.setdiag(rng.standard_normal(D))
    A = A.tocsc()
    diag, vv, starts, flat = pack_by_value(A)
    d_ref, buckets = opk.get_diags_and_offdiagonals_by_value(A)
    assert np.array_equal(diag, d_ref)
    assert len(vv) == len(buckets)
    for (v, f, s), (rv, rf, rs) in zip(_per_value_1based(vv, starts, flat), buckets):
        assert v == rv and f == list(rf + 1) and s == list(rs + 1)


# ------------------------------------------------ Quantum Sun, spin S and U(1)
def _kron_qsun(L_loc, L_grain, g0, S, use_U1, S_z, gamma, draws):
    """Independent dense construction from textbook spin-S matrices and
    Kronecker products (site 0 = the last, fastest factor)."""
    disorder, neighbors, couplings, goe = draws
    d = int(round(2 * S)) + 1
    L = L_loc + L_grain
    ms = np.arange(d) - S
    Sp = np.zeros((d, d))
    for k in range(d - 1):
        Sp[k + 1, k] = np.sqrt(S * (S + 1) - ms[k] * (ms[k] + 1))
    ops = {"z": np.diag(ms), "x": (Sp + Sp.T) / 2, "y": (Sp - Sp.T) / 2j}

    def at(name, site):
        out = np.ones((1, 1))
        for s in reversed(range(L)):
            out = np.kron(out, ops[name] if s == site else np.eye(d))
        return out

    N = d ** L
    H = np.zeros((N, N), dtype=complex)
    for j in range(L_loc):
        H += disorder[j] * at("z", L_grain + j)
        H += g0 * couplings[j] * at("x", L_grain + j) @ at("x", int(neighbors[j]))
        if use_U1:
            H += g0 * couplings[j] * at("y", L_grain + j) @ at("y", int(neighbors[j]))
    n = d ** L_grain
    G = np.array(goe, dtype=float)
    if use_U1:
        gm = np.array([sum((a // d ** s) % d - S for s in range(L_grain)) for a in range(n)])
        G[gm[:, None] != gm[None, :]] = 0.0
        mag = np.array([sum((i // d ** s) % d - S for s in range(L)) for i in range(N)])
        sel = np.flatnonzero(mag == S_z)
        R = np.kron(np.eye(d ** L_loc), G)[np.ix_(sel, sel)]
        R = gamma * R / np.sqrt(np.trace(R @ R) / len(sel))
        H = H[np.ix_(sel, sel)] + R
    else:
        H = H + np.kron(np.eye(d ** L_loc), gamma * G / np.sqrt(n + 1))
    assert np.abs(H.imag).max() < 1e-14
    return H.real


@pytest.mark.parametrize("L_loc,L_grain,S,use_U1,S_z", [
    (3, 2, 0.5, True, 0.5), (2, 2, 1.0, True, 0.0), (2, 1, 1.0, False, 0.0),
    (2, 1, 1.5, True, 0.5), (3, 2, 0.5, False, 0.0), (1, 2, 1.0, True, -1.0)])
def test_qsun_spin_matrix_equals_kronecker_construction(L_loc, L_grain, S, use_U1, S_z):
    """qsun_spin_matrix (bit-encoded basis, ladder amplitudes, sector ranks)
    against the dense tensor-product Hamiltonian restricted to the sector."""
    from polfed_amd import models
    d = int(round(2 * S)) + 1
    draws = models.qsun_spin_draws(L_loc, L_grain, 0.6, d, rng=np.random.default_rng(7))
    H = models.qsun_spin_matrix(L_loc, L_grain, 0.9, 0.6, S=S, use_U1=use_U1, S_z=S_z,
                                gamma=1.3, draws=draws)
    ref = _kron_qsun(L_loc, L_grain, 0.9, S, use_U1, S_z, 1.3, draws)
    assert H.shape == ref.shape
    assert np.abs(H.toarray() - ref).max() < 1e-13


def test_qsun_spin_matrix_reference_shapes_and_spin_half_oracle():
    """The sector sizes of test/test_qsun.jl:20-70, and the spin-1/2 full space
    equal to the oracle's Quantum Sun for the same draws."""
    from oracle import models as om
    from polfed_amd import models
    for kw, n in [(dict(L_loc=2, L_grain=2, use_U1=True, S_z=0.0), 6),
                  (dict(L_loc=2, L_grain=1, use_U1=True, S_z=0.5), 3),
                  (dict(L_loc=1, L_grain=1, S=1.5, use_U1=True, S_z=1.0), 3),
                  (dict(L_loc=2, L_grain=2), 16)]:
        H = models.qsun_spin_matrix(g0=1.0, alpha=0.5, rng=np.random.default_rng(3), **kw)
        assert H.shape == (n, n) and abs(H - H.T).max() == 0.0
    with pytest.raises(ValueError):
        models.qsun_spin_matrix(2, 2, 1.0, 0.5, use_U1=True, S_z=0.5)  # not a sector
    draws = models.qsun_spin_draws(5, 3, 0.75, 2, rng=np.random.default_rng(4))
    H = models.qsun_spin_matrix(5, 3, 1.0, 0.75, draws=draws)
    dis, nb, cp, goe = draws
    ref = om.qsun_hamiltonian_from_draws(5, 3, 1.0, dis, nb, cp, goe / np.sqrt(9))
    assert abs(H - ref).max() < 1e-14
