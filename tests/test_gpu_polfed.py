"""End-to-end POLFED on the GPU (FP64 drop-in path) against the CPU oracle and
exact spectra: same eigenvalue set, <= 1e-10 x spectral width, residuals <= 1e-10,
Ritz subspace angles <= 1e-8 per degenerate cluster."""
import numpy as np
import pytest

from oracle import clenshaw as oc
from oracle import models as om
from oracle import polfed as opf

pytestmark = pytest.mark.gpu


def _expected_set(ev, Emin, Emax, coeffs, nev):
    a, b = (Emax - Emin) / 2, (Emax + Emin) / 2
    p = oc.chebyshev_filter_dense((ev - b) / a, coeffs)
    return np.sort(ev[np.argsort(-p, kind="stable")[:nev]])


def _max_angle(V1, V2):
    """largest principal angle between the column spans, from its sine
    ||(I - Q2 Q2^T) Q1||_2 (arccos of the cosines loses everything below ~1e-8)."""
    Q1, _ = np.linalg.qr(V1)
    Q2, _ = np.linalg.qr(V2)
    return float(np.arcsin(min(1.0, np.linalg.norm(Q1 - Q2 @ (Q2.T @ Q1), 2))))


def test_polfed_block_fixed_order_matches_exact(dev):
    from polfed_amd import DoSConfig, FactorizationConfig, MappingConfig, TransformConfig, polfed
    from polfed_amd.models import xxz_hamiltonian
    L, nev, s = 12, 20, 4
    rng = np.random.default_rng(3)
    h = om.resolve_fields(L, 0.0, 2.0, rng=rng)
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h)
    ev, V = np.linalg.eigh(H.toarray())
    q, _ = np.linalg.qr(np.random.default_rng(4).standard_normal((op.D, s)))
    vals, vecs, rep = polfed(op, q, nev, ("rescaled", 0.1), produce_report=True,
                             mapping=MappingConfig(Emin=ev[0], Emax=ev[-1]),
                             transform=TransformConfig(order=200),
                             fact=FactorizationConfig(eigentol=1e-10), dos=DoSConfig(N=64, R=20))
    assert rep.order == 200
    assert len(vals) == nev
    width = ev[-1] - ev[0]
    exp = _expected_set(ev, ev[0], ev[-1], oc.filter_coefficients(0.1, 200), nev)
    assert np.max(np.abs(np.sort(vals) - exp)) <= 1e-10 * width
    psi = dev.to_host(vecs)
    res = np.linalg.norm(H @ psi - psi * vals, axis=0)
    assert res.max() <= 1e-10
    # subspace angle against the exact eigenvectors of the same eigenvalues
    idx = [int(np.argmin(np.abs(ev - v))) for v in vals]
    assert _max_angle(psi, V[:, idx]) <= 1e-8
    # the oracle (reference restatement) on the same problem returns the same set
    ref = opf.polfed(lambda X: H @ X, q, nev, ("rescaled", 0.1), Emin=ev[0], Emax=ev[-1], order=200,
                     eigentol=1e-10, N=64, R=20, rng=np.random.default_rng(5))
    assert len(ref.vals) == nev
    assert np.max(np.abs(np.sort(vals) - ref.vals)) <= 1e-10 * width


def test_polfed_vector_defaults_matches_oracle(dev):
    """Config-1 shape (L=14, D=3432, single-vector Lanczos, nev=50, middle
    target, default interval/order selection) with the same KPM probes."""
    from polfed_amd import FactorizationConfig, polfed
    from polfed_amd.blockops import fill_probes
    from polfed_amd.models import xxz_hamiltonian
    L = 14
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(0))
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    x0 = np.random.default_rng(100).standard_normal(op.D)
    x0 /= np.linalg.norm(x0)
    P = fill_probes(dev, op.D, 300, seed=77)
    vals, vecs, rep = polfed(op, x0, 50, ("rescaled", 0.0), produce_report=True,
                             fact=FactorizationConfig(eigentol=1e-10), probes=P)
    ref = opf.polfed(lambda X: H @ X, x0, 50, ("rescaled", 0.0), eigentol=1e-10,
                     probes=dev.to_host(P))
    ev = np.linalg.eigvalsh(H.toarray())
    width = ev[-1] - ev[0]
    assert abs(rep.Emin - ev[0]) < 1e-9 * width and abs(rep.Emax - ev[-1]) < 1e-9 * width
    assert abs(rep.order - ref.plan.order) <= 1
    assert len(vals) == 50 and len(ref.vals) == 50
    exp = _expected_set(ev, rep.Emin, rep.Emax, oc.filter_coefficients(rep.target, rep.order), 50)
    assert np.max(np.abs(np.sort(vals) - exp)) <= 1e-10 * width
    if rep.order == ref.plan.order:
        assert np.max(np.abs(vals - ref.vals)) <= 1e-10 * width
    psi = dev.to_host(vecs)
    assert np.linalg.norm(H @ psi - psi * vals, axis=0).max() <= 1e-10


def test_polfed_qsun_block(dev):
    """Quantum Sun, block size 2: the reference's own algorithm exhausts its
    iterations on this problem and returns 13 of 16 pairs; the GPU path must
    return the same pairs."""
    from polfed_amd import FactorizationConfig, polfed
    from polfed_amd.blockops import fill_probes
    from polfed_amd.models import qsun_hamiltonian
    draws = om.qsun_draws(7, 3, 0.75, rng=np.random.default_rng(8))
    op = qsun_hamiltonian(7, 3, 1.0, 0.75, draws=draws, device=dev)
    H = om.qsun_hamiltonian_from_draws(7, 3, 1.0, *draws)
    ev = np.linalg.eigvalsh(H.toarray())
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((op.D, 2)))
    P = fill_probes(dev, op.D, 300, seed=91)
    vals, vecs, rep = polfed(op, q, 16, "middle", produce_report=True,
                             fact=FactorizationConfig(eigentol=1e-10), probes=P)
    ref = opf.polfed(lambda X: H @ X, q, 16, "middle", eigentol=1e-10, probes=dev.to_host(P))
    width = ev[-1] - ev[0]
    assert abs(rep.order - ref.plan.order) <= 1
    assert len(vals) == len(ref.vals) >= 10
    if rep.order == ref.plan.order:
        assert np.max(np.abs(vals - ref.vals)) <= 1e-10 * width
    exp = _expected_set(ev, rep.Emin, rep.Emax, oc.filter_coefficients(rep.target, rep.order),
                        len(vals))
    assert np.max(np.abs(np.sort(vals) - exp)) <= 1e-10 * width
    psi = dev.to_host(vecs)
    assert np.linalg.norm(H @ psi - psi * vals, axis=0).max() <= 1e-10


def test_polfed_mixed_precision_meets_residual_tolerance(dev):
    """FP64 filter until the wanted Ritz pairs' residual estimate is small,
    then FP32: same eigenvalues as the FP64 run (<= 1e-10 x width), every
    residual <= 1e-10, and a nonzero number of FP32 iterations.  The pure
    FP32 filter reproduces the eigenvalues but not the residual level."""
    from polfed_amd import FactorizationConfig, MappingConfig, TransformConfig, polfed
    from polfed_amd.blockops import fill_probes
    from polfed_amd.models import xxz_hamiltonian
    L, nev, s = 14, 24, 4
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(2))
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    ev = np.linalg.eigvalsh(H.toarray())
    width = ev[-1] - ev[0]
    q, _ = np.linalg.qr(np.random.default_rng(7).standard_normal((op.D, s)))
    P = fill_probes(dev, op.D, 100, seed=5)
    common = dict(produce_report=True, mapping=MappingConfig(Emin=ev[0], Emax=ev[-1]),
                  fact=FactorizationConfig(eigentol=1e-10), probes=P)
    v64, _, r64 = polfed(op, q, nev, "middle", transform=TransformConfig(dtype="f64"), **common)
    vmx, psi, rmx = polfed(op, q, nev, "middle", transform=TransformConfig(dtype="mixed"), **common)
    assert len(v64) >= nev - 4 and len(vmx) == len(v64)
    assert rmx.lanczos.fast_iterations > 0
    assert np.max(np.abs(vmx - v64)) <= 1e-10 * width
    psi = dev.to_host(psi)
    assert np.linalg.norm(H @ psi - psi * vmx, axis=0).max() <= 1e-10
    v32, _, r32 = polfed(op, q, nev, "middle", transform=TransformConfig(dtype="f32"), **common)
    lz = r32.lanczos
    assert lz.fast_iterations == lz.numiter - 1
    allv = np.sort(lz.all_vals)
    assert len(allv) >= len(v64)
    # every FP64 eigenvalue is reproduced by an FP32-filter Ritz value; the FP32
    # operator coefficients perturb H~ by ~eps32, so agreement is ~1e-8 x width
    assert max(np.min(np.abs(allv - v)) for v in v64) <= 1e-6 * width
    assert lz.all_residuals.max() < 1e-3


def test_fp32_filter_polish(dev):
    """The FP64 polish of runs whose Krylov basis holds FP32-filter products:
    a window filter in FP64 and Rayleigh-Ritz with the unscaled H.  With an
    early switch to FP32 (and with FP32 throughout) the polished pairs have
    residuals at least 10x below the plain Ritz pairs' (recomputed here with
    the oracle's H), orthonormal vectors, Rayleigh quotients equal to the
    returned values, and every pair meeting 1e-10 matches an eigenvalue of the
    dense spectrum to 1e-10 x width."""
    from polfed_amd import FactorizationConfig, MappingConfig, TransformConfig, polfed
    from polfed_amd.blockops import fill_probes
    from polfed_amd.models import xxz_hamiltonian
    L, nev, s = 14, 24, 4
    h = om.resolve_fields(L, 0.0, 5.0, rng=np.random.default_rng(2))
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 5.0, fields=h)
    ev = np.linalg.eigvalsh(H.toarray())
    width = ev[-1] - ev[0]
    q, _ = np.linalg.qr(np.random.default_rng(7).standard_normal((op.D, s)))
    P = fill_probes(dev, op.D, 100, seed=5)
    common = dict(produce_report=True, mapping=MappingConfig(Emin=ev[0], Emax=ev[-1]),
                  fact=FactorizationConfig(eigentol=1e-10), probes=P)
    for mode, sw, gain in (("mixed", 1e-2, 10.0), ("f32", 1e-6, 3.0)):
        res = {}
        for polish in (False, True):
            tc = TransformConfig(dtype=mode, switch_tol=sw, polish=polish)
            _, _, rep = polfed(op, q, nev, "middle", transform=tc, **common)
            lz = rep.lanczos
            assert lz.fast_iterations > 0
            res[polish] = (lz.all_vals, lz.all_residuals)
        vals, r = res[True]
        assert np.max(r) * gain <= np.max(res[False][1]), mode
        # the returned pairs, checked on the host with the oracle's H
        vals_p, psi, rep = polfed(op, q, nev, "middle",
                                  transform=TransformConfig(dtype=mode, switch_tol=sw), **common)
        if len(vals_p):
            psi = dev.to_host(psi)
            Hpsi = H @ psi
            assert np.linalg.norm(Hpsi - psi * vals_p, axis=0).max() <= 1e-10
            assert np.max(np.abs(np.einsum("ij,ij->j", psi, Hpsi) - vals_p)) <= 1e-12 * width
            assert np.linalg.norm(psi.T @ psi - np.eye(len(vals_p))) <= 1e-12
            assert max(np.min(np.abs(ev - v)) for v in vals_p) <= 1e-10 * width


def test_lanczos_extrema_and_report(dev):
    """lanczos_extrema (src/Lanczos/Lanczos.jl:146-212; test/test_polfed.jl:94-107)
    on a matrix, a GPU operator and a callable: the extremal eigenvalues, the
    caller's start untouched, and the argument errors; display_report on a
    run's report."""
    from polfed_amd import DoSConfig, MappingConfig, TransformConfig, display_report, lanczos_extrema, polfed
    from polfed_amd.models import xxz_hamiltonian
    mat = np.diag([-3.0, -1.25, -0.2, 0.55, 1.8, 3.2])
    x0 = np.ones(6)
    Emin, Emax = lanczos_extrema(mat, x0)
    assert abs(Emin + 3.0) < 1e-10 and abs(Emax - 3.2) < 1e-10
    assert np.array_equal(x0, np.ones(6))
    L = 12
    h = om.resolve_fields(L, 0.0, 2.0, rng=np.random.default_rng(3))
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h, device=dev)
    ev = np.linalg.eigvalsh(om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h).toarray())
    width = ev[-1] - ev[0]
    for e in (lanczos_extrema(op, seed=1),
              lanczos_extrema(lambda X, Y: op.apply(X, Y), np.random.default_rng(2).standard_normal(op.D))):
        assert abs(e[0] - ev[0]) < 1e-9 * width and abs(e[1] - ev[-1]) < 1e-9 * width
    for bad in [(mat, np.zeros(6)), (np.ones((2, 3)), np.ones(2)), (mat, np.ones(5))]:
        with pytest.raises(ValueError):
            lanczos_extrema(*bad)
    q = np.linalg.qr(np.random.default_rng(4).standard_normal((op.D, 2)))[0]
    _, _, rep = polfed(op, q, 4, "middle", produce_report=True,
                       mapping=MappingConfig(Emin=ev[0], Emax=ev[-1]),
                       transform=TransformConfig(order=60), dos=DoSConfig(N=32, R=8))
    text = display_report(rep)
    assert "order 60" in text and "iterations" in text


def test_polfed_rejects_blocks_wider_than_64(dev):
    """Blocks wider than the device block kernels take fail before any work,
    with a clear error (not inside the Lanczos loop after the DoS)."""
    from polfed_amd import polfed
    from polfed_amd.models import xxz_hamiltonian
    op = xxz_hamiltonian(10, 5, 1.0, 1.0, 1.0, rng=np.random.default_rng(3), device=dev)
    x0 = np.random.default_rng(4).standard_normal((op.D, 65))
    with pytest.raises(ValueError, match="at most 64 columns"):
        polfed(op, x0, 70, "middle")


def test_projected_eigensolver_on_the_device(dev):
    """From k = DEVICE_EIGH_MIN the projected eigenproblem runs on the device:
    same eigenvalues as the host banded solver, orthonormal eigenvectors with
    small residuals, and the index subset of the switch test."""
    from polfed_amd.lanczos import DEVICE_EIGH_MIN, _proj_eigh
    k, s = DEVICE_EIGH_MIN + 96, 16
    rng = np.random.default_rng(11)
    T = np.zeros((k, k))
    for i in range(0, k, s):
        A = rng.standard_normal((s, s))
        T[i:i + s, i:i + s] = A + A.T
        if i + s < k:
            R = np.triu(rng.standard_normal((s, s)))
            T[i:i + s, i + s:i + 2 * s] = R.T
            T[i + s:i + 2 * s, i:i + s] = R
    lam_h, _ = _proj_eigh(T, s)
    lam_d, phi_d = _proj_eigh(T, s, device=dev.torch_device)
    scale = np.max(np.abs(lam_h))
    assert np.max(np.abs(lam_d - lam_h)) <= 1e-12 * scale
    assert np.linalg.norm(phi_d.T @ phi_d - np.eye(k)) <= 1e-11
    assert np.max(np.linalg.norm(T @ phi_d - phi_d * lam_d, axis=0)) <= 1e-11 * scale
    ls, ps = _proj_eigh(T, s, [k - 20, k - 1], device=dev.torch_device)
    assert np.array_equal(ls, lam_d[k - 20:])
    assert np.array_equal(ps, phi_d[:, k - 20:])
