"""GPU parity at the full sizes of the workload configurations, through the
launch layouts the autotuner picks there (the small-size tests elsewhere
cannot reach them): XXZ L=20 block 8 (D=184756), XXZ L=22 block 16
(D=705432, FP64 and FP32), the 64 Quantum-Sun instances of D 8192..131072 in
one batched launch (graph mode, FP64 and FP32), and J1-J2 L=24 block 32
(D=2704156: 24-bit rank offsets and 32-bit byte offsets near their limits),
each against the CPU oracle's sparse matrix on the same seeded inputs."""
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


def _gershgorin(H):
    """An interval holding the spectrum (any such bounds keep |H~| <= 1)."""
    H = H.tocsr()
    d = H.diagonal()
    r = np.asarray(abs(H).sum(axis=1)).ravel() - np.abs(d)
    return float(np.min(d - r)), float(np.max(d + r))


def _chain_case(dev, L, J2, W, seed):
    from polfed_amd.operators import SpinChainOperator
    if J2 == 0.0:
        bonds, Js, Ds = om.xxz_bonds(L, 1.0, 1.0, "periodic")
    else:
        bonds, Js, Ds = om.j1j2_bonds(L, 1.0, J2, 1.0, 1.0, "periodic")
    h = om.resolve_fields(L, 0.0, W, rng=np.random.default_rng(seed))
    op = SpinChainOperator(L, L // 2, bonds, Js, Ds, h, device=dev)
    H = om.spin_half_hamiltonian(L, bonds, Js, Ds, h, L // 2).tocsr()
    assert op.D == H.shape[0]
    return op, H


def _filter_vs_oracle(dev, op, H, s, K, dtypes, target=0.1, seed=5):
    """Autotune (dtype, s) as polfed does, then compare the filter in both row
    orders with the oracle's Clenshaw on the rescaled sparse matrix."""
    Emin, Emax = _gershgorin(H)
    op.set_rescale(Emin, Emax)
    a, b = (Emax - Emin) / 2, (Emax + Emin) / 2
    Ht = ((H - b * sp.identity(H.shape[0], format="csr")) / a).tocsr()
    c = oc.filter_coefficients(target, K)
    X = np.random.default_rng(seed).standard_normal((op.D, s))
    ref = oc.clenshaw_apply(lambda Z: Ht @ Z, c, X)
    Xd = dev.to_device(X)
    for dtype, tol in dtypes:
        op.ensure_tuned(s, dtype)
        outs = {}
        for part in (0, 1):
            op.set_tuning(row_partition=part)
            Y = dev.to_host(op.filter(c, Xd, dtype=dtype))
            err = _rel(Y, ref)
            print(f"D={op.D} s={s} K={K} {dtype} partition={part}: rel err {err:.2e}", flush=True)
            assert err < tol, (dtype, part, err)
            outs[part] = Y
        # the partition order changes where rows are processed, not any sum
        assert np.array_equal(outs[0], outs[1]), dtype
        op.set_tuning(row_partition=-1)
    return Ht


def test_xxz_l20_block8_filter_and_solve(dev):
    """XXZ L=20, W=3 seed 1, block 8: the FP64 filter at K=200 through the
    autotuned layout (both row orders) against the oracle, then one complete
    solve (nev=100, target middle) whose pairs are checked with the oracle's
    H: residuals <= 1e-10, Rayleigh quotients, orthonormal vectors."""
    from polfed_amd import FactorizationConfig, polfed
    op, H = _chain_case(dev, 20, 0.0, 3.0, 1)
    _filter_vs_oracle(dev, op, H, 8, 200, [("f64", 1e-12), ("f32", 5e-4)])
    q, _ = np.linalg.qr(np.random.default_rng(101).standard_normal((op.D, 8)))
    op.set_tuning(reset=True)
    vals, vecs = polfed(op, dev.to_device(q), 100, "middle",
                        fact=FactorizationConfig(eigentol=1e-10), device=dev)
    assert len(vals) == 100
    psi = dev.to_host(vecs)
    Hpsi = H @ psi
    res = np.linalg.norm(Hpsi - psi * vals, axis=0)
    rq = np.einsum("ij,ij->j", psi, Hpsi)
    Emin, Emax = _gershgorin(H)
    print(f"L=20 solve: max residual {res.max():.2e}, max |E - <psi|H|psi>| "
          f"{np.max(np.abs(rq - vals)):.2e}", flush=True)
    assert res.max() <= 1e-10
    assert np.max(np.abs(rq - vals)) <= 1e-12 * (Emax - Emin)
    assert np.linalg.norm(psi.T @ psi - np.eye(100)) <= 1e-12


def test_xxz_l22_block16_filters(dev):
    """XXZ L=22, W=2 seed 2, block 16: FP64 and FP32 filters through the
    autotuned layouts, both row orders, against the oracle."""
    op, H = _chain_case(dev, 22, 0.0, 2.0, 2)
    _filter_vs_oracle(dev, op, H, 16, 60, [("f64", 1e-12), ("f32", 5e-4)])


def test_j1j2_l24_block32_apply_filter_kpm(dev):
    """J1-J2 L=24 (J2=0.5), block 32: FP32 apply and FP64/FP32 filters against
    the oracle, and KPM moments from random-phase probes in FP64 and FP32
    against the oracle's FP64 moments on the same probes."""
    from polfed_amd.blockops import fill_probes
    op, H = _chain_case(dev, 24, 0.5, 0.0, 0)
    X = np.random.default_rng(9).standard_normal((op.D, 32)).astype(np.float32)
    HX = H @ X.astype(np.float64)
    Y32 = dev.to_host(op.apply(dev.to_device(X, dtype=torch.float32), dtype="f32"))
    err = _rel(Y32.astype(np.float64), HX)
    print(f"L=24 FP32 apply: rel err {err:.2e}", flush=True)
    assert err < 1e-5
    Ht = _filter_vs_oracle(dev, op, H, 32, 8, [("f64", 1e-12), ("f32", 5e-4)])
    N, R = 64, 4
    P = fill_probes(dev, op.D, R, seed=77, kind="phase")
    mu_ref = odos.dos_moments(lambda v: Ht @ v, N, dev.to_host(P))
    for dtype, tol in (("f64", 1e-12), ("f32", 1e-6)):
        mu = op.kpm_moments(N, P, dtype=dtype)
        err = np.max(np.abs(mu - mu_ref)) / op.D
        print(f"L=24 KPM {dtype}: max |mu - mu_ref| / D = {err:.2e}", flush=True)
        assert err < tol, dtype


def _qsun_batch_cases(dev, n=64):
    """The 64 Quantum-Sun realizations of the batched workload: L_grain=3,
    L_loc = 10..14 round-robin, alpha=0.75, g0=1, numpy seeds 1000+i."""
    from polfed_amd.operators import QSunOperator
    ops, Hs = [], []
    for i in range(n):
        L_loc = 10 + i % 5
        draws = om.qsun_draws(L_loc, 3, 0.75, rng=np.random.default_rng(1000 + i))
        op = QSunOperator(L_loc, 3, 1.0, *draws, device=dev)
        H = om.qsun_hamiltonian_from_draws(L_loc, 3, 1.0, *draws).tocsr()
        op.set_rescale(*_gershgorin(H))
        ops.append(op)
        Hs.append(H)
    return ops, Hs


def test_qsun_64_instances_batched_graph_filter(dev):
    """All 64 instances (D 8192..131072) in one batched filter per step, HIP
    graph mode with several replays per application, uneven orders; FP64,
    FP32 and both kinds in one application against the oracle for every
    instance, with the autotuned launch form and then the other one (block
    rows staged in LDS or not)."""
    from polfed_amd.batch import QSunBatch
    ops, Hs = _qsun_batch_cases(dev)
    orders = [8 + (7 * i) % 41 for i in range(len(ops))]
    targets = [0.3 * np.sin(i) for i in range(len(ops))]
    coeffs = [oc.filter_coefficients(t, K) for t, K in zip(targets, orders)]
    batch = QSunBatch(ops, 1)
    batch.set_filter(coeffs)
    batch.autotune("f32", nsteps=20)
    batch.set_tuning(graph=1, graph_steps=16)
    rng = np.random.default_rng(12)
    Xs = [rng.standard_normal((op.D, 1)) for op in ops]
    refs = []
    for H, op, c, X in zip(Hs, ops, coeffs, Xs):
        a, b = (op.Emax - op.Emin) / 2, (op.Emax + op.Emin) / 2
        Ht = ((H - b * sp.identity(H.shape[0], format="csr")) / a).tocsr()
        refs.append(oc.clenshaw_apply(lambda Z: Ht @ Z, c, X))
    tuned = batch.tuning()
    forms = [tuned, {"lds_rows": 1 - tuned["lds_used"], "rows_per_lane": 4}]
    tol = {"f64": 1e-12, "f32": 5e-4}
    for form in forms:
        batch.set_tuning(lds_rows=form["lds_rows"], rows_per_lane=form["rows_per_lane"])
        used = batch.tuning()["lds_used"]
        for mode in ("f64", "f32", "mixed"):
            dts = [("f64", "f32")[i % 2] if mode == "mixed" else mode for i in range(len(ops))]
            Xd = [dev.to_device(X) for X in Xs]
            Yd = [torch.full_like(x, np.nan) for x in Xd]
            batch.filter([(x, y, dt) for x, y, dt in zip(Xd, Yd, dts)])
            errs = [_rel(dev.to_host(y), ref) / tol[dt] for y, ref, dt in zip(Yd, refs, dts)]
            print(f"64-instance batch {mode} (LDS rows {used}): max err/tol {max(errs):.2e}",
                  flush=True)
            assert max(errs) < 1, (mode, used, int(np.argmax(errs)))
    assert {f["lds_rows"] for f in forms} == {0, 1}
    batch.close()
