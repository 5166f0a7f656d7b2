"""GPU parity of the batched multi-instance Quantum-Sun filter (pf_batch_*)
and of polfed_batch, against the CPU oracle and the single-instance path."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from oracle import clenshaw as oc
from oracle import models as om

pytestmark = pytest.mark.gpu

# (L_loc, L_grain, seed) of uneven sizes: D = 2^(L_loc+L_grain)
INSTANCES = [(5, 3, 11), (7, 3, 12), (4, 3, 13), (6, 2, 14), (8, 3, 15)]


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _instances(dev, specs=INSTANCES):
    from polfed_amd.operators import QSunOperator
    ops, Hs = [], []
    for L_loc, L_grain, seed in specs:
        draws = om.qsun_draws(L_loc, L_grain, 0.75, rng=np.random.default_rng(seed))
        op = QSunOperator(L_loc, L_grain, 1.0, *draws, device=dev)
        H = om.qsun_hamiltonian_from_draws(L_loc, L_grain, 1.0, *draws)
        w = np.linalg.eigvalsh(H.toarray())
        op.set_rescale(w[0] - 0.01, w[-1] + 0.02)
        ops.append(op)
        Hs.append(H)
    return ops, Hs


def _oracle_filter(H, op, c, X, polytype="Chebyshev"):
    a, b = (op.Emax - op.Emin) / 2, (op.Emax + op.Emin) / 2
    Ht = (H - b * sp.identity(H.shape[0])) / a
    return oc.clenshaw_apply(lambda Z: Ht @ Z, c, X, polytype)


@pytest.mark.parametrize("s,rows_per_lane,shuffle", [(1, 1, 1), (3, 1, 1), (1, 16, 0), (2, 64, 1),
                                                      (2, 4, 0)])
@pytest.mark.parametrize("graph", [1, 0])
def test_batch_filter_matches_oracle(dev, s, rows_per_lane, shuffle, graph):
    from polfed_amd.batch import QSunBatch
    ops, Hs = _instances(dev)
    orders = [37, 120, 0, 1, 64]      # uneven orders, including 0 and 1
    targets = [0.1, -0.3, 0.0, 0.5, 0.25]
    coeffs = [oc.filter_coefficients(t, K) if K > 0 else np.array([0.7])
              for t, K in zip(targets, orders)]
    batch = QSunBatch(ops, s)
    batch.set_filter(coeffs)
    batch.set_tuning(graph=graph, graph_steps=16,   # several replays per application
                     rows_per_lane=rows_per_lane, grain_shuffle=shuffle)
    rng = np.random.default_rng(3)
    Xs = [rng.standard_normal((op.D, s)) for op in ops]
    refs = [_oracle_filter(H, op, c, X) for H, op, c, X in zip(Hs, ops, coeffs, Xs)]
    for rep in range(2):  # a second application reuses the captured graph
        Xd = [dev.to_device(X) for X in Xs]
        Yd = [torch.full_like(x, np.nan) for x in Xd]
        batch.filter([(x, y, "f64") for x, y in zip(Xd, Yd)])
        for i, (y, ref) in enumerate(zip(Yd, refs)):
            assert _rel(dev.to_host(y), ref) < 1e-12, (rep, i)


@pytest.mark.parametrize("s,vec,rows_per_lane,used", [
    (1, 1, 4, 1), (1, 1, 8, 1), (2, 2, 2, 1), (4, 1, 4, 1), (8, 2, 2, 1), (4, 4, 1, 1),
    (1, 1, 2, 0), (3, 1, 4, 0)])
def test_batch_filter_lds_rows_matches_oracle(dev, s, vec, rows_per_lane, used):
    """Block rows staged in LDS: blocks of 128..2048 rows, so the larger
    instances gather some flips through memory and read the others from LDS,
    and the smaller ones lie inside one block.  Rows per lane that are not
    whole FP32 row groups, or ncols=3 (not a power of two), keep the other
    form.  FP64, FP32 and both kinds in one application."""
    from polfed_amd.batch import QSunBatch
    ops, Hs = _instances(dev)
    orders = [37, 120, 0, 1, 64]
    coeffs = [oc.filter_coefficients(0.1 * i - 0.2, K) if K > 0 else np.array([0.7])
              for i, K in enumerate(orders)]
    batch = QSunBatch(ops, s)
    batch.set_filter(coeffs)
    batch.set_tuning(vec=vec, rows_per_lane=rows_per_lane, lds_rows=1, graph_steps=16)
    t = batch.tuning()
    assert t["lds_rows"] == 1 and t["lds_used"] == used
    rng = np.random.default_rng(5)
    Xs = [rng.standard_normal((op.D, s)) for op in ops]
    refs = [_oracle_filter(H, op, c, X) for H, op, c, X in zip(Hs, ops, coeffs, Xs)]
    tol = {"f64": 1e-12, "f32": 5e-4}
    for dts in (["f64"] * 5, ["f32"] * 5, ["f32", "f64", "f64", "f32", "f64"]):
        Xd = [dev.to_device(X) for X in Xs]
        Yd = [torch.full_like(x, np.nan) for x in Xd]
        batch.filter([(x, y, dt) for x, y, dt in zip(Xd, Yd, dts)])
        for i, (y, ref, dt) in enumerate(zip(Yd, refs, dts)):
            assert _rel(dev.to_host(y), ref) < tol[dt], (dts, i)


@pytest.mark.parametrize("row_vectors,rows_per_lane,used", [
    (1, 1, 1), (1, 2, 1), (1, 4, 1), (2, 2, 1), (2, 4, 1), (2, 1, 0)])
def test_batch_filter_row_vectors_matches_oracle(dev, row_vectors, rows_per_lane, used):
    """Row-vector form (one column; a lane loads 16 bytes of consecutive rows
    and permutes them per flip): blocks of 1024..4096 rows, so the small
    instances lie inside one block (rows past the end read row mod D and
    store nothing) and the larger ones gather their far flips from memory;
    grain sizes 2, 4 and 8 take both grain paths.  One row per lane with two
    vectors at a time keeps the other form.  FP64, FP32 and both
    kinds in one application, graph and stream mode."""
    from polfed_amd.batch import QSunBatch
    specs = INSTANCES + [(9, 3, 16), (10, 2, 17), (9, 1, 18)]   # D up to 4096, grain 2
    ops, Hs = _instances(dev, specs)
    orders = [37, 120, 0, 1, 64, 90, 45, 70]
    coeffs = [oc.filter_coefficients(0.1 * i - 0.3, K) if K > 0 else np.array([0.7])
              for i, K in enumerate(orders)]
    batch = QSunBatch(ops, 1)
    batch.set_filter(coeffs)
    batch.set_tuning(vec=1, rows_per_lane=rows_per_lane, row_vectors=row_vectors,
                     graph_steps=16)
    t = batch.tuning()
    assert t["row_vectors_used"] == used
    rng = np.random.default_rng(9)
    Xs = [rng.standard_normal((op.D, 1)) for op in ops]
    refs = [_oracle_filter(H, op, c, X) for H, op, c, X in zip(Hs, ops, coeffs, Xs)]
    tol = {"f64": 1e-12, "f32": 5e-4}
    n = len(ops)
    for graph in (1, 0):
        batch.set_tuning(graph=graph)
        for dts in (["f64"] * n, ["f32"] * n, [("f32", "f64")[i % 2] for i in range(n)]):
            Xd = [dev.to_device(X) for X in Xs]
            Yd = [torch.full_like(x, np.nan) for x in Xd]
            batch.filter([(x, y, dt) for x, y, dt in zip(Xd, Yd, dts)])
            for i, (y, ref, dt) in enumerate(zip(Yd, refs, dts)):
                assert _rel(dev.to_host(y), ref) < tol[dt], (graph, dts[i], i)


def test_batch_filter_inactive_and_mixed_dtypes(dev):
    """Idle instances keep their outputs untouched; per-instance FP32/FP64."""
    from polfed_amd.batch import QSunBatch
    ops, Hs = _instances(dev)
    coeffs = [oc.filter_coefficients(0.05 * i, 30 + 17 * i) for i in range(len(ops))]
    batch = QSunBatch(ops, 2)
    batch.set_filter(coeffs)
    rng = np.random.default_rng(4)
    Xs = [rng.standard_normal((op.D, 2)) for op in ops]
    refs = [_oracle_filter(H, op, c, X) for H, op, c, X in zip(Hs, ops, coeffs, Xs)]
    dts = ["f32", "f64", None, "f64", "f32"]
    Xd = [dev.to_device(X) for X in Xs]
    Yd = [torch.full_like(x, 7.0) for x in Xd]
    batch.filter([None if dt is None else (x, y, dt) for x, y, dt in zip(Xd, Yd, dts)])
    for i, dt in enumerate(dts):
        Y = dev.to_host(Yd[i])
        if dt is None:
            assert np.all(Y == 7.0)
        elif dt == "f64":
            assert _rel(Y, refs[i]) < 1e-12
        else:
            assert _rel(Y, refs[i]) < 5e-4


def test_batch_other_polynomial_and_single_filter_agree(dev):
    """Legendre batch against the single-instance pf_filter of each operator."""
    from polfed_amd.batch import QSunBatch
    ops, Hs = _instances(dev, INSTANCES[:3])
    rng = np.random.default_rng(6)
    coeffs = [rng.standard_normal(K + 1) * 0.1 for K in (9, 4, 12)]
    batch = QSunBatch(ops, 4)
    batch.set_filter(coeffs, "Legendre")
    Xd = [dev.to_device(rng.standard_normal((op.D, 4))) for op in ops]
    Yd = [torch.empty_like(x) for x in Xd]
    batch.filter([(x, y, "f64") for x, y in zip(Xd, Yd)])
    for op, H, c, x, y in zip(ops, Hs, coeffs, Xd, Yd):
        single = dev.to_host(op.filter(c, x, polytype="Legendre"))
        assert _rel(dev.to_host(y), single) < 1e-13
        assert _rel(dev.to_host(y), _oracle_filter(H, op, c, dev.to_host(x), "Legendre")) < 1e-12


def test_batch_autotune_keeps_results(dev):
    from polfed_amd.batch import QSunBatch
    ops, Hs = _instances(dev)
    coeffs = [oc.filter_coefficients(0.2, 50) for _ in ops]
    batch = QSunBatch(ops, 1)
    batch.set_filter(coeffs)
    us = batch.autotune("f32", nsteps=20)
    assert 0 < us < 1e5
    rng = np.random.default_rng(8)
    Xs = [rng.standard_normal((op.D, 1)) for op in ops]
    Xd = [dev.to_device(X) for X in Xs]
    Yd = [torch.empty_like(x) for x in Xd]
    batch.filter([(x, y, "f64") for x, y in zip(Xd, Yd)])
    for H, op, c, X, y in zip(Hs, ops, coeffs, Xs, Yd):
        assert _rel(dev.to_host(y), _oracle_filter(H, op, c, X)) < 1e-12


def test_batch_rejects_spin_chains(dev):
    from polfed_amd.batch import QSunBatch
    from polfed_amd.operators import SpinChainOperator
    bonds, Js, Ds = om.xxz_bonds(8, 1.0, 1.0, "periodic")
    op = SpinChainOperator(8, 4, bonds, Js, Ds, np.zeros(8), device=dev)
    with pytest.raises(TypeError):
        QSunBatch([op], 1)


@pytest.mark.parametrize("dtype", ["f64", "mixed"])
def test_polfed_batch_matches_per_instance_polfed(dev, dtype):
    """Every instance of polfed_batch returns what polfed returns for it alone
    (same probes, same bounds, same order): the batch changes only how the
    filters are launched."""
    from polfed_amd import FactorizationConfig, MappingConfig, TransformConfig, polfed, polfed_batch
    from polfed_amd.blockops import fill_probes
    specs = [(5, 3, 21), (6, 3, 22), (4, 3, 23), (7, 3, 24)]
    ops, Hs = _instances(dev, specs)
    evs = [np.linalg.eigvalsh(H.toarray()) for H in Hs]
    x0s = [np.random.default_rng(30 + i).standard_normal(op.D) for i, op in enumerate(ops)]
    probes = [fill_probes(dev, op.D, 100, seed=40 + i) for i, op in enumerate(ops)]
    fact = FactorizationConfig(eigentol=1e-10)
    tr = TransformConfig(dtype=dtype)
    nev = 8
    outs, brep = polfed_batch(ops, x0s, nev, "middle", produce_report=True, fact=fact,
                              transform=tr, probes=probes)
    assert brep.n == len(ops) and brep.filter_calls > 0
    for i, op in enumerate(ops):
        vals_b, vecs_b, rep_b = outs[i]
        mp = MappingConfig(Emin=rep_b.Emin, Emax=rep_b.Emax)
        vals_s, vecs_s, rep_s = polfed(op, x0s[i], nev, "middle", produce_report=True, fact=fact,
                                       transform=tr, mapping=mp, probes=probes[i])
        width = evs[i][-1] - evs[i][0]
        assert rep_b.order == rep_s.order
        assert len(vals_b) == len(vals_s) == nev
        assert np.max(np.abs(vals_b - vals_s)) <= 1e-12 * width
        # every returned value is an eigenvalue, residuals meet the tolerance
        assert max(np.min(np.abs(evs[i] - v)) for v in vals_b) <= 1e-10 * width
        psi = dev.to_host(vecs_b)
        assert np.linalg.norm(Hs[i] @ psi - psi * vals_b, axis=0).max() <= 1e-10
