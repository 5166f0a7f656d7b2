"""Hybrid Krylov basis (device columns first, then pinned host columns streamed
through the device; src/Lanczos/OrthonormalBasis/HybridMatrixBasis.jl:8-114):
for every split of the basis, including all columns on the host, POLFED
returns the eigenpairs of the all-device basis (<= 1e-12 x width), with
residuals <= 1e-10 and eigenvector angles <= 1e-8; without host columns the
run is bit-identical to the all-device one."""
import numpy as np
import pytest

from oracle import models as om

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def chain(dev):
    from polfed_amd.models import xxz_hamiltonian
    L = 12
    h = om.resolve_fields(L, 0.0, 2.0, rng=np.random.default_rng(3))
    op = xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h, device=dev)
    H = om.xxz_hamiltonian(L, L // 2, 1.0, 1.0, 2.0, fields=h)
    ev, V = np.linalg.eigh(H.toarray())
    return op, H, ev, V


def _solve(dev, op, x0, nev, ev, basistype):
    from polfed_amd import DoSConfig, FactorizationConfig, MappingConfig, TransformConfig, polfed
    vals, vecs, rep = polfed(op, x0, nev, ("rescaled", 0.1), produce_report=True,
                             mapping=MappingConfig(Emin=ev[0], Emax=ev[-1]),
                             transform=TransformConfig(order=200),
                             fact=FactorizationConfig(eigentol=1e-10, basistype=basistype),
                             dos=DoSConfig(N=64, R=20))
    return vals, dev.to_host(vecs), rep


def _max_angle(V1, V2):
    Q1, _ = np.linalg.qr(V1)
    Q2, _ = np.linalg.qr(V2)
    return float(np.arcsin(min(1.0, np.linalg.norm(Q1 - Q2 @ (Q2.T @ Q1), 2))))


@pytest.mark.parametrize("s,nd", [(4, 0), (4, 8), (4, 40), (1, 0), (1, 13)])
def test_hybrid_basis_matches_device_basis(dev, chain, s, nd):
    from polfed_amd import HybridMatrixBasis, MatrixBasis
    op, H, ev, V = chain
    nev = 20
    rng = np.random.default_rng(4)
    x0 = np.linalg.qr(rng.standard_normal((op.D, s)))[0] if s > 1 else rng.standard_normal(op.D)
    v_ref, psi_ref, rep_ref = _solve(dev, op, x0, nev, ev, MatrixBasis)
    # staging buffers of two blocks (two columns at s = 1): many chunks per sweep
    hyb = HybridMatrixBasis(max_device_vectors=nd, chunk_bytes=8 * op.D * max(2 * s, 2))
    vals, psi, rep = _solve(dev, op, x0, nev, ev, hyb)
    lz = rep.lanczos
    assert lz.basis_device_vectors == nd and lz.basis_host_vectors == lz.maxiter * s - nd
    assert lz.krylovdim > nd + 2 * s, "the run must reorthogonalize against host columns"
    assert len(vals) == len(v_ref) >= nev - 2
    width = ev[-1] - ev[0]
    assert np.max(np.abs(vals - v_ref)) <= 1e-12 * width
    assert np.linalg.norm(H @ psi - psi * vals, axis=0).max() <= 1e-10
    idx = [int(np.argmin(np.abs(ev - v))) for v in vals]
    assert _max_angle(psi, V[:, idx]) <= 1e-8


def test_hybrid_basis_without_host_columns_is_bit_identical(dev, chain):
    from polfed_amd import HybridMatrixBasis, MatrixBasis
    op, H, ev, V = chain
    x0 = np.linalg.qr(np.random.default_rng(5).standard_normal((op.D, 4)))[0]
    v_ref, psi_ref, _ = _solve(dev, op, x0, 20, ev, MatrixBasis)
    vals, psi, rep = _solve(dev, op, x0, 20, ev, HybridMatrixBasis(max_device_vectors=10 ** 6))
    assert rep.lanczos.basis_host_vectors == 0
    assert np.array_equal(vals, v_ref) and np.array_equal(psi, psi_ref)
    # the default sizing (90 % of the free device memory) holds this basis on the device
    vals, psi, rep = _solve(dev, op, x0, 20, ev, HybridMatrixBasis)
    assert rep.lanczos.basis_host_vectors == 0
    assert np.array_equal(vals, v_ref)
