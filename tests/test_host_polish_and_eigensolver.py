"""Host logic of the FP64 polish (refine.py) and the projected eigensolver (lanczos.py)."""
import numpy as np
import pytest

from oracle import clenshaw as oc
from oracle import dos as odos


@pytest.mark.parametrize("target,degree", [(0.0, 64), (0.37, 300), (-0.8, 1000)])
def test_window_coefficients(target, degree):
    from polfed_amd import refine
    c = refine.window_coefficients(target, degree)
    # the Jackson-damped delta expansion, term by term (oracle kernel and
    # default coefficients), normalized to 1 at the target
    raw = np.array([odos.jackson(n, degree) * oc.coefficients(target, n)
                    for n in range(degree + 1)])
    peak = oc.clenshaw_scalar(lambda n: raw[n], degree, target)
    assert np.allclose(c, raw / peak, rtol=1e-12, atol=1e-15)
    # a window: 1 at the target, small a few widths away
    x = np.linspace(-1, 1, 4001)
    q = np.polynomial.chebyshev.chebval(x, c)
    assert abs(np.polynomial.chebyshev.chebval(target, c) - 1.0) < 1e-12
    far = np.abs(x - target) > 20.0 / degree
    assert np.max(np.abs(q[far])) < 0.02


def test_polish_degree():
    from polfed_amd import refine
    assert refine.polish_degree(28000) == 875
    assert refine.polish_degree(100) == 32
    assert refine.polish_degree(28000, 0.25) == 7000


def _block_tridiagonal(k, s, seed):
    rng = np.random.default_rng(seed)
    T = np.zeros((k, k))
    for i in range(0, k, s):
        A = rng.standard_normal((s, s))
        T[i:i + s, i:i + s] = A + A.T
        if i + s < k:
            R = np.triu(rng.standard_normal((s, s)))
            T[i:i + s, i + s:i + 2 * s] = R.T
            T[i + s:i + 2 * s, i:i + s] = R
    return T


@pytest.mark.parametrize("k,s", [(60, 1), (96, 4), (128, 16)])
def test_projected_eigensolver_host_paths(k, s):
    """The banded (and, for blocks of one, tridiagonal) host solvers against
    numpy's dense eigh, full and by index subset."""
    from polfed_amd.lanczos import _proj_eigh
    T = _block_tridiagonal(k, s, 3)
    w = np.linalg.eigvalsh(T)
    lam, phi = _proj_eigh(T, s)
    assert np.max(np.abs(lam - w)) <= 1e-12 * np.max(np.abs(w))
    assert np.linalg.norm(T @ phi - phi * lam) <= 1e-11 * np.max(np.abs(w)) * k
    lo, hi = k - 10, k - 1
    ls, ps = _proj_eigh(T, s, [lo, hi])
    assert np.max(np.abs(ls - w[lo:hi + 1])) <= 1e-12 * np.max(np.abs(w))
    assert np.linalg.norm(T @ ps - ps * ls) <= 1e-11 * np.max(np.abs(w)) * k
