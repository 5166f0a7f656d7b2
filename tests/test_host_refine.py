"""Host logic of the FP64 polish (refine.py): the window polynomial."""
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
    assert refine.polish_degree(28000) == 1750
    assert refine.polish_degree(100) == 32
    assert refine.polish_degree(28000, 0.25) == 7000
