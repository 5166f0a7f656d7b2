"""pf_filter_noalloc, the filter entry point that allocates nothing (its work
blocks live in the library's fixed per-GPU scratch array), called through the
C ABI against pf_filter on the same inputs: two calls in a row, the second on
a smaller problem, FP64 and FP32, plus the size limit's error path."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import clenshaw as oc
from oracle import models as om

pytestmark = pytest.mark.gpu


def _chain(dev, L, seed, j1j2=False):
    from polfed_amd.operators import SpinChainOperator
    if j1j2:
        bonds, Js, Ds = om.j1j2_bonds(L, 1.0, 0.5, 1.0, 1.0, "periodic")
    else:
        bonds, Js, Ds = om.xxz_bonds(L, 1.0, 1.0, "periodic")
    h = np.random.default_rng(seed).uniform(-3, 3, L)
    op = SpinChainOperator(L, L // 2, bonds, Js, Ds, h, device=dev)
    op.set_rescale(-2.0 * L, 2.0 * L)  # encloses the spectrum (|H| <= sum of |terms|)
    return op


def _filter_call(dev, name, op, coeffs, X, Y, dtype):
    """One call of pf_filter / pf_filter_noalloc through the C ABI."""
    from polfed_amd import _lib
    from polfed_amd.device import colmajor
    c = np.ascontiguousarray(coeffs, dtype=np.float64)
    px, ldx, D, n = colmajor(X)
    py, ldy, _, _ = colmajor(Y)
    code = _lib.PF_F64 if dtype == "f64" else _lib.PF_F32
    return getattr(dev.lib, name)(op.op, code, 0, c.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                  c.size - 1, px, ldx, py, ldy, n, dev.stream)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_noalloc_filter_equals_pf_filter(dev, dtype):
    cases = [(_chain(dev, 20, 1), 8, 0.1, 300),             # D = 184756, block 8
             (_chain(dev, 12, 2, j1j2=True), 3, -0.4, 90)]  # then a smaller problem
    for op, s, target, K in cases:
        c = oc.filter_coefficients(target, K)
        X = torch.randn((s, op.D), dtype=torch.float64, device=dev.torch_device)
        Y_ref = torch.full_like(X, np.nan)
        Y = torch.full_like(X, np.nan)
        assert _filter_call(dev, "pf_filter", op, c, X, Y_ref, dtype) == 0
        assert _filter_call(dev, "pf_filter_noalloc", op, c, X, Y, dtype) == 0
        torch.cuda.synchronize()
        print(f"noalloc {dtype} D={op.D} s={s} K={K}: identical={torch.equal(Y, Y_ref)}",
              flush=True)
        assert torch.isfinite(Y).all()
        assert torch.equal(Y, Y_ref)


def test_noalloc_filter_rejects_blocks_beyond_the_scratch(dev):
    from polfed_amd import _lib
    op = _chain(dev, 20, 3)
    c = oc.filter_coefficients(0.0, 10)
    # 4 x 184756 x 32 x 8 bytes = 189 MB > 128 MiB
    X = torch.zeros((32, op.D), dtype=torch.float64, device=dev.torch_device)
    Y = torch.empty_like(X)
    st = _filter_call(dev, "pf_filter_noalloc", op, c, X, Y, "f64")
    assert st == _lib.PF_ERR_INVALID
    assert b"no-allocation scratch" in dev.lib.pf_last_error()
    # the same block fits once split below the limit: 16 columns = 94.6 MB
    assert _filter_call(dev, "pf_filter_noalloc", op, c, X[:16], Y[:16], "f64") == 0
