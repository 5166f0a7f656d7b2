"""Oracle restatement of the packed value-bucket operator (test infrastructure).

Restates ``src/Optimization/optimization.jl`` (packing) and
``src/Optimization/cpu/{single_row_mappings,mapping,clenshaw_recurrence_relation,
clenshaw_final_sum}.jl`` (row kernels).  Indices here are 0-based; the
reference's ``flat``/``starts`` are 1-based (add 1 to compare).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def round_key(x: float, digits: int = 15) -> float:
    """``round_key`` (src/Optimization/optimization.jl:24): Julia's
    ``round(x; digits)`` is ``round(x * 10^digits) / 10^digits`` with ties to
    even, in FP64 (base/floatfuncs.jl ``_round_invstep``); not Python's
    correctly rounded decimal ``round``, which differs in the last bit for
    some values."""
    sc = 10.0 ** digits
    return float(np.rint(float(x) * sc) / sc)


def get_diags_and_offdiagonals_by_value(mat, tol=1e-13, round_digits=15):
    """``get_diags_and_offdiagonals_by_value`` (src/Optimization/optimization.jl:134-226).

    Returns ``(diagonals, [(value, flat, starts), ...])`` with buckets sorted by
    value (``pack_offdiagonals`` :259-265).  ``flat`` holds, row by row, the
    column indices j with H[i, j] in the bucket, ascending in j (the CSC
    column-major traversal of :149-179 visits j in increasing order).  0-based.
    """
    if sp.issparse(mat):
        A = sp.csc_matrix(mat)
    else:
        A = sp.csc_matrix(np.asarray(mat))
    dim = A.shape[0]
    diagonals = np.array([round_key(v, round_digits) for v in A.diagonal()])
    coo = A.tocoo()
    r, c, v = coo.row, coo.col, coo.data
    keep = (r != c) & (np.abs(v) >= tol)
    r, c, v = r[keep], c[keep], v[keep]
    keys = np.array([round_key(x, round_digits) for x in v])
    buckets = []
    for key in sorted(set(keys.tolist())):
        sel = keys == key
        rr, cc = r[sel], c[sel]
        order = np.lexsort((cc, rr))           # row-major, ascending column
        rr, cc = rr[order], cc[order]
        counts = np.bincount(rr, minlength=dim)
        starts = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int64)
        buckets.append((float(key), cc.astype(np.int64), starts))
    return diagonals, buckets


def _bucket_sums(X, flat, starts):
    """Per-row sum over a bucket's neighbor list (``mapping_offdiagonals_state_i``,
    src/Optimization/cpu/single_row_mappings.jl:47-63), ascending column order."""
    dim = X.shape[0]
    ends = np.concatenate([starts[1:], [flat.size]])
    out = np.zeros_like(X)
    counts = ends - starts
    maxc = int(counts.max()) if counts.size else 0
    for t in range(maxc):
        rows = np.nonzero(counts > t)[0]
        out[rows] += X[flat[starts[rows] + t]]
    return out


def optimized_mapping(diagonals, buckets, X):
    """``mapping!`` (src/Optimization/cpu/mapping.jl:27-120) / ``mapping_state_i``
    (single_row_mappings.jl:9-27, 66-79): Y = diag*X + sum_v v * sum_{j in N_v} X_j."""
    X = np.asarray(X)
    d = diagonals.reshape((-1,) + (1,) * (X.ndim - 1))
    Y = d * X
    for val, flat, starts in buckets:
        Y = Y + val * _bucket_sums(X, flat, starts)
    return Y


def clenshaw_recurrence(diagonals, buckets, b2, b3, c, X):
    """``map_with_clenshaw_recurrence_relation!``
    (src/Optimization/cpu/clenshaw_recurrence_relation.jl:104-145):
    b1 = c X + 2 H b2 - b3."""
    return c * X + 2 * optimized_mapping(diagonals, buckets, b2) - b3


def clenshaw_final_sum(diagonals, buckets, b1, b2, c, X):
    """``map_with_clenshaw_final_sum!`` (src/Optimization/cpu/clenshaw_final_sum.jl:25-60):
    Y = c X + H b1 - b2."""
    return c * X + optimized_mapping(diagonals, buckets, b1) - b2


def rescaled_packing(diagonals, buckets, Emin, Emax):
    """Rescaling step of ``optimize_spectral_transform``
    (src/Optimization/optimization.jl:98-103): diag -> (diag - b)/a, v -> v/a."""
    spread = (Emax - Emin) / 2
    center = (Emax + Emin) / 2
    return (diagonals - center) / spread, [(v / spread, f, s) for v, f, s in buckets]
