"""Value-bucket packing of a sparse matrix (host side of the packed operator).

Restates ``get_diags_and_offdiagonals_by_value``
(src/Optimization/optimization.jl:134-182) with its helpers ``round_key``
(:24-25), ``allocate_offdiagonal_buffers`` (:234-251) and
``pack_offdiagonals`` (:259-265): the diagonal and every off-diagonal entry
are rounded to ``round_digits`` decimals, entries with ``|v| < tol`` are
dropped, the distinct rounded values are sorted ascending, and for each value
the columns of each row are listed in ascending order.  The layout is the
C ABI's: 0-based columns, one concatenated ``flat`` array and one
``(D+1)``-entry table of absolute start offsets per value (the reference keeps
a 1-based ``(value, flat, starts)`` tuple per value, ``starts`` of length D).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp


def round_key(x, digits: int = 15):
    """``round(x; digits=digits)`` of the reference (optimization.jl:24).
    Julia computes ``round(x * 10^d) / 10^d`` (ties to even) in FP64, and so
    does ``numpy.round`` -- not the correctly rounded decimal rounding of
    Python's ``round``, which differs in the last bit for some values."""
    return np.round(np.asarray(x, dtype=np.float64), digits)


def pack_by_value(H, tol: float = 1e-13, round_digits: int = 15):
    """``(diag, vals, starts, flat)`` of a square real matrix (scipy sparse or
    dense): ``diag`` (D,), ``vals`` (nvals,) ascending, ``starts``
    (nvals, D+1) int64 and ``flat`` (nnz,) int32, with the columns of value v
    in row r at ``flat[starts[v, r]:starts[v, r+1]]``."""
    if sp.issparse(H):
        A = sp.csr_matrix(H, dtype=np.float64)
        A.sum_duplicates()
        coo = A.tocoo()
        r, c, v = coo.row.astype(np.int64), coo.col.astype(np.int64), coo.data
        diag_raw = A.diagonal()
    else:
        A = np.asarray(H, dtype=np.float64)
        r, c = np.nonzero(A)
        r, c = r.astype(np.int64), c.astype(np.int64)
        v = A[r, c]
        diag_raw = np.diagonal(A).copy()
    D = A.shape[0]
    if A.shape != (D, D):
        raise ValueError(f"matrix must be square, got {A.shape}")
    if np.iscomplexobj(v) or np.iscomplexobj(diag_raw):
        raise TypeError("the packed operator is real-only")
    diag = round_key(diag_raw, round_digits)
    keep = (r != c) & (np.abs(v) >= tol)
    r, c, v = r[keep], c[keep], v[keep]
    keys = round_key(v, round_digits)
    vals, inv = np.unique(keys, return_inverse=True)
    inv = inv.reshape(-1)
    nvals = len(vals)
    order = np.lexsort((c, r, inv))  # by value, then row, then ascending column
    flat = c[order].astype(np.int32)
    counts = np.zeros((nvals, D), dtype=np.int64)
    np.add.at(counts, (inv, r), 1)
    starts = np.zeros((nvals, D + 1), dtype=np.int64)
    starts[:, 1:] = np.cumsum(counts, axis=1)
    base = np.concatenate([[0], np.cumsum(counts.sum(axis=1))[:-1]]) if nvals else np.zeros(0)
    starts += base.astype(np.int64)[:, None]
    return diag, vals.astype(np.float64), starts, flat
