"""Krylov basis storage of the block Lanczos.

``DeviceBasis`` keeps every basis column in device memory (the reference's
``MatrixBasis``, src/Lanczos/OrthonormalBasis/MatrixBasis.jl).

``HybridBasis`` follows ``HybridMatrixBasis``
(src/Lanczos/OrthonormalBasis/HybridMatrixBasis.jl:8-114, sized as in
``createbasis``, OrthonormalBasis.jl:19-59): the first ``nd`` columns live in
device memory, the rest in pinned host memory, so the Krylov dimension is not
bounded by HBM.  The reference projects the host part on the CPU (the block is
copied to the host, BLAS GEMMs there, copied back;
Factorization/FactorizationSteps/reorthogonalization.jl:62-74).  Here every
projection stays on the device: host columns are streamed through two device
staging buffers on a copy stream, the copy of chunk i+1 overlapping the FP64
MFMA work on chunk i, and only the basis crosses PCIe (D x 8 B per column and
pass), never the residual block.

Block life cycle in the hybrid form: a block with index >= nd is written into
a ring of three device blocks (the newest block and the two the three-term
step reads: V_last, V_prev), copied to its host columns as soon as its QR is
final, and read back from the host only once it is older than the last two
blocks -- exactly the blocks the reorthogonalization covers.  A ring slot is
rewritten only after the copy of its previous block has completed.

Reorthogonalization with host columns runs, per pass, one classical
Gram-Schmidt sweep over the device columns and then one sweep per host chunk
(block modified Gram-Schmidt across the parts, as the reference does across
its two tiers); with no host columns in the range it is the same single CGS2
call as ``DeviceBasis``, so both forms give bit-identical iterations until the
first host column enters the reorthogonalization.
"""
from __future__ import annotations

import torch

from . import blockops as bo
from ._lib import PF_ERR_INVALID, PolfedHIPError
from .configs import HybridMatrixBasis, MatrixBasis
from .device import Device

_F64 = 8


class DeviceBasis:
    """All maxdim basis columns in one (maxdim, D) device block."""

    def __init__(self, dev: Device, maxdim: int, D: int, s: int, device):
        self.dev, self.s = dev, s
        self.V = torch.empty((maxdim, D), dtype=torch.float64, device=device)
        self.device_vectors, self.host_vectors = maxdim, 0

    def block(self, k: int) -> torch.Tensor:
        """Device view of basis columns k .. k+s-1 (k a multiple of s)."""
        return self.V[k:k + self.s]

    def prepare(self, k: int) -> None:
        """Called before block k is written."""

    def commit(self, k: int) -> None:
        """Called once block k is final (after its QR)."""

    def reorth(self, W: torch.Tensor, nb: int, passes: int) -> None:
        """W -= V_{0:nb} (V_{0:nb}^T W), ``passes`` sweeps."""
        bo.reorth(self.dev, self.V[0:nb], W, passes=passes)

    def combine(self, Phi: torch.Tensor, out: torch.Tensor, k: int) -> None:
        """out = V_{0:k} Phi (Phi an (n, k) device tensor = column-major k x n)."""
        bo.update(self.dev, self.V[0:k], Phi, out, alpha=1.0, beta=0.0)

    def close(self) -> None:
        """Orders every pending copy before later work on the compute stream."""


class HybridBasis:
    """nd device columns, then maxdim - nd pinned host columns."""

    def __init__(self, dev: Device, maxdim: int, D: int, s: int, device, nd: int,
                 chunk_bytes: int):
        if nd < 0 or nd % s:
            raise PolfedHIPError(PF_ERR_INVALID, "device columns must be whole blocks")
        self.dev, self.s, self.D = dev, s, D
        self.nd = min(nd, maxdim)
        self.nh = maxdim - self.nd
        self.device_vectors, self.host_vectors = self.nd, self.nh
        f64 = dict(dtype=torch.float64, device=device)
        self.Vd = torch.empty((self.nd, D), **f64)
        self.Vh = self.ring = None
        self.stg = []
        if self.nh:
            self.Vh = torch.empty((self.nh, D), dtype=torch.float64, pin_memory=True)
            self.ring = torch.empty((3 * s, D), **f64)
            cc = max(s, (chunk_bytes // (_F64 * D)) // s * s)
            self.cc = min(cc, self.nh)
            self.stg = [torch.empty((self.cc, D), **f64) for _ in range(2)]
            self.cs = torch.cuda.Stream(device=device)  # default priority
            self.committed = [None] * 3   # copy event of each ring slot's block
            self.stg_free = [None] * 2    # last use of each staging buffer

    def _slot(self, k: int) -> int:
        return (k - self.nd) // self.s % 3

    def block(self, k: int) -> torch.Tensor:
        if k < self.nd:
            return self.Vd[k:k + self.s]
        i = self._slot(k)
        return self.ring[i * self.s:(i + 1) * self.s]

    def prepare(self, k: int) -> None:
        if k >= self.nd:
            ev = self.committed[self._slot(k)]
            if ev is not None:  # the block held three blocks ago has reached the host
                torch.cuda.current_stream(self.Vd.device).wait_event(ev)

    def commit(self, k: int) -> None:
        if k < self.nd:
            return
        comp = torch.cuda.current_stream(self.Vd.device)
        ready = torch.cuda.Event()
        ready.record(comp)
        with torch.cuda.stream(self.cs):
            self.cs.wait_event(ready)
            h = k - self.nd
            self.Vh[h:h + self.s].copy_(self.block(k), non_blocking=True)
            done = torch.cuda.Event()
            done.record(self.cs)
        self.committed[self._slot(k)] = done

    def _stream(self, hi: int, fn) -> None:
        """fn(chunk, h0) on the compute stream for host columns [0, hi), each
        chunk copied to a staging buffer first (the copy of the next chunk
        overlaps fn on this one)."""
        comp = torch.cuda.current_stream(self.Vd.device)
        for i, h0 in enumerate(range(0, hi, self.cc)):
            h1 = min(h0 + self.cc, hi)
            j = i % 2
            buf = self.stg[j][:h1 - h0]
            with torch.cuda.stream(self.cs):
                if self.stg_free[j] is not None:
                    self.cs.wait_event(self.stg_free[j])
                buf.copy_(self.Vh[h0:h1], non_blocking=True)
                ready = torch.cuda.Event()
                ready.record(self.cs)
            comp.wait_event(ready)
            fn(buf, h0)
            free = torch.cuda.Event()
            free.record(comp)
            self.stg_free[j] = free

    def reorth(self, W: torch.Tensor, nb: int, passes: int) -> None:
        if nb <= self.nd:
            bo.reorth(self.dev, self.Vd[0:nb], W, passes=passes)
            return
        for _ in range(passes):
            if self.nd:
                bo.reorth(self.dev, self.Vd, W, passes=1)
            self._stream(nb - self.nd, lambda C, h0: bo.reorth(self.dev, C, W, passes=1))

    def combine(self, Phi: torch.Tensor, out: torch.Tensor, k: int) -> None:
        kd = min(k, self.nd)
        if kd:
            bo.update(self.dev, self.Vd[0:kd], Phi[:, 0:kd], out, alpha=1.0, beta=0.0)
        if k > self.nd:
            def acc(C, h0):
                c0 = self.nd + h0
                bo.update(self.dev, C, Phi[:, c0:c0 + C.shape[0]], out, alpha=1.0,
                          beta=1.0 if (kd or h0) else 0.0)
            self._stream(k - self.nd, acc)

    def close(self) -> None:
        if self.nh:  # the last host copies finish before the storage can be reused
            torch.cuda.current_stream(self.Vd.device).wait_stream(self.cs)


def make_basis(dev: Device, basistype, maxdim: int, D: int, s: int, device):
    """Basis storage for ``FactorizationConfig.basistype``: ``MatrixBasis``
    (default) or ``HybridMatrixBasis`` (the class or an instance carrying
    ``max_device_vectors`` / ``chunk_bytes``)."""
    if basistype is None or basistype is MatrixBasis or isinstance(basistype, MatrixBasis):
        return DeviceBasis(dev, maxdim, D, s, device)
    if basistype is HybridMatrixBasis:
        basistype = HybridMatrixBasis()
    if not isinstance(basistype, HybridMatrixBasis):
        raise PolfedHIPError(PF_ERR_INVALID, f"unknown basis type {basistype!r}")
    nd = basistype.max_device_vectors
    if nd is None:
        # createbasis (OrthonormalBasis.jl:23-36): 90 % of the free device
        # memory, less this storage's own ring and staging buffers
        free, _ = torch.cuda.mem_get_info(device)
        chunk = max(s, (basistype.chunk_bytes // (_F64 * D)) // s * s)
        nd = int((0.9 * free) // (_F64 * D)) - 3 * s - 2 * chunk
    nd = max(0, min(int(nd), maxdim)) // s * s
    return HybridBasis(dev, maxdim, D, s, device, nd, basistype.chunk_bytes)
