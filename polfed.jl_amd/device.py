"""Device context and block conventions of the host mirror.

PyTorch only provides device memory and streams here.  A "block" is a torch
tensor of shape (ncols, D), float64 unless stated, contiguous in its last
dimension: that is exactly a column-major D x ncols matrix with leading
dimension ``t.stride(0)`` -- the layout of the C ABI and of Julia.  A 1-D
tensor of length D is a single column.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from ._lib import PolfedHIPError, check


def _require_gpu():
    if not torch.cuda.is_available():
        raise PolfedHIPError(_lib.PF_ERR_UNSUPPORTED,
                             "no HIP device visible: the POLFED hot path runs only on the GPU")


class Device:
    """One GPU: owns a ``pf_handle`` (scratch buffers of the dense kernels)."""

    def __init__(self, index: int = 0):
        _require_gpu()
        self.index = index
        self.torch_device = torch.device("cuda", index)
        self.lib = _lib.load()
        h = ctypes.c_void_p()
        with torch.cuda.device(index):
            check(self.lib.pf_create(ctypes.byref(h), index))
        self.h = h

    @property
    def stream(self) -> ctypes.c_void_p:
        return ctypes.c_void_p(torch.cuda.current_stream(self.index).cuda_stream)

    def close(self):
        if getattr(self, "h", None):
            self.lib.pf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---------------------------------------------------------- allocation
    def block(self, D: int, ncols: int, dtype=torch.float64) -> torch.Tensor:
        return torch.empty((ncols, D), dtype=dtype, device=self.torch_device)

    def zeros(self, D: int, ncols: int, dtype=torch.float64) -> torch.Tensor:
        return torch.zeros((ncols, D), dtype=dtype, device=self.torch_device)

    def to_device(self, arr, dtype=torch.float64) -> torch.Tensor:
        """numpy (D,) or (D, ncols) column-major semantics -> device block (ncols, D)."""
        t = torch.as_tensor(arr, dtype=dtype)
        if t.dim() == 2:
            t = t.t()
        return t.contiguous().to(self.torch_device)

    @staticmethod
    def to_host(t: torch.Tensor):
        """device block (ncols, D) -> numpy (D, ncols) (or (D,) for 1-D)."""
        a = t.detach().to("cpu").numpy()
        return a.T.copy() if a.ndim == 2 else a.copy()


_DEFAULT: Device | None = None


def default_device() -> Device:
    global _DEFAULT
    if _DEFAULT is None:
        _DEFAULT = Device(torch.cuda.current_device() if torch.cuda.is_available() else 0)
    return _DEFAULT


def colmajor(t: torch.Tensor, dtype=torch.float64):
    """(pointer, ld, D, ncols) of a block tensor, validating layout and dtype."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise PolfedHIPError(_lib.PF_ERR_INVALID, "expected a device tensor")
    if t.dtype != dtype:
        raise PolfedHIPError(_lib.PF_ERR_INVALID, f"expected dtype {dtype}, got {t.dtype}")
    if t.dim() == 1:
        if t.stride(0) != 1:
            raise PolfedHIPError(_lib.PF_ERR_INVALID, "vector must be contiguous")
        return ctypes.c_void_p(t.data_ptr()), t.shape[0], t.shape[0], 1
    if t.dim() != 2 or (t.stride(1) != 1 and t.shape[1] > 1):
        raise PolfedHIPError(_lib.PF_ERR_INVALID,
                             "block must be a (ncols, D) tensor contiguous along D")
    # a single column may carry any stride in its (size-1) leading dimension
    ld = t.stride(0) if t.shape[0] > 1 else t.shape[1]
    return ctypes.c_void_p(t.data_ptr()), ld, t.shape[1], t.shape[0]
