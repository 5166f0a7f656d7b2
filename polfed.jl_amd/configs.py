"""Configuration objects and defaults, mirroring the reference's keyword API.

``PolfedDefaults`` (src/Polfed/Structs/PolfedDefaults.jl:61-217),
``MappingConfig`` (src/Polfed/Structs/MappingConfig.jl:82-103),
``TransformConfig`` (src/Polfed/Structs/TransformConfig.jl:77-100),
``FactorizationConfig`` (src/Polfed/Structs/FactorizationConfig.jl:49-70),
``DoSConfig`` (src/Polfed/Structs/DensitiesOfStatesConfig.jl:39-51), the
parallelization tags (src/Polfed/Structs/Parallelization.jl) and the
reorthogonalization tags (src/Lanczos/ReOrthTechnics/ReOrthTechnics.jl).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Optional


# ------------------------------------------------------------------ defaults
class PolfedDefaults:
    """Default constants of the reference (PolfedDefaults.jl)."""

    produce_report = False
    optimize_mapping = False
    polynomialtype = "Chebyshev"
    cutoff = 0.17
    normalization = 1.0
    order_safety_factor = 0.97
    overestimate_iters = 1.25
    tol = 1e-14
    eigentol = 1e-9
    which = "LR"
    kernel = "Jackson"
    N = 250
    R = 300

    @staticmethod
    def coefficients(lam: float, n: int) -> float:
        """c_0 = 1, c_n = 2 cos(n acos(lam)) (PolfedDefaults.jl:104; the code,
        not its docstring, which states the reverse)."""
        return float((2 - (n == 0)) * math.cos(n * math.acos(lam)))

    @staticmethod
    def expectedkrylovdim(howmany: int, blocksize: int, eta: float) -> int:
        """ceil((20.427 s + 1.696 howmany) eta) (PolfedDefaults.jl:61-62)."""
        return math.ceil((20.427 * blocksize + 1.696 * howmany) * eta)


# ---------------------------------------------------------- strategy tags
class Parallelization:
    """Tag types of the reference; on the GPU every call is ``NoParallel``
    (src/Polfed/Plan/Plan.jl:64-67): one HIP launch covers the whole block."""


class NoParallel(Parallelization):
    pass


class MulColsParallel(Parallelization):
    def __init__(self, nt_per_col: int = 1):
        self.nt_per_col = nt_per_col


class TwoLevelParallel(Parallelization):
    def __init__(self, nt_per_col: int = 1):
        self.nt_per_col = nt_per_col


class ReOrthTechnique:
    pass


class FullRO(ReOrthTechnique):
    """Full reorthogonalization every iteration.  ``passes`` classical
    Gram-Schmidt sweeps (the reference applies one; 2 = CGS2)."""

    def __init__(self, passes: int = 2):
        self.passes = passes


class PartialRO(ReOrthTechnique):
    """Periodic reorthogonalization (reorthogonalization.jl:33-48)."""

    def __init__(self, skip: int, passes: int = 2):
        self.skip = skip
        self.passes = passes


# ----------------------------------------------------------------- configs
@dataclass
class MappingConfig:
    """Mapping stage options (MappingConfig.jl:14-61).

    f_rescaled(Y, X), clenshaw_recurrence(b1, b2, b3, c, [k,] X) and
    clenshaw_finalsum(b2, b3, c0, Y, X) are device-tensor callbacks with the
    reference's signatures; Emin/Emax skip the extremal Lanczos probes.
    """

    parallel_strategy: Optional[Parallelization] = None
    optimize_mapping: bool = PolfedDefaults.optimize_mapping
    f_rescaled: Optional[Callable] = None
    clenshaw_recurrence: Optional[Callable] = None
    clenshaw_finalsum: Optional[Callable] = None
    Emin: Optional[float] = None
    Emax: Optional[float] = None


@dataclass
class TransformConfig:
    """Polynomial filter options (TransformConfig.jl:6-100).  ``dtype``
    selects the filter arithmetic on the GPU: "f64" (plain drop-in), "f32"
    (every filter in FP32: eigenvalues at FP64 accuracy, residuals limited by
    FP32 rounding), or "mixed" (FP64 until the wanted Ritz pairs' residual
    estimate drops below ``switch_tol``, then FP32, falling back to FP64 if the
    final residuals miss eigentol).  With ``polish`` (default) the pairs of a
    run that used FP32 filters are extracted by the FP64 polish of
    ``refine.polish`` (an FP64 window filter of degree ``polish_fraction`` x
    order on the Ritz vectors, then Rayleigh-Ritz with the unscaled H).
    Extensions of the reference, which filters in FP64 only."""

    coefficients: Callable = PolfedDefaults.coefficients
    normalization: float = PolfedDefaults.normalization
    cutoff: float = PolfedDefaults.cutoff
    left: Optional[float] = None
    right: Optional[float] = None
    order: Optional[int] = None
    order_safety_factor: float = PolfedDefaults.order_safety_factor
    polynomialtype: str = PolfedDefaults.polynomialtype
    dtype: str = "f64"
    switch_tol: float = 1e-4
    polish: bool = True
    polish_fraction: float = 1.0 / 32.0


class MatrixBasis:
    """Basis storage tag: every Krylov vector in device memory
    (src/Lanczos/OrthonormalBasis/MatrixBasis.jl; PolfedDefaults.jl:170)."""


@dataclass
class HybridMatrixBasis:
    """Basis storage tag: device columns first, the rest in pinned host memory
    (src/Lanczos/OrthonormalBasis/HybridMatrixBasis.jl:8-114).  Pass the class
    or an instance.  ``max_device_vectors`` None sizes the device part at 90 %
    of the free device memory, as ``createbasis`` does
    (OrthonormalBasis.jl:19-59); ``chunk_bytes`` is the size of each of the two
    device staging buffers through which host columns are streamed."""

    max_device_vectors: Optional[int] = None
    chunk_bytes: int = 512 << 20


@dataclass
class FactorizationConfig:
    """Krylov factorization options (FactorizationConfig.jl:49-70).
    ``basistype``: ``MatrixBasis`` (default) or ``HybridMatrixBasis``."""

    rot: ReOrthTechnique = None
    which: str = PolfedDefaults.which
    tol: float = PolfedDefaults.tol
    eigentol: Optional[float] = PolfedDefaults.eigentol
    overestimate_iters: float = PolfedDefaults.overestimate_iters
    basistype: object = None

    def __post_init__(self):
        if self.rot is None:
            self.rot = FullRO()
        if self.basistype is None:
            self.basistype = MatrixBasis


@dataclass
class DoSConfig:
    """Density-of-states options (DensitiesOfStatesConfig.jl:39-51).

    ``probe_kind`` "gaussian" (reference) or "phase"; ``seed`` drives the
    counter-based probe generator; ``dtype`` the KPM arithmetic ("f64"/"f32").
    """

    N: int = PolfedDefaults.N
    R: int = PolfedDefaults.R
    kernel: str = PolfedDefaults.kernel
    probe_kind: str = "gaussian"
    seed: int = 0x5EED
    dtype: str = "f64"
