"""polfed_amd -- MI355X-native hot path of the POLFED eigensolver.

Import name ``polfed_amd`` (the directory is ``polfed.jl_amd``; the test,
bench and build entry points register it under that name).  The GPU work runs
in ``lib/libpolfed_hip.so`` (C ABI in ``include/polfed_hip.h``); this package
is the host-side mirror of the reference's Julia API around it.
"""
from ._lib import NumericBreakdown, PolfedHIPError  # noqa: F401
from .configs import (DoSConfig, FactorizationConfig, FullRO, HybridMatrixBasis,  # noqa: F401
                      MappingConfig, MatrixBasis, MulColsParallel, NoParallel, PartialRO,
                      PolfedDefaults, TransformConfig, TwoLevelParallel)

__version__ = "0.1.0"


def polfed(*args, **kwargs):
    """See ``polfed_amd.core.polfed``."""
    from .core import polfed as _polfed
    return _polfed(*args, **kwargs)


def lanczos_extrema(*args, **kwargs):
    """See ``polfed_amd.core.lanczos_extrema``."""
    from .core import lanczos_extrema as _lanczos_extrema
    return _lanczos_extrema(*args, **kwargs)


def display_report(*args, **kwargs):
    """See ``polfed_amd.core.display_report``."""
    from .core import display_report as _display_report
    return _display_report(*args, **kwargs)


def polfed_batch(*args, **kwargs):
    """See ``polfed_amd.core.polfed_batch``."""
    from .core import polfed_batch as _polfed_batch
    return _polfed_batch(*args, **kwargs)
