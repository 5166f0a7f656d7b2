"""polfed_amd -- MI355X-native hot path of the POLFED eigensolver.

Import name ``polfed_amd`` (the directory is ``polfed.jl_amd``; entry points
register it with ``polfed_amd_loader``-style path setup).  The GPU work runs in
``lib/libpolfed_hip.so`` (C ABI in ``include/polfed_hip.h``); this package is
the host-side mirror of the reference's Julia API around it.
"""
from ._lib import PolfedHIPError, NumericBreakdown  # noqa: F401

__version__ = "0.1.0"
