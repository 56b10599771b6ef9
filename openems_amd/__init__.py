"""openems_amd -- MI355X (gfx950) FDTD timestep engine.

The engine is the C-ABI library liboe.so (declared in include/oe/oe.h,
kernels in openems_amd/csrc).  This package holds its build script and a
thin ctypes mirror of the reference engine API for Python callers and tests.
"""
from ._ffi import OEError, lib  # noqa: F401
from .engine import (  # noqa: F401
    ADE_DEBYE, ADE_DRUDE, ADE_LORENTZ, MODE_COMPAT, MODE_PHYSICAL, CpmlBoundaries, EngineBatch, FarFieldResult,
    HipEngine, Nf2ff, Operator, ScheduledExcitation, Sweep, SweepResult, Waveform,
)

__all__ = [
    "OEError", "lib", "Operator", "HipEngine", "EngineBatch", "ScheduledExcitation", "Waveform",
    "CpmlBoundaries", "Nf2ff", "FarFieldResult", "Sweep", "SweepResult", "MODE_COMPAT", "MODE_PHYSICAL", "ADE_LORENTZ", "ADE_DRUDE", "ADE_DEBYE",
]
