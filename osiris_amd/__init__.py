"""osiris_amd -- the OSIRIS per-timestep PIC hot path on MI355X (gfx950).

Gather + Boris/Vay push + charge-conserving deposit (splines of order 1..3) + Yee FDTD,
as hand-written HIP kernels behind a C ABI (include/osiris_amd.h); this package is the
Python host side over that ABI.
"""
from ._lib import OSR_BORIS, OSR_VAY, OsrError, lib  # noqa: F401
from .sim import (Simulation, advance_deposit, copy_probe, device_count, smooth_setup,  # noqa: F401
                  smooth_vdf, yee_advance)

__all__ = ["Simulation", "advance_deposit", "yee_advance", "smooth_setup", "smooth_vdf",
           "device_count", "copy_probe", "OSR_BORIS", "OSR_VAY", "OsrError", "lib"]
