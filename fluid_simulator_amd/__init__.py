"""MI355X-native FLIP/APIC Simulator step (HIP kernels for gfx950 behind a C-ABI).

The compute path is libfsim_hip.so (csrc/); `capi` is its ctypes binding and
`scenes` builds the synthetic benchmark scenes.  The C++ host mirror of the
reference's Simulator/MacGrid/HashedParticles interface lives in cpp/.
"""
from .capi import (APIC, FLIP, PIC, WATER, AIR, SOLID, FsimError, Grid, Simulation, load, make_obstacle,
                   make_params, scene_dam_break)

__all__ = ["APIC", "FLIP", "PIC", "WATER", "AIR", "SOLID", "FsimError", "Grid", "Simulation", "load",
           "make_obstacle", "make_params", "scene_dam_break"]
