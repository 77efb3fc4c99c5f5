"""MI355X-native Landau–Lifshitz spin dynamics (FP64 RK4 on gfx950).

Host-side mirror of the reference's lattice / Hamiltonian / simulator API for
the molecular_dynamics, pump_probe and 2dcs paths, on top of the C ABI of
libcsl_hip.so (include/csl_gpu.h).
"""
from ._capi import CslError, device_count  # noqa: F401
from .model import (UnitCell, HoneyComb, Lattice, build_kitaev_honeycomb,  # noqa: F401
                    kitaev_frame, kitaev_bond_matrices)
from .gpu import (GPULattice, molecular_dynamics, single_pulse_drive,  # noqa: F401
                  double_pulse_drive, pump_probe_spectroscopy, nonlinear_signal)
from . import seeding  # noqa: F401

__version__ = "0.1.0"
