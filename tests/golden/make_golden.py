"""Regenerate the golden fixtures of tests/golden/ from the CPU oracle.

  python tests/golden/make_golden.py

kitaev_6x6_seed1234.npz — Kitaev honeycomb 6x6x1 (K=-1, G=0.25, G'=-0.02,
h=0.1 along (1,1,1) unnormalised), seeded state (splitmix64 seed 1234):
  init [72,3], H [72,3] (local field), dsdt [72,3], energy,
  md_times/md_mags/md_energy/md_final of 200 RK4 steps (dt=0.01, Boost grid),
  pulse_times/pulse_mags of a damped (alpha=0.02) single pulse drive
  (A=0.5, w=0.5, x-polarised, t_B=0.5) on [-1, 1.005].
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle  # noqa: E402
from classical_mcmc_spin_amd import model, seeding  # noqa: E402


def make():
    uc = model.build_kitaev_honeycomb(field_strength=0.1, field_direction=(1.0, 1.0, 1.0))
    lat = model.Lattice(uc, 6, 6, 1)
    o = pyoracle.OracleLattice(uc, 6, 6, 1)
    s = seeding.random_spins(lat.lattice_size, 1234)
    flat = s.reshape(-1)
    H = np.array([o.local_field(flat, i) for i in range(lat.lattice_size)])
    d = o.rhs(flat).reshape(-1, 3)
    E = o.energy(flat)
    md = o.integrate_md(flat, 0.0, 2.005, 0.01)
    f = np.zeros((2, 3))
    f[:, 0] = 1.0
    o.set_alpha(0.02)
    pt, pm = o.single_pulse_drive(flat, f, 0.5, 0.5, 0.5, 0.0, -1.0, 1.005, 0.01)
    np.savez_compressed(os.path.join(HERE, "kitaev_6x6_seed1234.npz"), init=s, H=H, dsdt=d, energy=E,
                        md_times=md["times"], md_mags=md["mags"], md_energy=md["energy"],
                        md_final=md["final"].reshape(-1, 3), pulse_times=pt, pulse_mags=pm)


if __name__ == "__main__":
    make()
    print("written", os.path.join(HERE, "kitaev_6x6_seed1234.npz"))
