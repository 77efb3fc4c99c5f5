"""Seeded synthetic inputs of the standard workloads (lattices, couplings and
initial states), shared by the benchmark, the tests and user scripts.

  kitaev_md             Kitaev honeycomb 24x24x1, K=-1, G=0.25, G'=-0.02, no
                        field, spins uniform on the sphere from a seed (plain
                        LL dynamics, RK4 dt = 0.01, 10^4 undamped steps)
  kitaev_2dcs           Kitaev honeycomb 36x36x1, K=-1, G=0.25, G'=-0.02,
                        h = 0.1 along (1,1,1)/sqrt3, relaxed + tilted start,
                        2 pulses along x (A = 0.5, w = 0.5, omega = 0),
                        256 delays tau = linspace(0, 64, 256) -> 513 trajectories
  pyrochlore_dynamics   pyrochlore 16^3 conventional cells (65 536 sites),
                        Jzz=1, J+-=0.05, J+-+-=0.02, Jz+-=-0.01, 64 samples of
                        all-in-all-out + N(0, sigma^2) tilt, sigma ~ U[0.05, 0.5],
                        4096 RK4 steps of dt = 0.02
"""
from __future__ import annotations

import numpy as np

from . import model, seeding

PYRO_COUPLINGS = dict(Jzz=1.0, Jpm=0.05, Jpmpm=0.02, Jzpm=-0.01)


def kitaev_md(L1: int = 24, L2: int = 24, seed: int = 1234):
    """(unit cell, lattice, initial state [N, 3]) of the plain MD workload:
    zero field, spins uniform on the unit sphere from the seed."""
    uc = model.build_kitaev_honeycomb(K=-1.0, Gamma=0.25, Gammap=-0.02, field_strength=0.0)
    lat = model.Lattice(uc, L1, L2, 1)
    return uc, lat, seeding.random_spins(lat.lattice_size, seed)


def kitaev_2dcs(L1: int = 36, L2: int = 36, seed: int = 1234, tilt_seed: int = 4321, sweeps: int = 200):
    """(unit cell, lattice, initial state [N, 3], pulse direction [n_atoms, 3])
    of the 2DCS workload: a deterministic quench of a seeded random state,
    then N(0, 0.01^2) tilt per component, renormalised."""
    uc = model.build_kitaev_honeycomb(K=-1.0, Gamma=0.25, Gammap=-0.02, field_strength=0.1,
                                      field_direction=np.ones(3) / np.sqrt(3.0))
    lat = model.Lattice(uc, L1, L2, 1)
    relaxed = seeding.quench(lat, seeding.random_spins(lat.lattice_size, seed), sweeps=sweeps)
    init = seeding.tilt(relaxed, 0.01, seed=tilt_seed)
    field_in = np.zeros((lat.N_atoms, 3))
    field_in[:, 0] = 1.0
    return uc, lat, init, field_in


def pump_probe_replicas(taus, amp: float):
    """(t_pulse [B, 2], amp [B, 2]) of the 2DCS batch: M0 (pulse at 0),
    M1(tau) (pulse at tau), M01(tau) (pulses at 0 and tau); B = 1 + 2 n_tau."""
    taus = np.asarray(taus, dtype=float)
    n = len(taus)
    tp = np.zeros((1 + 2 * n, 2))
    am = np.zeros((1 + 2 * n, 2))
    am[0, 0] = amp
    tp[1:1 + n, 0] = taus
    am[1:1 + n, 0] = amp
    tp[1 + n:, 1] = taus
    am[1 + n:, :] = amp
    return tp, am


def pyrochlore_lattice(L: int = 16, **couplings) -> model.Lattice:
    cp = dict(PYRO_COUPLINGS)
    cp.update(couplings)
    return model.Lattice(model.build_pyrochlore_conventional(**cp), L, L, L)


def pyrochlore_tilted_aiao(lat: model.Lattice, n_samples: int, seed: int = 1234):
    """All-in-all-out state + independent N(0, sigma_b^2) tilt per sample,
    sigma_b uniform on [0.05, 0.5]; returns (states [B, N, 3], sigma [B])."""
    aiao = model.all_in_all_out(lat)
    sig = 0.05 + 0.45 * seeding.uniform01(n_samples, seed, stream=11)
    out = np.empty((n_samples,) + aiao.shape)
    for b in range(n_samples):
        out[b] = seeding.tilt(aiao, float(sig[b]), seed=seed + 1000 * (b + 1), spin_length=lat.spin_length)
    return out, sig
