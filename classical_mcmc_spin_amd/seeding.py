"""Seeded spin configurations (counter-based, platform independent).

The reference draws initial spins from its Lehmer generator seeded by the
clock (lattice.h:640, simple_linear_alg.h:192-226), so its states are not
reproducible; the spin states used here are generated from splitmix64
counters instead, which gives bit-identical inputs on every machine.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(counter: np.ndarray, seed: int) -> np.ndarray:
    """splitmix64 finaliser of (seed + golden * (counter + 1))."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (counter.astype(np.uint64) + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(n: int, seed: int, stream: int = 0) -> np.ndarray:
    """n doubles in (0, 1) from the top 53 bits of splitmix64."""
    c = np.arange(n, dtype=np.uint64) + np.uint64(stream) * np.uint64(1 << 40)
    z = splitmix64(c, seed)
    return ((z >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(n: int, seed: int, stream: int = 0) -> np.ndarray:
    """n standard normals by Box-Muller on two uniform streams."""
    u1 = uniform01(n, seed, 2 * stream)
    u2 = uniform01(n, seed, 2 * stream + 1)
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def random_spins(n_sites: int, seed: int, spin_dim: int = 3, spin_length: float = 1.0) -> np.ndarray:
    """Spins uniform on the sphere of radius spin_length: normalised Gaussian vectors."""
    g = normal(n_sites * spin_dim, seed).reshape(n_sites, spin_dim)
    return spin_length * g / np.linalg.norm(g, axis=1, keepdims=True)


def tilt(spins: np.ndarray, sigma, seed: int, spin_length: float = 1.0) -> np.ndarray:
    """Add N(0, sigma^2) noise to every component and renormalise."""
    s = np.asarray(spins, dtype=float)
    g = normal(s.size, seed, stream=7).reshape(s.shape)
    t = s + np.asarray(sigma)[..., None] * g if np.ndim(sigma) else s + sigma * g
    return spin_length * t / np.linalg.norm(t, axis=-1, keepdims=True)


def quench(lattice, spins: np.ndarray, sweeps: int = 200, tol: float = 1e-13) -> np.ndarray:
    """Deterministic zero-temperature relaxation: every spin is set antiparallel
    to its local field (Lattice::deterministic_sweep, lattice.h:3334-3349),
    colour class by colour class so the sweep order is fixed."""
    S = np.array(spins, dtype=float).reshape(lattice.lattice_size, lattice.spin_dim)
    color = lattice.coloring()
    classes = [np.nonzero(color == c)[0] for c in range(color.max() + 1)]
    prev = None
    for _ in range(sweeps):
        for idx in classes:
            H = lattice.local_fields_host(S)[idx]
            nrm = np.linalg.norm(H, axis=1, keepdims=True)
            ok = nrm[:, 0] > 1e-15
            S[idx[ok]] = -H[ok] / nrm[ok] * lattice.spin_length
        if prev is not None and np.max(np.abs(S - prev)) < tol:
            break
        prev = S.copy()
    return S
