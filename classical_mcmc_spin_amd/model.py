"""Hamiltonian description: unit cells, builders and the lattice bond tables.

Mirrors the reference's model layer:
  * UnitCell / HoneyComb / Pyrochlore        include/classical_spin/core/unitcell.h:102-356
  * build_kitaev_honeycomb                   src/core/unitcell_builders.cpp:178-283
  * Lattice(uc, d1, d2, d3) bond expansion   include/classical_spin/lattice/lattice.h:601-855
    and its CSR mirror                       lattice.h:868-903

The expansion reproduces the reference's per-site bond ORDER (sites in
flatten order, unit-cell bonds of the source atom in insertion order, the
forward bond J appended to the source and the reverse bond J^T appended to the
partner at the moment the bond is visited), so that local-field sums run in
the reference's order.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field as dc_field
from typing import List, Sequence

import numpy as np


@dataclass
class Bond:
    J: np.ndarray        # D x D
    source: int
    partner: int
    offset: tuple        # (di, dj, dk) cell offset source -> partner


class UnitCell:
    """Unit cell with per-atom field, onsite matrix, frame and AFM sign
    (unitcell.h:102-230).  Bonds are kept in insertion order."""

    def __init__(self, spin_dim: int, n_atoms: int, positions=None, lattice_vectors=None):
        self.N = int(spin_dim)
        self.N_atoms = int(n_atoms)
        D = self.N
        self.lattice_pos = (np.zeros((n_atoms, 3)) if positions is None
                            else np.asarray(positions, dtype=float).reshape(n_atoms, 3))
        self.lattice_vectors = (np.eye(3) if lattice_vectors is None
                                else np.asarray(lattice_vectors, dtype=float).reshape(3, 3))
        self.field = np.zeros((n_atoms, D))
        self.onsite_interaction = np.zeros((n_atoms, D, D))
        self.sublattice_frames = np.tile(np.eye(D), (n_atoms, 1, 1))
        self.afm_sublattice_signs = np.ones(n_atoms)
        self.bilinear: List[Bond] = []

    def set_field(self, f, index: int):
        f = np.asarray(f, dtype=float)
        if f.shape != (self.N,):
            raise ValueError("Field dimension mismatch")
        self.field[index] = f

    def set_bilinear_interaction(self, J, source: int, partner: int, offset):
        J = np.asarray(J, dtype=float)
        if J.shape != (self.N, self.N):
            raise ValueError("Bilinear matrix dimension mismatch")
        if not (0 <= source < self.N_atoms and 0 <= partner < self.N_atoms):
            raise IndexError("Atom index out of range")
        self.bilinear.append(Bond(J.copy(), int(source), int(partner), tuple(int(o) for o in offset)))

    def set_onsite_interaction(self, A, index: int):
        A = np.asarray(A, dtype=float)
        if A.shape != (self.N, self.N):
            raise ValueError("Onsite matrix dimension mismatch")
        self.onsite_interaction[index] = A

    def set_sublattice_frame(self, R, index: int):
        R = np.asarray(R, dtype=float)
        if R.shape != (self.N, self.N):
            raise ValueError("Frame matrix dimension mismatch")
        self.sublattice_frames[index] = R

    def set_afm_sublattice_signs(self, signs):
        signs = np.asarray(signs, dtype=float)
        if signs.shape != (self.N_atoms,):
            raise ValueError("AFM signs size must match N_atoms")
        self.afm_sublattice_signs = signs.copy()


class HoneyComb(UnitCell):
    """Two-site honeycomb cell (unitcell.h:283-289)."""

    def __init__(self, spin_dim: int = 3):
        super().__init__(spin_dim, 2,
                         [[0, 0, 0], [0, 1 / math.sqrt(3.0), 0]],
                         [[1, 0, 0], [0.5, math.sqrt(3.0) / 2, 0], [0, 0, 1]])


def kitaev_frame() -> np.ndarray:
    """Local Kitaev frame, columns x'=(1,1,-2)/sqrt6, y'=(-1,1,0)/sqrt2,
    z'=(1,1,1)/sqrt3 (unitcell_builders.cpp:264-267, kitaev_bonds.h:35-41)."""
    s6, s2, s3 = math.sqrt(6.0), math.sqrt(2.0), math.sqrt(3.0)
    return np.array([[1.0 / s6, -1.0 / s2, 1.0 / s3],
                     [1.0 / s6, 1.0 / s2, 1.0 / s3],
                     [-2.0 / s6, 0.0, 1.0 / s3]])


def kitaev_bond_matrices(K: float, Gamma: float, Gammap: float, J: float = 0.0):
    """(Jx, Jy, Jz) of the K-Gamma-Gamma' model (unitcell_builders.cpp:197-213)."""
    Jx = np.array([[J + K, Gammap, Gammap], [Gammap, J, Gamma], [Gammap, Gamma, J]])
    Jy = np.array([[J, Gammap, Gamma], [Gammap, J + K, Gammap], [Gamma, Gammap, J]])
    Jz = np.array([[J, Gamma, Gammap], [Gamma, J, Gammap], [Gammap, Gammap, J + K]])
    return Jx, Jy, Jz


def build_kitaev_honeycomb(K=-1.0, Gamma=0.25, Gammap=-0.02, J=0.0, field_strength=0.0,
                           field_direction=(0.0, 1.0, 0.0), g_factor=(1.0, 1.0, 1.0),
                           J2_A=0.0, J2_B=0.0, J3=0.0, J_perp=0.0, Ka=0.0, Kb=0.0, Kc=0.0) -> UnitCell:
    """build_kitaev_honeycomb (unitcell_builders.cpp:178-283).  Defaults are the
    SpinConfig defaults (spin_config.h:257-259); the field direction is used as
    given, without normalisation (:274-280)."""
    uc = HoneyComb(3)
    Jx, Jy, Jz = kitaev_bond_matrices(K, Gamma, Gammap, J)
    uc.set_bilinear_interaction(Jx, 0, 1, (0, -1, 0))
    uc.set_bilinear_interaction(Jy, 0, 1, (1, -1, 0))
    uc.set_bilinear_interaction(Jz, 0, 1, (0, 0, 0))
    if abs(J2_A) > 1e-12 or abs(J2_B) > 1e-12:
        for o in ((1, 0, 0), (0, 1, 0), (1, -1, 0)):
            uc.set_bilinear_interaction(J2_A * np.eye(3), 0, 0, o)
        for o in ((1, 0, 0), (0, 1, 0), (1, -1, 0)):
            uc.set_bilinear_interaction(J2_B * np.eye(3), 1, 1, o)
    if abs(J3) > 1e-12:
        for o in ((1, -2, 0), (-1, 0, 0), (1, 0, 0)):
            uc.set_bilinear_interaction(J3 * np.eye(3), 0, 1, o)
    if abs(J_perp) > 1e-12:
        uc.set_bilinear_interaction(J_perp * np.eye(3), 0, 0, (0, 0, 1))
        uc.set_bilinear_interaction(J_perp * np.eye(3), 1, 1, (0, 0, 1))
    if abs(Ka) > 1e-12 or abs(Kb) > 1e-12 or abs(Kc) > 1e-12:
        Km = np.diag([Ka, Kb, Kc])
        uc.set_onsite_interaction(Km, 0)
        uc.set_onsite_interaction(Km, 1)
    R = kitaev_frame()
    uc.set_sublattice_frame(R, 0)
    uc.set_sublattice_frame(R, 1)
    g = np.asarray(g_factor, dtype=float)
    d = np.asarray(field_direction, dtype=float)
    f = np.array([g[0] * field_strength * d[0], g[1] * field_strength * d[1], g[2] * field_strength * d[2]])
    uc.set_field(f, 0)
    uc.set_field(f, 1)
    return uc


_PYRO_POS = np.array([[0.125, 0.125, 0.125], [0.125, -0.125, -0.125],
                      [-0.125, 0.125, -0.125], [-0.125, -0.125, 0.125]])
_FCC = np.array([[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0]])
# bond angle theta of each sublattice pair: the "Jz" (0), "Jx" (2pi/3) and
# "Jy" (4pi/3) bond types of build_pyrochlore_non_kramer
# (unitcell_builders.cpp:420-433)
_PYRO_THETA = {(0, 1): 0.0, (0, 2): 2 * math.pi / 3, (0, 3): 4 * math.pi / 3,
               (1, 2): 4 * math.pi / 3, (1, 3): 2 * math.pi / 3, (2, 3): 0.0}


def pyrochlore_frames() -> np.ndarray:
    """Local frames of the four pyrochlore sublattices, columns (x_a, y_a, z_a)
    with z_a along the local <111> axis (unitcell.h:322-351)."""
    s2, s3, s6 = math.sqrt(2.0), math.sqrt(3.0), math.sqrt(6.0)
    z = np.array([[1, 1, 1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]) / s3
    y = np.array([[0, -1, 1], [0, 1, -1], [0, -1, -1], [0, 1, 1]]) / s2
    x = np.array([[-2, 1, 1], [-2, -1, -1], [2, 1, -1], [2, -1, 1]]) / s6
    return np.stack([np.stack([x[a], y[a], z[a]], axis=1) for a in range(4)])


class Pyrochlore(UnitCell):
    """Primitive 4-site FCC pyrochlore cell with its local frames
    (unitcell.h:312-356)."""

    def __init__(self, spin_dim: int = 3):
        super().__init__(spin_dim, 4, _PYRO_POS, [[0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
        if spin_dim == 3:
            for a, R in enumerate(pyrochlore_frames()):
                self.set_sublattice_frame(R, a)


def pyrochlore_local_exchange(theta: float, Jzz: float, Jpm: float, Jpmpm: float,
                              Jzpm: float = 0.0) -> np.ndarray:
    """Nearest-neighbour exchange in the two sites' local frames,
        Jzz s^z s^z - Jpm (s+ s- + h.c.) + Jpmpm (g s+ s+ + h.c.)
          + Jzpm (s^z (z s+ + z* s-) + i<->j),   g = exp(i theta), z = -g*,
    written as a 3x3 matrix on (s^x, s^y, s^z).  With Jzpm = 0 it is the
    J(theta) of build_pyrochlore_non_kramer (unitcell_builders.cpp:406-413);
    the Jzpm column/row is this build's extension for the Jzz/J+-/J+-+-/Jz+-
    model."""
    c, s = math.cos(theta), math.sin(theta)
    return np.array([[-2 * Jpm + 2 * Jpmpm * c, -2 * Jpmpm * s, -2 * Jzpm * c],
                     [-2 * Jpmpm * s, -2 * Jpm - 2 * Jpmpm * c, -2 * Jzpm * s],
                     [-2 * Jzpm * c, -2 * Jzpm * s, Jzz]])


def build_pyrochlore_conventional(Jzz=1.0, Jpm=0.05, Jpmpm=0.02, Jzpm=-0.01,
                                  field_strength=0.0, field_direction=(0.0, 0.0, 1.0)) -> UnitCell:
    """16-site conventional cubic pyrochlore cell (lattice constant 1) with
    global-frame nearest-neighbour bonds.

    Sites are the 4 FCC translations x the 4 tetrahedron sublattices of
    unitcell.h:312-320 (atom = 4*fcc + sublattice); each site has 6 nearest
    neighbours at distance sqrt(2)/4.  Spins are stored in the global frame,
    so a bond a->b carries J_ab = R_a J_loc(theta_ab) R_b^T with R the
    sublattice frames and J_loc from `pyrochlore_local_exchange`; frames of
    the cell are the identity (observables are global-frame sums).  The
    reference only builds the primitive 4-site cell; an L^3 supercell of this
    cell is the cubic L x L x L lattice of 16 L^3 sites.  The field is
    h*direction on every site (direction used as given)."""
    R = pyrochlore_frames()
    pos = np.array([_FCC[f] + _PYRO_POS[a] for f in range(4) for a in range(4)])
    uc = UnitCell(3, 16, pos, np.eye(3))
    dnn = math.sqrt(2.0) / 4.0
    offsets = [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)]
    for s in range(16):
        for t in range(s + 1, 16):
            a, b = s % 4, t % 4
            for o in offsets:
                d = pos[t] + np.asarray(o, float) - pos[s]
                if abs(np.linalg.norm(d) - dnn) > 1e-9:
                    continue
                key = (min(a, b), max(a, b))
                Jl = pyrochlore_local_exchange(_PYRO_THETA[key], Jzz, Jpm, Jpmpm, Jzpm)
                uc.set_bilinear_interaction(R[a] @ Jl @ R[b].T, s, t, o)
    d = np.asarray(field_direction, dtype=float)
    for s in range(16):
        uc.set_field(field_strength * d, s)
    return uc


def all_in_all_out(lat: "Lattice") -> np.ndarray:
    """All-in-all-out state of a pyrochlore lattice built from
    `build_pyrochlore_conventional`: every spin along its sublattice's local
    z axis (pointing out of the up tetrahedra)."""
    z = pyrochlore_frames()[:, :, 2]
    sub = (np.arange(lat.lattice_size) % lat.N_atoms) % 4
    return z[sub] * lat.spin_length


class Lattice:
    """Periodic L1 x L2 x L3 expansion of a unit cell with the reference's
    flat bond tables (lattice.h:601-903).

    Attributes (numpy):
      field [N, D], onsite [N, D, D], bi_offset [N+1], bi_partner [nnz],
      bi_J [nnz, D, D] (reference CSR order), frames [n_atoms, D, D],
      afm_signs [n_atoms], positions [N, 3], spins [N, D] (initial state).
    """

    def __init__(self, uc: UnitCell, dim1: int, dim2: int, dim3: int, spin_length: float = 1.0):
        self.unit_cell = uc
        self.spin_dim = D = uc.N
        self.N_atoms = A = uc.N_atoms
        self.dim1, self.dim2, self.dim3 = int(dim1), int(dim2), int(dim3)
        self.spin_length = float(spin_length)
        ncell = self.dim1 * self.dim2 * self.dim3
        self.lattice_size = N = ncell * A
        cells = np.arange(ncell)
        ci = cells // (self.dim2 * self.dim3)
        cj = (cells // self.dim3) % self.dim2
        ck = cells % self.dim3
        site = np.arange(N)
        atom = site % A
        self.field = uc.field[atom].copy()
        self.onsite = uc.onsite_interaction[atom].copy()
        self.frames = uc.sublattice_frames.copy()
        self.afm_signs = uc.afm_sublattice_signs.copy()
        cell_of = site // A
        self.positions = (uc.lattice_pos[atom] + ci[cell_of, None] * uc.lattice_vectors[0]
                          + cj[cell_of, None] * uc.lattice_vectors[1]
                          + ck[cell_of, None] * uc.lattice_vectors[2])

        # Event list in the constructor's visiting order (lattice.h:744-807):
        # site s = flatten(i,j,k,atom) ascending, bonds of atom in insertion order.
        # The reference wraps a partner coordinate once (lattice.h:750-756), so
        # an offset outside [-L, L] would index outside the lattice there.
        for b in uc.bilinear:
            for o, L in zip(b.offset, (self.dim1, self.dim2, self.dim3)):
                if not -L <= o <= L:
                    raise ValueError(f"bond offset {b.offset} exceeds the lattice dimensions "
                                     f"{(self.dim1, self.dim2, self.dim3)}")
        per_atom = [[b for b in uc.bilinear if b.source == a] for a in range(A)]
        owners, keys, partners, mats = [], [], [], []
        ev = 0
        nb_atom = np.array([len(per_atom[a]) for a in range(A)])
        # vectorised over cells: for each (atom, bond-rank) the event index of
        # site (cell, atom) is  prefix(cell, atom) + rank
        events_per_cell = int(nb_atom.sum())
        atom_prefix = np.concatenate([[0], np.cumsum(nb_atom)[:-1]])
        for a in range(A):
            for r, b in enumerate(per_atom[a]):
                di, dj, dk = b.offset
                pi = _wrap(ci + di, self.dim1)
                pj = _wrap(cj + dj, self.dim2)
                pk = _wrap(ck + dk, self.dim3)
                src = cells * A + a
                par = ((pi * self.dim2 + pj) * self.dim3 + pk) * A + b.partner
                e = cells * events_per_cell + atom_prefix[a] + r
                owners += [src, par]
                keys += [2 * e, 2 * e + 1]
                partners += [par, src]
                mats += [np.broadcast_to(b.J, (ncell, D, D)), np.broadcast_to(b.J.T, (ncell, D, D))]
                ev += 1
        if owners:
            owner = np.concatenate(owners)
            key = np.concatenate(keys)
            partner = np.concatenate(partners)
            Jall = np.concatenate(mats)
            order = np.lexsort((key, owner))
            owner, partner, Jall = owner[order], partner[order], Jall[order]
            counts = np.bincount(owner, minlength=N)
        else:
            partner = np.zeros(0, dtype=np.int64)
            Jall = np.zeros((0, D, D))
            counts = np.zeros(N, dtype=np.int64)
        self.bi_offset = np.concatenate([[0], np.cumsum(counts)]).astype(np.uintp)
        self.bi_partner = partner.astype(np.uintp)
        self.bi_J = np.ascontiguousarray(Jall, dtype=np.float64)
        self.num_bi = int(counts.max()) if N else 0
        self.spins = np.zeros((N, D))

    def flatten_index(self, i, j, k, atom):
        """lattice.h:1057-1059"""
        return ((i * self.dim2 + j) * self.dim3 + k) * self.N_atoms + atom

    # --- reference padded layout (Lattice::ensure_gpu_data_initialized, lattice.h:5440-5479)
    def padded_tables(self):
        N, D, nb = self.lattice_size, self.spin_dim, self.num_bi
        bil = np.zeros((N, nb, D, D))
        par = np.zeros((N, nb), dtype=np.uintp)
        cnt = np.diff(self.bi_offset.astype(np.int64)).astype(np.uintp)
        for n in range(nb):
            has = cnt > n
            idx = self.bi_offset[:-1].astype(np.int64)[has] + n
            bil[has, n] = self.bi_J[idx]
            par[has, n] = self.bi_partner[idx]
        return bil, par, cnt

    # --- spin configuration text format (lattice.h:4687-4743)
    def save_spin_config(self, path: str, spins=None):
        s = self.spins if spins is None else np.asarray(spins).reshape(self.lattice_size, self.spin_dim)
        with open(path, "w") as f:
            for row in s:
                f.write(" ".join(f"{v:.16e}" for v in row) + "\n")

    def load_spin_config(self, path: str):
        rows = []
        with open(path) as f:
            for line in f:
                if len(rows) >= self.lattice_size:
                    break
                vals = line.split()
                if len(vals) < self.spin_dim:
                    raise ValueError(f"Incomplete spin data at line {len(rows)}")
                rows.append([float(v) for v in vals[: self.spin_dim]])
        if len(rows) != self.lattice_size:
            raise ValueError(f"File contained {len(rows)} spins, expected {self.lattice_size}")
        self.spins = np.array(rows)
        return self.spins

    # --- host-side helpers that are not on the time-stepping path
    def local_fields_host(self, spins) -> np.ndarray:
        """Local field H (no drive) of every site, vectorised numpy; used to
        prepare relaxed initial states (not a replacement for the device RHS)."""
        S = np.asarray(spins, dtype=float).reshape(self.lattice_size, self.spin_dim)
        H = -self.field + 2.0 * np.einsum("nde,ne->nd", self.onsite, S)
        owner = np.repeat(np.arange(self.lattice_size), np.diff(self.bi_offset.astype(np.int64)))
        contrib = np.einsum("kde,ke->kd", self.bi_J, S[self.bi_partner.astype(np.int64)])
        np.add.at(H, owner, contrib)
        return H

    def coloring(self) -> np.ndarray:
        """Greedy colouring of the bond graph (no two bonded sites share a colour)."""
        N = self.lattice_size
        color = -np.ones(N, dtype=np.int64)
        off = self.bi_offset.astype(np.int64)
        par = self.bi_partner.astype(np.int64)
        for i in range(N):
            used = set(color[par[off[i]:off[i + 1]]].tolist())
            c = 0
            while c in used:
                c += 1
            color[i] = c
        return color


def _wrap(c, n):
    """periodic_boundary (lattice.h:1064-1071): one wrap of an offset."""
    c = np.asarray(c)
    return np.where(c < 0, c + n, np.where(c >= n, c - n, c))
