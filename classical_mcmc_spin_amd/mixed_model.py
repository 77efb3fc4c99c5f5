"""Mixed SU(2) + SU(3) Hamiltonians: unit cell, lattice tables and the
TmFeO3-like builder.

Mirrors the reference's mixed model layer:
  * MixedUnitCell (SU(2) cell, SU(3) cell, SU(2)-SU(3) bilinears keyed by
    the SU(2) source atom)                      unitcell.h:240-275
  * MixedLattice: each species expanded like Lattice (forward J at the
    source, J^T at the partner; build_sublattice, mixed_lattice.h:591-700),
    then the mixed bonds: the SU(2) site gets J (3x8), the SU(3) partner
    gets J^T (8x3), visited SU(2) site by SU(2) site
                                                mixed_lattice.h:745-798
  * flat state [N2*3 | N3*8]                     mixed_lattice.h:3878-3900
  * TmFeO3 Fe / Tm cells                         unitcell.h:359-385
    Fe-Fe bond geometry                          unitcell_builders.cpp:594-633
    Tm crystal-field field                       unitcell_builders.cpp:680-712
    Fe-Tm bond pairs                             unitcell_builders.cpp:794-818
"""
from __future__ import annotations

import math
from typing import List

import numpy as np

from . import seeding
from .model import Bond, Lattice, UnitCell, _wrap


class MixedUnitCell:
    """SU(2) cell + SU(3) cell + SU(2)-SU(3) bilinears (J is 3 x 8)."""

    def __init__(self, su2: UnitCell, su3: UnitCell):
        self.SU2_cell = su2
        self.SU3_cell = su3
        self.bilinear_SU2_SU3: List[Bond] = []

    def set_mixed_bilinear(self, J, source: int, partner: int, offset):
        J = np.asarray(J, dtype=float)
        if J.shape != (self.SU2_cell.N, self.SU3_cell.N):
            raise ValueError("mixed bilinear must be spin_dim_SU2 x spin_dim_SU3")
        if not (0 <= source < self.SU2_cell.N_atoms and 0 <= partner < self.SU3_cell.N_atoms):
            raise IndexError("atom index out of range")
        self.bilinear_SU2_SU3.append(Bond(J.copy(), int(source), int(partner), tuple(int(o) for o in offset)))


class MixedLattice:
    """L1 x L2 x L3 expansion of a MixedUnitCell.

    Attributes: su2, su3 (Lattice objects of the two species with their own
    bond tables), and the mixed CSR tables
      m2_offset [N2+1], m2_partner (SU(3) site), m2_J [nnz, 3, 8]
      m3_offset [N3+1], m3_partner (SU(2) site), m3_J [nnz, 8, 3]
    in the reference's per-site order."""

    def __init__(self, muc: MixedUnitCell, dim1: int, dim2: int, dim3: int):
        self.unit_cell = muc
        self.dims = (int(dim1), int(dim2), int(dim3))
        self.su2 = Lattice(muc.SU2_cell, dim1, dim2, dim3)
        self.su3 = Lattice(muc.SU3_cell, dim1, dim2, dim3)
        A2, A3 = muc.SU2_cell.N_atoms, muc.SU3_cell.N_atoms
        D2, D3 = muc.SU2_cell.N, muc.SU3_cell.N
        N2, N3 = self.su2.lattice_size, self.su3.lattice_size
        d1, d2, d3 = self.dims
        ncell = d1 * d2 * d3
        for b in muc.bilinear_SU2_SU3:
            for o, L in zip(b.offset, self.dims):
                if not -L <= o <= L:
                    raise ValueError(f"mixed bond offset {b.offset} exceeds the lattice dimensions {self.dims}")
        cells = np.arange(ncell)
        ci = cells // (d2 * d3)
        cj = (cells // d3) % d2
        ck = cells % d3
        per_atom = [[b for b in muc.bilinear_SU2_SU3 if b.source == a] for a in range(A2)]
        nb_atom = np.array([len(p) for p in per_atom])
        epc = int(nb_atom.sum())
        prefix = np.concatenate([[0], np.cumsum(nb_atom)[:-1]]) if A2 else np.zeros(0, int)
        src2, par3, ev, J23 = [], [], [], []
        for a in range(A2):
            for r, b in enumerate(per_atom[a]):
                di, dj, dk = b.offset
                pi, pj, pk = _wrap(ci + di, d1), _wrap(cj + dj, d2), _wrap(ck + dk, d3)
                src2.append(cells * A2 + a)
                par3.append(((pi * d2 + pj) * d3 + pk) * A3 + b.partner)
                ev.append(cells * epc + prefix[a] + r)
                J23.append(np.broadcast_to(b.J, (ncell, D2, D3)))
        if src2:
            s2 = np.concatenate(src2)
            p3 = np.concatenate(par3)
            e = np.concatenate(ev)
            J = np.concatenate(J23)
            o2 = np.lexsort((e, s2))
            o3 = np.lexsort((e, p3))
            c2 = np.bincount(s2, minlength=N2)
            c3 = np.bincount(p3, minlength=N3)
            self.m2_partner = p3[o2].astype(np.uintp)
            self.m2_J = np.ascontiguousarray(J[o2])
            self.m3_partner = s2[o3].astype(np.uintp)
            self.m3_J = np.ascontiguousarray(np.transpose(J[o3], (0, 2, 1)))
        else:
            c2 = np.zeros(N2, int)
            c3 = np.zeros(N3, int)
            self.m2_partner = np.zeros(0, np.uintp)
            self.m2_J = np.zeros((0, D2, D3))
            self.m3_partner = np.zeros(0, np.uintp)
            self.m3_J = np.zeros((0, D3, D2))
        self.m2_offset = np.concatenate([[0], np.cumsum(c2)]).astype(np.uintp)
        self.m3_offset = np.concatenate([[0], np.cumsum(c3)]).astype(np.uintp)

    @property
    def N2(self):
        return self.su2.lattice_size

    @property
    def N3(self):
        return self.su3.lattice_size

    def flat_state(self, s2, s3) -> np.ndarray:
        """[N2*3 | N3*8] (spins_to_state, mixed_lattice.h:3878-3900)."""
        return np.concatenate([np.asarray(s2, float).reshape(-1), np.asarray(s3, float).reshape(-1)])

    def split_state(self, x):
        x = np.asarray(x, float).reshape(-1)
        n2 = self.N2 * self.su2.spin_dim
        return x[:n2].reshape(self.N2, -1), x[n2:].reshape(self.N3, -1)


# --------------------------------------------------------------------------
# SU(3) helpers

def gell_mann() -> np.ndarray:
    """The 8 Gell-Mann matrices lambda_1..lambda_8 (complex 3x3)."""
    l = np.zeros((8, 3, 3), dtype=complex)
    l[0][0, 1] = l[0][1, 0] = 1
    l[1][0, 1], l[1][1, 0] = -1j, 1j
    l[2][0, 0], l[2][1, 1] = 1, -1
    l[3][0, 2] = l[3][2, 0] = 1
    l[4][0, 2], l[4][2, 0] = -1j, 1j
    l[5][1, 2] = l[5][2, 1] = 1
    l[6][1, 2], l[6][2, 1] = -1j, 1j
    l[7] = np.diag([1, 1, -2]) / math.sqrt(3.0)
    return l


def random_su3_states(n: int, seed: int) -> np.ndarray:
    """n unit 8-vectors along the Gell-Mann expectation values
    lambda_a = <psi| l_a |psi> of seeded random qutrit states psi."""
    g = seeding.normal(6 * n, seed).reshape(n, 2, 3)
    psi = g[:, 0] + 1j * g[:, 1]
    psi /= np.linalg.norm(psi, axis=1, keepdims=True)
    lm = gell_mann()
    v = np.real(np.einsum("ni,aij,nj->na", psi.conj(), lm, psi))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


# --------------------------------------------------------------------------
# TmFeO3-like Pbnm cell

_FE_POS = [[0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0, 0], [0, 0.5, 0]]
_TM_POS = [[0.02111, 0.92839, 0.75], [0.52111, 0.57161, 0.25],
           [0.47889, 0.42839, 0.75], [0.97889, 0.07161, 0.25]]
# 16 inversion-related Fe-Tm bond pairs: (Fe atom, Tm atom, Tm cell - Fe cell)
# of the even and the odd member (unitcell_builders.cpp:794-818)
_FE_TM_PAIRS = [
    ((0, 3, (-1, 0, 0)), (0, 0, (0, 0, 0))), ((0, 2, (0, 0, 0)), (0, 1, (-1, 0, 0))),
    ((0, 1, (0, 0, 0)), (0, 2, (-1, 0, 0))), ((0, 0, (0, -1, 0)), (0, 3, (-1, 1, 0))),
    ((1, 2, (0, 0, 0)), (1, 1, (0, -1, 0))), ((1, 3, (0, 0, 0)), (1, 0, (0, -1, 0))),
    ((1, 0, (1, -1, 0)), (1, 3, (-1, 0, 0))), ((1, 1, (0, 0, 0)), (1, 2, (0, -1, 0))),
    ((2, 1, (0, -1, 0)), (2, 2, (0, 0, -1))), ((2, 0, (0, -1, -1)), (2, 3, (0, 0, 0))),
    ((2, 3, (-1, 0, 0)), (2, 0, (1, -1, -1))), ((2, 2, (0, -1, -1)), (2, 1, (0, 0, 0))),
    ((3, 0, (0, 0, -1)), (3, 3, (-1, 0, 0))), ((3, 1, (-1, 0, 0)), (3, 2, (0, 0, -1))),
    ((3, 2, (-1, 0, -1)), (3, 1, (0, 0, 0))), ((3, 3, (-1, 1, 0)), (3, 0, (0, -1, -1))),
]


def _j_dm(J, Dy, Dz):
    """J delta_ab + eps_abc D_c with D = (0, Dy, Dz) (unitcell_builders.cpp:570-576)."""
    return np.array([[J, Dz, -Dy], [-Dz, J, 0.0], [Dy, 0.0, J]])


def build_tmfeo3_mixed(J=1.0, D=0.05, K=0.01, fe_tm_sigma=0.02, seed=2025,
                       e1=2.067834, e2=4.9628, field_strength=0.0,
                       field_direction=(0.0, 0.0, 1.0)) -> MixedUnitCell:
    """TmFeO3-like Pbnm mixed cell: 4 Fe SU(2) + 4 Tm SU(3) sites.

    Fe-Fe: the reference's nearest-neighbour geometry (in-plane J1 bonds
    with the DM vector D_y = +D / -D on the z = 1/2 / z = 0 planes, J1 along
    c), all with Heisenberg J; single-ion anisotropy diag(-K, 0, -K); AFM
    signs (+,-,+,-).  Tm: crystal-field field B_3 = e1, B_8 = (2 e2 - e1)/sqrt3.
    Fe-Tm: the 32 bonds of the 16 inversion-related pairs, each with its
    own 3x8 matrix of N(0, fe_tm_sigma^2) entries from `seed`."""
    fe = UnitCell(3, 4, _FE_POS, np.eye(3))
    tm = UnitCell(8, 4, _TM_POS, np.eye(3))
    Ja, Ja23, Jc = _j_dm(J, D, 0.0), _j_dm(J, -D, 0.0), J * np.eye(3)
    for o in ((0, 0, 0), (1, -1, 0), (0, -1, 0), (1, 0, 0)):
        fe.set_bilinear_interaction(Ja, 1, 0, o)
    for o in ((0, 0, 0), (1, -1, 0), (0, -1, 0), (1, 0, 0)):
        fe.set_bilinear_interaction(Ja23, 2, 3, o)
    for (s, p) in ((0, 3), (1, 2)):
        fe.set_bilinear_interaction(Jc, s, p, (0, 0, 0))
        fe.set_bilinear_interaction(Jc, s, p, (0, 0, 1))
    Kmat = np.diag([-K, 0.0, -K])
    h = field_strength * np.asarray(field_direction, float)
    for a in range(4):
        fe.set_onsite_interaction(Kmat, a)
        fe.set_field(h, a)
    fe.set_afm_sublattice_signs(np.array([1.0, -1.0, 1.0, -1.0]))
    cef = np.zeros(8)
    cef[2] = e1
    cef[7] = (2.0 * e2 - e1) / math.sqrt(3.0)
    for a in range(4):
        tm.set_field(cef, a)
    muc = MixedUnitCell(fe, tm)
    g = seeding.normal(32 * 24, seed, stream=5).reshape(32, 3, 8) * fe_tm_sigma
    q = 0
    for even, odd in _FE_TM_PAIRS:
        for (f, t, o) in (even, odd):
            muc.set_mixed_bilinear(g[q], f, t, o)
            q += 1
    return muc


def scale_fe_tm(muc: MixedUnitCell, scale: float) -> MixedUnitCell:
    """Copy of a mixed cell with every SU(2)-SU(3) matrix multiplied by scale."""
    out = MixedUnitCell(muc.SU2_cell, muc.SU3_cell)
    for b in muc.bilinear_SU2_SU3:
        out.set_mixed_bilinear(scale * b.J, b.source, b.partner, b.offset)
    return out
