"""Pyrochlore builders: conventional 16-site cubic cell with global-frame
bonds (the 16^3 / 65 536-site lattice) and the primitive 4-site cell.

Pins (none of them restates the builder):
  * geometry: 6 nearest neighbours per site at sqrt(2)/4 (minimum image),
    always on another sublattice, 3N bonds, the reverse bond J^T present;
  * the 3x3 local exchange equals the operator form written with complex
    s+ / s- (independent route to the same bilinear form);
  * the conventional cell (neighbours found by a distance search) and the
    primitive cell with the reference's explicit bond list
    (unitcell_builders.cpp:320-333) give the same energy per site and the
    same local field per sublattice for sublattice-uniform states;
  * all-in-all-out energy = 3 Jzz per site (every bond has s^z = 1 on both
    ends in the local frames).
"""
import math

import numpy as np
import pytest

from classical_mcmc_spin_amd import model, seeding

RNG = np.random.default_rng(2024)
COUPLINGS = dict(Jzz=1.0, Jpm=0.05, Jpmpm=0.02, Jzpm=-0.01)


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def test_geometry_and_bond_tables():
    uc = model.build_pyrochlore_conventional(**COUPLINGS)
    assert uc.N_atoms == 16 and len(uc.bilinear) == 48
    L = 2
    lat = model.Lattice(uc, L, L, L)
    N = lat.lattice_size
    assert N == 16 * L ** 3
    off = lat.bi_offset.astype(np.int64)
    par = lat.bi_partner.astype(np.int64)
    assert np.all(np.diff(off) == 6)
    pairs = {}
    for i in range(N):
        for k in range(off[i], off[i + 1]):
            j = par[k]
            d = lat.positions[j] - lat.positions[i]
            d -= L * np.round(d / L)
            assert abs(np.linalg.norm(d) - math.sqrt(2) / 4) < 1e-12
            assert (i % 16) % 4 != (j % 16) % 4
            pairs[(i, j)] = lat.bi_J[k]
    assert len(pairs) == 6 * N
    for (i, j), J in pairs.items():
        assert np.array_equal(pairs[(j, i)], J.T)


def _op_form(si, sj, theta, Jzz, Jpm, Jpmpm, Jzpm):
    g = np.exp(1j * theta)
    z = -np.conj(g)
    spi, smi = si[0] + 1j * si[1], si[0] - 1j * si[1]
    spj, smj = sj[0] + 1j * sj[1], sj[0] - 1j * sj[1]
    e = (Jzz * si[2] * sj[2] - Jpm * (spi * smj + smi * spj)
         + Jpmpm * (g * spi * spj + np.conj(g) * smi * smj)
         + Jzpm * (si[2] * (z * spj + np.conj(z) * smj) + sj[2] * (z * spi + np.conj(z) * smi)))
    assert abs(e.imag) < 1e-14
    return e.real


@pytest.mark.parametrize("theta", [0.0, 2 * math.pi / 3, 4 * math.pi / 3])
def test_local_exchange_operator_form(theta):
    for _ in range(5):
        c = RNG.normal(size=4)
        si, sj = RNG.normal(size=3), RNG.normal(size=3)
        J = model.pyrochlore_local_exchange(theta, *c)
        assert np.allclose(J, J.T)
        assert abs(si @ J @ sj - _op_form(si, sj, theta, *c)) < 1e-13


def test_non_kramer_limit():
    """Jzpm = 0 reduces to J(theta) of build_pyrochlore_non_kramer
    (unitcell_builders.cpp:406-413), entries written out here."""
    Jpm, Jpmpm, Jzz = 0.3, -0.2, 1.1
    for th in (0.0, 2 * math.pi / 3, 4 * math.pi / 3):
        ref = np.array([[-2 * Jpm + 2 * Jpmpm * math.cos(th), -2 * Jpmpm * math.sin(th), 0],
                        [-2 * Jpmpm * math.sin(th), -2 * Jpm - 2 * Jpmpm * math.cos(th), 0],
                        [0, 0, Jzz]])
        assert np.allclose(model.pyrochlore_local_exchange(th, Jzz, Jpm, Jpmpm, 0.0), ref, atol=1e-15)


def test_frames_orthonormal_right_handed():
    for R in model.pyrochlore_frames():
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-15)
        assert abs(np.linalg.det(R) - 1.0) < 1e-14


def _primitive(couplings, h=0.0, d=(0, 0, 1)):
    """Primitive cell with the reference's explicit nearest-neighbour list
    (unitcell_builders.cpp:320-333) and global-frame matrices."""
    uc = model.Pyrochlore(3)
    R = model.pyrochlore_frames()
    theta = {(0, 1): 0.0, (0, 2): 2 * math.pi / 3, (0, 3): 4 * math.pi / 3,
             (1, 2): 4 * math.pi / 3, (1, 3): 2 * math.pi / 3, (2, 3): 0.0}
    inter = {(0, 1): (1, 0, 0), (0, 2): (0, 1, 0), (0, 3): (0, 0, 1),
             (1, 2): (-1, 1, 0), (1, 3): (-1, 0, 1), (2, 3): (0, 1, -1)}
    for o in ("intra", "inter"):
        for (a, b), th in theta.items():
            J = R[a] @ model.pyrochlore_local_exchange(th, **couplings) @ R[b].T
            uc.set_bilinear_interaction(J, a, b, (0, 0, 0) if o == "intra" else inter[(a, b)])
    for a in range(4):
        uc.set_field(h * np.asarray(d, float), a)
        uc.set_sublattice_frame(np.eye(3), a)
    return uc


def test_conventional_matches_primitive(po):
    cp = dict(Jzz=1.0, Jpm=0.13, Jpmpm=-0.07, Jzpm=0.11)
    conv = po.OracleLattice(model.build_pyrochlore_conventional(**cp, field_strength=0.3,
                                                                field_direction=(1, 2, 3)), 1, 1, 1)
    prim = po.OracleLattice(_primitive(cp, 0.3, (1, 2, 3)), 2, 2, 2)
    for seed in range(3):
        sub = seeding.random_spins(4, 50 + seed)
        s_conv = sub[np.arange(conv.N) % 4]
        s_prim = sub[np.arange(prim.N) % 4]
        assert abs(conv.energy(s_conv.reshape(-1)) / conv.N - prim.energy(s_prim.reshape(-1)) / prim.N) < 1e-14
        for a in range(4):
            hc = conv.local_field(s_conv.reshape(-1), a)
            hp = prim.local_field(s_prim.reshape(-1), a)
            assert np.max(np.abs(np.asarray(hc) - np.asarray(hp))) < 1e-14


def test_all_in_all_out_energy(po):
    uc = model.build_pyrochlore_conventional(**COUPLINGS)
    lat = model.Lattice(uc, 2, 2, 2)
    s = model.all_in_all_out(lat)
    o = po.OracleLattice(uc, 2, 2, 2)
    assert abs(o.energy(s.reshape(-1)) / lat.lattice_size - 3 * COUPLINGS["Jzz"]) < 1e-12
    # the AIAO state is stationary in zero field: every local field is
    # parallel to its spin
    H = lat.local_fields_host(s)
    assert np.max(np.abs(np.cross(H, s))) < 1e-14


def test_tilted_aiao_states():
    from classical_mcmc_spin_amd.workloads import pyrochlore_tilted_aiao
    uc = model.build_pyrochlore_conventional(**COUPLINGS)
    lat = model.Lattice(uc, 2, 2, 2)
    st, sig = pyrochlore_tilted_aiao(lat, 4, seed=7)
    assert st.shape == (4, lat.lattice_size, 3)
    assert np.all((sig >= 0.05) & (sig <= 0.5))
    assert np.allclose(np.linalg.norm(st, axis=2), 1.0, atol=1e-14)
    st2, _ = pyrochlore_tilted_aiao(lat, 4, seed=7)
    assert np.array_equal(st, st2)
