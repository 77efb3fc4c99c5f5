"""Trajectories on lattices of different sizes in one persistent launch
(csl_integrate_multi).  Each trajectory's records must match the same
trajectory run alone on its lattice (csl_integrate_batched) to 1e-12 — the
site ownership differs between the two launch shapes, so the sums behind the
magnetisations are taken in another order — and the oracle to 1e-10.  The
second case holds more trajectories than workgroups fit the GPU at once, so
the work split (pieces continued from saved states) runs across lattices."""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, seeding
from classical_mcmc_spin_amd.gpu import integrate_multi
from helpers import kitaev, unit_x_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _drive(gl, lat, amp, width, freq):
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    gl.set_pulse(np.concatenate([fd, fd]), amp, width, freq, 0.0, 0.0)
    return f


def _run_case(po, sizes, counts, n, check_split=False):
    dt = 0.01
    T1 = (n + 0.5) * dt
    lats = [kitaev(L, L, h=0.1 + 0.05 * i) for i, L in enumerate(sizes)]
    B = sum(counts)
    tp = np.zeros((B, 2))
    tp[:, 1] = 0.013 * np.arange(B)
    am = 0.3 + 0.1 * (np.arange(2 * B).reshape(B, 2) % 3)
    inits = [np.stack([seeding.random_spins(lat.lattice_size, 100 * i + k) for k in range(c)])
             for i, ((_, lat), c) in enumerate(zip(lats, counts))]
    gls = [GPULattice(lat) for _, lat in lats]
    try:
        fields = [_drive(gl, lat, 0.5, 0.3 + 0.1 * i, 1.0 + 0.5 * i) for i, (gl, (_, lat)) in enumerate(zip(gls, lats))]
        multi = integrate_multi(gls, inits, tp, am, 0.0, T1, dt, tgrid="boost")
        grid = gls[0].last_grid()
        print(f"lattices {sizes} x {counts}: {grid} workgroups", flush=True)
        if check_split:
            assert grid < B, f"{grid} workgroups hold all {B} trajectories: no split"
        off = 0
        for i, (gl, c) in enumerate(zip(gls, counts)):
            alone = gl.integrate_batched(inits[i], tp[off:off + c], am[off:off + c], 0.0, T1, dt, tgrid="boost")
            assert np.array_equal(multi["times"], alone["times"])
            assert np.max(np.abs(multi["mags"][:, off:off + c] - alone["mags"])) <= 1e-12, i
            assert np.max(np.abs(multi["energy"][:, off:off + c] - alone["energy"])) <= 1e-12, i
            off += c
    finally:
        for gl in gls:
            gl.close()
    # the last trajectory of every lattice against the oracle
    off = 0
    for i, ((uc, lat), c, L) in enumerate(zip(lats, counts, sizes)):
        b = off + c - 1
        o = po.OracleLattice(uc, L, L, 1)
        f = fields[i]
        o.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, 0.3 + 0.1 * i, 1.0 + 0.5 * i)
        ref = o.integrate_md(inits[i][c - 1].reshape(-1), 0.0, T1, dt)
        assert np.max(np.abs(multi["mags"][:, b] - ref["mags"])) <= 1e-10, (i, b)
        off += c


def test_three_sizes_one_launch(po):
    _run_case(po, (8, 12, 6), (3, 2, 4), 200)


def test_split_across_lattices(po):
    """36x36 (130 KB of LDS state: one workgroup per CU) with 12x12: 300
    trajectories over the workgroups that fit."""
    _run_case(po, (36, 12), (200, 100), 120, check_split=True)


def test_rejects_mismatched_inputs():
    uc, lat = kitaev(6, 6, h=0.1)
    with GPULattice(lat) as gl:
        with pytest.raises(ValueError):
            integrate_multi([gl], [], np.zeros((1, 2)))


def test_mixed_coordination_one_launch(po):
    """Kitaev 8x8 (3 bonds per site, driven) with pyrochlore 2x2x2 (6 bonds
    per site, no drive directions set) in one launch: the bond loop runs at
    the larger coordination, each site over its own bond count."""
    from classical_mcmc_spin_amd import workloads
    uc, kl = kitaev(8, 8, h=0.1)
    pl = workloads.pyrochlore_lattice(2)
    Bk, Bp = 3, 3
    ki = np.stack([seeding.random_spins(kl.lattice_size, 40 + k) for k in range(Bk)])
    pi, _ = workloads.pyrochlore_tilted_aiao(pl, Bp, seed=21)
    B = Bk + Bp
    tp = np.zeros((B, 2))
    tp[:, 1] = 0.1 * np.arange(B)
    am = np.full((B, 2), 0.4)
    dt, n = 0.01, 200
    T1 = (n + 0.5) * dt
    with GPULattice(kl) as gk, GPULattice(pl) as gp:
        f = _drive(gk, kl, 0.4, 0.3, 1.0)
        multi = integrate_multi([gk, gp], [ki, pi], tp, am, 0.0, T1, dt, tgrid="boost")
        ak = gk.integrate_batched(ki, tp[:Bk], am[:Bk], 0.0, T1, dt, tgrid="boost")
        ap = gp.integrate_batched(pi, tp[Bk:], am[Bk:], 0.0, T1, dt, tgrid="boost")
    assert np.max(np.abs(multi["mags"][:, :Bk] - ak["mags"])) <= 1e-12
    assert np.max(np.abs(multi["mags"][:, Bk:] - ap["mags"])) <= 1e-12
    assert np.max(np.abs(multi["energy"][:, Bk:] - ap["energy"])) <= 1e-12
    ok = po.OracleLattice(uc, 8, 8, 1)
    b = Bk - 1
    ok.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, 0.3, 1.0)
    ref = ok.integrate_md(ki[b].reshape(-1), 0.0, T1, dt)
    assert np.max(np.abs(multi["mags"][:, b] - ref["mags"])) <= 1e-10
    op = po.OracleLattice(pl.unit_cell, 2, 2, 2)
    for k in range(Bp):
        ref = op.integrate_md(pi[k].reshape(-1), 0.0, T1, dt)
        assert np.max(np.abs(multi["mags"][:, Bk + k] - ref["mags"])) <= 1e-10, k
        assert np.max(np.abs(multi["energy"][:, Bk + k] - ref["energy"])) <= 1e-10, k


def _perturbed(L1, L2, seed):
    """Kitaev lattice with per-bond distinct matrices and per-site fields (no
    translation-invariant tables) and the oracle built from the same CSR."""
    _, lat = kitaev(L1, L2, h=0.1)
    rng = np.random.default_rng(seed)
    J = lat.bi_J + 0.05 * rng.standard_normal(lat.bi_J.shape)
    off = lat.bi_offset.astype(np.int64)
    par = lat.bi_partner.astype(np.int64)
    seen = {}
    for i in range(lat.lattice_size):
        for k in range(off[i], off[i + 1]):
            seen.setdefault((min(i, par[k]), max(i, par[k])), []).append(k)
    for ks in seen.values():
        if len(ks) == 2:
            J[ks[1]] = J[ks[0]].T
    lat.bi_J = J
    lat.field = lat.field + 0.01 * rng.standard_normal(lat.field.shape)
    return lat, (lat.bi_offset, lat.bi_partner, J, lat.field, lat.onsite, lat.frames, lat.afm_signs)


def test_general_tables_one_launch(po):
    """Two lattices on the general (per-site table) path in one launch."""
    shapes = ((5, 4), (4, 3))
    lats = [_perturbed(L1, L2, 5 + i) for i, (L1, L2) in enumerate(shapes)]
    inits = [np.stack([seeding.random_spins(lat.lattice_size, 60 + 10 * i + k) for k in range(2)])
             for i, (lat, _) in enumerate(lats)]
    B = 4
    dt, n = 0.01, 200
    T1 = (n + 0.5) * dt
    tp = np.zeros((B, 2))
    with GPULattice(lats[0][0]) as g0, GPULattice(lats[1][0]) as g1:
        assert not g0.info()["uniform"] and not g1.info()["uniform"]
        multi = integrate_multi([g0, g1], inits, tp, None, 0.0, T1, dt, tgrid="boost")
        alone = [g.integrate_batched(x, tp[:2], None, 0.0, T1, dt, tgrid="boost") for g, x in zip((g0, g1), inits)]
    for i in range(2):
        assert np.max(np.abs(multi["mags"][:, 2 * i:2 * i + 2] - alone[i]["mags"])) <= 1e-12, i
        assert np.max(np.abs(multi["energy"][:, 2 * i:2 * i + 2] - alone[i]["energy"])) <= 1e-12, i
        (L1, L2), (lat, csr) = shapes[i], lats[i]
        o = po.OracleLattice.from_csr(3, 2, (L1, L2, 1), *csr)
        ref = o.integrate_md(inits[i][1].reshape(-1), 0.0, T1, dt)
        assert np.max(np.abs(multi["mags"][:, 2 * i + 1] - ref["mags"])) <= 1e-10, i
        assert np.max(np.abs(multi["energy"][:, 2 * i + 1] - ref["energy"])) <= 1e-10, i
