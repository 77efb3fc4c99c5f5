"""GPU parity of the pyrochlore dynamics (6 neighbours per site, 16-site
conventional cell) against the CPU oracle, on both device paths: one
workgroup per trajectory with the lattice in LDS (block = 0) and the
grid-wide stage kernels (block = -1; the path the 65 536-site lattice takes).

Bar: max|dM|, |dE|, |dS| <= 1e-10 over <= 500 RK4 steps.  At the full
65 536-site / 64-sample / 4096-step size: two samples against the oracle over
300 steps, and for the whole run the size-independent properties: every
sample equals the same sample integrated alone (bit for bit), spin lengths
and energies drift by less than 1e-3.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, model, workloads

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _t1(T0, n, dt):
    return T0 + (n + 0.5) * dt


@pytest.mark.parametrize("block", [0, -1])
def test_pyrochlore_small_both_paths(po, block):
    lat = workloads.pyrochlore_lattice(2)
    uc = lat.unit_cell
    B = 3
    st, _ = workloads.pyrochlore_tilted_aiao(lat, B, seed=21)
    T0, dt, n = 0.0, 0.02, 300
    with GPULattice(lat) as gl:
        assert gl.info()["uniform"]
        res = gl.integrate_batched(st, np.zeros((B, 2)), np.zeros((B, 2)), T0, _t1(T0, n, dt), dt,
                                   tgrid="boost", block=block, want_final=True)
    o = po.OracleLattice(uc, 2, 2, 2)
    for b in range(B):
        ref = o.integrate_md(st[b].reshape(-1), T0, _t1(T0, n, dt), dt)
        assert np.array_equal(res["times"], ref["times"])
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL


def test_pyrochlore_full_size_vs_oracle(po):
    lat = workloads.pyrochlore_lattice(16)
    assert lat.lattice_size == 65536
    B = 64
    st, _ = workloads.pyrochlore_tilted_aiao(lat, B, seed=1234)
    T0, dt, n = 0.0, 0.02, 300
    with GPULattice(lat) as gl:
        res = gl.integrate_batched(st, np.zeros((B, 2)), np.zeros((B, 2)), T0, _t1(T0, n, dt), dt,
                                   tgrid="boost", want_final=True)
        assert gl.last_launch_config()["block"] == -1      # too large for one workgroup's LDS
    o = po.OracleLattice(lat.unit_cell, 16, 16, 16)
    for b in (0, B - 1):
        ref = o.integrate_md(st[b].reshape(-1), T0, _t1(T0, n, dt), dt)
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL


def test_pyrochlore_full_workload_properties():
    """64 samples x 4096 steps of dt = 0.02 on 65 536 sites."""
    lat = workloads.pyrochlore_lattice(16)
    B = 64
    st, _ = workloads.pyrochlore_tilted_aiao(lat, B, seed=1234)
    T0, dt, n = 0.0, 0.02, 4096
    with GPULattice(lat) as gl:
        res = gl.integrate_batched(st, np.zeros((B, 2)), np.zeros((B, 2)), T0, _t1(T0, n, dt), dt,
                                   tgrid="boost", want_final=True)
        one = gl.integrate_batched(st[37], np.zeros((1, 2)), np.zeros((1, 2)), T0, _t1(T0, n, dt), dt,
                                   tgrid="boost", want_final=True)
    assert res["mags"].shape == (n + 1, B, 3, 3)
    assert np.all(np.isfinite(res["mags"])) and np.all(np.isfinite(res["final"]))
    assert np.array_equal(one["final"][0], res["final"][37])
    assert np.array_equal(one["mags"][:, 0], res["mags"][:, 37])
    assert np.array_equal(one["energy"][:, 0], res["energy"][:, 37])
    # RK4 at |H| dt ~ 0.13 keeps spin lengths and the energy to well below
    # 1e-3 over 4096 steps; a broken kernel does not
    norms = np.linalg.norm(res["final"], axis=2)
    assert np.max(np.abs(norms - 1.0)) <= 1e-3
    e = res["energy"]
    assert np.max(np.abs(e[-1] - e[0]) / np.abs(e[0])) <= 1e-3
