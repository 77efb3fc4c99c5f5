"""The MD and pump-probe drivers with HDF5 output: the files written by
libcsl_h5.so hold the GPU results (bit for bit) in the reference writers'
layout, and those results match the oracle to the trajectory bar."""
import os

import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, h5io, seeding
from classical_mcmc_spin_amd import gpu as G
from helpers import kitaev, random_state, unit_x_field

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-10


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


@pytest.fixture(scope="module", autouse=True)
def _h5lib():
    try:
        h5io.lib()
    except h5io.H5Error as e:
        pytest.skip(f"libcsl_h5.so unavailable: {e}")


def test_molecular_dynamics_trajectory_h5(po, tmp_path):
    uc, lat = kitaev(4, 4, h=0.1)
    o = po.OracleLattice(uc, 4, 4, 1)
    s = random_state(lat, 21)
    T0, T1, dt, every = 0.0, 1.005, 0.01, 10
    ref = o.integrate_md(s.reshape(-1), T0, T1, dt, every)
    with GPULattice(lat) as gl:
        res = G.molecular_dynamics(gl, s, T0, T1, dt, save_interval=every, out_dir=str(tmp_path))
    p = os.path.join(str(tmp_path), "trajectory.h5")
    assert np.array_equal(res["times"], ref["times"])
    assert np.max(np.abs(res["mags"] - ref["mags"])) <= TOL_TRAJ
    assert np.array_equal(h5io.read_dataset(p, "/trajectory/times"), res["times"])
    for m, name in enumerate(("antiferro", "local", "global")):
        assert np.array_equal(h5io.read_dataset(p, f"/trajectory/magnetization_{name}"), res["mags"][:, m])
    spins = h5io.read_dataset(p, "/trajectory/spins")
    assert spins.shape == (len(res["times"]), lat.lattice_size, 3)
    assert np.array_equal(spins, res["states"])
    assert np.max(np.abs(spins[-1] - ref["final"].reshape(-1, 3))) <= TOL_TRAJ
    assert h5io.read_attr(p, "/metadata", "save_interval") == every
    assert np.array_equal(h5io.read_dataset(p, "/metadata/positions"), lat.positions)


def test_pump_probe_spectroscopy_h5(po, tmp_path):
    uc, lat = kitaev(6, 6, h=0.1)
    o = po.OracleLattice(uc, 6, 6, 1)
    g = seeding.tilt(seeding.quench(lat, random_state(lat, 3), sweeps=50), 0.01, seed=9)
    f = unit_x_field(lat)
    args = (0.5, 0.5, 0.0, 0.0, 1.0, 0.25, -1.0, 2.005, 0.01)
    ref = o.pump_probe(g.reshape(-1), f, *args, reuse_m0=False, n_outer=1)
    with GPULattice(lat) as gl:
        res = G.pump_probe_spectroscopy(gl, g, f, *args, dir_name=str(tmp_path))
        E = gl.energy() / lat.lattice_size  # handle holds the ground state after the write
    p = os.path.join(str(tmp_path), "pump_probe_spectroscopy.h5")
    for k in ("M0", "M1", "M01"):
        assert np.max(np.abs(res[k] - ref[k])) <= TOL_TRAJ, k
    assert np.array_equal(h5io.read_dataset(p, "/tau_scan/tau_values"), res["tau"])
    assert np.array_equal(h5io.read_dataset(p, "/reference/times"), res["times"])
    for m, name in enumerate(("antiferro", "local", "global")):
        assert np.array_equal(h5io.read_dataset(p, f"/reference/M_{name}"), res["M0"][:, m])
        for t in range(len(res["tau"])):
            assert np.array_equal(h5io.read_dataset(p, f"/tau_scan/tau_{t}/M1_{name}"), res["M1"][t, :, m])
            assert np.array_equal(h5io.read_dataset(p, f"/tau_scan/tau_{t}/M01_{name}"), res["M01"][t, :, m])
    assert h5io.read_attr(p, "/metadata", "ground_state_energy") == E
    m_loc = g.reshape(-1, 3).mean(axis=0)
    assert abs(h5io.read_attr(p, "/metadata", "ground_state_magnetization") - np.linalg.norm(m_loc)) <= 1e-14
    assert np.array_equal(h5io.read_dataset(p, "/metadata/pulse_field_direction"), np.asarray(f, dtype=float))
