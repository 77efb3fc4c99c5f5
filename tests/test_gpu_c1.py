"""Plain LL dynamics at the plain-MD workload's size: Kitaev honeycomb 24x24x1
(1152 sites), K = -1, Gamma = 0.25, Gamma' = -0.02, no field, seeded spins
uniform on the sphere, RK4 dt = 0.01, undamped (molecular_dynamics with
method "rk4", lattice_md.cpp:341-453), on every engine of csl_integrate:
0 (auto: the persistent workgroup here), 1 (grid-wide stage kernels),
2 (persistent workgroup), 3 (stage kernels replayed from a hipGraph).

Bars:
* 500 steps: M(t), E(t)/N and the final spins within 1e-10 of the oracle;
* the whole 10^4-step run: the first 500 records within 1e-10 of the oracle,
  and the drift of the conserved quantities (max |E(t) - E(0)| over the run,
  max ||S_i| - 1| at the end) of the same order as the oracle's drift on the
  same run (at most twice it plus 1e-12): the trajectories themselves part
  after a few thousand steps (chaotic dynamics), the conservation does not.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, workloads
from classical_mcmc_spin_amd import gpu as G

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT = 0.01


@pytest.fixture(scope="module")
def c1(oracle):
    uc, lat, s = workloads.kitaev_md(24, 24, seed=1234)
    o = oracle.OracleLattice(uc, 24, 24, 1)
    return uc, lat, s, o


@pytest.fixture(scope="module")
def ref_full(c1):
    uc, lat, s, o = c1
    print("[C1] oracle: 10^4 RK4 steps of 1152 sites", flush=True)
    return o.integrate_md(s.reshape(-1), 0.0, (10000 + 0.5) * DT, DT)


@pytest.mark.parametrize("engine", [0, 1, 2, 3])
def test_c1_500_steps(c1, engine):
    uc, lat, s, o = c1
    print(f"[C1] engine {engine}: 500 steps", flush=True)
    T1 = (500 + 0.5) * DT
    ref = o.integrate_md(s.reshape(-1), 0.0, T1, DT)
    with GPULattice(lat) as gl:
        res = G.molecular_dynamics(gl, s, 0.0, T1, DT, engine=engine)
        fin = gl.get_spins()
        cfg = gl.last_launch_config()
    assert len(ref["times"]) == 501 and np.array_equal(res["times"], ref["times"])
    assert np.max(np.abs(res["mags"] - ref["mags"])) <= TOL
    assert np.max(np.abs(res["energy"] - ref["energy"])) <= TOL
    assert np.max(np.abs(fin - ref["final"].reshape(-1, 3))) <= TOL
    if engine in (0, 2):
        assert cfg["block"] > 0          # one persistent workgroup holds the lattice
    else:
        assert cfg["block"] == -1        # grid-wide stage kernels


def _drift(energy, final):
    return float(np.max(np.abs(energy - energy[0]))), float(np.max(np.abs(np.linalg.norm(final, axis=-1) - 1.0)))


@pytest.mark.parametrize("engine", [0, 1])
def test_c1_full_run_conservation(c1, ref_full, engine):
    uc, lat, s, o = c1
    print(f"[C1] engine {engine}: 10^4 steps", flush=True)
    with GPULattice(lat) as gl:
        res = G.molecular_dynamics(gl, s, 0.0, (10000 + 0.5) * DT, DT, engine=engine)
        fin = gl.get_spins()
    ref = ref_full
    assert len(res["times"]) == 10001 and np.array_equal(res["times"], ref["times"])
    assert np.max(np.abs(res["mags"][:501] - ref["mags"][:501])) <= TOL
    assert np.max(np.abs(res["energy"][:501] - ref["energy"][:501])) <= TOL
    dE, dS = _drift(res["energy"], fin)
    dE_o, dS_o = _drift(ref["energy"], ref["final"].reshape(-1, 3))
    print(f"[C1] drift: E {dE:.3e} (oracle {dE_o:.3e}), |S| {dS:.3e} (oracle {dS_o:.3e})", flush=True)
    assert dE <= 2.0 * dE_o + 1e-12
    assert dS <= 2.0 * dS_o + 1e-12
    # RK4 at dt = 0.01 keeps both far below the trajectories' own scale
    assert dE_o < 1e-6 and dS_o < 1e-6
