"""GPU parity of the persistent batched kernel of the mixed SU(2) + SU(3)
lattice (k_mixed_persist: one workgroup integrates one trajectory over the
whole time grid, stage values in LDS plus two SU(3) components in global
memory, x in the trajectory's state slot).

Bar: records and final states against the mixed oracle <= 1e-10 over <= 300
steps with every term of the Hamiltonian, Gilbert and Bloch damping and both
pulses on; against the grid-wide stage kernels on the same batch (same
arithmetic per site, records summed in another order) <= 1e-12; a trajectory
alone equals the same trajectory in a batch bit for bit, also when the batch
has more trajectories than workgroups and the trajectory ran in two pieces.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import gpu_mixed as GM
from classical_mcmc_spin_amd import mixed_model as mm
from test_mixed_oracle import random_flat, rich_cell

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _t1(T0, n, dt):
    return T0 + (n + 0.5) * dt


def _drive(ncomp, comp):
    f = np.zeros((4, ncomp))
    f[:, comp] = 1.0
    return f


def test_rich_cell_damped_driven_vs_oracle(po):
    """Tm-Tm bonds, SU(3) onsite and field, frames, Gilbert and Bloch damping,
    two pulses per species, per-trajectory states, pulse times and scales."""
    muc = rich_cell()
    dims = (3, 2, 2)
    ml = mm.MixedLattice(muc, *dims)
    B = 3
    inits = np.stack([random_flat(ml, 90 + b) for b in range(B)])
    tp = np.array([[0.0, 0.1], [0.15, 0.05], [0.05, 0.3]])
    sc = np.array([[1.0, 0.5], [0.7, 1.3], [1.0, 0.0]])
    rng = np.random.default_rng(11)
    d2a, d2b = rng.normal(size=(4, 3)), rng.normal(size=(4, 3))
    d3a, d3b = rng.normal(size=(4, 8)), rng.normal(size=(4, 8))
    gamma = np.array([0.02, 0.0, 0.05, 0.0, 0.0, 0.01, 0.0, 0.03])
    eq = rng.normal(size=(ml.N3, 8)) * 0.1
    T0, dt, n = -0.3, 0.005, 300
    with GM.GPUMixedLattice(ml) as gm:
        gm.set_damping(0.02, gamma, eq)
        GM._set_pulses(gm, d2a, d2b, d3a, d3b, 0.0, 0.0, 0.5, 0.3, 2.0, 0.4, 0.35, 1.1)
        res = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
        path = gm.last_path()
    assert path["persistent"] and path["split"] == 0, path
    for b in range(B):
        o = po.MixedOracle(muc, *dims)
        o.set_damping(0.02, gamma, eq)
        o.set_pulse(2, d2a * sc[b, 0], tp[b, 0], d2b * sc[b, 1], tp[b, 1], 0.5, 0.3, 2.0)
        o.set_pulse(3, d3a * sc[b, 0], tp[b, 0], d3b * sc[b, 1], tp[b, 1], 0.4, 0.35, 1.1)
        ref = o.integrate_md(inits[b], T0, _t1(T0, n, dt), dt)
        assert np.array_equal(res["times"], ref["times"])
        assert np.max(np.abs(res["records"][:, b] - ref["records"])) <= TOL, b
        assert np.max(np.abs(res["final"][b] - ref["final"])) <= TOL, b


def test_save_interval_and_undriven(po):
    """Records every 7th step and after the last one, no drive."""
    muc = mm.build_tmfeo3_mixed(seed=8)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    x = random_flat(ml, 95)
    T0, dt, n, si = 0.0, 0.005, 200, 7
    with GM.GPUMixedLattice(ml) as gm:
        res = gm.integrate_batched(x, np.zeros((1, 2)), np.zeros((1, 2)), T0, _t1(T0, n, dt), dt,
                                   save_interval=si, want_final=True)
        assert gm.last_path()["persistent"]
    o = po.MixedOracle(muc, 2, 2, 2)
    ref = o.integrate_md(x, T0, _t1(T0, n, dt), dt)
    keep = list(range(0, n, si)) + [n]
    assert res["records"].shape[0] == len(keep)
    assert np.max(np.abs(res["records"][:, 0] - ref["records"][keep])) <= TOL
    assert np.max(np.abs(res["final"][0] - ref["final"])) <= TOL


def test_full_size_equals_stage_kernels():
    """8x8x8 cells (2048 + 2048 sites, two sites of each species per thread
    of 1024): the persistent kernel against the stage kernels on one batch."""
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, 8, 8, 8)
    B = 6
    inits = np.stack([random_flat(ml, 100 + b) for b in range(B)])
    tp = np.stack([np.linspace(0.0, 0.5, B), np.linspace(0.3, 0.0, B)], axis=1)
    sc = np.ones((B, 2))
    sc[::2, 1] = 0.0
    T0, dt, n = -0.5, 0.005, 400
    f2, f3 = _drive(3, 0), _drive(8, 1)
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.5, 0.5, 0.0)
        print("[mixed-persist] 8x8x8 persistent", flush=True)
        got = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
        assert gm.last_path()["persistent"]
        gm.set_persistent(-1)
        print("[mixed-persist] 8x8x8 stage kernels", flush=True)
        ref = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
        assert not gm.last_path()["persistent"]
    assert np.max(np.abs(got["records"] - ref["records"])) <= 1e-12
    assert np.max(np.abs(got["final"] - ref["final"])) <= 1e-12


def test_work_split_and_alone():
    """More trajectories than workgroups: with at most 8 workgroups, 16
    trajectories run as two whole trajectories per workgroup and 17 as equal
    shares of the trajectory-steps (seven trajectories run in two pieces by
    two workgroups, the second piece starting from the state the first left);
    with 32 workgroups, 33 trajectories (30 split).  All equal the stage kernels on the
    same batch; a trajectory alone equals the batched one bit for bit."""
    muc = mm.build_tmfeo3_mixed(seed=9)
    ml = mm.MixedLattice(muc, 4, 4, 4)
    rng = np.random.default_rng(12)
    f2, f3 = _drive(3, 2), _drive(8, 4)
    T0, dt, n = -0.2, 0.005, 150
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.5, 0.5, 0.0)
        for cap, B, split in ((8, 16, 0), (8, 17, 7), (32, 33, 30)):
            inits = np.stack([random_flat(ml, 200 + b) for b in range(B)])
            tp = rng.uniform(0.0, 0.4, size=(B, 2))
            sc = rng.uniform(0.5, 1.5, size=(B, 2))
            print(f"[mixed-persist] cap {cap}, B {B}", flush=True)
            gm.set_persistent(cap)
            got = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
            path = gm.last_path()
            assert (path["persistent"], path["workgroups"], path["split"]) == (True, cap, split), path
            assert path["kernel_ms"] > 0.0
            gm.set_persistent(-1)
            ref = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
            assert np.max(np.abs(got["records"] - ref["records"])) <= 1e-12, (cap, B)
            assert np.max(np.abs(got["final"] - ref["final"])) <= 1e-12, (cap, B)
            gm.set_persistent(0)
            for b in (0, 2, B - 2, B - 1):
                one = gm.integrate_batched(inits[b], tp[b:b + 1], sc[b:b + 1], T0, _t1(T0, n, dt), dt,
                                           want_final=True)
                assert gm.last_path()["persistent"]
                assert np.array_equal(one["records"][:, 0], got["records"][:, b]), (cap, b)
                assert np.array_equal(one["final"][0], got["final"][b]), (cap, b)


def test_cooperative_small_batch_equals_stage_kernels():
    """Mode -2: two 8x8x8 trajectories on the cooperative stage kernel (16
    workgroups each, one launch, a grid-wide barrier per stage) against the
    stage kernels on the same batch."""
    muc = mm.build_tmfeo3_mixed(seed=10)
    ml = mm.MixedLattice(muc, 8, 8, 8)
    B = 2
    inits = np.stack([random_flat(ml, 300 + b) for b in range(B)])
    tp = np.array([[0.0, 0.2], [0.1, 0.0]])
    sc = np.array([[1.0, 0.5], [1.0, 1.0]])
    f2, f3 = _drive(3, 1), _drive(8, 2)
    T0, dt, n = -0.3, 0.005, 300
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.5, 0.5, 0.0)
        gm.set_persistent(-2)
        print("[mixed-coop] 8x8x8, B 2", flush=True)
        got = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, save_interval=3, want_final=True)
        path = gm.last_path()
        assert path["cooperative"] and not path["persistent"] and path["workgroups"] == 32, path
        gm.set_persistent(-1)
        ref = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, save_interval=3, want_final=True)
        assert not (gm.last_path()["persistent"] or gm.last_path()["cooperative"])
    assert got["records"].shape == ref["records"].shape
    assert np.max(np.abs(got["records"] - ref["records"])) <= 1e-12
    assert np.max(np.abs(got["final"] - ref["final"])) <= 1e-12
