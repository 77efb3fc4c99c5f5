"""GPU parity of the mixed SU(2) + SU(3) path (libcsl_hip.so through
include/csl_mixed.h) against the mixed oracle.

Bar: right-hand side, energy and records of one state to ~1e-13; records
and states of trajectories to <= 1e-10 over <= 500 RK4 steps.  At the full
8x8x8-cell size (2048 + 2048 sites, 257 trajectories): two trajectories
against the oracle over 200 steps, and for 8192 steps the size-independent
properties (every trajectory equals the same trajectory integrated alone,
bit for bit; SU(2) lengths and SU(3) norms within 1e-6 of 1 without damping).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import gpu_mixed as GM
from classical_mcmc_spin_amd import mixed_model as mm
from classical_mcmc_spin_amd import seeding
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


def _pulses(seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(4, 3)), rng.normal(size=(4, 3)), rng.normal(size=(4, 8)), rng.normal(size=(4, 8))


def test_rhs_energy_record(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 3, 2)
    o = po.MixedOracle(muc, 2, 3, 2)
    x = random_flat(ml, 21)
    d2a, d2b, d3a, d3b = _pulses()
    gamma = np.linspace(0.0, 0.2, 8)
    eq = np.random.default_rng(2).normal(size=(ml.N3, 8))
    with GM.GPUMixedLattice(ml) as gm:
        info = gm.info()
        assert info["N2"] == ml.N2 and info["N3"] == ml.N3 and all(info["uniform"])
        gm.set_spins(x)
        assert np.array_equal(gm.get_spins(), x)
        gm.set_damping(0.03, gamma, eq)
        gm.set_pulse(2, np.stack([gm.drive_dirs(2, d2a), gm.drive_dirs(2, d2b)]), 0.4, 0.6, 1.1, 0.2, 0.5)
        gm.set_pulse(3, np.stack([gm.drive_dirs(3, d3a), gm.drive_dirs(3, d3b)]), 0.3, 0.8, 0.7, 0.2, 0.5)
        d = gm.rhs(0.35)
        r = gm.record()
        E = gm.energy()
    o.set_damping(0.03, gamma, eq)
    o.set_pulse(2, d2a, 0.2, d2b, 0.5, 0.4, 0.6, 1.1)
    o.set_pulse(3, d3a, 0.2, d3b, 0.5, 0.3, 0.8, 0.7)
    assert np.max(np.abs(d - o.rhs(x, 0.35))) < 1e-13
    assert np.max(np.abs(r - o.record(x))) < 1e-13
    assert abs(E - o.energy(x)) < 1e-12 * max(1.0, abs(E))


def test_md_damped_driven_300_steps(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 3)
    o = po.MixedOracle(muc, 2, 2, 3)
    x = random_flat(ml, 23)
    d2a, d2b, d3a, d3b = _pulses(4)
    gamma = np.array([0.02, 0.0, 0.05, 0.0, 0.0, 0.01, 0.0, 0.03])
    eq = np.random.default_rng(3).normal(size=(ml.N3, 8)) * 0.1
    T0, dt, n = -0.5, 0.005, 300
    with GM.GPUMixedLattice(ml) as gm:
        gm.set_damping(0.02, gamma, eq)
        gm.set_pulse(2, np.stack([gm.drive_dirs(2, d2a), gm.drive_dirs(2, d2b)]), 0.5, 0.3, 2.0, 0.0, 0.4)
        gm.set_pulse(3, np.stack([gm.drive_dirs(3, d3a), gm.drive_dirs(3, d3b)]), 0.5, 0.3, 2.0, 0.1, 0.4)
        gm.set_spins(x)
        res = gm.integrate(T0, _t1(T0, n, dt), dt, 1, tgrid="boost", want_states=True)
        final = gm.get_spins()
    o.set_damping(0.02, gamma, eq)
    o.set_pulse(2, d2a, 0.0, d2b, 0.4, 0.5, 0.3, 2.0)
    o.set_pulse(3, d3a, 0.1, d3b, 0.4, 0.5, 0.3, 2.0)
    ref = o.integrate_md(x, T0, _t1(T0, n, dt), dt)
    assert np.array_equal(res["times"], ref["times"])
    assert np.max(np.abs(res["records"] - ref["records"])) <= TOL
    assert np.max(np.abs(final - ref["final"])) <= TOL
    assert np.max(np.abs(res["states"][-1] - ref["final"])) <= TOL


def test_batched_per_replica_states_and_pulses(po):
    muc = mm.build_tmfeo3_mixed(seed=5)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    B = 4
    inits = np.stack([random_flat(ml, 40 + b) for b in range(B)])
    tp = np.array([[0.0, 0.0], [0.2, 0.0], [0.0, 0.3], [0.1, 0.25]])
    sc = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.5, 2.0]])
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    T0, dt, n = -0.4, 0.005, 250
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.3, 0.4, 1.5)
        res = gm.integrate_batched(inits, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
    for b in range(B):
        o = po.MixedOracle(muc, 2, 2, 2)
        o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], 0.5, 0.5, 0.0)
        o.set_pulse(3, f3 * sc[b, 0], tp[b, 0], f3 * sc[b, 1], tp[b, 1], 0.3, 0.4, 1.5)
        ref = o.integrate_md(inits[b], T0, _t1(T0, n, dt), dt)
        assert np.array_equal(res["times"], ref["times"])
        assert np.max(np.abs(res["records"][:, b] - ref["records"])) <= TOL, b
        assert np.max(np.abs(res["final"][b] - ref["final"])) <= TOL, b


def test_pump_probe_small(po):
    muc = mm.build_tmfeo3_mixed(seed=6)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    g = random_flat(ml, 60)
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 4] = 1.0
    T0, T1, dt = -0.5, 0.8025, 0.005
    args = (0.5, 0.3, 0.0, 0.2, 0.3, 1.0)
    with GM.GPUMixedLattice(ml) as gm:
        res = GM.pump_probe_spectroscopy(gm, g, f2, f3, *args, 0.0, 0.2, 0.1, T0, T1, dt)
    assert np.allclose(res["tau"], [0.0, 0.1, 0.2])
    o = po.MixedOracle(muc, 2, 2, 2)
    z2, z3 = np.zeros((4, 3)), np.zeros((4, 8))

    def oracle_run(t1, b2_2, b3_2, t2):
        o.set_pulse(2, f2, t1, b2_2, t2, 0.5, 0.3, 0.0)
        o.set_pulse(3, f3, t1, b3_2, t2, 0.2, 0.3, 1.0)
        r = o.integrate_md(g, T0, T1, dt)
        keep = GM._gate(r["times"], T0, T1, dt)
        return r["times"][keep], r["records"][keep]
    t0, m0 = oracle_run(0.0, z2, z3, 0.0)
    assert np.array_equal(res["times"], t0)
    assert res["times"][0] > T0
    rec = lambda d: np.concatenate([d["M2"].reshape(d["M2"].shape[:-2] + (9,)),
                                    d["M3"].reshape(d["M3"].shape[:-2] + (24,)), d["energy"][..., None]], -1)
    assert np.max(np.abs(rec(res["M0"]) - m0)) <= TOL
    for i, tau in enumerate(res["tau"]):
        _, m1 = oracle_run(tau, z2, z3, 0.0)
        _, m01 = oracle_run(0.0, f2, f3, tau)
        assert np.max(np.abs(rec(res["M1"])[i] - m1)) <= TOL
        assert np.max(np.abs(rec(res["M01"])[i] - m01)) <= TOL


def test_tmfeo3_full_size_vs_oracle(po):
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, 8, 8, 8)
    assert ml.N2 == 2048 and ml.N3 == 2048
    g = random_flat(ml, 80)
    B = 257
    taus = 0.05 * np.arange(128)
    tp = np.zeros((B, 2))
    sc = np.zeros((B, 2))
    sc[0, 0] = 1.0
    tp[1:129, 0] = taus
    sc[1:129, 0] = 1.0
    tp[129:, 1] = taus
    sc[129:, :] = 1.0
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    T0, dt, n = -1.0, 0.005, 200
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.5, 0.5, 0.0)
        res = gm.integrate_batched(g, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
    for b in (0, 129 + 10):
        o = po.MixedOracle(muc, 8, 8, 8)
        o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], 0.5, 0.5, 0.0)
        o.set_pulse(3, f3 * sc[b, 0], tp[b, 0], f3 * sc[b, 1], tp[b, 1], 0.5, 0.5, 0.0)
        ref = o.integrate_md(g, T0, _t1(T0, n, dt), dt)
        assert np.max(np.abs(res["records"][:, b] - ref["records"])) <= TOL, b
        assert np.max(np.abs(res["final"][b] - ref["final"])) <= TOL, b


def test_tmfeo3_full_workload_properties(po):
    """Full length (8192 steps): trajectory 200 alone equals the batched one
    bit for bit; its SU(2) lengths, SU(3) norms and energy (after the pulse)
    drift no more than ten times the oracle's on the same trajectory (the
    dynamics is chaotic, so the oracle fixes the order of the drift, not the
    pointwise values at this length)."""
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, 8, 8, 8)
    g = random_flat(ml, 81)
    B = 257
    tp = np.zeros((B, 2))
    tp[1:129, 0] = 0.05 * np.arange(128)
    tp[129:, 1] = 0.05 * np.arange(128)
    sc = np.ones((B, 2))
    sc[:129, 1] = 0.0
    T0, dt, n = -1.0, 0.005, 8192
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.5, 0.0, 0.0, 0.5, 0.0)
        res = gm.integrate_batched(g, tp, sc, T0, _t1(T0, n, dt), dt, want_final=True)
        one = gm.integrate_batched(g, tp[200:201], sc[200:201], T0, _t1(T0, n, dt), dt, want_final=True)
    assert res["records"].shape == (n + 1, B, 34)
    assert np.all(np.isfinite(res["records"]))
    assert np.array_equal(one["final"][0], res["final"][200])
    assert np.array_equal(one["records"][:, 0], res["records"][:, 200])
    s2, s3 = ml.split_state(res["final"][0])
    assert np.max(np.abs(np.linalg.norm(s2, axis=1) - 1.0)) < 1e-6
    assert np.max(np.abs(np.linalg.norm(s3, axis=1) - 1.0)) < 1e-6
    b = 200
    print(f"[tmfeo3-full] oracle, trajectory {b}, {n} steps", flush=True)
    o = po.MixedOracle(muc, 8, 8, 8)
    o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], 0.5, 0.5, 0.0)
    o.set_pulse(3, f3, tp[b, 0], f3, tp[b, 1], 0.5, 0.5, 0.0)
    ref = o.integrate_md(g, T0, _t1(T0, n, dt), dt)
    g2, g3 = ml.split_state(res["final"][b])
    r2, r3 = ml.split_state(ref["final"])
    for got, want in ((g2, r2), (g3, r3)):
        d_gpu = np.max(np.abs(np.linalg.norm(got, axis=1) - 1.0))
        d_ref = np.max(np.abs(np.linalg.norm(want, axis=1) - 1.0))
        assert d_gpu <= 10.0 * d_ref + 1e-13, (d_gpu, d_ref)
    i0 = int(np.ceil((max(tp[b]) + 5.0 - T0) / dt))
    e_gpu = GM.split_records(res["records"][i0:, b])["energy"]
    e_ref = GM.split_records(ref["records"][i0:])["energy"]
    assert np.max(np.abs(e_gpu - e_gpu[0])) <= 10.0 * np.max(np.abs(e_ref - e_ref[0])) + 1e-12
