"""GPU parity: libcsl_hip.so (through its C ABI) against the CPU oracle.

Bar (stated per test): local fields / dS/dt / energy to ~1e-13 relative
(one FP64 evaluation, FMA vs separate multiply-add); trajectories to
max|dM|, max|dS| <= 1e-10 over <= 500 RK4 steps (dt = 0.01), the horizon the
reference's own equivalence test uses to stay below chaotic divergence
(test_mixed_gpu_cpu_equivalence.cpp:198-205).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, model, seeding
from classical_mcmc_spin_amd import gpu as G
from helpers import kitaev, random_state, unit_x_field

pytestmark = pytest.mark.gpu

TOL_TRAJ = 1e-10


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _pair(po, L1=6, L2=6, h=0.1, **kw):
    uc, lat = kitaev(L1, L2, h=h, **kw)
    return uc, lat, po.OracleLattice(uc, L1, L2, 1)


@pytest.mark.parametrize("padded", [False, True])
def test_local_field_rhs_energy(po, padded):
    uc, lat, o = _pair(po, 6, 5, h=0.1, Ka=0.05, Kb=-0.02, Kc=0.07)
    s = random_state(lat, 11)
    with GPULattice(lat, padded=padded) as gl:
        info = gl.info()
        assert info["uniform"] and info["N"] == lat.lattice_size
        gl.set_spins(s)
        H = gl.local_field()
        Ho = np.array([o.local_field(s.reshape(-1), i) for i in range(lat.lattice_size)])
        assert np.max(np.abs(H - Ho)) <= 1e-13 * max(1.0, np.max(np.abs(Ho)))
        d = gl.rhs(0.3)
        do = o.rhs(s.reshape(-1), 0.3).reshape(-1, 3)
        assert np.max(np.abs(d - do)) <= 1e-13
        E = gl.energy()
        Eo = o.energy(s.reshape(-1))
        assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))


def test_rhs_with_drive_and_damping(po):
    uc, lat, o = _pair(po, 4, 4, h=0.2)
    s = random_state(lat, 5)
    f = unit_x_field(lat)
    o.set_pulse(f, 0.1, f * 0.5, 1.3, 0.7, 0.4, 0.9)
    o.set_alpha(0.05)
    with GPULattice(lat) as gl:
        fd = np.concatenate([np.einsum("aij,aj->ai", lat.frames, f).reshape(-1),
                             np.einsum("aij,aj->ai", lat.frames, 0.5 * f).reshape(-1)])
        gl.set_pulse(fd, 0.7, 0.4, 0.9, 0.1, 1.3)
        gl.set_damping(0.05)
        gl.set_spins(s)
        for t in (0.0, 0.25, 1.1):
            d = gl.rhs(t)
            do = o.rhs(s.reshape(-1), t).reshape(-1, 3)
            assert np.max(np.abs(d - do)) <= 1e-13


@pytest.mark.parametrize("engine", [1, 2])
def test_md_trajectory_500_steps(po, engine):
    """molecular_dynamics (rk4, Boost grid): M(t), E(t) and final spins vs the oracle."""
    uc, lat, o = _pair(po, 6, 6, h=0.1)
    s = random_state(lat, 1234)
    T0, T1, dt = 0.0, 5.005, 0.01   # 500 steps
    ref = o.integrate_md(s.reshape(-1), T0, T1, dt)
    with GPULattice(lat) as gl:
        res = G.molecular_dynamics(gl, s, T0, T1, dt, engine=engine)
        fin = gl.get_spins()
    assert len(ref["times"]) == 501 and np.array_equal(res["times"], ref["times"])
    assert np.max(np.abs(res["mags"] - ref["mags"])) <= TOL_TRAJ
    assert np.max(np.abs(res["energy"] - ref["energy"])) <= TOL_TRAJ
    assert np.max(np.abs(fin - ref["final"].reshape(-1, 3))) <= TOL_TRAJ


def test_damped_single_and_double_pulse(po):
    uc, lat, o = _pair(po, 6, 6, h=0.05)
    s = random_state(lat, 77)
    f = unit_x_field(lat)
    o.set_alpha(0.02)
    T0, T1, step = -1.0, 3.005, 0.01
    t_ref, m_ref = o.single_pulse_drive(s.reshape(-1), f, 0.5, 0.5, 0.5, 0.0, T0, T1, step)
    t2_ref, m2_ref = o.double_pulse_drive(s.reshape(-1), f, 0.0, f, 1.25, 0.5, 0.5, 0.0, T0, T1, step)
    with GPULattice(lat) as gl:
        gl.set_damping(0.02)
        t, m = G.single_pulse_drive(gl, s, f, 0.5, 0.5, 0.5, 0.0, T0, T1, step)
        t2, m2 = G.double_pulse_drive(gl, s, f, 0.0, f, 1.25, 0.5, 0.5, 0.0, T0, T1, step)
    assert np.array_equal(t, t_ref) and t[0] == pytest.approx(T0 + step)  # t = T_start dropped
    assert np.max(np.abs(m - m_ref)) <= TOL_TRAJ
    assert np.array_equal(t2, t2_ref)
    assert np.max(np.abs(m2 - m2_ref)) <= TOL_TRAJ


def test_batched_per_replica_pulses_and_states(po):
    """integrate_batched with per-trajectory initial states and pulse times vs
    one oracle double-pulse run per trajectory."""
    uc, lat, o = _pair(po, 6, 6, h=0.1)
    B = 5
    inits = np.stack([random_state(lat, 100 + b) for b in range(B)])
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    tp = np.array([[0.0, 0.0], [0.3, 0.0], [0.0, 0.5], [0.2, 0.9], [0.1, 0.1]])
    am = np.array([[0.5, 0.0], [0.5, 0.0], [0.5, 0.5], [0.3, 0.7], [0.0, 0.0]])
    T0, T1, dt = 0.0, 3.005, 0.01
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 1.0, 0.5, 0.3, 0.0, 0.0)
        res = gl.integrate_batched(inits, tp, am, T0, T1, dt, tgrid="boost", want_final=True)
    for b in range(B):
        o.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, 0.5, 0.3)
        ref = o.integrate_md(inits[b].reshape(-1), T0, T1, dt)
        assert np.array_equal(res["times"], ref["times"])
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL_TRAJ
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL_TRAJ
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL_TRAJ
    o.clear_pulse()


def test_batched_per_replica_hamiltonians(po):
    """One Kitaev parameter set per trajectory (parameter starts), damped and
    driven, vs one oracle run per parameter set."""
    params = [(-1.0, 0.25, -0.02, 0.05), (-1.1, 0.22, -0.018, 0.045), (-0.9, 0.27, -0.021, 0.055)]
    lats, ucs = [], []
    for K, G, Gp, h in params:
        uc, lat = kitaev(6, 6, h=h, K=K, Gamma=G, Gammap=Gp)
        ucs.append(uc)
        lats.append(lat)
    s = random_state(lats[0], 17)
    f = unit_x_field(lats[0])
    fd = np.einsum("aij,aj->ai", lats[0].frames, f).reshape(-1)
    B = len(params)
    T0, T1, dt = -1.0, 2.005, 0.01
    with GPULattice(lats[0]) as gl:
        gl.set_damping(0.02)
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
        gl.set_replica_hamiltonians(lats)
        res = gl.integrate_batched(s, np.zeros((B, 2)), np.tile([0.5, 0.0], (B, 1)), T0, T1, dt,
                                   tgrid="boost", want_final=True)
        gl.set_replica_hamiltonians(None)
    for b in range(B):
        o = po.OracleLattice(ucs[b], 6, 6, 1)
        o.set_alpha(0.02)
        o.set_pulse(f, 0.0, f * 0.0, 0.0, 0.5, 0.5, 0.0)
        ref = o.integrate_md(s.reshape(-1), T0, T1, dt)
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL_TRAJ
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL_TRAJ
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL_TRAJ


def test_pump_probe_2dcs_small(po):
    """2DCS driver (M0, M1(tau), M01(tau)) vs the oracle pump_probe_spectroscopy."""
    uc, lat, o = _pair(po, 6, 6, h=0.1)
    g = seeding.tilt(seeding.quench(lat, random_state(lat, 3), sweeps=50), 0.01, seed=9)
    f = unit_x_field(lat)
    args = (0.5, 0.5, 0.0, 0.0, 1.0, 0.25, -1.0, 2.005, 0.01)
    ref = o.pump_probe(g.reshape(-1), f, *args, reuse_m0=False, n_outer=1)
    with GPULattice(lat) as gl:
        res = G.pump_probe_spectroscopy(gl, g, f, *args)
    assert np.array_equal(res["tau"], ref["tau"]) and len(res["tau"]) == 5
    assert np.array_equal(res["times"], ref["times"])
    for k in ("M0", "M1", "M01"):
        assert np.max(np.abs(res[k] - ref[k])) <= TOL_TRAJ, k


def test_pump_probe_w1_reuse_m0(po):
    """W1 (reuse_m0_for_m1): on a stationary ground state (3000 greedy
    sweeps, max|dS/dt| ~ 1e-13) M1(tau) is the time shift of M0, matching the
    oracle's synthesize_M1_from_M0 and, to integration accuracy, the
    integrated M1 (the window starts at -5 so that the pulse of width 0.5 is
    off at T_start, which the time shift presumes); on a non-stationary state
    the guard falls back."""
    uc, lat, o = _pair(po, 6, 6, h=0.1)
    g = seeding.quench(lat, random_state(lat, 3), sweeps=3000)
    f = unit_x_field(lat)
    args = (0.5, 0.5, 0.0, 0.0, 1.0, 0.25, -5.0, 2.005, 0.01)
    ref = o.pump_probe(g.reshape(-1), f, *args, reuse_m0=True, stat_tol=1e-6, n_outer=1)
    assert ref["used_w1"]
    with GPULattice(lat) as gl:
        res = G.pump_probe_spectroscopy(gl, g, f, *args, reuse_m0_for_m1=True, stationarity_tol=1e-6)
        full = G.pump_probe_spectroscopy(gl, g, f, *args)
        tilted = seeding.tilt(g, 0.01, seed=9)
        res2 = G.pump_probe_spectroscopy(gl, tilted, f, *args, reuse_m0_for_m1=True, stationarity_tol=1e-6)
        full2 = G.pump_probe_spectroscopy(gl, tilted, f, *args)
    assert res["w1"] and res["max_dsdt"] <= 1e-10
    assert abs(res["max_dsdt"] - o.max_dsdt_no_drive(g.reshape(-1))) <= 1e-15
    assert np.array_equal(res["times"], ref["times"])
    for k in ("M0", "M1", "M01"):
        assert np.max(np.abs(res[k] - ref[k])) <= TOL_TRAJ, k
    # synthesised vs integrated M1: after tau the shifted M0; before it the
    # ground state, whose antiferro entry the reference takes with alternating
    # site signs (magnetization_local_antiferro) rather than the observer's
    # sublattice signs, so only M_local and M_global compare there
    t = res["times"]
    for i, tau in enumerate(res["tau"]):
        after = t >= tau + t[0]
        assert np.max(np.abs(res["M1"][i][after] - full["M1"][i][after])) <= 1e-8
        assert np.max(np.abs(res["M1"][i][~after][:, 1:] - full["M1"][i][~after][:, 1:]), initial=0.0) <= 1e-8
    # batch sizes differ (1 + n_tau vs 1 + 2 n_tau replicas): launch shapes,
    # hence the order of the record reductions, may differ
    for k in ("M0", "M01"):
        assert np.max(np.abs(res[k] - full[k])) <= 1e-13, k
    assert not res2["w1"] and res2["max_dsdt"] > 1e-6
    for k in ("M0", "M1", "M01"):
        assert np.max(np.abs(res2[k] - full2[k])) <= 1e-13, k


def test_general_tables_non_uniform(po):
    """Per-bond distinct matrices and per-site fields take the general
    (non-compressed) path; checked against the oracle built from the same CSR."""
    uc, lat = kitaev(5, 4, h=0.1)
    rng = np.random.default_rng(5)
    J = lat.bi_J + 0.05 * rng.standard_normal(lat.bi_J.shape)
    # keep the bond pair (i->p, p->i) transposes of each other
    off = lat.bi_offset.astype(np.int64)
    par = lat.bi_partner.astype(np.int64)
    seen = {}
    for i in range(lat.lattice_size):
        for k in range(off[i], off[i + 1]):
            key = (min(i, par[k]), max(i, par[k]))
            seen.setdefault(key, []).append(k)
    for ks in seen.values():
        if len(ks) == 2:
            J[ks[1]] = J[ks[0]].T
    lat.bi_J = J
    lat.field = lat.field + 0.01 * rng.standard_normal(lat.field.shape)
    o = po.OracleLattice.from_csr(3, 2, (5, 4, 1), lat.bi_offset, lat.bi_partner, J, lat.field, lat.onsite,
                                  lat.frames, lat.afm_signs)
    s = random_state(lat, 8)
    ref = o.integrate_md(s.reshape(-1), 0.0, 2.005, 0.01)
    with GPULattice(lat) as gl:
        assert not gl.info()["uniform"]
        res = G.molecular_dynamics(gl, s, 0.0, 2.005, 0.01)
    assert np.max(np.abs(res["mags"] - ref["mags"])) <= TOL_TRAJ
    assert np.max(np.abs(res["energy"] - ref["energy"])) <= TOL_TRAJ


def test_quench_matches_host_colour_sweeps():
    """Device greedy quench vs the host colour-class sweeps (same colouring,
    same update) over 40 sweeps."""
    uc, lat = kitaev(6, 6, h=0.1)
    s0 = np.stack([random_state(lat, 30 + b) for b in range(3)])
    with GPULattice(lat) as gl:
        q, done = gl.quench(s0, max_sweeps=40, rel_tol=0.0)
        assert np.all(done <= 40)
        for b in range(3):
            ref = seeding.quench(lat, s0[b], sweeps=40, tol=-1.0)
            assert np.max(np.abs(q[b] - ref)) <= 1e-10


@pytest.mark.parametrize("L,h", [(6, 0.1), (12, 0.05)])
def test_quench_vs_oracle_greedy_quench(po, L, h):
    """csl_quench against the oracle's Lattice::greedy_quench (lattice.h:
    3352-3368) sweeping the sites in the device's colour-class order (sites
    of one class are not bonded, so updating a class at once equals visiting
    its sites one after another): same sweep count under the reference's
    energy criterion, same states; the converged states are fixed points
    S_i = -H_i/|H_i| with lower energy than the start."""
    uc, lat, o = _pair(po, L, L, h=h)
    order = np.argsort(lat.coloring(), kind="stable")
    s0 = np.stack([random_state(lat, 60 + b) for b in range(3)])
    with GPULattice(lat) as gl:
        q, done = gl.quench(s0, max_sweeps=3000, rel_tol=1e-12)
        for b in range(3):
            _, n = o.greedy_quench(s0[b].reshape(-1), order, rel_tol=1e-12, max_sweeps=3000)
            assert abs(int(done[b]) - n) <= 1, (b, int(done[b]), n)   # the stop test sits on rounding
            ref, m = o.greedy_quench(s0[b].reshape(-1), order, rel_tol=0.0, max_sweeps=int(done[b]))
            assert m == int(done[b])
            assert np.max(np.abs(q[b].reshape(-1) - ref)) <= 1e-10
            gl.set_spins(s0[b])
            e0 = gl.energy()
            gl.set_spins(q[b])
            assert gl.energy() < e0
            H = gl.local_field()
            fixed = -H / np.linalg.norm(H, axis=1, keepdims=True)
            assert np.max(np.abs(q[b] - fixed)) <= 1e-5


@pytest.mark.parametrize("n_tau,n_t", [(5, 41), (8, 64), (1, 10)])
def test_nonlinear_spectrum_vs_numpy(n_tau, n_t):
    """GPU M_NL 2-D spectrum against numpy.fft on the same arrays, with the
    reference's windowing and padding (reader_pyrochlore.py:760-830)."""
    rng = np.random.default_rng(n_tau * 100 + n_t)
    res = {"M0": rng.normal(size=(n_t, 3, 3)), "M1": rng.normal(size=(n_tau, n_t, 3, 3)),
           "M01": rng.normal(size=(n_tau, n_t, 3, 3)), "times": 0.1 * np.arange(n_t),
           "tau": 0.25 * np.arange(n_tau)}
    out = G.nonlinear_spectrum(res, mag_index=2)
    dom = n_t // 2
    mnl = res["M01"][:, -dom:, 2] - res["M0"][None, -dom:, 2] - res["M1"][:, -dom:, 2]
    P1 = 2 ** int(np.ceil(np.log2(n_tau)))
    P2 = 2 ** int(np.ceil(np.log2(dom)))
    ref = np.stack([np.abs(np.fft.fftshift(np.fft.fft2(mnl[:, :, d], s=(P1, P2)))) for d in range(3)], -1)
    assert out["M_NL_FF"].shape == ref.shape
    assert np.max(np.abs(out["M_NL_FF"] - ref)) <= 1e-12 * max(1.0, np.max(ref))
    assert np.allclose(out["freq_T"], np.fft.fftshift(np.fft.fftfreq(P2, d=0.1)))


def test_graph_replay_matches_eager_launches(po):
    """engine 3 (stage launches captured once into a hipGraph, then replayed)
    against engine 1 (the same kernels launched eagerly): three initial
    states, three rounds of replays with eager integrations and host
    synchronisation between them, every replay bit-identical; the eager
    result itself against the oracle."""
    uc, lat, o = _pair(po, 6, 6, h=0.1)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    seeds = (1, 2, 3)
    with GPULattice(lat) as gl:
        gl.set_damping(0.01)
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.3, 0.8)
        eager = []
        for sd in seeds:
            gl.set_spins(random_state(lat, sd))
            r = gl.integrate(0.0, 1.505, 0.01, 1, tgrid="boost", engine=1)
            eager.append((r, gl.get_spins()))
        for rnd in range(3):
            for i, sd in enumerate(seeds):
                gl.set_spins(random_state(lat, sd))
                r = gl.integrate(0.0, 1.505, 0.01, 1, tgrid="boost", engine=3)
                assert np.array_equal(r["times"], eager[i][0]["times"])
                assert np.array_equal(r["mags"], eager[i][0]["mags"]), (rnd, i)
                assert np.array_equal(r["energy"], eager[i][0]["energy"]), (rnd, i)
                assert np.array_equal(gl.get_spins(), eager[i][1]), (rnd, i)
                gl.set_spins(random_state(lat, 50 + i))
                gl.integrate(0.0, 0.505, 0.01, 1, tgrid="boost", engine=1)
    o.set_alpha(0.01)
    o.set_pulse(f, 0.3, f, 0.8, 0.5, 0.5, 0.0)
    ref = o.integrate_md(random_state(lat, 1).reshape(-1), 0.0, 1.505, 0.01)
    assert np.max(np.abs(eager[0][0]["mags"] - ref["mags"])) <= TOL_TRAJ
