"""Dormand-Prince 5(4) on the GPU (csl_integrate_dopri5) against the oracle.

Both sides run the same tableau, error norm and step controller, so the
step sequences coincide (same accepted / rejected counts).  On the fixed and
integrate_const grids the records sit at the same times and agree to
max|dM| <= 1e-10, the bar of the RK4 path.  With integrate_adaptive the
record times are the accepted step times: the error estimate is a difference
of nearly equal stage combinations, so FMA-level rounding differences in the
stages move it by ~1e-10 relative at tol 1e-6 and each step size by ~1e-11;
there the times compare to 1e-9 and M to 1e-8.  The physics check (tight
tolerance vs converged RK4) pins what the oracle's own tests pin on the CPU.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, seeding
from classical_mcmc_spin_amd import gpu as G
from classical_mcmc_spin_amd._capi import CslError, CSL_ERR_UNSUPPORTED
from helpers import kitaev, random_state, unit_x_field

pytestmark = pytest.mark.gpu

TOL = 1e-10
MAX_ATTEMPTS = 200000


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _pair(po, L1, L2, h=0.1):
    uc, lat = kitaev(L1, L2, h=h)
    return lat, po.OracleLattice(uc, L1, L2, 1)


def _cmp(r, ref, b=0, adaptive=True):
    assert r["n_rec"][b] == len(ref["times"])
    assert r["accepted"][b] == ref["accepted"] and r["rejected"][b] == ref["rejected"]
    t_tol, m_tol = (1e-9, 1e-8) if adaptive else (0.0, TOL)
    assert np.max(np.abs(r["times"][b] - ref["times"])) <= t_tol
    assert np.max(np.abs(r["mags"][b] - ref["mags"])) <= m_tol
    if r["energy"] is not None:
        assert np.max(np.abs(r["energy"][b] - ref["energy"])) <= m_tol


@pytest.mark.parametrize("L1,L2", [(4, 4), (16, 16), (24, 24), (32, 32)])  # SPT 1, 2, 3, 4
def test_adaptive_md_vs_oracle(po, L1, L2):
    """integrate_adaptive (control 1), damped, tol 1e-6 (the reference's
    md_abs_tol / md_rel_tol defaults), save_interval 2."""
    print(f"dopri5 control 1, {L1}x{L2}", flush=True)
    lat, o = _pair(po, L1, L2)
    s = random_state(lat, 7 + L1)
    o.set_alpha(0.02)
    T1 = 1.0 if L1 <= 16 else 0.5
    ref = o.integrate_dopri5(s.reshape(-1), 0.0, T1, 0.01, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=2)
    with GPULattice(lat) as gl:
        gl.set_damping(0.02)
        r = gl.integrate_dopri5(s, 0.0, T1, 0.01, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=2,
                                max_attempts=MAX_ATTEMPTS)
    _cmp(r, ref)
    assert np.max(np.abs(r["final"][0] - ref["final"].reshape(-1, 3))) <= 1e-8
    o.set_alpha(0.0)


def test_controlled_const_and_fixed_with_pulses_vs_oracle(po):
    """control 2 (integrate_const + controller) and control 0 (fixed steps)
    with a two-pulse drive, against the oracle with the same pulse."""
    lat, o = _pair(po, 6, 6)
    s = random_state(lat, 3)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    o.set_pulse(f, 0.2, f * 0.7, 0.9, 0.6, 0.35, 1.3)
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, 0.7 * fd]), 1.0, 0.35, 1.3, 0.0, 0.0)
        for control in (2, 0):
            print(f"dopri5 control {control}", flush=True)
            ref = o.integrate_dopri5(s.reshape(-1), -0.5, 1.505, 0.05, control=control, abs_tol=1e-9, rel_tol=1e-9)
            r = gl.integrate_dopri5(s, -0.5, 1.505, 0.05, control=control, abs_tol=1e-9, rel_tol=1e-9,
                                    t_pulse=[[0.2, 0.9]], pulse_amp=[[0.6, 0.6]], max_attempts=MAX_ATTEMPTS)
            _cmp(r, ref, adaptive=False)
    o.clear_pulse()


def test_batched_replicas_own_step_sequences(po):
    """Five trajectories with different initial states and pulse times in one
    launch; each runs its own step-size sequence."""
    lat, o = _pair(po, 6, 6)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    B = 5
    inits = np.stack([random_state(lat, 50 + b) for b in range(B)])
    tp = np.array([[0.0, 0.0], [0.3, 0.0], [0.0, 0.5], [0.2, 0.9], [0.1, 0.1]])
    am = np.array([[0.5, 0.0], [0.5, 0.0], [0.5, 0.5], [0.3, 0.7], [0.0, 0.0]])
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 1.0, 0.4, 0.8, 0.0, 0.0)
        r = gl.integrate_dopri5(inits, 0.0, 1.5, 0.01, control=1, abs_tol=1e-8, rel_tol=1e-8, t_pulse=tp,
                                pulse_amp=am, max_attempts=MAX_ATTEMPTS)
    steps = set()
    for b in range(B):
        o.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, 0.4, 0.8)
        ref = o.integrate_dopri5(inits[b].reshape(-1), 0.0, 1.5, 0.01, control=1, abs_tol=1e-8, rel_tol=1e-8)
        _cmp(r, ref, b)
        assert np.max(np.abs(r["final"][b] - ref["final"].reshape(-1, 3))) <= 1e-8
        steps.add(r["accepted"][b])
    assert len(steps) > 1
    o.clear_pulse()


def test_pulse_window_segments_vs_oracle(po):
    """control 2 by pulse-window segments (W3 chunking): every trajectory
    with its own pulse centres and segment list, one integrate_const call
    per segment; the seams are observed twice as in the reference's runs."""
    lat, o = _pair(po, 6, 6)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    width, T0, T1, st = 0.05, -0.5, 3.0, 0.02
    tp = np.array([[0.0, 0.0], [0.8, 0.0], [0.0, 0.3], [0.0, 2.9]])
    am = np.array([[0.6, 0.0], [0.6, 0.0], [0.6, 0.6], [0.6, 0.6]])
    centers = [[0.0], [0.8], [0.0, 0.3], [0.0, 2.9]]
    segs = [G.pulse_segments(T0, T1, c, 9 * width, st) for c in centers]
    assert [len(q) for q in segs] == [3, 3, 3, 4]
    inits = np.stack([random_state(lat, 60 + b) for b in range(len(tp))])
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 1.0, width, 3.0, 0.0, 0.0)
        r = gl.integrate_dopri5(inits, T0, T1, st, control=2, abs_tol=1e-9, rel_tol=1e-9, t_pulse=tp,
                                pulse_amp=am, segments=segs, max_attempts=MAX_ATTEMPTS)
    for b in range(len(tp)):
        print(f"segments, trajectory {b}", flush=True)
        o.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, width, 3.0)
        ref = o.integrate_dopri5(inits[b].reshape(-1), T0, T1, st, control=2, abs_tol=1e-9, rel_tol=1e-9,
                                 segments=segs[b])
        _cmp(r, ref, b, adaptive=False)
    o.clear_pulse()


def test_record_buffer_retry_and_limits(po):
    lat, o = _pair(po, 4, 4)
    s = random_state(lat, 9)
    with GPULattice(lat) as gl:
        small = gl.integrate_dopri5(s, 0.0, 2.0, 0.01, control=1, abs_tol=1e-10, rel_tol=1e-10, max_rec=4,
                                    max_attempts=MAX_ATTEMPTS)
        big = gl.integrate_dopri5(s, 0.0, 2.0, 0.01, control=1, abs_tol=1e-10, rel_tol=1e-10, max_rec=100000,
                                  max_attempts=MAX_ATTEMPTS)
        assert small["n_rec"] == big["n_rec"] and small["n_rec"][0] > 4
        assert np.array_equal(small["mags"][0], big["mags"][0])
        with pytest.raises(CslError):  # attempt cap reached
            gl.integrate_dopri5(s, 0.0, 2.0, 0.01, control=1, abs_tol=1e-10, rel_tol=1e-10, max_attempts=5)
    uc, big_lat = kitaev(40, 40)  # 3200 sites: more than one workgroup holds
    with GPULattice(big_lat) as gl:
        with pytest.raises(CslError) as e:
            gl.integrate_dopri5(random_state(big_lat, 1), 0.0, 0.1, 0.01)
        assert e.value.code == CSL_ERR_UNSUPPORTED


def test_drivers_dopri5(po):
    """molecular_dynamics and the pump-probe driver with method "dopri5":
    the MD run against the oracle, the pump-probe M0 / M1 / M01 (with and
    without pulse-window chunking) and single_pulse_drive against oracle runs
    of the controlled integrate_const, and the tight-tolerance MD run
    against converged RK4."""
    lat, o = _pair(po, 6, 6)
    s = random_state(lat, 13)
    ref = o.integrate_dopri5(s.reshape(-1), 0.0, 2.0, 0.01, control=1, abs_tol=1e-6, rel_tol=1e-6)
    g = seeding.tilt(seeding.quench(lat, random_state(lat, 3), sweeps=50), 0.01, seed=9)
    f = unit_x_field(lat)
    with GPULattice(lat) as gl:
        md = G.molecular_dynamics(gl, s, 0.0, 2.0, 0.01, method="dopri5")
        assert md["accepted"] == ref["accepted"]
        assert np.max(np.abs(md["mags"] - ref["mags"])) <= 1e-8
        tight = G.molecular_dynamics(gl, s, 0.0, 2.0, 0.01, method="dopri5", abs_tol=1e-11, rel_tol=1e-11)
        conv = o.integrate_md(s.reshape(-1), 0.0, 2.0 + 1e-4, 1e-3)
        assert np.max(np.abs(tight["mags"][-1] - conv["mags"][-1])) <= 1e-8
        # pulse width 0.05: windows of 0.45 around the centres leave free
        # segments in [-0.5, 1.005]
        pp = {c: G.pump_probe_spectroscopy(gl, g, f, 0.5, 0.05, 2.0, 0.0, 0.5, 0.25, -0.5, 1.005, 0.01,
                                           method="dopri5", abs_tol=1e-8, rel_tol=1e-8, pulse_window_chunking=c)
              for c in (True, False)}
        t_sp, m_sp = G.single_pulse_drive(gl, g, f, 0.25, 0.5, 0.05, 2.0, -0.5, 1.005, 0.01, method="dopri5")
    fd = np.asarray(f, dtype=float)
    for chunk in (True, False):
        for name, t1, t2, a2 in (("M0", 0.0, 0.0, 0.0), ("M1", 0.25, 0.0, 0.0), ("M01", 0.0, 0.25, 0.5)):
            print(f"pump-probe dopri5 {name}, chunking {chunk}", flush=True)
            if name == "M1":
                o.set_pulse(fd, 0.25, fd * 0.0, 0.0, 0.5, 0.05, 2.0)
                c = [0.25]
            else:
                o.set_pulse(fd, 0.0, fd * (1.0 if a2 else 0.0), t2, 0.5, 0.05, 2.0)
                c = [0.0, 0.25] if a2 else [0.0]
            segs = pyoracle_segments(c) if chunk else None
            r = o.integrate_dopri5(g.reshape(-1), -0.5, 1.005, 0.01, control=2, abs_tol=1e-8, rel_tol=1e-8,
                                   segments=segs)
            keep = G._gate(r["times"], -0.5, 1.005, 0.01)
            res = pp[chunk]
            got = res[name] if name == "M0" else res[name][1]
            assert len(got) == len(keep), name
            assert np.max(np.abs(got - r["mags"][keep])) <= TOL, name
            if name == "M0":
                assert np.max(np.abs(res["times"] - r["times"][keep])) <= 1e-12
            if name == "M1" and chunk:  # single_pulse_drive at t_B = 0.25, chunked by default
                # its grid times start from its own segments: equal to M0's up to rounding
                assert np.max(np.abs(t_sp - res["times"])) <= 1e-12 and np.max(np.abs(m_sp - got)) <= TOL
    o.clear_pulse()


def pyoracle_segments(centers):
    import pyoracle
    return pyoracle.pulse_segments(-0.5, 1.005, centers, 9 * 0.05, 0.01, 20.0)
