"""Dormand-Prince 5(4) on the mixed SU(2) + SU(3) lattice
(csl_mixed_integrate_dopri5) against the oracle's mixed Dormand-Prince.

Both sides run the same tableau, error norm, controller and controls, so
they take the same accepted / rejected steps.  On the integrate_const grid
(controls 0 and 2, with and without pulse-window segments) the record times
are equal and the records agree to 1e-10; with integrate_adaptive the times
are the accepted step times, which FMA-level rounding moves by ~1e-11, so
times compare to 1e-9 and records to 1e-8 (as on the SU(2) path).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import gpu as G
from classical_mcmc_spin_amd import gpu_mixed as GM
from classical_mcmc_spin_amd import mixed_model as mm
from classical_mcmc_spin_amd._capi import CslError
from test_mixed_oracle import random_flat, rich_cell

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _damping(gm, o, ml):
    gamma = np.array([0.05, 0.0, 0.02, 0.0, 0.03, 0.0, 0.0, 0.01])
    eq = np.tile(mm.random_su3_states(1, 99)[0], (ml.N3, 1))
    gm.set_damping(0.03, gamma, eq)
    o.set_damping(0.03, gamma, eq)


def _cmp(r, ref, b, adaptive):
    assert r["n_rec"][b] == len(ref["times"]), b
    assert (r["accepted"][b], r["rejected"][b]) == (ref["accepted"], ref["rejected"]), b
    t_tol, m_tol = (1e-9, 1e-8) if adaptive else (0.0, TOL)
    assert np.max(np.abs(r["times"][b] - ref["times"])) <= t_tol, b
    assert np.max(np.abs(r["records"][b] - ref["records"])) <= m_tol, b


def test_adaptive_damped_vs_oracle(po):
    """integrate_adaptive (control 1), Gilbert and Bloch damping, two
    trajectories with their own initial states, save_interval 2."""
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = po.MixedOracle(muc, 2, 2, 2)
    inits = np.stack([random_flat(ml, 11 + b) for b in range(2)])
    with GM.GPUMixedLattice(ml) as gm:
        _damping(gm, o, ml)
        r = gm.integrate_dopri5(inits, np.zeros((2, 2)), None, 0.0, 1.0, 0.01, control=1, abs_tol=1e-6,
                                rel_tol=1e-6, save_interval=2, want_final=True)
    for b in range(2):
        print(f"mixed dopri5 control 1, trajectory {b}", flush=True)
        ref = o.integrate_dopri5(inits[b], 0.0, 1.0, 0.01, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=2)
        _cmp(r, ref, b, adaptive=True)
        assert np.max(np.abs(r["final"][b] - ref["final"])) <= 1e-8


def test_controlled_const_segments_and_fixed_vs_oracle(po):
    """control 2 by pulse-window segments (each trajectory its own pulse
    centres and scales), control 2 over the whole window, and control 0."""
    muc = rich_cell(seed=4)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 2] = 1.0
    w2, w3 = 0.05, 0.04
    T0, T1, st = -0.3, 1.5, 0.02
    tp = np.array([[0.0, 0.0], [0.4, 0.0], [0.0, 0.2], [0.1, 1.3]])
    sc = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 1.0], [0.5, 2.0]])
    centers = [[0.0], [0.4], [0.0, 0.2], [0.1, 1.3]]
    segs = [G.pulse_segments(T0, T1, c, 9 * max(w2, w3), st) for c in centers]
    inits = np.stack([random_flat(ml, 50 + b) for b in range(4)])
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, w2, 3.0, 0.4, w3, 2.0)
        runs = {
            "segments": gm.integrate_dopri5(inits, tp, sc, T0, T1, st, control=2, abs_tol=1e-9, rel_tol=1e-9,
                                            segments=segs, want_final=True),
            "window": gm.integrate_dopri5(inits, tp, sc, T0, T1, st, control=2, abs_tol=1e-9, rel_tol=1e-9),
            "fixed": gm.integrate_dopri5(inits, tp, sc, T0, T1, st, control=0),
        }
    for b in range(4):
        o = po.MixedOracle(muc, 2, 2, 2)
        o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], 0.5, w2, 3.0)
        o.set_pulse(3, f3 * sc[b, 0], tp[b, 0], f3 * sc[b, 1], tp[b, 1], 0.4, w3, 2.0)
        for name, kw in (("segments", dict(control=2, segments=segs[b])), ("window", dict(control=2)),
                         ("fixed", dict(control=0))):
            print(f"mixed dopri5 {name}, trajectory {b}", flush=True)
            ref = o.integrate_dopri5(inits[b], T0, T1, st, abs_tol=1e-9, rel_tol=1e-9, **kw)
            _cmp(runs[name], ref, b, adaptive=False)
            if name == "segments":
                assert np.max(np.abs(runs[name]["final"][b] - ref["final"])) <= TOL


def test_drivers_dopri5(po):
    """molecular_dynamics and single_pulse_drive with method "dopri5"
    (pulse-window chunking on by default) against the oracle."""
    muc = rich_cell(seed=6)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = po.MixedOracle(muc, 2, 2, 2)
    x0 = random_flat(ml, 70)
    f2 = np.zeros((4, 3))
    f2[:, 1] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 0] = 1.0
    with GM.GPUMixedLattice(ml) as gm:
        md = GM.molecular_dynamics(gm, x0, 0.0, 2.0, 0.01, save_interval=3, method="dopri5")
        sp = GM.single_pulse_drive(gm, x0, f2, f3, 0.5, 0.6, 0.1, 2.0, 0.3, 0.08, 1.0, -0.5, 2.0, 0.01,
                                   method="dopri5")
    ref = o.integrate_dopri5(x0, 0.0, 2.0, 0.01, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=3)
    assert md["accepted"] == ref["accepted"]
    assert np.max(np.abs(md["times"] - ref["times"])) <= 1e-9
    assert np.max(np.abs(md["M2"].reshape(len(ref["times"]), -1) - ref["records"][:, :9])) <= 1e-8
    o.set_pulse(2, f2, 0.5, f2 * 0.0, 0.0, 0.6, 0.1, 2.0)
    o.set_pulse(3, f3, 0.5, f3 * 0.0, 0.0, 0.3, 0.08, 1.0)
    segs = G.pulse_segments(-0.5, 2.0, [0.5], 9 * 0.1, 0.01)
    r = o.integrate_dopri5(x0, -0.5, 2.0, 0.01, control=2, abs_tol=1e-8, rel_tol=1e-8, segments=segs)
    keep = G._gate(r["times"], -0.5, 2.0, 0.01)
    assert np.array_equal(sp["times"], r["times"][keep])
    got = np.concatenate([sp["M2"].reshape(len(keep), -1), sp["M3"].reshape(len(keep), -1)], axis=1)
    assert np.max(np.abs(got - r["records"][keep][:, :33])) <= TOL


def test_full_size_segments_vs_oracle(po):
    """8x8x8 cells (2048 + 2048 sites), three trajectories of the two-pulse
    scan with their segment lists, 100 observation intervals."""
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, 8, 8, 8)
    g = random_flat(ml, 80)
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    tp = np.array([[0.0, 0.0], [0.15, 0.0], [0.0, 0.15]])
    sc = np.array([[1.0, 0.0], [1.0, 0.0], [1.0, 1.0]])
    T0, T1, st = -0.2, 0.3, 0.005
    segs = [G.pulse_segments(T0, T1, c, 9 * 0.02, st) for c in ([0.0], [0.15], [0.0, 0.15])]
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.02, 0.0, 0.5, 0.02, 0.0)
        r = gm.integrate_dopri5(g, tp, sc, T0, T1, st, control=2, abs_tol=1e-8, rel_tol=1e-8, segments=segs)
    for b in (1, 2):
        print(f"mixed dopri5 8x8x8, trajectory {b}", flush=True)
        o = po.MixedOracle(muc, 8, 8, 8)
        o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], 0.5, 0.02, 0.0)
        o.set_pulse(3, f3 * sc[b, 0], tp[b, 0], f3 * sc[b, 1], tp[b, 1], 0.5, 0.02, 0.0)
        ref = o.integrate_dopri5(g, T0, T1, st, control=2, abs_tol=1e-8, rel_tol=1e-8, segments=segs[b])
        _cmp(r, ref, b, adaptive=False)


def test_record_retry_and_attempt_cap(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    x0 = random_flat(ml, 3)
    with GM.GPUMixedLattice(ml) as gm:
        small = gm.integrate_dopri5(x0, [[0.0, 0.0]], None, 0.0, 0.5, 0.01, control=1, abs_tol=1e-10,
                                    rel_tol=1e-10, max_rec=3)
        big = gm.integrate_dopri5(x0, [[0.0, 0.0]], None, 0.0, 0.5, 0.01, control=1, abs_tol=1e-10,
                                  rel_tol=1e-10, max_rec=100000)
        assert small["n_rec"] == big["n_rec"] and small["n_rec"][0] > 3
        assert np.array_equal(small["records"][0], big["records"][0])
        with pytest.raises(CslError):
            gm.integrate_dopri5(x0, [[0.0, 0.0]], None, 0.0, 0.5, 0.01, control=1, abs_tol=1e-10, rel_tol=1e-10,
                                max_attempts=5)
