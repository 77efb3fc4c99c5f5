"""The oracle's Dormand-Prince 5(4) over the mixed SU(2) + SU(3) state
(orc_mixed_integrate_dopri5).  The stepper, controller and controls are the
ones pinned on the SU(2) lattice (tests/test_oracle_dopri5.py); here the
mixed right-hand side under them is pinned against converged RK4 and by the
reference's own pulse-drive tolerance regression
(src/apps/test_pulse_drive_tolerance.cpp: a TmFeO3 lattice driven by a short
pulse through single_pulse_drive with "dopri5" and W3 chunking; tolerance
1e-8 against 1e-10 within 5e-6 at every sample)."""
import numpy as np
import pytest

import pyoracle
from classical_mcmc_spin_amd import mixed_model as mm
from test_mixed_oracle import random_flat, rich_cell


def _gate(times, T0, T1, step):
    keep, last = [], T0
    for i, t in enumerate(times):
        if t - last >= step - 1e-10 or t >= T1 - 1e-10:
            keep.append(i)
            last = t
    return np.array(keep)


def _setup(o, ml, rng):
    gamma = np.array([0.05, 0.0, 0.02, 0.0, 0.03, 0.0, 0.0, 0.01])
    eq = np.tile(mm.random_su3_states(1, 99)[0], (ml.N3, 1))
    o.set_damping(0.03, gamma, eq)
    f2 = rng.normal(size=(4, 3))
    f3 = rng.normal(size=(4, 8))
    o.set_pulse(2, f2, 0.3, 0.5 * f2, 0.6, 0.4, 0.25, 1.2)
    o.set_pulse(3, f3, 0.3, 0.7 * f3, 0.6, 0.3, 0.25, 0.8)


def test_mixed_dopri5_vs_converged_rk4():
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = pyoracle.MixedOracle(muc, 2, 2, 2)
    _setup(o, ml, np.random.default_rng(1))
    x0 = random_flat(ml, 21)
    ref = o.integrate_md(x0, 0.0, 1.0 + 1e-5, 1e-3)
    a = o.integrate_dopri5(x0, 0.0, 1.0, 0.05, control=1, abs_tol=1e-11, rel_tol=1e-11)
    assert a["times"][-1] == 1.0 and a["rejected"] < a["accepted"]
    assert np.max(np.abs(a["records"][-1] - ref["records"][-1])) <= 1e-8
    assert np.max(np.abs(a["final"] - ref["final"])) <= 1e-8
    c = o.integrate_dopri5(x0, 0.0, 1.0 + 1e-5, 0.05, control=2, abs_tol=1e-11, rel_tol=1e-11)
    assert len(c["times"]) == 21
    assert np.max(np.abs(c["records"] - ref["records"][::50])) <= 1e-8
    # looser tolerance: error grows, and stays near the tolerance
    d = o.integrate_dopri5(x0, 0.0, 1.0 + 1e-5, 0.05, control=2, abs_tol=1e-6, rel_tol=1e-6)
    err = np.max(np.abs(d["records"] - ref["records"][::50]))
    assert 1e-10 < err < 1e-4


def test_mixed_pulse_drive_tolerance_regression():
    """test_pulse_drive_tolerance.cpp restated on the package's TmFeO3-like
    cell (2x2x2 cells): t_B = 1, SU(2) pulse 0.5 / width 0.2 / freq 4, SU(3)
    pulse 0.3 / 0.2 / 5, T in [-1, 8] with T_step 0.05, segments of 9 widths
    around t_B."""
    muc = mm.build_tmfeo3_mixed(seed=7)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = pyoracle.MixedOracle(muc, 2, 2, 2)
    f2 = np.tile([1.0, 0.0, 0.0], (4, 1))
    f3 = np.zeros((4, 8))
    f3[:, 2] = 1.0
    o.set_pulse(2, f2, 1.0, f2 * 0.0, 0.0, 0.5, 0.2, 4.0)
    o.set_pulse(3, f3, 1.0, f3 * 0.0, 0.0, 0.3, 0.2, 5.0)
    x0 = random_flat(ml, 33)
    T0, T1, st = -1.0, 8.0, 0.05
    segs = pyoracle.pulse_segments(T0, T1, [1.0], 9 * 0.2, st, 20.0)
    runs = {}
    for tol in (1e-10, 1e-8):
        r = o.integrate_dopri5(x0, T0, T1, st, control=2, abs_tol=tol, rel_tol=tol, segments=segs)
        k = _gate(r["times"], T0, T1, st)
        runs[tol] = (r["times"][k], r["records"][k])
    (ta, ra), (tb, rb) = runs[1e-10], runs[1e-8]
    assert len(ta) == len(tb) >= 179 and np.max(np.abs(ta - tb)) <= 1e-12
    d = np.max(np.abs(ra[:, :33] - rb[:, :33]))
    assert 0.0 < d <= 5e-6
