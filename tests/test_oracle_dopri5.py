"""Pinning of the oracle's Dormand-Prince 5(4) stepper and controller.

Boost.Odeint (runge_kutta_dopri5, controlled_runge_kutta) is not in the
reference tree, so the accept/reject sequence cannot be compared with Boost
itself.  What is pinned instead:
  * the tableau, by convergence orders on a problem with a closed form and a
    time-dependent forcing (a wrong a_ij, b_i or c_i drops the order of the
    propagated solution below 5; a wrong error weight drops the order of the
    embedded error estimate below 5 per step);
  * the controller, by tolerance-proportional accuracy, exact landing on the
    end time and the observer count of integrate_adaptive;
  * the lattice integration, against RK4 at a step size where RK4 is
    converged.
"""
import math

import numpy as np
import pytest

import pyoracle
from helpers import kitaev, random_state


def exact(lam, om, y0, T0, t):
    A = -lam / (lam * lam + om * om)
    B = om / (lam * lam + om * om)
    yp = lambda s: A * math.cos(om * s) + B * math.sin(om * s)
    return (y0 - yp(T0)) * math.exp(lam * (t - T0)) + yp(t)


LAM, OM, Y0, T0 = -0.7, 2.3, 0.4, 0.3


def test_fixed_step_global_order_five():
    errs = []
    for n in (20, 40, 80):
        dt = 2.0 / n
        y, acc, rej, _ = pyoracle.dopri5_scalar(LAM, OM, Y0, T0, T0 + 2.0 + 0.25 * dt, dt, control=0)
        assert acc == n and rej == 0
        errs.append(abs(y - exact(LAM, OM, Y0, T0, T0 + 2.0)))
    for a, b in zip(errs, errs[1:]):
        assert 4.6 < math.log2(a / b) < 5.4


def test_one_step_local_orders():
    """Propagated 5th-order solution: local error O(h^6); embedded estimate
    (5th minus 4th order): O(h^5)."""
    loc, est = [], []
    for h in (0.2, 0.1, 0.05):
        _, _, _, (xerr, y1) = pyoracle.dopri5_scalar(LAM, OM, Y0, T0, T0 + h, h, control=0)
        loc.append(abs(y1 - exact(LAM, OM, Y0, T0, T0 + h)))
        est.append(abs(xerr))
    for a, b in zip(loc, loc[1:]):
        assert 5.5 < math.log2(a / b) < 6.6
    for a, b in zip(est, est[1:]):
        assert 4.6 < math.log2(a / b) < 5.4
    # the estimate tracks the error of the embedded 4th-order solution:
    # |y4 - y| = |y1 - xerr - y| ~ |xerr| for small h
    y4_err = abs(y1 - xerr - exact(LAM, OM, Y0, T0, T0 + 0.05))
    assert abs(y4_err - est[-1]) <= 0.05 * est[-1]


def test_adaptive_error_tracks_tolerance():
    T1 = T0 + 10.0
    res = {}
    for tol in (1e-5, 1e-8, 1e-11):
        y, acc, rej, _ = pyoracle.dopri5_scalar(LAM, OM, Y0, T0, T1, 0.01, control=1, abs_tol=tol, rel_tol=tol)
        res[tol] = (abs(y - exact(LAM, OM, Y0, T0, T1)), acc, rej)
    assert res[1e-5][0] < 1e-4 and res[1e-8][0] < 1e-7 and res[1e-11][0] < 1e-10
    assert res[1e-5][0] > res[1e-8][0] > res[1e-11][0]
    # step counts grow roughly like tol^(-1/5): x1000 tolerance -> ~x4 steps
    r = res[1e-11][1] / res[1e-8][1]
    assert 2.5 < r < 6.5
    # the 0.9 safety factor and the err^(-1/5) growth keep rejections a
    # minority of the attempts on a smooth problem
    assert res[1e-8][2] <= 0.3 * res[1e-8][1] and res[1e-11][2] <= 0.3 * res[1e-11][1]
    # a large initial dt is rejected and shrunk by the controller
    _, acc, rej, _ = pyoracle.dopri5_scalar(LAM, OM, Y0, T0, T1, 5.0, control=1, abs_tol=1e-9, rel_tol=1e-9)
    assert rej >= 1


@pytest.fixture(scope="module")
def lat_pair(oracle):
    uc, lat = kitaev(4, 4, h=0.1)
    return lat, oracle.OracleLattice(uc, 4, 4, 1)


def test_lattice_adaptive_vs_converged_rk4(lat_pair):
    lat, o = lat_pair
    s = random_state(lat, 41)
    o.set_alpha(0.02)
    T1 = 2.0
    ref = o.integrate_md(s.reshape(-1), 0.0, T1 + 1e-4, 1e-3)  # RK4, dt = 1e-3: error ~1e-12
    for tol, bound in ((1e-6, 3e-5), (1e-10, 1e-8)):
        r = o.integrate_dopri5(s.reshape(-1), 0.0, T1, 0.01, control=1, abs_tol=tol, rel_tol=tol)
        assert r["times"][0] == 0.0 and r["times"][-1] == T1        # lands on T_end exactly
        assert len(r["times"]) == r["accepted"] + 1                  # observer after every step
        assert np.all(np.diff(r["times"]) > 0)
        assert np.max(np.abs(r["final"] - ref["final"])) <= bound
    o.set_alpha(0.0)


def test_lattice_controlled_const_grid(lat_pair):
    """control 2 observes on the integrate_const grid of RK4 and matches the
    converged trajectory there; control 0 is the plain fixed-step method."""
    lat, o = lat_pair
    s = random_state(lat, 42)
    f = np.zeros((2, 3))
    f[:, 0] = 1.0
    o.set_pulse(f, 0.5, f * 0.0, 0.0, 0.7, 0.3, 1.1)
    ref = o.integrate_md(s.reshape(-1), 0.0, 1.5 + 1e-4, 1e-3)
    grid = o.integrate_md(s.reshape(-1), 0.0, 1.505, 0.05)
    r = o.integrate_dopri5(s.reshape(-1), 0.0, 1.505, 0.05, control=2, abs_tol=1e-10, rel_tol=1e-10)
    assert np.array_equal(r["times"], grid["times"])
    assert r["accepted"] >= len(r["times"]) - 1
    ref_on_grid = ref["mags"][::50]
    assert np.max(np.abs(r["mags"] - ref_on_grid)) <= 1e-8
    r0 = o.integrate_dopri5(s.reshape(-1), 0.0, 1.505, 0.05, control=0)
    assert r0["accepted"] == len(r0["times"]) - 1 and r0["rejected"] == 0
    assert np.max(np.abs(r0["mags"] - ref_on_grid)) <= 1e-6
    o.clear_pulse()
