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


def test_controlled_const_carries_step_size(lat_pair):
    """integrate_const with the controlled stepper hands the step size of one
    observation interval to the next (Boost passes dt by reference through
    its integrate_adaptive calls): with an interval four times the step the
    tolerance allows, a controller restarted at dt would reject at least once
    per interval; the carried step rejects only a few times in total."""
    lat, o = lat_pair
    s = random_state(lat, 8)
    n_int = 40
    r = o.integrate_dopri5(s.reshape(-1), 0.0, 0.4 * n_int + 1e-3, 0.4, control=2, abs_tol=1e-10, rel_tol=1e-10)
    assert len(r["times"]) == n_int + 1
    assert r["accepted"] >= 3 * n_int  # several steps per interval
    assert r["rejected"] <= n_int // 4


def test_single_segment_equals_window(lat_pair):
    """One segment covering the window is the unchunked integrate_const call."""
    lat, o = lat_pair
    s = random_state(lat, 9)
    f = np.zeros((2, 3))
    f[:, 1] = 1.0
    o.set_pulse(f, 0.4, f * 0.0, 0.0, 0.6, 0.25, 2.0)
    a = o.integrate_dopri5(s.reshape(-1), -0.5, 1.0, 0.05, control=2, abs_tol=1e-8, rel_tol=1e-8)
    b = o.integrate_dopri5(s.reshape(-1), -0.5, 1.0, 0.05, control=2, abs_tol=1e-8, rel_tol=1e-8,
                           segments=[[-0.5, 1.0]])
    o.clear_pulse()
    for k in ("times", "mags", "final"):
        assert np.array_equal(a[k], b[k])
    assert (a["accepted"], a["rejected"]) == (b["accepted"], b["rejected"])


def _gate(times, T0, T1, step):
    keep, last = [], T0
    for i, t in enumerate(times):
        if t - last >= step - 1e-10 or t >= T1 - 1e-10:
            keep.append(i)
            last = t
    return np.array(keep)


def test_chunked_pulse_drive_regression(lat_pair):
    """W3 chunking with the controlled stepper (pulse_chunking.h; the
    reference's test_pulse_drive_tolerance.cpp on the Kitaev lattice): the
    chunked runs keep the unchunked sample grid, every seam is observed
    twice before the observer gate, the chunked and unchunked trajectories
    agree to the tolerance level, and tolerance 1e-8 agrees with 1e-10 to
    the reference's 5e-6 ceiling at every sample."""
    lat, o = lat_pair
    s = random_state(lat, 10)
    f = np.zeros((2, 3))
    f[:, 0] = 1.0
    t_B, amp, width, freq = 1.0, 0.5, 0.2, 4.0
    T0, T1, st = -1.0, 8.0, 0.05
    o.set_pulse(f, t_B, f * 0.0, 0.0, amp, width, freq)
    segs = pyoracle.pulse_segments(T0, T1, [t_B], 9 * width, st, 20.0)
    assert len(segs) == 3

    def run(tol, chunk):
        r = o.integrate_dopri5(s.reshape(-1), T0, T1, st, control=2, abs_tol=tol, rel_tol=tol,
                               segments=segs if chunk else None)
        k = _gate(r["times"], T0, T1, st)
        return r, r["times"][k], r["mags"][k]

    r_c, t_c, m_c = run(1e-10, True)
    r_u, t_u, m_u = run(1e-10, False)
    r_8, t_8, m_8 = run(1e-8, True)
    o.clear_pulse()
    assert len(r_c["times"]) == len(r_u["times"]) + len(segs) - 1
    # samples -0.95 .. 7.95: the step to 8.0 fails Boost's epsilon test
    # (-1 + 179 * 0.05) + 0.05 - 8 <= eps in both runs
    assert len(t_c) == len(t_u) == len(t_8) == 179
    assert np.max(np.abs(t_c - t_u)) <= 1e-12 and np.max(np.abs(t_8 - t_c)) <= 1e-12
    assert np.max(np.abs(m_c - m_u)) <= 1e-8
    d = np.max(np.abs(m_8 - m_c))
    assert 0.0 < d <= 5e-6
