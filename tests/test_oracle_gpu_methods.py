"""Pins of the oracle's restatement of the reference GPU stepper
gpu::ode::step (integrator.cuh:91-458) and its integrate loop
(integrator.cuh:470-487), against mathematics rather than its own formulas:

* global convergence order of every method on a driven, damped Kitaev lattice
  (a nonlinear, time-dependent problem, so the stage times c_s and the
  non-linear order conditions are exercised), measured against a converged
  solution of the separately pinned Boost-RK4 oracle at dt = 2^-13;
* aliases and the fallback of the reference's name table;
* "rk4" of the GPU stepper equals the Boost-order RK4 up to rounding;
* the same stepper over the mixed SU(2)+SU(3) RHS converges with the
  method's order.

A planted wrong coefficient, weight, stage time or buffer in any method
changes its observed order.
"""
import math

import numpy as np
import pytest

from helpers import kitaev, random_state, unit_x_field

# rk78: the reference's fifth stage puts -25/16 on k4 and +25/16 on k3
# (integrator.cuh:273-275), the transpose of Fehlberg's a53 = -25/16,
# a54 = 25/16; with that table the 7th-order weights give a 6th-order method
# (measured here; the same scheme with Fehlberg's a5j converges with order 7).
# The oracle and the device reproduce the reference's coefficients.
ORDERS = {"euler": 1, "rk2": 2, "rk4": 4, "ssprk53": 3, "rk5": 5, "dopri5": 5, "rk78": 6,
          "bulirsch_stoer": 2}
# step pairs (2^-a, 2^-(a+1)) that keep each method's error far above rounding
STEPS = {"euler": 8, "rk2": 6, "rk4": 4, "ssprk53": 5, "rk5": 3, "dopri5": 4, "rk78": 3, "bulirsch_stoer": 5}


def _driven_lattice(oracle):
    uc, lat = kitaev(2, 2, h=0.4, direction=(0.3, -0.2, 1.0), J=0.2)
    o = oracle.OracleLattice(uc, 2, 2, 1)
    f = unit_x_field(lat)
    f[:, 1] = 0.5
    o.set_pulse(f, 0.45, f, 0.8, 0.9, 0.3, 1.7)
    o.set_alpha(0.05)
    return o, random_state(lat, 77).reshape(-1)


@pytest.fixture(scope="module")
def converged(oracle):
    o, s = _driven_lattice(oracle)
    T1 = 1.0
    dt = 2.0 ** -13
    # Boost integrate_const grid 0, dt, ..., T1 (2^13 steps, exact in binary)
    ref = o.integrate_md(s, 0.0, T1, dt)
    assert ref["times"][-1] == T1
    return ref["final"]


@pytest.mark.parametrize("method", sorted(ORDERS))
def test_global_order(oracle, converged, method):
    o, s = _driven_lattice(oracle)
    a = STEPS[method]
    errs = []
    for k in (a, a + 1):
        r = o.integrate_gpu(s, 0.0, 1.0, 2.0 ** -k, method)
        assert r["known"] and r["times"][-1] == 1.0
        errs.append(np.max(np.abs(r["final"] - converged)))
    assert errs[1] > 1e-12, f"{method}: error {errs[1]:.2e} too close to rounding to fit an order"
    p = math.log2(errs[0] / errs[1])
    assert abs(p - ORDERS[method]) < 0.35, f"{method}: observed order {p:.2f}"


@pytest.mark.parametrize("alias,name", [("midpoint", "rk2"), ("rkck54", "rk5"), ("rk54", "dopri5"),
                                        ("rkf54", "dopri5"), ("rkf78", "rk78"), ("bs", "bulirsch_stoer")])
def test_aliases(oracle, alias, name):
    o, s = _driven_lattice(oracle)
    a = o.integrate_gpu(s, 0.0, 0.5, 0.0625, alias)
    b = o.integrate_gpu(s, 0.0, 0.5, 0.0625, name)
    assert a["known"] and np.array_equal(a["final"], b["final"])


def test_unknown_method_runs_ssprk53(oracle):
    o, s = _driven_lattice(oracle)
    a = o.integrate_gpu(s, 0.0, 0.5, 0.0625, "no-such-method")
    b = o.integrate_gpu(s, 0.0, 0.5, 0.0625, "ssprk53")
    assert not a["known"] and b["known"]
    assert np.array_equal(a["final"], b["final"])


def test_rk4_matches_boost_order(oracle):
    """The GPU stepper's rk4 (x += dt/6 (k1 + 2k2 + 2k3 + k4)) and Boost's
    runge_kutta4 (x + dt/6 k1 + dt/3 k2 + ...) differ only by rounding."""
    o, s = _driven_lattice(oracle)
    dt = 2.0 ** -6
    g = o.integrate_gpu(s, 0.0, 1.0, dt, "rk4")
    b = o.integrate_md(s, 0.0, 1.0, dt)
    assert b["times"][-1] == 1.0
    assert np.max(np.abs(g["final"] - b["final"])) < 1e-13
    assert np.max(np.abs(g["mags"] - b["mags"])) < 1e-13


def test_loop_records(oracle):
    """integrate loop: observer at step 0, every save_interval steps and once
    after the loop, t accumulated by t += dt while t < T1."""
    o, s = _driven_lattice(oracle)
    dt, T1 = 0.1, 1.0
    r = o.integrate_gpu(s, 0.0, T1, dt, "euler", save_interval=3, want_states=True)
    t, n = 0.0, 0
    times = []
    while t < T1:
        if n % 3 == 0:
            times.append(t)
        t += dt
        n += 1
    times.append(t)
    assert n == 11  # accumulated 0.1 steps stay below 1.0 after ten additions
    assert np.array_equal(r["times"], np.array(times))
    assert np.array_equal(r["states"][0], s)
    assert np.array_equal(r["states"][-1], r["final"])
    # the recorded states are the loop's states: Euler replayed by hand with
    # the oracle's RHS (x + dt f(t, x) at the loop's accumulated times)
    x, t = s.copy(), 0.0
    for _ in range(3):
        x = x + dt * o.rhs(x, t)
        t += dt
    assert np.max(np.abs(r["states"][1] - x)) < 1e-15
    assert r["energy"][0] == pytest.approx(o.energy(s) / (len(s) // 3), rel=1e-14)


def test_mixed_stepper_order(oracle):
    from classical_mcmc_spin_amd import mixed_model as mm
    from test_mixed_oracle import random_flat, rich_cell
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 1, 2, 1)
    o = oracle.MixedOracle(muc, 1, 2, 1)
    x = random_flat(ml, 3)
    rng = np.random.default_rng(5)
    o.set_pulse(2, rng.normal(size=(4, 3)), 0.3, rng.normal(size=(4, 3)), 0.6, 0.4, 0.3, 1.1)
    o.set_pulse(3, rng.normal(size=(4, 8)), 0.3, rng.normal(size=(4, 8)), 0.6, 0.3, 0.4, 0.7)
    dt_ref = 2.0 ** -12
    ref = o.integrate_md(x, 0.0, 1.0, dt_ref)["final"]
    for method, a in (("dopri5", 4), ("ssprk53", 6), ("rk2", 7)):
        errs = [np.max(np.abs(o.integrate_gpu(x, 0.0, 1.0, 2.0 ** -k, method)["final"] - ref)) for k in (a, a + 1)]
        p = math.log2(errs[0] / errs[1])
        assert abs(p - ORDERS[method]) < 0.35, f"mixed {method}: observed order {p:.2f}"
    g = o.integrate_gpu(x, 0.0, 1.0, 2.0 ** -6, "rk4")
    b = o.integrate_md(x, 0.0, 1.0, 2.0 ** -6)
    assert np.max(np.abs(g["final"] - b["final"])) < 1e-13
    assert np.max(np.abs(g["records"] - b["records"])) < 1e-13
