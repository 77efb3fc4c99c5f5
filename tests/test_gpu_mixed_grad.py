"""Gradient of a loss on the mixed-lattice records with respect to the scale
of the Fe-Tm (SU(2)-SU(3)) couplings (csl_mixed_scale_gradient).

The tangent records d rec / d theta are checked against central differences
of the device forward run at theta = 1 +- h (h = 1e-5: truncation ~h^2, far
below the 1e-6 relative bar), the loss gradient against central differences
of the oracle's forward runs with the scaled couplings, and at the full
8x8x8-cell size against device central differences.  The forward records of
the gradient run equal those of csl_mixed_integrate_batched on the stage
kernels bit for bit (same kernels), and those of its persistent kernel to
1e-12 (records summed in another order).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import gpu_mixed as GM
from classical_mcmc_spin_amd import mixed_model as mm
from test_mixed_oracle import random_flat, rich_cell

pytestmark = pytest.mark.gpu

H = 1e-5
REL = 1e-6


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _t1(T0, n, dt):
    return T0 + (n + 0.5) * dt


def _setup(gm, ml, damped=True):
    rng = np.random.default_rng(11)
    f2 = rng.normal(size=(4, 3))
    f3 = rng.normal(size=(4, 8))
    GM._set_pulses(gm, f2, 0.5 * f2, f3, 0.7 * f3, 0.0, 0.0, 0.6, 0.4, 0.8, 0.5, 0.3, 1.2)
    if damped:
        gamma = np.array([0.05, 0.0, 0.02, 0.0, 0.03, 0.0, 0.0, 0.01])
        eq = np.tile(mm.random_su3_states(1, 99)[0], (ml.N3, 1))
        gm.set_damping(0.03, gamma, eq)
    return f2, f3


def _records(ml_cell, dims, init, tp, sc, T0, T1, dt, damped=True):
    ml = mm.MixedLattice(ml_cell, *dims)
    with GM.GPUMixedLattice(ml) as gm:
        _setup(gm, ml, damped)
        return gm.integrate_batched(init, tp, sc, T0, T1, dt)["records"]


def test_tangent_records_vs_central_differences():
    muc = rich_cell()
    dims = (2, 2, 2)
    ml = mm.MixedLattice(muc, *dims)
    B = 3
    inits = np.stack([random_flat(ml, 70 + b) for b in range(B)])
    tp = np.array([[0.0, 0.0], [0.0, 0.15], [0.1, 0.3]])
    sc = np.array([[1.0, 0.0], [1.0, 1.0], [0.5, 2.0]])
    T0, dt, n = -0.3, 0.005, 200
    T1 = _t1(T0, n, dt)
    nr = n + 1
    with GM.GPUMixedLattice(ml) as gm:
        _setup(gm, ml)
        g = gm.scale_gradient(inits, tp, np.zeros((nr, 34)), np.r_[np.ones(33), 0.0], sc, T0, T1, dt,
                              want_records=True)
        fwd_persistent = gm.integrate_batched(inits, tp, sc, T0, T1, dt)["records"]
        assert gm.last_path()["persistent"]
        gm.set_persistent(-1)
        fwd = gm.integrate_batched(inits, tp, sc, T0, T1, dt)["records"]
    assert np.array_equal(g["records"], fwd)
    assert np.max(np.abs(fwd_persistent - fwd)) <= 1e-12
    rp = _records(mm.scale_fe_tm(muc, 1.0 + H), dims, inits, tp, sc, T0, T1, dt)
    rm = _records(mm.scale_fe_tm(muc, 1.0 - H), dims, inits, tp, sc, T0, T1, dt)
    fd = (rp - rm) / (2 * H)
    tr = g["trecords"]
    assert np.all(tr[..., 33] == 0.0)
    scale = np.max(np.abs(tr[..., :33]))
    assert scale > 1e-3
    assert np.max(np.abs(tr[..., :33] - fd[..., :33])) <= REL * scale + 1e-9
    assert np.all(tr[0] == 0.0)  # the initial state does not depend on the scale


def test_loss_gradient_vs_oracle_differences(po):
    """Loss on M_global of both species against a run at scale 1.05; the
    oracle's forward runs at 1 +- h give the reference central difference."""
    muc = rich_cell(seed=4)
    dims = (2, 2, 2)
    ml = mm.MixedLattice(muc, *dims)
    init = random_flat(ml, 90)
    T0, dt, n = -0.2, 0.005, 240
    T1 = _t1(T0, n, dt)
    f2, f3 = np.zeros((4, 3)), np.zeros((4, 8))
    f2[:, 0] = 1.0
    f3[:, 2] = 1.0
    w = np.zeros(34)
    w[6:9] = 1.0
    w[25:33] = 1.0

    def oracle_records(scale):
        o = po.MixedOracle(mm.scale_fe_tm(muc, scale), *dims)
        o.set_pulse(2, f2, 0.0, f2 * 0.0, 0.0, 0.5, 0.4, 0.0)
        o.set_pulse(3, f3, 0.0, f3 * 0.0, 0.0, 0.4, 0.4, 0.0)
        return o.integrate_md(init, T0, T1, dt)["records"]

    target = oracle_records(1.05)
    loss = lambda r: float(np.sum(w * (r - target) ** 2))
    fd = (loss(oracle_records(1.0 + H)) - loss(oracle_records(1.0 - H))) / (2 * H)
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2 * 0.0, f3, f3 * 0.0, 0.0, 0.0, 0.5, 0.4, 0.0, 0.4, 0.4, 0.0)
        g = gm.scale_gradient(init, np.zeros((1, 2)), target, w, np.array([[1.0, 0.0]]), T0, T1, dt)
    assert abs(g["loss"][0] - loss(oracle_records(1.0))) <= 1e-10 * max(1.0, abs(g["loss"][0]))
    assert abs(g["grad"][0] - fd) <= REL * abs(fd)
    assert abs(fd) > 1e-6


def test_full_size_gradient_vs_device_differences():
    """8x8x8 cells (2048 Fe + 2048 Tm), three delays of the two-pulse scan,
    150 steps: per-trajectory gradients against device central differences."""
    muc = mm.build_tmfeo3_mixed()
    dims = (8, 8, 8)
    ml = mm.MixedLattice(muc, *dims)
    init = random_flat(ml, 80)
    tp = np.array([[0.0, 0.0], [0.0, 0.2], [0.0, 0.4]])
    sc = np.array([[1.0, 0.0], [1.0, 1.0], [1.0, 1.0]])
    T0, dt, n = -0.5, 0.005, 150
    T1 = _t1(T0, n, dt)
    w = np.zeros(34)
    w[0:9] = 1.0
    w[9:33] = 0.5
    ref = _records(mm.scale_fe_tm(muc, 0.95), dims, init, tp, sc, T0, T1, dt, damped=False)
    with GM.GPUMixedLattice(ml) as gm:
        _setup(gm, ml, damped=False)
        g = gm.scale_gradient(init, tp, ref, w, sc, T0, T1, dt)
    lp = np.sum(w * (_records(mm.scale_fe_tm(muc, 1.0 + H), dims, init, tp, sc, T0, T1, dt, False) - ref) ** 2,
                axis=(0, 2))
    lm = np.sum(w * (_records(mm.scale_fe_tm(muc, 1.0 - H), dims, init, tp, sc, T0, T1, dt, False) - ref) ** 2,
                axis=(0, 2))
    fd = (lp - lm) / (2 * H)
    assert np.all(np.abs(fd) > 1e-8)
    assert np.max(np.abs(g["grad"] - fd) / np.abs(fd)) <= REL
