"""Discrete-adjoint gradients on the GPU (csl_adjoint_gradient) against the
FP64 CPU tangent-linear oracle, central finite differences and an
initial-state dot-product check.

Bar: gradient relative error <= 1e-6 against the oracle tangent-linear
gradient and against central differences (step 1e-6) of the GPU loss; the
loss itself equals the oracle's forward records' loss to 1e-10 relative.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import adjoint as AD
from classical_mcmc_spin_amd import model, seeding
from classical_mcmc_spin_amd.gpu import GPULattice

pytestmark = pytest.mark.gpu

THETA_STAR = np.array([-1.0, 0.25, -0.02, 0.0, 0.05])
W = np.array([0, 0, 0, 0, 0, 0, 1.0, 1.0, 1.0])


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def _oracle_pair(po, theta, L1, L2, alpha, pulse):
    uc = model.build_kitaev_honeycomb(K=theta[0], Gamma=theta[1], Gammap=theta[2], J=theta[3],
                                      field_strength=theta[4], field_direction=(1.0, 1.0, 1.0))
    o = po.OracleLattice(uc, L1, L2, 1)
    o.set_alpha(alpha)
    amp, width, freq, tc = pulse
    f = np.zeros((2, 3))
    f[:, 0] = 1.0
    o.set_pulse(f, tc, f * 0.0, 0.0, amp, width, freq)
    return uc, o


def _unit_tangents():
    """per unit-cell bond dJ [n_bonds*9] and per-atom dB [2*3] of each parameter"""
    out = []
    for p in range(5):
        e = np.zeros(5)
        e[p] = 1.0
        uc = model.build_kitaev_honeycomb(K=e[0], Gamma=e[1], Gammap=e[2], J=e[3], field_strength=e[4],
                                          field_direction=(1.0, 1.0, 1.0))
        out.append((np.array([b.J for b in uc.bilinear]).reshape(-1), uc.field.reshape(-1)))
    return out


def _oracle_grad(po, theta, L1, L2, init, T0, T1, dt, target, alpha, pulse):
    uc, o = _oracle_pair(po, theta, L1, L2, alpha, pulse)
    g = np.zeros(5)
    loss = None
    for p, (dJ, dB) in enumerate(_unit_tangents()):
        m, dm = o.tangent_md(init.reshape(-1), np.zeros(init.size), dJ, dB, T0, T1, dt)
        d = m.reshape(-1, 9) - target
        loss = float(np.sum(W * d * d))
        g[p] = float(np.sum(2.0 * W * d * dm.reshape(-1, 9)))
    return loss, g


def _setup(po, L1, L2, n, dt, alpha, pulse, seed=5):
    lat = model.Lattice(model.build_kitaev_honeycomb(), L1, L2, 1)
    init = seeding.tilt(seeding.quench(lat, seeding.random_spins(lat.lattice_size, seed), sweeps=30), 0.05,
                        seed=seed + 1)
    T0 = 0.0
    T1 = T0 + (n + 0.5) * dt
    _, o = _oracle_pair(po, THETA_STAR, L1, L2, alpha, pulse)
    ref = o.integrate_md(init.reshape(-1), T0, T1, dt)
    target = ref["mags"].reshape(-1, 9)
    return init, T0, T1, target


def test_adjoint_small_vs_tangent_linear_and_fd(po):
    L1, L2, n, dt, alpha = 4, 4, 150, 0.01, 0.02
    pulse = (0.5, 0.5, 0.0, 0.3)
    init, T0, T1, target = _setup(po, L1, L2, n, dt, alpha, pulse)
    fit = AD.KitaevFit(L1, L2, init, target, T0, T1, dt, pulse=pulse, alpha=alpha, weights=W)
    rng = np.random.default_rng(3)
    thetas = THETA_STAR * (1.0 + 0.1 * rng.uniform(-1, 1, size=(3, 5)))
    thetas[:, 3] = rng.uniform(-0.05, 0.05, size=3)
    loss, grad = fit(thetas)
    for b in range(3):
        lo, go = _oracle_grad(po, thetas[b], L1, L2, init, T0, T1, dt, target, alpha, pulse)
        assert abs(loss[b] - lo) <= 1e-10 * max(lo, 1e-12)
        assert np.max(np.abs(grad[b] - go)) <= 1e-6 * np.max(np.abs(go))
    # central differences of the GPU loss, every parameter of start 0
    h = 1e-6
    for p in range(5):
        tp, tm = thetas[:1].copy(), thetas[:1].copy()
        tp[0, p] += h
        tm[0, p] -= h
        fd = (fit(tp)[0][0] - fit(tm)[0][0]) / (2 * h)
        assert abs(fd - grad[0, p]) <= 1e-6 * max(abs(grad[0, p]), np.max(np.abs(grad[0])) * 1e-3)
    fit.close()


def test_adjoint_initial_state_dot_product(po):
    """<dL/dx0, v> equals the central difference of L along v."""
    L1, L2, n, dt, alpha = 3, 4, 80, 0.01, 0.02
    pulse = (0.5, 0.5, 0.0, 0.2)
    init, T0, T1, target = _setup(po, L1, L2, n, dt, alpha, pulse, seed=9)
    lat = AD.kitaev_lattice(THETA_STAR * 1.05, L1, L2)
    with GPULattice(lat) as gl:
        gl.set_damping(alpha)
        f = np.zeros((2, 3))
        f[:, 0] = 1.0
        d = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
        gl.set_pulse(np.concatenate([d, d]), pulse[0], pulse[1], pulse[2], pulse[3], 0.0)
        dJ, dB = AD.kitaev_tangents(L1, L2)
        args = (np.array([[pulse[3], 0.0]]), np.array([[pulse[0], 0.0]]), T0, T1, dt, target, W, dJ, dB)
        r = AD.adjoint_gradient(gl, init, *args, want_grad_init=True)
        v = np.random.default_rng(4).normal(size=init.shape)
        eps = 1e-6
        lp = AD.adjoint_gradient(gl, init + eps * v, *args)["loss"][0]
        lm = AD.adjoint_gradient(gl, init - eps * v, *args)["loss"][0]
    fd = (lp - lm) / (2 * eps)
    an = float(np.sum(r["grad_init"][0] * v))
    assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-12)


def test_adjoint_checkpoint_interval_invariance(po):
    """The gradient does not depend on the checkpoint spacing (bit for bit:
    the recomputed segments repeat the forward arithmetic exactly)."""
    L1, L2, n, dt, alpha = 4, 4, 100, 0.01, 0.02
    pulse = (0.5, 0.5, 0.0, 0.3)
    init, T0, T1, target = _setup(po, L1, L2, n, dt, alpha, pulse, seed=11)
    th = THETA_STAR[None] * 1.07
    outs = []
    for ck in (0, 1, 7, 100):
        fit = AD.KitaevFit(L1, L2, init, target, T0, T1, dt, pulse=pulse, alpha=alpha, weights=W, ckpt_interval=ck)
        outs.append(fit(th))
        fit.close()
    for lo, g in outs[1:]:
        assert np.array_equal(lo, outs[0][0])
        assert np.array_equal(g, outs[0][1])


def test_adjoint_fit_full_size(po):
    """24x24 honeycomb, 2000 RK4 steps, alpha = 0.02, one pulse; 32 starts
    within +-10 % of the target parameters; two starts checked against the
    CPU tangent-linear gradient."""
    L1 = L2 = 24
    n, dt, alpha = 2000, 0.01, 0.02
    pulse = (0.5, 0.5, 0.0, 1.0)
    init, T0, T1, target = _setup(po, L1, L2, n, dt, alpha, pulse, seed=21)
    rng = np.random.default_rng(8)
    thetas = THETA_STAR * (1.0 + 0.1 * rng.uniform(-1, 1, size=(32, 5)))
    fit = AD.KitaevFit(L1, L2, init, target, T0, T1, dt, pulse=pulse, alpha=alpha, weights=W)
    loss, grad = fit(thetas)
    fit.close()
    assert np.all(np.isfinite(grad)) and np.all(loss > 0)
    for b in (0, 17):
        lo, go = _oracle_grad(po, thetas[b], L1, L2, init, T0, T1, dt, target, alpha, pulse)
        assert abs(loss[b] - lo) <= 1e-8 * lo
        assert np.max(np.abs(grad[b] - go)) <= 1e-6 * np.max(np.abs(go))


def test_torch_autograd_function():
    import torch
    L1, L2, n, dt = 3, 3, 60, 0.01
    lat = model.Lattice(model.build_kitaev_honeycomb(), L1, L2, 1)
    init = seeding.random_spins(lat.lattice_size, 2)
    target = np.zeros((n + 1, 9))
    fit = AD.KitaevFit(L1, L2, init, target, 0.0, (n + 0.5) * dt, dt, alpha=0.0)
    f = AD.torch_loss(fit)
    theta = torch.tensor([[-1.0, 0.25, -0.02, 0.0, 0.05], [-0.9, 0.2, -0.01, 0.01, 0.04]],
                         dtype=torch.float64, requires_grad=True)
    loss = f(theta)
    loss.sum().backward()
    l2, g2 = fit(theta.detach().numpy())
    assert np.allclose(loss.detach().numpy(), l2, rtol=0, atol=0)
    assert np.allclose(theta.grad.numpy(), g2, rtol=0, atol=0)
    fit.close()
