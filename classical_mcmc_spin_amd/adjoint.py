"""Gradients of trajectory losses by the device discrete adjoint
(csl_adjoint_gradient in include/csl_gpu.h), and a torch.autograd.Function
over it for fitting Kitaev couplings to a target M(t).

The forward model is the reference's MD path (Landau-Lifshitz RHS,
lattice_md.cpp:39-153, Boost runge_kutta4 steps, the observer's
[M_antiferro, M_local, M_global] records, lattice_md.cpp:383-436); the
reference itself has no gradient path.

Parameters theta = (K, Gamma, Gamma', J, h) of build_kitaev_honeycomb
(unitcell_builders.cpp:178-283) enter the bond matrices and the field
linearly, so d(tables)/d theta_p are the tables of the same builder with
theta = e_p (and every other coupling zero).
"""
from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from . import _capi, model
from ._capi import check, dptr, f64, lib
from .gpu import GPULattice

KITAEV_PARAMS = ("K", "Gamma", "Gammap", "J", "h")


def kitaev_lattice(theta, L1, L2, field_direction=(1.0, 1.0, 1.0)) -> model.Lattice:
    K, G, Gp, J, h = (float(v) for v in theta)
    uc = model.build_kitaev_honeycomb(K=K, Gamma=G, Gammap=Gp, J=J, field_strength=h,
                                      field_direction=field_direction)
    return model.Lattice(uc, L1, L2, 1)


def kitaev_tangents(L1, L2, field_direction=(1.0, 1.0, 1.0)):
    """(dJ [5, nnz, 3, 3], dB [5, N, 3]) of theta = (K, Gamma, Gamma', J, h)
    in the Lattice's CSR bond order."""
    dJ, dB = [], []
    for p in range(5):
        e = np.zeros(5)
        e[p] = 1.0
        lat = kitaev_lattice(e, L1, L2, field_direction)
        dJ.append(lat.bi_J)
        dB.append(lat.field)
    return np.stack(dJ), np.stack(dB)


def adjoint_gradient(gl: GPULattice, init, t_pulse, pulse_amp, T0, T1, dt, target, weights, dJ, dB,
                     tgrid="boost", ckpt_interval=0, want_records=False, want_grad_init=False):
    """csl_adjoint_gradient: returns dict(loss [B], grad [B, P], records
    [n, B, 3, 3], grad_init [B, N, 3]) for the B trajectories of t_pulse."""
    tp = f64(t_pulse).reshape(-1, 2)
    B = tp.shape[0]
    init = f64(init)
    per = 1 if init.ndim == 3 else 0
    dJ = f64(dJ)
    dB = f64(dB)
    P = dJ.shape[0] if dJ.size else 0
    target = f64(target).reshape(-1, 9)
    loss = np.empty(B)
    grad = np.empty(B * max(P, 1))
    rec = np.empty(target.shape[0] * B * 9) if want_records else None
    gi = np.empty(B * gl.N * 3) if want_grad_init else None
    check(lib().csl_adjoint_gradient(gl.h, B, dptr(init.reshape(-1)), per, dptr(tp.reshape(-1)),
                                     dptr(None if pulse_amp is None else f64(pulse_amp).reshape(-1)), T0, T1, dt,
                                     _capi.TGRID[tgrid], dptr(target.reshape(-1)), target.shape[0],
                                     dptr(f64(weights)), P, dptr(dJ.reshape(-1)), dptr(dB.reshape(-1)),
                                     ckpt_interval, dptr(loss), dptr(grad), dptr(rec), dptr(gi)))
    out = {"loss": loss, "grad": grad[:B * P].reshape(B, P)}
    if want_records:
        out["records"] = rec.reshape(-1, B, 3, 3)
    if want_grad_init:
        out["grad_init"] = gi.reshape(B, gl.N, 3)
    return out


class KitaevFit:
    """Loss L_b = sum_t sum_j w_j (m_bj(t) - m*_j(t))^2 of B Kitaev parameter
    starts theta_b (rows of a [B, 5] array) on one L1 x L2 lattice, with one
    pulse (amplitude, width, frequency, centre), Gilbert damping alpha, a
    shared initial state and the Boost RK4 grid on [T0, T1]."""

    def __init__(self, L1, L2, init, target, T0, T1, dt, pulse=(0.5, 0.5, 0.0, 0.0), alpha=0.0,
                 weights=(0, 0, 0, 0, 0, 0, 1, 1, 1), field_direction=(1.0, 1.0, 1.0), ckpt_interval=0):
        self.L1, self.L2 = L1, L2
        self.fd = field_direction
        self.base = kitaev_lattice((-1.0, 0.25, -0.02, 0.0, 0.0), L1, L2, field_direction)
        self.gl = GPULattice(self.base)
        self.gl.set_damping(alpha)
        amp, width, freq, tc = pulse
        f = np.zeros((self.base.N_atoms, 3))
        f[:, 0] = 1.0
        d = np.einsum("aij,aj->ai", self.base.frames, f).reshape(-1)
        self.gl.set_pulse(np.concatenate([d, d]), amp, width, freq, tc, 0.0)
        self.amp, self.tc = amp, tc
        self.init = f64(init).reshape(-1)
        self.target = f64(target).reshape(-1, 9)
        self.T0, self.T1, self.dt = T0, T1, dt
        self.weights = f64(weights)
        self.dJ, self.dB = kitaev_tangents(L1, L2, field_direction)
        self.ckpt = ckpt_interval

    def close(self):
        self.gl.close()

    def __call__(self, theta):
        """(loss [B], grad [B, 5]) at the parameter rows of theta."""
        theta = np.atleast_2d(f64(theta))
        B = theta.shape[0]
        # the Hamiltonian is linear in theta: the per-start bond matrices and
        # fields are theta . (dJ, dB) -- the same numbers the builder gives
        # (every entry is a sum of parameters times 0, +-1 or a fixed factor)
        # without building a lattice per start
        J = np.tensordot(theta, self.dJ, 1)
        field = np.tensordot(theta, self.dB, 1)
        self.gl.set_replica_hamiltonians(J=J, field=field)
        try:
            r = adjoint_gradient(self.gl, self.init, np.tile([self.tc, 0.0], (B, 1)),
                                 np.tile([self.amp, 0.0], (B, 1)), self.T0, self.T1, self.dt, self.target,
                                 self.weights, self.dJ, self.dB, ckpt_interval=self.ckpt)
        finally:
            self.gl.set_replica_hamiltonians(None)
        return r["loss"], r["grad"]


def torch_loss(fit: KitaevFit):
    """torch.autograd.Function computing fit(theta) on the GPU: forward gives
    the loss of every parameter row, backward the adjoint gradient."""
    import torch

    class _KitaevAdjointLoss(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta):
            loss, grad = fit(theta.detach().cpu().double().numpy())
            ctx.save_for_backward(torch.from_numpy(grad).to(theta.device, theta.dtype))
            return torch.from_numpy(loss).to(theta.device, theta.dtype)

        @staticmethod
        def backward(ctx, grad_out):
            (g,) = ctx.saved_tensors
            return grad_out[:, None] * g

    return _KitaevAdjointLoss.apply
