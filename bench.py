#!/usr/bin/env python3
"""Benchmark of the batched FP64 2DCS workload on one MI355X.

Workload (one "step" = one pass of the hot path over one batch):
  Kitaev honeycomb 36x36x1 (2592 SU(2) spins), K=-1, Gamma=0.25, Gamma'=-0.02,
  static h = 0.1 along [111] (unit direction), Gaussian pulses A=0.5, w=0.5
  along global x, omega = 0; pulse 1 at t=0, delays tau = 0, 0.25, ..., 63.75
  (256 points).  M0 + 256 M1(tau) + 256 M01(tau) = 513 trajectories of 8192
  RK4 steps (dt = 0.01, t in [-10, 71.92]) from one seeded relaxed state tilted
  by N(0, 0.01) noise and renormalised; M(t) recorded every step.

Prints one JSON line (rank 0).  `value` = trajectories/s of the batched
persistent kernel with the inputs and the M(t) records resident in HBM (the
timed steps do not copy the 302 MB of records to the host); the rate with that
copy (pcie_inclusive_traj_per_s), the plain one-trajectory path, the oracle CPU
baseline and the error against the oracle are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LL RK4 site-steps/s + 2DCS trajectories/s, batched vs plain, at fixed FP64 error"
HBM_PEAK = 8.0e12          # MI355X HBM3E, B/s
BYTES_PER_SITE_STEP = 192  # 4 RK stages x (read + write) x 3 doubles x 8 B
FLOPS_PER_SITE_STEP = 306  # z = 3 SU(2) site, RK4

L1 = L2 = 36
DT = 0.01
T0 = -10.0
N_STEPS = 8192
T1 = T0 + (N_STEPS + 0.5) * DT     # Boost grid: exactly N_STEPS steps
TAU0, TAU_STEP, N_TAU = 0.0, 0.25, 256
AMP, WIDTH, FREQ = 0.5, 0.5, 0.0


def build_workload():
    from classical_mcmc_spin_amd import workloads
    return workloads.kitaev_2dcs(L1, L2)


def replica_tables(n_tau):
    from classical_mcmc_spin_amd import workloads
    taus = TAU0 + TAU_STEP * np.arange(n_tau)
    tp, am = workloads.pump_probe_replicas(taus, AMP)
    return tp, am, taus


def cpu_baseline(uc, init, field_in, n_tau_sample):
    """Oracle (FP64 C++/OpenMP restatement, outer OpenMP over tau with serial
    inner RHS as the reference's W2 path) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.OracleLattice(uc, L1, L2, 1)
    threads = pyoracle.max_threads()
    t = time.perf_counter()
    o.pump_probe(init.reshape(-1), field_in, AMP, WIDTH, FREQ, TAU0, TAU0 + TAU_STEP * (n_tau_sample - 1),
                 TAU_STEP, T0, T1, DT, reuse_m0=False, n_outer=threads)
    el = time.perf_counter() - t
    n_traj = 1 + 2 * n_tau_sample
    return {"value": n_traj / el, "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": f"{n_traj} trajectories (M0 + {n_tau_sample} M1 + {n_tau_sample} M01) x {N_STEPS} "
                      f"RK4 steps of the same 36x36 lattice, {el:.1f} s on {threads} OpenMP threads "
                      f"({platform.processor() or platform.machine()})",
            "site_steps_per_s": n_traj * N_STEPS * L1 * L2 * 2 / el}


def parity_check(uc, lat, gl, init, field_in, n_steps=500):
    """Short-horizon error of the batched kernel against the oracle (3
    trajectories: M0, M1(tau=0.25), M01(tau=0.25))."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.OracleLattice(uc, L1, L2, 1)
    Tend = T0 + (n_steps + 0.5) * DT
    tp = np.array([[0.0, 0.0], [0.25, 0.0], [0.0, 0.25]])
    am = np.array([[AMP, 0.0], [AMP, 0.0], [AMP, AMP]])
    res = gl.integrate_batched(init, tp, am, T0, Tend, DT, tgrid="boost", want_final=True)
    err_m = err_s = 0.0
    f = field_in
    for b in range(3):
        o.set_pulse(f * (am[b, 0] != 0), tp[b, 0], f * (am[b, 1] != 0), tp[b, 1], AMP, WIDTH, FREQ)
        ref = o.integrate_md(init.reshape(-1), T0, Tend, DT)
        err_m = max(err_m, float(np.max(np.abs(res["mags"][:, b] - ref["mags"]))))
        err_s = max(err_s, float(np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3)))))
    return err_m, err_s


PYRO_L, PYRO_B, PYRO_STEPS, PYRO_DT = 16, 64, 4096, 0.02


def bench_pyrochlore(args):
    """64 samples x 4096 RK4 steps (dt = 0.02) of the 16^3-cell conventional
    pyrochlore lattice (65 536 sites, z = 6), all-in-all-out + tilt starts,
    M(t) and E(t) every step.  The lattice is too large for one workgroup's
    LDS, so the grid-wide stage kernels run (4 launches per step)."""
    from classical_mcmc_spin_amd import GPULattice, workloads
    lat = workloads.pyrochlore_lattice(PYRO_L)
    N = lat.lattice_size
    init, _ = workloads.pyrochlore_tilted_aiao(lat, PYRO_B, seed=1234)
    T1p = (PYRO_STEPS + 0.5) * PYRO_DT
    zero = np.zeros((PYRO_B, 2))
    gl = GPULattice(lat)

    def one_step():
        return gl.integrate_batched(init, zero, zero, 0.0, T1p, PYRO_DT, tgrid="boost")

    for _ in range(args.warmup):
        one_step()
    kernel_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        res = one_step()
        kernel_ms.append(gl.last_kernel_ms()[0])
    el = time.perf_counter() - t
    assert res["mags"].shape[0] == PYRO_STEPS + 1 and np.all(np.isfinite(res["mags"]))
    gl.close()
    kms = float(np.mean(kernel_ms))
    site_steps = PYRO_B * N * PYRO_STEPS
    samples_s = PYRO_B * args.steps / el
    bytes_launch = PYRO_B * N * BYTES_PER_SITE_STEP / 4       # one RK stage of every site
    line = {
        "metric": METRIC, "value": samples_s, "unit": "trajectories/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"pyrochlore_16x16x16_conventional_64samples_{PYRO_STEPS}steps_rk4_dt0.02",
                   "sites": N, "trajectories": PYRO_B, "rk4_steps": PYRO_STEPS, "record_every": 1},
        "site_steps_per_s": samples_s * N * PYRO_STEPS,
        "roofline": {"bound": "hbm", "achieved": bytes_launch / (kms * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": bytes_launch / (kms * 1e-3) / HBM_PEAK, "traffic": None,
                     "kernel": "k_stage", "kernel_ms": kms, "algorithmic_bytes_per_launch": bytes_launch},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.OracleLattice(lat.unit_cell, PYRO_L, PYRO_L, PYRO_L)
        n = 60
        t = time.perf_counter()
        o.integrate_md(init[0].reshape(-1), 0.0, (n + 0.5) * PYRO_DT, PYRO_DT)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / el / PYRO_STEPS, "unit": "trajectories/s", "cores": pyoracle.max_threads(),
            "kind": "port",
            "sample": f"1 sample x {n} RK4 steps of the same lattice ({el:.1f} s, OpenMP over sites "
                      f"inside each RHS as the reference's MD path)",
            "site_steps_per_s": n * N / el}
    print(json.dumps(line), flush=True)


MD_L, MD_STEPS, MD_DT = 24, 10000, 0.01


def bench_md(args):
    """Plain one-trajectory MD: Kitaev 24x24x1 (1152 sites), K=-1, G=0.25,
    G'=-0.02, h = 0, seeded random start, 10^4 RK4 steps of dt = 0.01, M(t)
    and E(t) every step.  value = trajectories/s of the fastest device engine;
    the three engines (grid-wide stage kernels launched eagerly, the same
    launch sequence replayed from a hipGraph, one persistent workgroup) are
    reported beside it."""
    from classical_mcmc_spin_amd import GPULattice, model, seeding
    uc = model.build_kitaev_honeycomb(K=-1.0, Gamma=0.25, Gammap=-0.02)
    lat = model.Lattice(uc, MD_L, MD_L, 1)
    s0 = seeding.random_spins(lat.lattice_size, 1234)
    T1m = (MD_STEPS + 0.5) * MD_DT
    gl = GPULattice(lat)
    eng = {}
    for name, e in (("stage_kernels", 1), ("stage_kernels_hipgraph", 3), ("persistent_workgroup", 2)):
        for _ in range(max(args.warmup, 1)):
            gl.set_spins(s0)
            gl.integrate(0.0, T1m, MD_DT, 1, tgrid="boost", engine=e)
        t = time.perf_counter()
        for _ in range(args.steps):
            gl.set_spins(s0)
            r = gl.integrate(0.0, T1m, MD_DT, 1, tgrid="boost", engine=e)
        el = (time.perf_counter() - t) / args.steps
        assert r["mags"].shape[0] == MD_STEPS + 1 and np.all(np.isfinite(r["mags"]))
        eng[name] = {"traj_per_s": 1.0 / el, "ms_per_traj": 1e3 * el,
                     "site_steps_per_s": MD_STEPS * lat.lattice_size / el, "kernel_ms": gl.last_kernel_ms()[0]}
    gl.close()
    best = max(eng, key=lambda k: eng[k]["traj_per_s"])
    v = eng[best]["traj_per_s"]
    line = {
        "metric": METRIC, "value": v, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 / v, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "md_kitaev_honeycomb_24x24x1_1traj_10000steps_rk4_dt0.01", "sites": lat.lattice_size,
                   "rk4_steps": MD_STEPS, "record_every": 1, "engine": best},
        "site_steps_per_s": eng[best]["site_steps_per_s"], "engines": eng,
        "roofline": {"bound": "hbm", "achieved": v * MD_STEPS * lat.lattice_size * BYTES_PER_SITE_STEP / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": v * MD_STEPS * lat.lattice_size * BYTES_PER_SITE_STEP / HBM_PEAK, "traffic": None,
                     "kernel": best},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.OracleLattice(uc, MD_L, MD_L, 1)
        n = 2000
        t = time.perf_counter()
        o.integrate_md(s0.reshape(-1), 0.0, (n + 0.5) * MD_DT, MD_DT)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / el / MD_STEPS, "unit": "trajectories/s", "cores": pyoracle.max_threads(), "kind": "port",
            "sample": f"{n} of the {MD_STEPS} RK4 steps ({el:.1f} s, OpenMP over sites inside each RHS as the "
                      f"reference's MD path)",
            "site_steps_per_s": n * lat.lattice_size / el}
    print(json.dumps(line), flush=True)


DP_T1, DP_TOL, DP_B = 100.0, 1e-6, 256


def bench_md_dopri5(args):
    """MD with the reference's default integrator (md_integrator "dopri5",
    md_abs_tol = md_rel_tol = 1e-6, spin_config.h:93,111-112):
    integrate_adaptive of the Kitaev 24x24x1 lattice (1152 sites) from t = 0
    to 100 (the horizon of the RK4 MD workload), M(t) and E(t) at every
    accepted step.  value = trajectories/s of one trajectory on the
    Dormand-Prince kernel; a batch of 256 trajectories (one workgroup per CU,
    seeded starts) is reported beside it."""
    from classical_mcmc_spin_amd import GPULattice, model, seeding
    uc = model.build_kitaev_honeycomb(K=-1.0, Gamma=0.25, Gammap=-0.02)
    lat = model.Lattice(uc, MD_L, MD_L, 1)
    s0 = seeding.random_spins(lat.lattice_size, 1234)
    gl = GPULattice(lat)
    run = lambda init: gl.integrate_dopri5(init, 0.0, DP_T1, MD_DT, control=1, abs_tol=DP_TOL, rel_tol=DP_TOL)
    for _ in range(max(args.warmup, 1)):
        r = run(s0)
    t = time.perf_counter()
    for _ in range(args.steps):
        r = run(s0)
    el = (time.perf_counter() - t) / args.steps
    kms = gl.last_kernel_ms()[0]
    acc, rej = r["accepted"][0], r["rejected"][0]
    assert np.all(np.isfinite(r["mags"][0]))
    inits = np.stack([seeding.random_spins(lat.lattice_size, 5000 + b) for b in range(DP_B)])
    rb = run(inits)
    t = time.perf_counter()
    rb = run(inits)
    elb = time.perf_counter() - t
    kmsb = gl.last_kernel_ms()[0]
    gl.close()
    v = 1.0 / el
    line = {
        "metric": METRIC, "value": v, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "md_kitaev_honeycomb_24x24x1_1traj_t100_dopri5_adaptive_tol1e-6",
                   "sites": lat.lattice_size, "T": DP_T1, "abs_tol": DP_TOL, "rel_tol": DP_TOL,
                   "record_every": "accepted step"},
        "accepted_steps": acc, "rejected_steps": rej, "kernel_ms": kms,
        "site_steps_per_s": acc * lat.lattice_size / el,
        "batched": {"trajectories": DP_B, "traj_per_s": DP_B / elb, "kernel_ms": kmsb,
                    "accepted_steps_mean": float(np.mean(rb["accepted"])),
                    "rejected_steps_mean": float(np.mean(rb["rejected"]))},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.OracleLattice(uc, MD_L, MD_L, 1)
        runs, t = 0, time.perf_counter()
        while True:  # whole trajectories, repeated for about 3 s
            ro = o.integrate_dopri5(s0.reshape(-1), 0.0, DP_T1, MD_DT, control=1, abs_tol=DP_TOL, rel_tol=DP_TOL,
                                    max_rec=4096)
            runs += 1
            elc = time.perf_counter() - t
            if elc >= 3.0:
                break
        line["cpu_baseline"] = {
            "value": runs / elc, "unit": "trajectories/s", "cores": pyoracle.max_threads(), "kind": "port",
            "sample": f"{runs} whole trajectories t = 0..{DP_T1:g} ({ro['accepted']} accepted steps each, {elc:.1f} s, "
                      f"OpenMP over sites inside each RHS)",
            "site_steps_per_s": runs * ro["accepted"] * lat.lattice_size / elc}
    print(json.dumps(line), flush=True)


ADJ_L, ADJ_STEPS, ADJ_B, ADJ_B_FULL = 24, 2000, 32, 256


def bench_adjoint(args):
    """Adjoint gradients: Kitaev 24x24 (1152 sites), one pulse (A = 0.5,
    w = 0.5 at t = 1), Gilbert alpha = 0.02, loss on M_global over 2000 RK4
    steps (dt = 0.01) against a target run at (K, G, G', J, h) =
    (-1, 0.25, -0.02, 0, 0.05); 32 parameter starts within +-10 %.
    value = gradients (starts) per second.  One workgroup per start and one
    per CU (the LDS state of 1152 sites takes 83 KB), so 32 starts use 32 of
    the 256 CUs; a batch of 256 starts is reported beside it."""
    from classical_mcmc_spin_amd import adjoint as AD
    from classical_mcmc_spin_amd import GPULattice, model, seeding
    lat = model.Lattice(model.build_kitaev_honeycomb(), ADJ_L, ADJ_L, 1)
    init = seeding.tilt(seeding.quench(lat, seeding.random_spins(lat.lattice_size, 21), sweeps=30), 0.05, seed=22)
    dt = 0.01
    T1a = (ADJ_STEPS + 0.5) * dt
    star = np.array([-1.0, 0.25, -0.02, 0.0, 0.05])
    tl = AD.kitaev_lattice(star, ADJ_L, ADJ_L)
    with GPULattice(tl) as g0:
        g0.set_damping(0.02)
        fdir = np.zeros((2, 3))
        fdir[:, 0] = 1.0
        d = np.einsum("aij,aj->ai", tl.frames, fdir).reshape(-1)
        g0.set_pulse(np.concatenate([d, d]), 0.5, 0.5, 0.0, 1.0, 0.0)
        tr = g0.integrate_batched(init, np.array([[1.0, 0.0]]), np.array([[0.5, 0.0]]), 0.0, T1a, dt,
                                  tgrid="boost", want_energy=False)
    target = tr["mags"][:, 0].reshape(-1, 9)
    rng = np.random.default_rng(8)
    thetas_full = star * (1.0 + 0.1 * rng.uniform(-1, 1, size=(ADJ_B_FULL, 5)))
    thetas = thetas_full[:ADJ_B]
    fit = AD.KitaevFit(ADJ_L, ADJ_L, init, target, 0.0, T1a, dt, pulse=(0.5, 0.5, 0.0, 1.0), alpha=0.02)
    for _ in range(args.warmup):
        fit(thetas)
    t = time.perf_counter()
    for _ in range(args.steps):
        loss, grad = fit(thetas)
    el = (time.perf_counter() - t) / args.steps
    kms = fit.gl.last_kernel_ms()[0]
    assert np.all(np.isfinite(grad))
    fit(thetas_full)
    t = time.perf_counter()
    _, grad_full = fit(thetas_full)
    el_full = time.perf_counter() - t
    kms_full = fit.gl.last_kernel_ms()[0]
    fit.close()
    assert np.all(np.isfinite(grad_full))
    v = ADJ_B / el
    # forward-equivalent site-steps: forward + recomputation + transposed steps
    line = {
        "metric": METRIC, "value": v, "unit": "gradients/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"adjoint_kitaev_honeycomb_{ADJ_L}x{ADJ_L}x1_{ADJ_B}starts_{ADJ_STEPS}steps_rk4_dt0.01",
                   "sites": lat.lattice_size, "starts": ADJ_B, "rk4_steps": ADJ_STEPS, "parameters": 5},
        "site_steps_per_s": v * ADJ_STEPS * lat.lattice_size, "kernel_ms": kms,
        "roofline": {"bound": "hbm",
                     "achieved": v * ADJ_STEPS * lat.lattice_size * BYTES_PER_SITE_STEP / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": v * ADJ_STEPS * lat.lattice_size * BYTES_PER_SITE_STEP / HBM_PEAK, "traffic": None,
                     "kernel": "k_adjoint"},
        "starts_256": {"gradients_per_s": ADJ_B_FULL / el_full, "kernel_ms": kms_full},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        uc = model.build_kitaev_honeycomb(K=thetas[0, 0], Gamma=thetas[0, 1], Gammap=thetas[0, 2], J=thetas[0, 3],
                                          field_strength=thetas[0, 4], field_direction=(1.0, 1.0, 1.0))
        o = pyoracle.OracleLattice(uc, ADJ_L, ADJ_L, 1)
        o.set_alpha(0.02)
        o.set_pulse(fdir, 1.0, fdir * 0.0, 0.0, 0.5, 0.5, 0.0)
        eK = model.build_kitaev_honeycomb(K=1.0, Gamma=0.0, Gammap=0.0)
        dJ = np.array([b.J for b in eK.bilinear]).reshape(-1)
        n = 400
        t = time.perf_counter()
        o.tangent_md(init.reshape(-1), np.zeros(init.size), dJ, np.zeros(6), 0.0, (n + 0.5) * dt, dt)
        el = time.perf_counter() - t
        # a 5-parameter gradient by tangent-linear sweeps costs 5 such runs
        line["cpu_baseline"] = {
            "value": n / (5 * el * ADJ_STEPS), "unit": "gradients/s", "cores": pyoracle.max_threads(),
            "kind": "port",
            "sample": f"one tangent-linear sweep over {n} of the {ADJ_STEPS} steps ({el:.1f} s); a 5-parameter "
                      f"gradient needs 5 sweeps",
            "site_steps_per_s": n * lat.lattice_size / el}
    print(json.dumps(line), flush=True)


MIX_L, MIX_NTAU, MIX_STEPS, MIX_DT = 8, 128, 8192, 0.005


MIXDP_INTERVALS, MIXDP_TOL = 2048, 1e-8


def bench_mixed_dopri5(args):
    """The mixed 2DCS batch of bench_mixed (257 trajectories, 8x8x8 cells)
    with the reference's default pulse-drive integrator: Dormand-Prince 5(4)
    through integrate_const with step control at T_step = 0.005, tolerance
    1e-8 (kDefaultPumpProbeAbsTol / RelTol), W3 pulse-window segments per
    trajectory (windows of 9 widths around its pulse centres), over 2048
    observation intervals from t = -1.  Every trajectory runs its own step
    sequence; value = trajectories/s."""
    from classical_mcmc_spin_amd import gpu as G
    from classical_mcmc_spin_amd import gpu_mixed as GM
    from classical_mcmc_spin_amd import mixed_model as mm
    from classical_mcmc_spin_amd import seeding
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, MIX_L, MIX_L, MIX_L)
    g = ml.flat_state(seeding.random_spins(ml.N2, 1234), mm.random_su3_states(ml.N3, 4321))
    taus = 0.05 * np.arange(MIX_NTAU)
    B = 1 + 2 * MIX_NTAU
    tp = np.zeros((B, 2))
    sc = np.zeros((B, 2))
    sc[0, 0] = 1.0
    tp[1:1 + MIX_NTAU, 0] = taus
    sc[1:1 + MIX_NTAU, 0] = 1.0
    tp[1 + MIX_NTAU:, 1] = taus
    sc[1 + MIX_NTAU:, :] = 1.0
    centers = [[0.0]] + [[t] for t in taus] + [[0.0, t] for t in taus]
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    T0 = -1.0
    T1 = T0 + (MIXDP_INTERVALS + 0.5) * MIX_DT
    segs = [G.pulse_segments(T0, T1, c, 9.0 * WIDTH, MIX_DT) for c in centers]
    gm = GM.GPUMixedLattice(ml)
    GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, AMP, WIDTH, 0.0, AMP, WIDTH, 0.0)

    def one_step():
        return gm.integrate_dopri5(g, tp, sc, T0, T1, MIX_DT, control=2, abs_tol=MIXDP_TOL, rel_tol=MIXDP_TOL,
                                   segments=segs)

    for _ in range(args.warmup):
        one_step()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = one_step()
    el = (time.perf_counter() - t) / args.steps
    kms, launches = gm.last_kernel_ms()
    gm.close()
    acc = np.array(r["accepted"])
    assert all(np.all(np.isfinite(x)) for x in r["records"])
    n_sites = ml.N2 + ml.N3
    line = {
        "metric": METRIC, "value": B / el, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * el, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"2dcs_tmfeo3_mixed_8x8x8_257traj_{MIXDP_INTERVALS}x0.005_dopri5_tol1e-8_w3",
                   "sites_su2": ml.N2, "sites_su3": ml.N3, "trajectories": B, "intervals": MIXDP_INTERVALS,
                   "abs_tol": MIXDP_TOL, "rel_tol": MIXDP_TOL, "segments_max": max(len(q) for q in segs)},
        "accepted_steps_mean": float(acc.mean()), "rejected_steps_mean": float(np.mean(r["rejected"])),
        "launch_ms_mean": kms, "launches": launches,
        "site_steps_per_s": float(acc.sum()) * n_sites / el,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.MixedOracle(muc, MIX_L, MIX_L, MIX_L)
        b = 1 + MIX_NTAU + 20  # an M01 trajectory
        o.set_pulse(2, f2, tp[b, 0], f2, tp[b, 1], AMP, WIDTH, 0.0)
        o.set_pulse(3, f3, tp[b, 0], f3, tp[b, 1], AMP, WIDTH, 0.0)
        n = 200
        Tc = T0 + (n + 0.5) * MIX_DT
        seg_c = G.pulse_segments(T0, Tc, centers[b], 9.0 * WIDTH, MIX_DT)
        t = time.perf_counter()
        ro = o.integrate_dopri5(g, T0, Tc, MIX_DT, control=2, abs_tol=MIXDP_TOL, rel_tol=MIXDP_TOL, segments=seg_c)
        elc = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / elc / MIXDP_INTERVALS, "unit": "trajectories/s", "cores": pyoracle.max_threads(),
            "kind": "port",
            "sample": f"1 trajectory x {n} of the {MIXDP_INTERVALS} intervals ({ro['accepted']} accepted steps, "
                      f"{elc:.1f} s, OpenMP over sites inside each RHS)"}
    print(json.dumps(line), flush=True)


def bench_mixed(args):
    """Mixed SU(2)+SU(3) 2DCS batch: TmFeO3-like 8x8x8 cells (2048 Fe SU(2)
    + 2048 Tm SU(3) sites), pulses A = 0.5, w = 0.5 along x (Fe) and
    lambda_2 (Tm) at 0 and tau = 0, 0.05, ..., 6.35 (128 delays):
    M0 + 128 M1 + 128 M01 = 257 trajectories x 8192 RK4 steps (dt = 0.005),
    records every step.  Grid-wide stage kernels (the 180 KB state exceeds
    one workgroup's LDS)."""
    from classical_mcmc_spin_amd import gpu_mixed as GM
    from classical_mcmc_spin_amd import mixed_model as mm
    from classical_mcmc_spin_amd import seeding
    muc = mm.build_tmfeo3_mixed()
    ml = mm.MixedLattice(muc, MIX_L, MIX_L, MIX_L)
    g = ml.flat_state(seeding.random_spins(ml.N2, 1234), mm.random_su3_states(ml.N3, 4321))
    taus = 0.05 * np.arange(MIX_NTAU)
    B = 1 + 2 * MIX_NTAU
    tp = np.zeros((B, 2))
    sc = np.zeros((B, 2))
    sc[0, 0] = 1.0
    tp[1:1 + MIX_NTAU, 0] = taus
    sc[1:1 + MIX_NTAU, 0] = 1.0
    tp[1 + MIX_NTAU:, 1] = taus
    sc[1 + MIX_NTAU:, :] = 1.0
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    T0 = -1.0
    T1m = T0 + (MIX_STEPS + 0.5) * MIX_DT
    gm = GM.GPUMixedLattice(ml)
    GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, AMP, WIDTH, 0.0, AMP, WIDTH, 0.0)

    def one_step():
        return gm.integrate_batched(g, tp, sc, T0, T1m, MIX_DT)

    for _ in range(args.warmup):
        one_step()
    kernel_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        res = one_step()
        kernel_ms.append(gm.last_kernel_ms()[0])
    el = time.perf_counter() - t
    assert res["records"].shape[0] == MIX_STEPS + 1 and np.all(np.isfinite(res["records"]))
    kms = float(np.mean(kernel_ms))
    # d loss / d (Fe-Tm coupling scale) for every trajectory of the batch:
    # forward and tangent-linear run in the same stage launches; loss on the
    # 33 magnetisation entries of the records against the M0 records x 1.01
    target = res["records"][:, 0] * 1.01
    w = np.r_[np.ones(33), 0.0]
    t = time.perf_counter()
    gr = gm.scale_gradient(g, tp, target, w, sc, T0, T1m, MIX_DT)
    el_g = time.perf_counter() - t
    kms_g = gm.last_kernel_ms()[0]
    assert np.all(np.isfinite(gr["grad"]))
    gm.close()
    traj_s = B * args.steps / el
    n_sites = ml.N2 + ml.N3
    # one stage: SU(2) sites read+write 3 doubles, SU(3) sites 8 doubles
    bytes_launch = B * (ml.N2 * 3 + ml.N3 * 8) * 2 * 8
    line = {
        "metric": METRIC, "value": traj_s, "unit": "trajectories/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"2dcs_tmfeo3_mixed_8x8x8_257traj_{MIX_STEPS}steps_rk4_dt0.005",
                   "sites_su2": ml.N2, "sites_su3": ml.N3, "trajectories": B, "rk4_steps": MIX_STEPS,
                   "record_every": 1},
        "site_steps_per_s": traj_s * n_sites * MIX_STEPS,
        "roofline": {"bound": "hbm", "achieved": bytes_launch / (kms * 1e-3) / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": bytes_launch / (kms * 1e-3) / HBM_PEAK, "traffic": None,
                     "kernel": "k_mixed_stage", "kernel_ms": kms, "algorithmic_bytes_per_launch": bytes_launch},
        "scale_gradient": {"gradients_per_s": B / el_g, "ms_per_batch": 1e3 * el_g,
                           "kernel_ms_per_launch": kms_g, "forward_plus_tangent_over_forward": el_g / (el / args.steps),
                           "grad_first_last": [float(gr["grad"][0]), float(gr["grad"][-1])]},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        o = pyoracle.MixedOracle(muc, MIX_L, MIX_L, MIX_L)
        o.set_pulse(2, f2, 0.0, f2 * 0.0, 0.0, AMP, WIDTH, 0.0)
        o.set_pulse(3, f3, 0.0, f3 * 0.0, 0.0, AMP, WIDTH, 0.0)
        n = 400
        t = time.perf_counter()
        o.integrate_md(g, T0, T0 + (n + 0.5) * MIX_DT, MIX_DT)
        el = time.perf_counter() - t
        line["cpu_baseline"] = {
            "value": n / el / MIX_STEPS, "unit": "trajectories/s", "cores": pyoracle.max_threads(),
            "kind": "port",
            "sample": f"1 trajectory x {n} RK4 steps of the same lattice ({el:.1f} s, OpenMP over sites "
                      f"inside each RHS as the reference's mixed MD path)",
            "site_steps_per_s": n * n_sites / el}
        # the CPU gradient by central differences: two forward runs per trajectory
        line["scale_gradient"]["cpu_baseline"] = {
            "value": n / el / MIX_STEPS / 2.0, "unit": "gradients/s", "cores": pyoracle.max_threads(),
            "kind": "port", "sample": "central difference = 2 forward runs, timed on the sample above"}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-tau", type=int, default=6, help="tau points in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--launch", default="", help="fixed persistent launch shape 'block,spt,dbuf' "
                                                   "instead of the autotuner's choice")
    ap.add_argument("--workload", default="2dcs",
                    choices=["2dcs", "pyrochlore", "mixed", "mixed_dopri5", "md", "md_dopri5", "adjoint"],
                    help="2dcs: the batched 2DCS workload (default); pyrochlore: 64 samples of the "
                         "65 536-site pyrochlore lattice; mixed: 257-trajectory SU(2)+SU(3) 2DCS batch "
                         "(both on the grid-wide stage kernels)")
    ap.add_argument("--rk-steps", type=int, default=0,
                    help="RK4 steps of the pyrochlore / mixed integrations (default 4096 / 8192); "
                         "shorter runs keep counter-collection passes within a few hundred dispatches")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed batched launches (for rocprofv3 runs): no autotune sweep, "
                         "plain paths, parity check or CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the batched 2DCS "
              "trajectories shard by batch across GPUs with no exchange; multi-GPU runs are not "
              "part of this build.", file=sys.stderr)
        sys.exit(2)

    global PYRO_STEPS, MIX_STEPS
    if args.rk_steps > 0:
        PYRO_STEPS = MIX_STEPS = args.rk_steps
    if args.workload == "pyrochlore":
        return bench_pyrochlore(args)
    if args.workload == "mixed":
        return bench_mixed(args)
    if args.workload == "mixed_dopri5":
        return bench_mixed_dopri5(args)
    if args.workload == "md":
        return bench_md(args)
    if args.workload == "md_dopri5":
        return bench_md_dopri5(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    from classical_mcmc_spin_amd import GPULattice

    uc, lat, init, field_in = build_workload()
    tp, am, taus = replica_tables(N_TAU)
    B = tp.shape[0]
    N = lat.lattice_size
    gl = GPULattice(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, field_in).reshape(-1)
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.0)

    if args.launch:
        blk, spt, db = (int(v) for v in args.launch.split(","))
        gl.set_launch_config(B, blk, spt, db)
        tuned_block, tuned_ms = blk, 0.0
    elif args.profile:
        tuned_block, tuned_ms = args.block, 0.0
    else:
        tuned_block, tuned_ms = gl.autotune(B, 64)
        print(f"[bench] autotuned block {tuned_block} ({tuned_ms:.2f} ms / 64 steps)", file=sys.stderr,
              flush=True)

    def one_step(records_to_host):
        return gl.integrate_batched(init, tp, am, T0, T1, DT, tgrid="boost", block=args.block,
                                    want_energy=False, want_mags=records_to_host)

    # warm-up steps bring the records back and check them; the timed steps
    # leave M(t) in HBM (inputs and outputs device-resident), and one more
    # step with the 302 MB record copy gives the PCIe-inclusive rate
    for _ in range(max(args.warmup, 1)):
        res = one_step(True)
    n_rec = res["mags"].shape[0]
    assert n_rec == N_STEPS + 1 and np.all(np.isfinite(res["mags"]))
    kernel_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        one_step(False)
        kernel_ms.append(gl.last_kernel_ms()[0])
    el = time.perf_counter() - t
    ms_step = 1e3 * el / args.steps
    traj_s = B * args.steps / el
    t = time.perf_counter()
    one_step(True)
    pcie_traj_s = B / (time.perf_counter() - t)

    # roofline of the dominant kernel (k_persistent, one launch per step)
    kms = float(np.mean(kernel_ms))
    site_steps_launch = B * N * N_STEPS
    achieved = site_steps_launch * BYTES_PER_SITE_STEP / (kms * 1e-3)
    launch_cfg = gl.last_launch_config()
    launch_cfg["workgroups"] = gl.last_grid()  # < B when the batch is split over the co-resident workgroups
    if args.profile:
        print(json.dumps({"profile": True, "kernel_ms": kms, "ms_per_step": ms_step, "launch": launch_cfg,
                          "traj_per_s": traj_s, "pcie_inclusive_traj_per_s": pcie_traj_s}), flush=True)
        gl.close()
        return

    # plain one-trajectory-per-launch paths (M01 at tau = 0.25), one trajectory
    gl.set_spins(init)
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.25)
    t = time.perf_counter()
    gl.integrate(T0, T1, DT, 1, tgrid="boost", engine=1, want_energy=False)
    plain_stage_s = time.perf_counter() - t
    gl.set_spins(init)
    t = time.perf_counter()
    gl.integrate(T0, T1, DT, 1, tgrid="boost", engine=2, want_energy=False)
    plain_persist_s = time.perf_counter() - t
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.0)

    err_m, err_s = parity_check(uc, lat, gl, init, field_in)

    line = {
        "metric": METRIC,
        "value": traj_s,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "2dcs_kitaev_honeycomb_36x36x1_513traj_8192steps_rk4_dt0.01",
                   "lattice": [L1, L2, 1], "sites": N, "trajectories": B, "rk4_steps": N_STEPS,
                   "record_every": 1, "launch": launch_cfg},
        "site_steps_per_s": traj_s * N * N_STEPS,
        "pcie_inclusive_traj_per_s": pcie_traj_s,
        "plain_stage_kernels_traj_per_s": 1.0 / plain_stage_s,
        "plain_persistent_traj_per_s": 1.0 / plain_persist_s,
        "batched_over_plain": traj_s * plain_stage_s,
        "max_abs_err_vs_oracle_500_steps": {"M": err_m, "S": err_s},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": None,
                     "kernel": "k_persistent", "kernel_ms": kms,
                     "algorithmic_bytes_per_launch": site_steps_launch * BYTES_PER_SITE_STEP,
                     "fp64_tflops": site_steps_launch * FLOPS_PER_SITE_STEP / (kms * 1e-3) / 1e12},
    }
    traffic_file = os.path.join(ROOT, "profiles", "traffic_k_persistent_2dcs_36x36.txt")
    if os.path.exists(traffic_file):
        for ln in open(traffic_file):
            if ln.startswith("traffic_bytes_per_launch:"):
                line["roofline"]["traffic"] = float(ln.split(":", 1)[1])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(uc, init, field_in, args.cpu_tau)
    gl.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
