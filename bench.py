#!/usr/bin/env python3
"""Benchmark of the FP64 Landau-Lifshitz RK4 path on one MI355X.

Default run: the batched 2DCS workload is the headline (one "step" = one
pass of the hot path over its batch); the four other workloads run after it
and are reported as sub-results of the same JSON line ("configs"), each with
its value, its error against the oracle, the roofline of its dominant kernel
and a CPU baseline at one thread and at every available thread.

  2dcs        Kitaev honeycomb 36x36x1 (2592 SU(2) spins), K=-1, Gamma=0.25,
              Gamma'=-0.02, static h = 0.1 along [111] (unit direction),
              Gaussian pulses A = 0.5 along global x, omega = 0, width w = 0.5
              in the reference's exp(-((t-t_c)/(2w))^2) (sigma = sqrt(2) w =
              0.707 in exp(-(t-t_c)^2/(2 sigma^2))); pulse 1 at t = 0, delays
              tau on the 256-point grid linspace(0, 64); M0 + 256 M1(tau) + 256
              M01(tau) = 513 trajectories of 8192 RK4 steps (dt = 0.01,
              t in [-10, 71.92]) from one seeded relaxed state tilted by
              N(0, 0.01) noise and renormalised; M(t) recorded every step.
  md          plain LL dynamics: Kitaev 24x24x1 (1152 sites), no field,
              seeded random start, 10^4 RK4 steps (dt = 0.01), M(t) and E(t)
              every step, one trajectory per launch.
  pyrochlore  64 samples of the 16^3-cell conventional pyrochlore lattice
              (65 536 sites), 4096 RK4 steps (dt = 0.02).
  adjoint     32 parameter starts of the discrete-adjoint gradient on the
              24x24 Kitaev lattice, 2000 RK4 steps, Gilbert alpha = 0.02.
  mixed       257-trajectory SU(2)+SU(3) 2DCS batch on the 8x8x8 TmFeO3-like
              lattice, 8192 RK4 steps (dt = 0.005), and the gradient w.r.t.
              the Fe-Tm coupling scale.

`value` = trajectories/s of the 2DCS batch with the inputs and the M(t)
records resident in HBM (the timed steps do not copy the 302 MB of records
to the host); the rate with that copy is pcie_inclusive_traj_per_s.
--workload NAME runs one workload alone (its own JSON line; used for
profiling).
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
HBM_PEAK = 8.0e12           # MI355X HBM3E, B/s
FP64_PEAK = 78.6e12         # MI355X FP64 (vector and matrix), FLOP/s
BYTES_PER_SITE_STEP = 192   # 4 RK stages x (read + write) x 3 doubles x 8 B
# Algorithmic FP64 flops per SU(2) site-step (FMA = 2 flops), z = 3 bonds:
# per RK stage 54 (three 3x3 bond mat-vecs) + 3 (field) + 9 (H x S) + 12
# (accumulator and next stage input), i.e. 78, the last stage 72; records of
# M and E every step 13  ->  4 * 78 - 6 + 13 = 319; a pulse drive adds 12 per
# stage (f1 d1 + f2 d2), 48 per step.
FLOPS_Z3 = 319
# the same count with z = 6 bonds: 4 stages x (18 z + 21) + 19
FLOPS_Z6 = 535
FLOPS_DRIVE = 48

DT = 0.01
L1 = L2 = 36
T0 = -10.0
N_STEPS = 8192
T1 = T0 + (N_STEPS + 0.5) * DT     # Boost grid: exactly N_STEPS steps
N_TAU = 256
TAUS = np.linspace(0.0, 64.0, N_TAU)
AMP, WIDTH, FREQ = 0.5, 0.5, 0.0


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    pyoracle.lib()
    return pyoracle


def cpu_baseline(unit, per_run, runs_nproc, runs_one, describe):
    """Times the oracle (the FP64 C++/OpenMP restatement of the reference CPU
    path, "port") at every available OpenMP thread and at one thread.
    per_run(n) runs a sample of size n and returns the work it stands for in
    `unit` (e.g. trajectories); runs_* are the sample sizes."""
    po = _oracle()
    full = po.max_threads()
    out = {}
    for threads, n in ((full, runs_nproc), (1, runs_one)):
        po.set_threads(threads)
        t = time.perf_counter()
        work = per_run(n)
        el = time.perf_counter() - t
        out[threads] = (work / el, el, n)
    po.set_threads(full)
    v, el, n = out[full]
    v1, el1, n1 = out[1]
    return {"value": v, "unit": unit, "cores": full, "kind": "port", "cpu_model": cpu_model(),
            "sample": describe(n, el, full),
            "single_thread": {"value": v1, "unit": unit, "cores": 1, "sample": describe(n1, el1, 1)}}


def fp64_roofline(kernel, flops_launch, kernel_ms, **extra):
    a = flops_launch / (kernel_ms * 1e-3)
    r = {"bound": "fp64", "achieved": a / 1e12, "peak": FP64_PEAK / 1e12, "unit": "TFLOP/s",
         "frac": a / FP64_PEAK, "traffic": None, "kernel": kernel, "kernel_ms": kernel_ms,
         "algorithmic_flops_per_launch": flops_launch}
    r.update(extra)
    return r


def hbm_roofline(kernel, bytes_launch, kernel_ms, **extra):
    a = bytes_launch / (kernel_ms * 1e-3)
    r = {"bound": "hbm", "achieved": a / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": a / HBM_PEAK,
         "traffic": None, "kernel": kernel, "kernel_ms": kernel_ms, "algorithmic_bytes_per_launch": bytes_launch}
    r.update(extra)
    return r


def profile_number(name, key):
    """A `key: value` line of a committed rocprofv3 summary in profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        for ln in open(path):
            if ln.startswith(key + ":"):
                return float(ln.split(":", 1)[1].split()[0])
    return None


# --------------------------------------------------------------------- 2DCS

def bench_2dcs(args):
    """The batched 2DCS workload (module docstring).  Parity: the timed
    launch shape at the full batch (513 trajectories, work split active)
    over 500 steps, M0, M1 and M01 trajectories and the trajectories whose
    state crosses a workgroup seam, against the oracle."""
    from classical_mcmc_spin_amd import GPULattice, workloads
    from classical_mcmc_spin_amd import gpu as G
    uc, lat, init, field_in = workloads.kitaev_2dcs(L1, L2)
    tp, am = workloads.pump_probe_replicas(TAUS, AMP)
    B = tp.shape[0]
    N = lat.lattice_size
    gl = GPULattice(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, field_in).reshape(-1)
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.0)
    if args.launch:
        blk, spt, fl = (int(v) for v in args.launch.split(","))
        gl.set_launch_config(B, blk, spt, fl)
    elif not args.profile:
        blk, ms = gl.autotune(B, 64)
        print(f"[bench] 2dcs: autotuned block {blk} ({ms:.2f} ms / 64 steps)", file=sys.stderr, flush=True)

    def one_step(records_to_host):
        return gl.integrate_batched(init, tp, am, T0, T1, DT, tgrid="boost", want_energy=False,
                                    want_mags=records_to_host)

    for _ in range(max(args.warmup, 1)):
        res = one_step(True)
    assert res["mags"].shape[0] == N_STEPS + 1 and np.all(np.isfinite(res["mags"]))
    kernel_ms = []
    t = time.perf_counter()
    for _ in range(args.steps):
        one_step(False)
        kernel_ms.append(gl.last_kernel_ms()[0])
    el = time.perf_counter() - t
    ms_step = 1e3 * el / args.steps
    traj_s = B * args.steps / el
    launch = gl.last_launch_config()
    launch["workgroups"] = gl.last_grid()
    kms = float(np.mean(kernel_ms))
    if args.profile:
        gl.close()
        return {"profile": True, "workload": "2dcs", "kernel_ms": kms, "ms_per_step": ms_step, "launch": launch,
                "traj_per_s": traj_s}
    t = time.perf_counter()
    one_step(True)
    pcie_traj_s = B / (time.perf_counter() - t)

    # plain one-trajectory-per-launch paths (M01 at tau = 0.25)
    gl.set_spins(init)
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.25)
    plain = {}
    for name, eng in (("stage_kernels", 1), ("persistent_workgroup", 2)):
        gl.set_spins(init)
        t = time.perf_counter()
        gl.integrate(T0, T1, DT, 1, tgrid="boost", engine=eng, want_energy=False)
        plain[name] = 1.0 / (time.perf_counter() - t)
    gl.set_pulse(np.concatenate([fd, fd]), AMP, WIDTH, FREQ, 0.0, 0.0)

    # parity on the timed launch shape: the full batch over 500 steps
    n_par = 500
    Tp = T0 + (n_par + 0.5) * DT
    gl.set_launch_config(B, launch["block"], launch["sites_per_thread"], launch["flags"])
    pr = gl.integrate_batched(init, tp, am, T0, Tp, DT, tgrid="boost", want_final=True)
    lp = gl.last_launch_config()
    same_shape = all(lp[k] == launch[k] for k in ("block", "sites_per_thread", "flags", "field_only", "atom_major"))
    Gp = gl.last_grid()
    seams = G.seam_trajectories(B, Gp, n_par) if Gp < B else []
    check = sorted({0, 1, N_TAU, N_TAU + 1, B - 1, *seams[:3]})
    po = _oracle()
    o = po.OracleLattice(uc, L1, L2, 1)
    err_m = err_s = 0.0
    for b in check:
        o.set_pulse(field_in * (am[b, 0] != 0), tp[b, 0], field_in * (am[b, 1] != 0), tp[b, 1], AMP, WIDTH, FREQ)
        ref = o.integrate_md(init.reshape(-1), T0, Tp, DT)
        err_m = max(err_m, float(np.max(np.abs(pr["mags"][:, b] - ref["mags"]))))
        err_s = max(err_s, float(np.max(np.abs(pr["final"][b] - ref["final"].reshape(-1, 3)))))
    gl.close()

    site_steps = B * N * N_STEPS
    flops_launch = site_steps * (FLOPS_Z3 + FLOPS_DRIVE)
    roof = fp64_roofline("k_persistent", flops_launch, kms,
                         algorithmic_flops_per_site_step=FLOPS_Z3 + FLOPS_DRIVE,
                         hbm_algorithmic_frac=site_steps * BYTES_PER_SITE_STEP / (kms * 1e-3) / HBM_PEAK)
    roof["traffic"] = profile_number("traffic_k_persistent_2dcs_36x36.txt", "traffic_bytes_per_launch")
    for key in ("valu_issue_frac", "lds_busy_frac"):
        v = profile_number("counters_k_persistent_2dcs_36x36.txt", key)
        if v is not None:
            roof[key] = v
    line = {
        "metric": METRIC, "value": traj_s, "unit": "trajectories/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "2dcs_kitaev_honeycomb_36x36x1_513traj_8192steps_rk4_dt0.01",
                   "lattice": [L1, L2, 1], "sites": N, "trajectories": B, "rk4_steps": N_STEPS, "record_every": 1,
                   "tau_grid": "linspace(0, 64, 256)",
                   "pulse": {"amp": AMP, "w": WIDTH, "convention": "exp(-((t-tc)/(2w))^2)",
                             "sigma_equivalent": float(np.sqrt(2.0) * WIDTH), "freq": FREQ},
                   "launch": launch},
        "site_steps_per_s": traj_s * N * N_STEPS,
        "pcie_inclusive_traj_per_s": pcie_traj_s,
        "plain_traj_per_s": plain,
        "batched_over_plain": traj_s / max(plain.values()),
        "batched_over_plain_stage_kernels": traj_s / plain["stage_kernels"],
        "max_abs_err_vs_oracle_500_steps": {"M": err_m, "S": err_s, "trajectories": check,
                                            "seam_trajectories": seams[:3], "workgroups": Gp,
                                            "timed_launch_shape": same_shape},
        "roofline": roof,
    }
    if not args.no_cpu:
        o = po.OracleLattice(uc, L1, L2, 1)
        tau_s = TAUS[:16]

        def run(n_tau):   # outer OpenMP over tau, serial RHS (the reference's W2 path)
            o.pump_probe(init.reshape(-1), field_in, AMP, WIDTH, FREQ, float(tau_s[0]), float(tau_s[n_tau - 1]),
                         float(tau_s[1] - tau_s[0]), T0, T1, DT, reuse_m0=False, n_outer=po.max_threads())
            return 1 + 2 * n_tau

        def run1(n_steps):  # one M01 trajectory, part of its steps
            o.set_pulse(field_in, 0.0, field_in, 0.25, AMP, WIDTH, FREQ)
            o.integrate_md(init.reshape(-1), T0, T0 + (n_steps + 0.5) * DT, DT)
            return n_steps / N_STEPS

        full = po.max_threads()
        t = time.perf_counter()
        w = run(16)
        el_n = time.perf_counter() - t
        po.set_threads(1)
        t = time.perf_counter()
        w1 = sum(run1(8192) for _ in range(4))
        el1 = time.perf_counter() - t
        po.set_threads(full)
        line["cpu_baseline"] = {
            "value": w / el_n, "unit": "trajectories/s", "cores": full, "kind": "port", "cpu_model": cpu_model(),
            "sample": f"33 trajectories (M0 + 16 M1 + 16 M01) x 8192 RK4 steps of the same 36x36 lattice, "
                      f"{el_n:.1f} s, outer OpenMP over tau on {full} threads",
            "site_steps_per_s": w / el_n * N * N_STEPS,
            "single_thread": {"value": w1 / el1, "unit": "trajectories/s", "cores": 1,
                              "sample": f"four whole M01 trajectories (8192 RK4 steps each), {el1:.1f} s"}}
    return line


# ----------------------------------------------------------------- plain MD

MD_L, MD_STEPS, MD_DT = 24, 10000, 0.01


def bench_md(args):
    """Plain one-trajectory MD (module docstring).  value = trajectories/s
    of the fastest engine; every engine is reported.  Error: 500 steps of
    the fastest engine against the oracle, and the drift of E and |S| over
    the whole run against the oracle's drift on the same run."""
    from classical_mcmc_spin_amd import GPULattice, workloads
    uc, lat, s0 = workloads.kitaev_md(MD_L, MD_L, seed=1234)
    N = lat.lattice_size
    T1m = (MD_STEPS + 0.5) * MD_DT
    gl = GPULattice(lat)
    eng = {}
    last = {}
    engines = (("stage_kernels", 1), ("stage_kernels_hipgraph", 3), ("persistent_workgroup", 2))
    if args.profile:   # one dispatch per integration: the persistent workgroup only
        engines = (("persistent_workgroup", 2),)
    for name, e in engines:
        for _ in range(max(args.warmup, 1)):
            gl.set_spins(s0)
            gl.integrate(0.0, T1m, MD_DT, 1, tgrid="boost", engine=e)
        t = time.perf_counter()
        for _ in range(args.steps):
            gl.set_spins(s0)
            r = gl.integrate(0.0, T1m, MD_DT, 1, tgrid="boost", engine=e)
        el = (time.perf_counter() - t) / args.steps
        assert r["mags"].shape[0] == MD_STEPS + 1 and np.all(np.isfinite(r["mags"]))
        eng[name] = {"traj_per_s": 1.0 / el, "ms_per_traj": 1e3 * el, "site_steps_per_s": MD_STEPS * N / el,
                     "kernel_ms": gl.last_kernel_ms()[0], "launches": gl.last_kernel_ms()[1]}
        last[name] = (r, gl.get_spins(), e)
    best = max(eng, key=lambda k: eng[k]["traj_per_s"])
    if args.profile:
        gl.close()
        return {"profile": True, "workload": "md", "engines": eng}
    r_full, fin_full, e_best = last[best]
    gl.set_spins(s0)
    r500 = gl.integrate(0.0, (500 + 0.5) * MD_DT, MD_DT, 1, tgrid="boost", engine=e_best)
    fin500 = gl.get_spins()
    gl.close()
    po = _oracle()
    o = po.OracleLattice(uc, MD_L, MD_L, 1)
    ref500 = o.integrate_md(s0.reshape(-1), 0.0, (500 + 0.5) * MD_DT, MD_DT)
    t = time.perf_counter()
    ref = o.integrate_md(s0.reshape(-1), 0.0, T1m, MD_DT)
    el_full = time.perf_counter() - t
    drift = lambda e, f: (float(np.max(np.abs(e - e[0]))), float(np.max(np.abs(np.linalg.norm(f, axis=-1) - 1))))
    dE, dS = drift(r_full["energy"], fin_full)
    dEo, dSo = drift(ref["energy"], ref["final"].reshape(-1, 3))
    v = eng[best]["traj_per_s"]
    kms = eng[best]["kernel_ms"] * eng[best]["launches"]
    out = {
        "value": v, "unit": "trajectories/s", "ms_per_step": 1e3 / v,
        "config": {"workload": "md_kitaev_honeycomb_24x24x1_1traj_10000steps_rk4_dt0.01", "sites": N,
                   "rk4_steps": MD_STEPS, "record_every": 1, "engine": best},
        "site_steps_per_s": eng[best]["site_steps_per_s"], "engines": eng,
        "max_abs_err_vs_oracle_500_steps": {
            "M": float(np.max(np.abs(r500["mags"] - ref500["mags"]))),
            "E": float(np.max(np.abs(r500["energy"] - ref500["energy"]))),
            "S": float(np.max(np.abs(fin500 - ref500["final"].reshape(-1, 3))))},
        "drift_10k_steps": {"E_per_site": dE, "spin_length": dS, "oracle_E_per_site": dEo,
                            "oracle_spin_length": dSo},
        "roofline": fp64_roofline(best, N * MD_STEPS * FLOPS_Z3, kms, algorithmic_flops_per_site_step=FLOPS_Z3,
                                  note="one trajectory: one workgroup (persistent) or a few workgroups per "
                                       "launch (stage kernels) of the 256 CUs"),
    }
    if not args.no_cpu:
        base = {"value": 1.0 / el_full, "unit": "trajectories/s", "cores": po.max_threads(), "kind": "port",
                "cpu_model": cpu_model(),
                "sample": f"the whole 10^4-step trajectory ({el_full:.1f} s, OpenMP over sites inside each RHS "
                          f"as the reference's MD path)",
                "site_steps_per_s": MD_STEPS * N / el_full}
        po.set_threads(1)
        t = time.perf_counter()
        o.integrate_md(s0.reshape(-1), 0.0, T1m, MD_DT)
        el1 = time.perf_counter() - t
        po.set_threads(base["cores"])
        base["single_thread"] = {"value": 1.0 / el1, "unit": "trajectories/s", "cores": 1,
                                 "sample": f"the whole 10^4-step trajectory, {el1:.1f} s"}
        out["cpu_baseline"] = base
    return out


def bench_md_dopri5(args):
    """MD with the reference's default integrator (md_integrator "dopri5",
    md_abs_tol = md_rel_tol = 1e-6, spin_config.h:93,111-112):
    integrate_adaptive of the Kitaev 24x24x1 lattice from t = 0 to 100, M(t)
    and E(t) at every accepted step; a batch of 256 trajectories (one
    workgroup per CU, seeded starts) is reported beside it."""
    from classical_mcmc_spin_amd import GPULattice, workloads, seeding
    uc, lat, s0 = workloads.kitaev_md(MD_L, MD_L, seed=1234)
    T1d, tol, Bd = 100.0, 1e-6, 256
    gl = GPULattice(lat)
    run = lambda init: gl.integrate_dopri5(init, 0.0, T1d, MD_DT, control=1, abs_tol=tol, rel_tol=tol)
    for _ in range(max(args.warmup, 1)):
        r = run(s0)
    t = time.perf_counter()
    for _ in range(args.steps):
        r = run(s0)
    el = (time.perf_counter() - t) / args.steps
    kms = gl.last_kernel_ms()[0]
    acc, rej = int(r["accepted"][0]), int(r["rejected"][0])
    inits = np.stack([seeding.random_spins(lat.lattice_size, 5000 + b) for b in range(Bd)])
    run(inits)
    t = time.perf_counter()
    rb = run(inits)
    elb = time.perf_counter() - t
    kmsb = gl.last_kernel_ms()[0]
    gl.close()
    return {"value": 1.0 / el, "unit": "trajectories/s", "ms_per_step": 1e3 * el,
            "config": {"workload": "md_kitaev_honeycomb_24x24x1_1traj_t100_dopri5_adaptive_tol1e-6",
                       "sites": lat.lattice_size, "T": T1d, "abs_tol": tol, "rel_tol": tol},
            "accepted_steps": acc, "rejected_steps": rej, "kernel_ms": kms,
            "batched": {"trajectories": Bd, "traj_per_s": Bd / elb, "kernel_ms": kmsb}}


# --------------------------------------------------------------- pyrochlore

PYRO_L, PYRO_B, PYRO_STEPS, PYRO_DT = 16, 64, 4096, 0.02


def bench_pyrochlore(args):
    """64 samples x 4096 RK4 steps of the 16^3-cell conventional pyrochlore
    lattice (65 536 sites, z = 6), all-in-all-out + tilt starts, M(t) and
    E(t) every step.  The lattice is too large for one workgroup's LDS: the
    default path is the partitioned persistent kernel (each sample spread
    over co-resident workgroups exchanging halo values every stage, one
    launch); the grid-wide stage kernels (4 launches per step) are timed
    beside it.  Error: two samples over 200 steps against the oracle."""
    from classical_mcmc_spin_amd import GPULattice, workloads
    lat = workloads.pyrochlore_lattice(PYRO_L)
    N = lat.lattice_size
    init, _ = workloads.pyrochlore_tilted_aiao(lat, PYRO_B, seed=1234)
    T1p = (PYRO_STEPS + 0.5) * PYRO_DT
    zero = np.zeros((PYRO_B, 2))
    gl = GPULattice(lat)
    one_step = lambda: gl.integrate_batched(init, zero, zero, 0.0, T1p, PYRO_DT, tgrid="boost")

    def timed(steps):
        kernel_ms = []
        t = time.perf_counter()
        for _ in range(steps):
            res = one_step()
            kernel_ms.append(gl.last_kernel_ms())
        el = time.perf_counter() - t
        assert res["mags"].shape[0] == PYRO_STEPS + 1 and np.all(np.isfinite(res["mags"]))
        return el, float(np.mean([k[0] for k in kernel_ms])), kernel_ms[-1][1]

    for _ in range(args.warmup):
        one_step()
    el, kms, _ = timed(args.steps)
    parts, groups = gl.last_partition()
    launch = gl.last_launch_config()
    if args.profile:
        gl.close()
        return {"profile": True, "workload": "pyrochlore", "kernel_ms": kms, "ms_per_step": 1e3 * el / args.steps,
                "parts": parts, "groups": groups}
    # the grid-wide stage kernels on the same batch
    gl.set_partition(-1)
    one_step()
    el_st, kms_st, launches_st = timed(args.steps)
    gl.set_partition(0)
    n_par = 200
    pr = gl.integrate_batched(init[[0, PYRO_B - 1]], np.zeros((2, 2)), np.zeros((2, 2)), 0.0,
                              (n_par + 0.5) * PYRO_DT, PYRO_DT, tgrid="boost", want_final=True)
    gl.close()
    po = _oracle()
    o = po.OracleLattice(lat.unit_cell, PYRO_L, PYRO_L, PYRO_L)
    err = {"M": 0.0, "E": 0.0, "S": 0.0}
    for j, b in enumerate((0, PYRO_B - 1)):
        ref = o.integrate_md(init[b].reshape(-1), 0.0, (n_par + 0.5) * PYRO_DT, PYRO_DT)
        err["M"] = max(err["M"], float(np.max(np.abs(pr["mags"][:, j] - ref["mags"]))))
        err["E"] = max(err["E"], float(np.max(np.abs(pr["energy"][:, j] - ref["energy"]))))
        err["S"] = max(err["S"], float(np.max(np.abs(pr["final"][j] - ref["final"].reshape(-1, 3)))))
    samples_s = PYRO_B * args.steps / el
    site_steps = PYRO_B * N * PYRO_STEPS
    roof = fp64_roofline("k_partitioned", site_steps * FLOPS_Z6, kms, algorithmic_flops_per_site_step=FLOPS_Z6,
                         hbm_algorithmic_bytes_per_launch=site_steps * BYTES_PER_SITE_STEP)
    roof["traffic"] = profile_number("traffic_k_partitioned_pyrochlore_16x16x16.txt", "traffic_bytes_per_launch")
    stage_roof = hbm_roofline("k_stage", PYRO_B * N * BYTES_PER_SITE_STEP / 4, kms_st)  # one RK stage of every site
    stage_roof["traffic"] = profile_number("traffic_l2_k_stage_pyrochlore_16x16x16.txt", "traffic_bytes_per_launch")
    out = {
        "value": samples_s, "unit": "trajectories/s", "ms_per_step": 1e3 * el / args.steps,
        "config": {"workload": f"pyrochlore_16x16x16_conventional_64samples_{PYRO_STEPS}steps_rk4_dt0.02",
                   "sites": N, "trajectories": PYRO_B, "rk4_steps": PYRO_STEPS, "record_every": 1,
                   "path": "partitioned persistent kernel", "parts_per_sample": parts, "sample_groups": groups,
                   "launch": launch},
        "site_steps_per_s": samples_s * N * PYRO_STEPS,
        "max_abs_err_vs_oracle_200_steps": err,
        "roofline": roof,
        "stage_kernels": {"value": PYRO_B * args.steps / el_st, "unit": "trajectories/s",
                          "ms_per_step": 1e3 * el_st / args.steps, "kernel_ms_per_launch": kms_st,
                          "launches_per_run": launches_st, "roofline": stage_roof},
        "speedup_over_stage_kernels": el_st / el,
    }
    if not args.no_cpu:
        def per_run(n):
            o.integrate_md(init[0].reshape(-1), 0.0, (n + 0.5) * PYRO_DT, PYRO_DT)
            return n / PYRO_STEPS

        desc = lambda n, el_, th: (f"1 sample x {n} of the {PYRO_STEPS} RK4 steps, {el_:.1f} s on {th} thread(s) "
                                   f"(OpenMP over sites inside each RHS as the reference's MD path)")
        out["cpu_baseline"] = cpu_baseline("trajectories/s", per_run, 300, 40, desc)
    return out


# ------------------------------------------------------------------ adjoint

ADJ_L, ADJ_STEPS, ADJ_B, ADJ_B_FULL = 24, 2000, 32, 256
# Algorithmic FP64 flops per site-step of one gradient (z = 3, Gilbert
# damping, pulse drive, 5 parameters): forward RK4 with records 465 (its
# stage inputs are kept, so nothing is integrated twice), transposed step
# 2119 (per stage: field at the stage input 69, H x S and the Gilbert factor
# 16, the cotangent map w, u 24, the direct term 46, the transposed bond
# gather 54, five parameter derivatives u . (sum dJ S - dB) 310,
# combinations 10; plus the loss term 3).
ADJ_FLOPS_PER_SITE_STEP = 465 + 2119
# mixed TmFeO3 cell, per (SU(2) + SU(3)) site pair and RK4 step: SU(2) site
# 2A S 18, 6 SU(2) bonds 108, 8 SU(3) bonds 384, drive 12, H x S 9, update 12
# = 543; SU(3) site 8 SU(2) bonds 384, drive 32, f_abc 108, update 32 = 556
MIX_FLOPS_PAIR_STEP = 4 * (543 + 556)


def bench_adjoint(args):
    """Adjoint gradients: Kitaev 24x24 (1152 sites), one pulse (A = 0.5,
    w = 0.5 at t = 1), Gilbert alpha = 0.02, loss on M_global over 2000 RK4
    steps (dt = 0.01) against a target run at (K, G, G', J, h) =
    (-1, 0.25, -0.02, 0, 0.05); 32 parameter starts within +-10 %.
    value = gradients (starts) per second; a batch of 256 starts beside it.
    Error: two starts against the oracle's tangent-linear gradient."""
    from classical_mcmc_spin_amd import adjoint as AD
    from classical_mcmc_spin_amd import GPULattice, model, seeding
    lat = model.Lattice(model.build_kitaev_honeycomb(), ADJ_L, ADJ_L, 1)
    init = seeding.tilt(seeding.quench(lat, seeding.random_spins(lat.lattice_size, 21), sweeps=30), 0.05, seed=22)
    dt = 0.01
    T1a = (ADJ_STEPS + 0.5) * dt
    star = np.array([-1.0, 0.25, -0.02, 0.0, 0.05])
    tl = AD.kitaev_lattice(star, ADJ_L, ADJ_L)
    fdir = np.zeros((2, 3))
    fdir[:, 0] = 1.0
    with GPULattice(tl) as g0:
        g0.set_damping(0.02)
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
    if args.profile:
        fit.close()
        return {"profile": True, "workload": "adjoint", "kernel_ms": kms, "ms_per_step": 1e3 * el}
    fit(thetas_full)
    t = time.perf_counter()
    _, grad_full = fit(thetas_full)
    el_full = time.perf_counter() - t
    kms_full = fit.gl.last_kernel_ms()[0]
    fit.close()
    assert np.all(np.isfinite(grad_full))
    po = _oracle()
    N = lat.lattice_size

    # per unit-cell bond dJ and per-atom dB of each of the 5 parameters
    unit = []
    for p in range(5):
        e = np.zeros(5)
        e[p] = 1.0
        ue = model.build_kitaev_honeycomb(K=e[0], Gamma=e[1], Gammap=e[2], J=e[3], field_strength=e[4],
                                          field_direction=(1.0, 1.0, 1.0))
        unit.append((np.array([bd.J for bd in ue.bilinear]).reshape(-1), ue.field.reshape(-1)))

    def oracle_grad(th, n_steps):
        """Tangent-linear gradient of the loss on M_global: one forward-mode
        sweep of the oracle per parameter."""
        uc = model.build_kitaev_honeycomb(K=th[0], Gamma=th[1], Gammap=th[2], J=th[3], field_strength=th[4],
                                          field_direction=(1.0, 1.0, 1.0))
        o = po.OracleLattice(uc, ADJ_L, ADJ_L, 1)
        o.set_alpha(0.02)
        o.set_pulse(fdir, 1.0, fdir * 0.0, 0.0, 0.5, 0.5, 0.0)
        Tn = (n_steps + 0.5) * dt
        g = np.zeros(5)
        lo = 0.0
        for p, (dJ, dB) in enumerate(unit):
            m, dm = o.tangent_md(init.reshape(-1), np.zeros(init.size), dJ, dB, 0.0, Tn, dt)
            diff = (m.reshape(-1, 9) - target[:n_steps + 1])[:, 6:9]
            lo = float(np.sum(diff ** 2))
            g[p] = float(np.sum(2.0 * diff * dm.reshape(-1, 9)[:, 6:9]))
        return lo, g

    err = 0.0
    for b in (0, ADJ_B - 1):
        lo, go = oracle_grad(thetas[b], ADJ_STEPS)
        err = max(err, float(np.max(np.abs(grad[b] - go)) / np.max(np.abs(go))))
    v = ADJ_B / el
    out = {
        "value": v, "unit": "gradients/s", "ms_per_step": 1e3 * el,
        "config": {"workload": f"adjoint_kitaev_honeycomb_{ADJ_L}x{ADJ_L}x1_{ADJ_B}starts_{ADJ_STEPS}steps_rk4_dt0.01",
                   "sites": N, "starts": ADJ_B, "rk4_steps": ADJ_STEPS, "parameters": 5, "alpha": 0.02},
        "kernel_ms": kms,
        "max_rel_err_vs_oracle_tangent_linear": err,
        "roofline": fp64_roofline("k_adj_fwd + k_adj_bwd + k_adj_par", ADJ_B * N * ADJ_STEPS * ADJ_FLOPS_PER_SITE_STEP,
                                  kms, algorithmic_flops_per_site_step=ADJ_FLOPS_PER_SITE_STEP,
                                  note=f"forward and backward sweeps: one workgroup per start ({ADJ_B} of the 256 CUs); "
                                       "parameter terms on the other CUs, overlapped with the sweep"),
        "starts_256": {"gradients_per_s": ADJ_B_FULL / el_full, "kernel_ms": kms_full,
                       "fp64_frac": ADJ_B_FULL * N * ADJ_STEPS * ADJ_FLOPS_PER_SITE_STEP / (kms_full * 1e-3)
                       / FP64_PEAK},
    }
    if not args.no_cpu:
        def per_run(n):
            oracle_grad(thetas[0], n)
            return n / ADJ_STEPS

        desc = lambda n, el_, th: (f"tangent-linear gradient (5 forward-mode sweeps) over {n} of the {ADJ_STEPS} "
                                   f"steps, {el_:.1f} s on {th} thread(s)")
        out["cpu_baseline"] = cpu_baseline("gradients/s", per_run, 2000, 300, desc)
    return out


# -------------------------------------------------------------------- mixed

MIX_L, MIX_NTAU, MIX_STEPS, MIX_DT = 8, 128, 8192, 0.005
MIXDP_INTERVALS, MIXDP_TOL = 2048, 1e-8


def _mixed_setup():
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
    return muc, ml, g, taus, B, tp, sc, f2, f3


def bench_mixed(args):
    """Mixed SU(2)+SU(3) 2DCS batch: TmFeO3-like 8x8x8 cells (2048 Fe SU(2)
    + 2048 Tm SU(3) sites), pulses A = 0.5, w = 0.5 along x (Fe) and
    lambda_2 (Tm) at 0 and tau = 0, 0.05, ..., 6.35 (128 delays):
    M0 + 128 M1 + 128 M01 = 257 trajectories x 8192 RK4 steps (dt = 0.005),
    records every step.  Default path: the persistent kernel, one launch,
    the 257 trajectories' steps shared equally by the co-resident workgroups
    (whole trajectories plus one split with a neighbour); the grid-wide stage
    kernels alone are timed beside it.  Then d loss / d (Fe-Tm scale) for every trajectory.  Error:
    two trajectories over 200 steps against the mixed oracle."""
    from classical_mcmc_spin_amd import gpu_mixed as GM
    muc, ml, g, taus, B, tp, sc, f2, f3 = _mixed_setup()
    T0m = -1.0
    T1m = T0m + (MIX_STEPS + 0.5) * MIX_DT
    gm = GM.GPUMixedLattice(ml)
    GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, AMP, WIDTH, 0.0, AMP, WIDTH, 0.0)
    one_step = lambda: gm.integrate_batched(g, tp, sc, T0m, T1m, MIX_DT)

    def timed(steps):
        paths, kernel_ms = [], []
        t = time.perf_counter()
        for _ in range(steps):
            res = one_step()
            paths.append(gm.last_path())
            kernel_ms.append(gm.last_kernel_ms())
        el = time.perf_counter() - t
        assert res["records"].shape[0] == MIX_STEPS + 1 and np.all(np.isfinite(res["records"]))
        return el, res, paths, kernel_ms

    for _ in range(args.warmup):
        one_step()
    el, res, paths, kernel_ms = timed(args.steps)
    path = paths[-1]
    kms = float(np.mean([p["kernel_ms"] for p in paths]))   # the persistent launch
    kms_all = float(np.mean([k[0] for k in kernel_ms]))      # the whole integration
    if args.profile:
        gm.close()
        return {"profile": True, "workload": "mixed", "kernel_ms": kms, "integration_ms": kms_all,
                "ms_per_step": 1e3 * el / args.steps, "path": path}
    # the grid-wide stage kernels on the same batch
    gm.set_persistent(-1)
    one_step()
    el_st, _, _, kernel_ms_st = timed(args.steps)
    kms_st = float(np.mean([k[0] for k in kernel_ms_st]))
    gm.set_persistent(0)
    target = res["records"][:, 0] * 1.01
    w = np.r_[np.ones(33), 0.0]
    t = time.perf_counter()
    gr = gm.scale_gradient(g, tp, target, w, sc, T0m, T1m, MIX_DT)
    el_g = time.perf_counter() - t
    kms_g = gm.last_kernel_ms()[0]
    assert np.all(np.isfinite(gr["grad"]))
    n_par = 200
    Tp = T0m + (n_par + 0.5) * MIX_DT
    pick = [0, B - 1]
    pr = gm.integrate_batched(g, tp[pick], sc[pick], T0m, Tp, MIX_DT, want_final=True)
    gm.close()
    po = _oracle()
    err = {"records": 0.0, "state": 0.0}
    for j, b in enumerate(pick):
        o = po.MixedOracle(muc, MIX_L, MIX_L, MIX_L)
        o.set_pulse(2, f2 * sc[b, 0], tp[b, 0], f2 * sc[b, 1], tp[b, 1], AMP, WIDTH, 0.0)
        o.set_pulse(3, f3 * sc[b, 0], tp[b, 0], f3 * sc[b, 1], tp[b, 1], AMP, WIDTH, 0.0)
        ref = o.integrate_md(g, T0m, Tp, MIX_DT)
        err["records"] = max(err["records"], float(np.max(np.abs(pr["records"][:, j] - ref["records"]))))
        err["state"] = max(err["state"], float(np.max(np.abs(pr["final"][j] - ref["final"]))))
    traj_s = B * args.steps / el
    n_sites = ml.N2 + ml.N3
    roof = fp64_roofline("k_mixed_persist", MIX_FLOPS_PAIR_STEP * ml.N2 * B * MIX_STEPS, kms,
                         algorithmic_flops_per_pair_step=MIX_FLOPS_PAIR_STEP, trajectories_in_launch=B)
    roof["traffic"] = profile_number("traffic_k_mixed_persist_tmfeo3_8x8x8.txt", "traffic_bytes_per_launch")
    bytes_stage = B * (ml.N2 * 3 + ml.N3 * 8) * 2 * 8   # one stage: read + write of every component
    stage_roof = hbm_roofline("k_mixed_stage", bytes_stage, kms_st)
    stage_roof["traffic"] = profile_number("traffic_l2_k_mixed_stage_tmfeo3_8x8x8.txt", "traffic_bytes_per_launch")
    out = {
        "value": traj_s, "unit": "trajectories/s", "ms_per_step": 1e3 * el / args.steps,
        "config": {"workload": f"2dcs_tmfeo3_mixed_8x8x8_257traj_{MIX_STEPS}steps_rk4_dt0.005",
                   "sites_su2": ml.N2, "sites_su3": ml.N3, "trajectories": B, "rk4_steps": MIX_STEPS,
                   "record_every": 1, "path": "persistent kernel (trajectory-steps shared by the workgroups)",
                   "workgroups": path["workgroups"], "split_trajectories": path["split"]},
        "site_steps_per_s": traj_s * n_sites * MIX_STEPS,
        "max_abs_err_vs_oracle_200_steps": err,
        "roofline": roof,
        "integration_ms": kms_all,
        "stage_kernels": {"value": B * args.steps / el_st, "unit": "trajectories/s",
                          "ms_per_step": 1e3 * el_st / args.steps, "kernel_ms_per_launch": kms_st,
                          "roofline": stage_roof},
        "speedup_over_stage_kernels": el_st / el,
        "scale_gradient": {"gradients_per_s": B / el_g, "ms_per_batch": 1e3 * el_g, "kernel_ms_per_launch": kms_g,
                           "forward_plus_tangent_over_forward": el_g / (el / args.steps)},
    }
    if not args.no_cpu:
        o = po.MixedOracle(muc, MIX_L, MIX_L, MIX_L)
        o.set_pulse(2, f2, 0.0, f2 * 0.0, 0.0, AMP, WIDTH, 0.0)
        o.set_pulse(3, f3, 0.0, f3 * 0.0, 0.0, AMP, WIDTH, 0.0)

        def per_run(n):
            o.integrate_md(g, T0m, T0m + (n + 0.5) * MIX_DT, MIX_DT)
            return n / MIX_STEPS

        desc = lambda n, el_, th: (f"1 trajectory x {n} of the {MIX_STEPS} RK4 steps, {el_:.1f} s on {th} "
                                   f"thread(s) (OpenMP over sites inside each RHS as the reference's mixed MD path)")
        out["cpu_baseline"] = cpu_baseline("trajectories/s", per_run, 3000, 400, desc)
        out["scale_gradient"]["cpu_baseline_gradients_per_s"] = out["cpu_baseline"]["value"] / 2.0
    return out


def bench_mixed_dopri5(args):
    """The mixed 2DCS batch with the reference's default pulse-drive
    integrator: Dormand-Prince 5(4) through integrate_const with step
    control at T_step = 0.005, tolerance 1e-8, W3 pulse-window segments per
    trajectory, over 2048 observation intervals from t = -1."""
    from classical_mcmc_spin_amd import gpu as G
    from classical_mcmc_spin_amd import gpu_mixed as GM
    muc, ml, g, taus, B, tp, sc, f2, f3 = _mixed_setup()
    centers = [[0.0]] + [[t] for t in taus] + [[0.0, t] for t in taus]
    T0m = -1.0
    T1m = T0m + (MIXDP_INTERVALS + 0.5) * MIX_DT
    segs = [G.pulse_segments(T0m, T1m, c, 9.0 * WIDTH, MIX_DT) for c in centers]
    gm = GM.GPUMixedLattice(ml)
    GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, AMP, WIDTH, 0.0, AMP, WIDTH, 0.0)
    one_step = lambda: gm.integrate_dopri5(g, tp, sc, T0m, T1m, MIX_DT, control=2, abs_tol=MIXDP_TOL,
                                           rel_tol=MIXDP_TOL, segments=segs)
    for _ in range(args.warmup):
        one_step()
    t = time.perf_counter()
    for _ in range(args.steps):
        r = one_step()
    el = (time.perf_counter() - t) / args.steps
    kms, launches = gm.last_kernel_ms()
    gm.close()
    acc = np.array(r["accepted"])
    return {"value": B / el, "unit": "trajectories/s", "ms_per_step": 1e3 * el,
            "config": {"workload": f"2dcs_tmfeo3_mixed_8x8x8_257traj_{MIXDP_INTERVALS}x0.005_dopri5_tol1e-8_w3",
                       "trajectories": B, "intervals": MIXDP_INTERVALS},
            "accepted_steps_mean": float(acc.mean()), "rejected_steps_mean": float(np.mean(r["rejected"])),
            "launch_ms_mean": kms, "launches": launches}


# --------------------------------------------------------------------- main

SUBS = (("plain_md_kitaev_24x24", bench_md), ("pyrochlore_16x16x16_dssf", bench_pyrochlore),
        ("adjoint_kitaev_24x24", bench_adjoint), ("mixed_tmfeo3_8x8x8", bench_mixed))
WORKLOADS = {"2dcs": bench_2dcs, "md": bench_md, "md_dopri5": bench_md_dopri5, "pyrochlore": bench_pyrochlore,
             "adjoint": bench_adjoint, "mixed": bench_mixed, "mixed_dopri5": bench_mixed_dopri5}


def _standalone(name, args, r):
    """A workload run alone: its result as a full bench line."""
    if r.get("profile"):
        return r
    line = {"metric": METRIC, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic"}
    line.update(r)
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--launch", default="", help="fixed persistent launch shape 'block,spt,flags' for the 2DCS "
                                                   "batch instead of the autotuner's choice")
    ap.add_argument("--workload", default="all", choices=["all"] + sorted(WORKLOADS),
                    help="all (default): the 2DCS headline plus every other workload as a sub-result")
    ap.add_argument("--sub-steps", type=int, default=2, help="timed steps of each sub-result workload")
    ap.add_argument("--rk-steps", type=int, default=0,
                    help="RK4 steps of the pyrochlore / mixed integrations (default 4096 / 8192); shorter runs "
                         "keep counter-collection passes within a few hundred dispatches")
    ap.add_argument("--profile", action="store_true",
                    help="only the timed launches of the chosen workload (for rocprofv3 runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the batched trajectories shard by "
              "batch across GPUs with no exchange; multi-GPU runs are not part of this build.", file=sys.stderr)
        sys.exit(2)
    global PYRO_STEPS, MIX_STEPS
    if args.rk_steps > 0:
        PYRO_STEPS = MIX_STEPS = args.rk_steps
    if args.workload != "all":
        print(json.dumps(_standalone(args.workload, args, WORKLOADS[args.workload](args))), flush=True)
        return
    line = bench_2dcs(args)
    sub_args = argparse.Namespace(**vars(args))
    sub_args.steps = max(1, min(args.sub_steps, args.steps))
    sub_args.warmup = min(args.warmup, 1)
    line["configs"] = {}
    for name, fn in SUBS:
        t = time.perf_counter()
        print(f"[bench] {name}", file=sys.stderr, flush=True)
        r = fn(sub_args)
        r["wall_s"] = time.perf_counter() - t
        line["configs"][name] = r
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
