"""Device handle of a mixed SU(2) + SU(3) lattice and the reference's
MixedLattice simulator entry points on it.

GPUMixedLattice wraps one csl_mixed handle (include/csl_mixed.h).  The
drivers mirror src/core/mixed_lattice_md.cpp and mixed_lattice.h:

  molecular_dynamics        MixedLattice::molecular_dynamics_cpu, "rk4" or "dopri5"
                            (mixed_lattice_md.cpp:1208-1333)
  single_pulse_drive        MixedLattice::single_pulse_drive (mixed_lattice.h:4754-4881)
  double_pulse_drive        MixedLattice::double_pulse_drive (mixed_lattice.h:4883-)
  pump_probe_spectroscopy   MixedLattice::pump_probe_spectroscopy (mixed_lattice_md.cpp:1400-1750),
                            every M0 / M1(tau) / M01(tau) trajectory in one batched call

Records are the observer's [M_SU2 (antiferro, local, global), M_SU3
(antiferro, local, global)] and E/(N2+N3); times follow Boost integrate_const
with dt = step_size and the pulse drivers' observer gate (the t = T_start
sample is dropped, mixed_lattice.h:4797-4799).  The reference's pulse-window
chunking integrates grid-snapped segments with the same step, so for RK4 it
gives the same samples as one unchunked run; with "dopri5" every segment is
its own integrate_const call of the controlled stepper.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import check, dptr, f64, lib, szptr
from .gpu import _gate
from .mixed_model import MixedLattice

REC = 34


def split_records(r):
    """[..., 34] -> dict(M2 [..., 3, 3], M3 [..., 3, 8], energy [...])."""
    r = np.asarray(r)
    return {"M2": r[..., :9].reshape(r.shape[:-1] + (3, 3)),
            "M3": r[..., 9:33].reshape(r.shape[:-1] + (3, 8)),
            "energy": r[..., 33]}


class GPUMixedLattice:
    """Owns a device copy of a MixedLattice's tables (csl_mixed_create)."""

    def __init__(self, ml: MixedLattice):
        self.lattice = ml
        s2, s3 = ml.su2, ml.su3
        if s2.spin_dim != 3 or s3.spin_dim != 8:
            raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "mixed path handles spin_dim 3 + 8")
        self.A2, self.A3 = s2.N_atoms, s3.N_atoms
        self.N2, self.N3 = s2.lattice_size, s3.lattice_size
        n_cells = self.N2 // self.A2
        u = lambda a: np.ascontiguousarray(a, dtype=np.uintp)
        keep = [f64(s2.field), f64(s2.onsite), u(s2.bi_offset), u(s2.bi_partner), f64(s2.bi_J),
                f64(s3.field), f64(s3.onsite), u(s3.bi_offset), u(s3.bi_partner), f64(s3.bi_J),
                u(ml.m2_offset), u(ml.m2_partner), f64(ml.m2_J),
                u(ml.m3_offset), u(ml.m3_partner), f64(ml.m3_J)]
        h = C.c_void_p()
        args = [szptr(a) if a.dtype == np.uintp else dptr(a) for a in keep]
        check(lib().csl_mixed_create(self.A2, self.A3, n_cells, *args, C.byref(h)))
        self.h = h
        self.pulse_centres = (0.0, 0.0)
        check(lib().csl_mixed_set_frames(self.h, dptr(f64(s2.frames)), dptr(f64(s2.afm_signs)),
                                         dptr(f64(s3.frames))))

    # -- lifetime
    def close(self):
        if getattr(self, "h", None):
            lib().csl_mixed_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- state and parameters
    @property
    def n_state(self):
        return self.N2 * 3 + self.N3 * 8

    def info(self):
        n2, n3 = C.c_size_t(), C.c_size_t()
        u = [C.c_int() for _ in range(4)]
        check(lib().csl_mixed_info(self.h, C.byref(n2), C.byref(n3), *(C.byref(x) for x in u)))
        return {"N2": n2.value, "N3": n3.value, "uniform": tuple(bool(x.value) for x in u)}

    def set_spins(self, flat):
        check(lib().csl_mixed_set_spins(self.h, dptr(f64(flat).reshape(-1))))

    def get_spins(self) -> np.ndarray:
        out = np.empty(self.n_state)
        check(lib().csl_mixed_get_spins(self.h, dptr(out)))
        return out

    def set_pulse(self, species: int, drive, amp, width, freq, t1, t2):
        """set_gpu_pulse_SU2 / SU3: drive [2, n_atoms, D] already F^T B."""
        check(lib().csl_mixed_set_pulse(self.h, species, dptr(f64(drive).reshape(-1)), amp, width, freq, t1, t2))
        self.pulse_centres = (float(t1), float(t2))

    def set_damping(self, alpha=0.0, gamma=None, eq=None):
        check(lib().csl_mixed_set_damping(self.h, alpha, dptr(None if gamma is None else f64(gamma)),
                                          dptr(None if eq is None else f64(eq).reshape(-1))))

    def record(self):
        out = np.empty(REC)
        check(lib().csl_mixed_record(self.h, dptr(out)))
        return out

    def energy(self) -> float:
        e = C.c_double()
        check(lib().csl_mixed_energy(self.h, C.byref(e)))
        return e.value

    def rhs(self, t: float) -> np.ndarray:
        out = np.empty(self.n_state)
        check(lib().csl_mixed_rhs(self.h, t, dptr(out)))
        return out

    def set_stage_graph(self, enable: bool):
        """csl_mixed_set_stage_graph: replay the batched launch sequence from
        a hipGraph."""
        check(lib().csl_mixed_set_stage_graph(self.h, int(bool(enable))))

    def set_persistent(self, mode: int):
        """csl_mixed_set_persistent: 0 the persistent batched kernel when the
        lattice qualifies (default), -1 the stage kernels, -2 the cooperative
        stage kernel (the whole batch's stage workgroups in one launch over
        the time grid), n > 0 the persistent kernel with at most n
        workgroups."""
        check(lib().csl_mixed_set_persistent(self.h, int(mode)))

    def last_path(self):
        """csl_mixed_last_path of the last integrate_batched:
        {"persistent": bool, "cooperative": bool, "workgroups": int, "split": int,
        "kernel_ms": float} (split: trajectories run in two pieces by two
        workgroups; kernel_ms: the persistent or cooperative launch)."""
        p, w, t, ms = C.c_int(), C.c_int(), C.c_size_t(), C.c_double()
        check(lib().csl_mixed_last_path(self.h, C.byref(p), C.byref(w), C.byref(t), C.byref(ms)))
        return {"persistent": p.value == 1, "cooperative": p.value == 2, "workgroups": w.value,
                "split": t.value, "kernel_ms": ms.value}

    def normalize(self, len2=1.0, len3=1.0):
        check(lib().csl_mixed_normalize(self.h, len2, len3))

    def step(self, t, dt, method="rk4"):
        check(lib().csl_mixed_step(self.h, t, dt, method.encode()))

    def drive_dirs(self, species: int, field_in) -> np.ndarray:
        """F^T B per atom (set_pulse_SU2/SU3, mixed_lattice_md.cpp:130-135, 208-213)."""
        lat = self.lattice.su2 if species == 2 else self.lattice.su3
        fi = np.asarray(field_in, dtype=float).reshape(lat.N_atoms, lat.spin_dim)
        return np.einsum("aji,aj->ai", lat.frames, fi)

    # -- integration
    def integrate(self, T0, T1, dt, save_interval=1, method="rk4", tgrid="boost", want_states=False):
        _, nr = _count(T0, T1, dt, save_interval, tgrid)
        times = np.empty(nr)
        recs = np.empty(nr * REC)
        states = np.empty(nr * self.n_state) if want_states else None
        n = C.c_size_t()
        check(lib().csl_mixed_integrate(self.h, T0, T1, dt, save_interval, method.encode(), _capi.TGRID[tgrid], nr,
                                        dptr(times), dptr(states), dptr(recs), C.byref(n)))
        out = {"times": times, "records": recs.reshape(nr, REC)}
        if want_states:
            out["states"] = states.reshape(nr, self.n_state)
        return out

    def integrate_batched(self, init, t_pulse, pulse_scale=None, T0=0.0, T1=1.0, dt=0.01, save_interval=1,
                          method="rk4", tgrid="boost", want_final=False):
        tp = f64(t_pulse).reshape(-1, 2)
        B = tp.shape[0]
        init = f64(init)
        per = 1 if init.ndim == 2 else 0
        if per and init.shape[0] != B:
            raise ValueError("init has a different batch size than t_pulse")
        sc = None if pulse_scale is None else f64(pulse_scale).reshape(B, 2)
        _, nr = _count(T0, T1, dt, save_interval, tgrid)
        times = np.empty(nr)
        recs = np.empty(nr * B * REC)
        final = np.empty(B * self.n_state) if want_final else None
        n = C.c_size_t()
        check(lib().csl_mixed_integrate_batched(self.h, B, dptr(init.reshape(-1)), per, dptr(tp.reshape(-1)),
                                                dptr(None if sc is None else sc.reshape(-1)), T0, T1, dt,
                                                save_interval, method.encode(), _capi.TGRID[tgrid], nr, dptr(times),
                                                dptr(recs), dptr(final), C.byref(n)))
        return {"times": times, "records": recs.reshape(nr, B, REC),
                "final": None if final is None else final.reshape(B, self.n_state)}

    def integrate_dopri5(self, init, t_pulse, pulse_scale=None, T0=0.0, T1=1.0, dt=0.01, control=1,
                         abs_tol=1e-6, rel_tol=1e-6, save_interval=1, tgrid="boost", segments=None,
                         max_rec=None, max_attempts=0, want_final=False):
        """csl_mixed_integrate_dopri5: Dormand-Prince 5(4) for B trajectories
        (each its own step-size sequence), controls as
        GPULattice.integrate_dopri5; segments: B sequences of (t0, t1[, dt])
        rows for control 2.  Returns dict(times, records: lists of
        per-trajectory arrays [n_b], [n_b, 34], final [B, n_state], n_rec,
        accepted, rejected)."""
        from .gpu import _const_calls
        tp = f64(t_pulse).reshape(-1, 2)
        B = tp.shape[0]
        init = f64(init)
        per = 1 if init.ndim == 2 else 0
        if per and init.shape[0] != B:
            raise ValueError("init has a different batch size than t_pulse")
        sc = None if pulse_scale is None else f64(pulse_scale).reshape(B, 2)
        sg = nsg = None
        ms = 0
        if segments is not None:
            rows = [np.asarray(r, dtype=float).reshape(-1, np.shape(r)[-1])[:, :2] for r in segments]
            if len(rows) != B:
                raise ValueError("segments needs one sequence per trajectory")
            ms = max(1, max(len(r) for r in rows))
            sg = np.zeros((B, ms, 2))
            for b, r in enumerate(rows):
                sg[b, :len(r)] = r
            nsg = np.array([len(r) for r in rows], dtype=np.uintp)
        if max_rec is None:
            if sg is not None:
                max_rec = max(sum(_const_calls(t0, t1, dt) for t0, t1 in r) for r in rows)
            else:
                _, nr = _count(T0, T1, dt, 1, tgrid)
                max_rec = nr + 1 if control != 1 else max(64, 2 * nr)
        while True:
            times = np.empty(B * max_rec)
            recs = np.empty(B * max_rec * REC)
            final = np.empty(B * self.n_state) if want_final else None
            nrec = np.zeros(B, dtype=np.uintp)
            nacc = np.zeros(B, dtype=np.uintp)
            nrej = np.zeros(B, dtype=np.uintp)
            rc = lib().csl_mixed_integrate_dopri5(
                self.h, B, dptr(init.reshape(-1)), per, dptr(tp.reshape(-1)),
                dptr(None if sc is None else sc.reshape(-1)), T0, T1, dt, control, abs_tol, rel_tol, save_interval,
                _capi.TGRID[tgrid], dptr(None if sg is None else sg.reshape(-1)), szptr(nsg), ms, max_attempts,
                max_rec, dptr(times), dptr(recs), dptr(final), szptr(nrec), szptr(nacc), szptr(nrej))
            if rc == _capi.CSL_ERR_CAPACITY and int(nrec.max()) > max_rec:
                max_rec = int(nrec.max())
                continue
            check(rc)
            break
        n = [int(v) for v in nrec]
        t2 = times.reshape(B, max_rec)
        r2 = recs.reshape(B, max_rec, REC)
        return {"times": [t2[b, :n[b]] for b in range(B)], "records": [r2[b, :n[b]] for b in range(B)],
                "final": None if final is None else final.reshape(B, self.n_state), "n_rec": n,
                "accepted": [int(v) for v in nacc], "rejected": [int(v) for v in nrej]}

    def scale_gradient(self, init, t_pulse, target, weights, pulse_scale=None, T0=0.0, T1=1.0, dt=0.01,
                       save_interval=1, tgrid="boost", want_records=False):
        """csl_mixed_scale_gradient: per trajectory the loss
        sum_t sum_j w_j (rec_j(t) - target_j(t))^2 and its derivative with
        respect to a common scale of the Fe-Tm couplings (at 1), exact for
        the discrete RK4 scheme.  target [n_rec, 34] shared or [n_rec, B, 34];
        weights [34] (entry 33, the energy, must be 0).  Returns dict(loss [B],
        grad [B], times[, records, trecords [n_rec, B, 34]])."""
        tp = f64(t_pulse).reshape(-1, 2)
        B = tp.shape[0]
        init = f64(init)
        per = 1 if init.ndim == 2 else 0
        if per and init.shape[0] != B:
            raise ValueError("init has a different batch size than t_pulse")
        sc = None if pulse_scale is None else f64(pulse_scale).reshape(B, 2)
        ns, nr = _count(T0, T1, dt, save_interval, tgrid)
        tgt = f64(target)
        per_t = 1 if tgt.ndim == 3 else 0
        if tgt.shape != ((nr, B, REC) if per_t else (nr, REC)):
            raise ValueError(f"target must be [{nr}, 34] or [{nr}, {B}, 34]")
        w = f64(weights).reshape(REC)
        loss, grad = np.empty(B), np.empty(B)
        recs = np.empty(nr * B * REC) if want_records else None
        trecs = np.empty(nr * B * REC) if want_records else None
        n = C.c_size_t()
        check(lib().csl_mixed_scale_gradient(self.h, B, dptr(init.reshape(-1)), per, dptr(tp.reshape(-1)),
                                             dptr(None if sc is None else sc.reshape(-1)), T0, T1, dt, save_interval,
                                             _capi.TGRID[tgrid], dptr(tgt.reshape(-1)), per_t, dptr(w), dptr(loss),
                                             dptr(grad), nr, dptr(recs), dptr(trecs), C.byref(n)))
        # record times on the Boost grid t_k = T0 + k dt (other grids: None)
        times = (np.array([T0 + min(r * save_interval, ns) * dt for r in range(nr)]) if tgrid == "boost"
                 else None)
        out = {"loss": loss, "grad": grad, "times": times}
        if want_records:
            out["records"] = recs.reshape(nr, B, REC)
            out["trecords"] = trecs.reshape(nr, B, REC)
        return out

    def last_kernel_ms(self):
        ms, n = C.c_double(), C.c_size_t()
        check(lib().csl_mixed_last_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


def _count(T0, T1, dt, save_interval, tgrid):
    st, nr = C.c_size_t(), C.c_size_t()
    check(lib().csl_count_records(T0, T1, dt, save_interval, _capi.TGRID[tgrid], C.byref(st), C.byref(nr)))
    return st.value, nr.value


def _set_pulses(gm: GPUMixedLattice, f2a, f2b, f3a, f3b, t1, t2, amp2, w2, fr2, amp3, w3, fr3):
    d2 = np.stack([gm.drive_dirs(2, f2a), gm.drive_dirs(2, f2b)])
    d3 = np.stack([gm.drive_dirs(3, f3a), gm.drive_dirs(3, f3b)])
    gm.set_pulse(2, d2, amp2, w2, fr2, t1, t2)
    gm.set_pulse(3, d3, amp3, w3, fr3, t1, t2)


def _reset(gm: GPUMixedLattice):
    gm.set_pulse(2, np.zeros((2, gm.A2, 3)), 0.0, 1.0, 0.0, 0.0, 0.0)
    gm.set_pulse(3, np.zeros((2, gm.A3, 8)), 0.0, 1.0, 0.0, 0.0, 0.0)


def molecular_dynamics(gm: GPUMixedLattice, flat, T_start, T_end, dt, save_interval=1, method="rk4",
                       abs_tol=-1.0, rel_tol=-1.0):
    """MixedLattice::molecular_dynamics_cpu (mixed_lattice_md.cpp:1208-1333)
    without the HDF5 file: "rk4" records every save_interval integrate_const
    calls; "dopri5" (the reference default) runs integrate_adaptive with
    the controlled Dormand-Prince stepper and records every save_interval-th
    accepted step; tolerances <= 0 fall back to 1e-6
    (get_integration_tolerances, mixed_lattice.h:4296-4301)."""
    _reset(gm)
    if method == "dopri5":
        a = abs_tol if abs_tol > 0.0 else 1e-6
        r_ = rel_tol if rel_tol > 0.0 else 1e-6
        r = gm.integrate_dopri5(np.asarray(flat, float).reshape(-1), [[0.0, 0.0]], None, T_start, T_end, dt,
                                control=1, abs_tol=a, rel_tol=r_, save_interval=save_interval)
        out = split_records(r["records"][0])
        out["times"] = r["times"][0]
        out["accepted"], out["rejected"] = r["accepted"][0], r["rejected"][0]
        return out
    if method != "rk4":
        raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "methods: 'rk4', 'dopri5'")
    gm.set_spins(flat)
    r = gm.integrate(T_start, T_end, dt, save_interval, method)
    out = split_records(r["records"])
    out["times"] = r["times"]
    return out


def single_pulse_drive(gm, flat, field_in_SU2, field_in_SU3, t_B, amp2, w2, fr2, amp3, w3, fr3,
                       T_start, T_end, step_size, method="rk4", pulse_window_chunking=True,
                       abs_tol=1e-8, rel_tol=1e-8):
    """MixedLattice::single_pulse_drive (mixed_lattice.h:4754-4881): returns
    dict(times, M2, M3, energy).  "dopri5": integrate_const with the
    controlled stepper, by pulse-window segments (window 9 max(w2, w3))
    with pulse_window_chunking."""
    z2, z3 = np.zeros((gm.A2, 3)), np.zeros((gm.A3, 8))
    _set_pulses(gm, field_in_SU2, z2, field_in_SU3, z3, t_B, 0.0, amp2, w2, fr2, amp3, w3, fr3)
    return _gated_run(gm, flat, T_start, T_end, step_size, method, (t_B,), max(w2, w3), pulse_window_chunking,
                      abs_tol, rel_tol)


def double_pulse_drive(gm, flat, field_in_SU2_1, field_in_SU3_1, t_B_1, field_in_SU2_2, field_in_SU3_2, t_B_2,
                       amp2, w2, fr2, amp3, w3, fr3, T_start, T_end, step_size, method="rk4",
                       pulse_window_chunking=True, abs_tol=1e-8, rel_tol=1e-8):
    """MixedLattice::double_pulse_drive: returns dict(times, M2, M3, energy);
    the segments surround both pulse centres."""
    _set_pulses(gm, field_in_SU2_1, field_in_SU2_2, field_in_SU3_1, field_in_SU3_2, t_B_1, t_B_2, amp2, w2, fr2,
                amp3, w3, fr3)
    return _gated_run(gm, flat, T_start, T_end, step_size, method, (t_B_1, t_B_2), max(w2, w3),
                      pulse_window_chunking, abs_tol, rel_tol)


def _gated_run(gm, flat, T_start, T_end, step, method="rk4", centers=(), width=1.0, chunking=True,
               abs_tol=1e-8, rel_tol=1e-8):
    if method == "dopri5":
        from .gpu import PULSE_WINDOW_WIDTHS, pulse_segments
        segs = None
        if chunking:
            segs = pulse_segments(T_start, T_end, centers, PULSE_WINDOW_WIDTHS * width, step)
            if len(segs) == 0:
                _reset(gm)
                out = split_records(np.empty((0, REC)))
                out["times"] = np.empty(0)
                return out
            segs = [segs]
        # the handle's pulse centres are its own t_pulse (set by _set_pulses)
        tp = [[gm.pulse_centres[0], gm.pulse_centres[1]]]
        r = gm.integrate_dopri5(np.asarray(flat, float).reshape(-1), tp, None, T_start, T_end, step, control=2,
                                abs_tol=abs_tol, rel_tol=rel_tol, segments=segs)
        times, recs = r["times"][0], r["records"][0]
    elif method == "rk4":
        gm.set_spins(flat)
        r = gm.integrate(T_start, T_end, step, 1, "rk4")
        times, recs = r["times"], r["records"]
    else:
        raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "methods: 'rk4', 'dopri5'")
    _reset(gm)
    keep = _gate(times, T_start, T_end, step)
    out = split_records(recs[keep])
    out["times"] = times[keep]
    return out


def pump_probe_spectroscopy(gm, ground, field_in_SU2, field_in_SU3, amp2, w2, fr2, amp3, w3, fr3,
                            tau_start, tau_end, tau_step, T_start, T_end, T_step):
    """MixedLattice::pump_probe_spectroscopy without HDF5 output: M0 (pump
    at 0), M1(tau) (pump at tau) and M01(tau) (pumps at 0 and tau) for
    n_tau = floor(|tau_end - tau_start| / tau_step) + 1 delays, integrated as
    1 + 2 n_tau trajectories in one batched call.  Returns dict(times, tau,
    M0, M1, M01) with each M a split_records dict."""
    n_tau = int(abs((tau_end - tau_start) / tau_step)) + 1
    taus = tau_start + tau_step * np.arange(n_tau)
    _set_pulses(gm, field_in_SU2, field_in_SU2, field_in_SU3, field_in_SU3, 0.0, 0.0, amp2, w2, fr2,
                amp3, w3, fr3)
    B = 1 + 2 * n_tau
    tp = np.zeros((B, 2))
    sc = np.zeros((B, 2))
    sc[0, 0] = 1.0
    tp[1:1 + n_tau, 0] = taus
    sc[1:1 + n_tau, 0] = 1.0
    tp[1 + n_tau:, 1] = taus
    sc[1 + n_tau:, :] = 1.0
    r = gm.integrate_batched(np.asarray(ground, float).reshape(-1), tp, sc, T_start, T_end, T_step)
    _reset(gm)
    keep = _gate(r["times"], T_start, T_end, T_step)
    rec = r["records"][keep]
    return {"times": r["times"][keep], "tau": taus, "M0": split_records(rec[:, 0]),
            "M1": split_records(np.swapaxes(rec[:, 1:1 + n_tau], 0, 1)),
            "M01": split_records(np.swapaxes(rec[:, 1 + n_tau:], 0, 1))}
