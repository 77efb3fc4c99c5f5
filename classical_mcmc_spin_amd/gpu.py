"""Device lattice handle and the reference's simulator entry points on it.

GPULattice wraps one csl_lattice handle (include/csl_gpu.h).  The driver
functions below mirror the reference's Lattice simulator API for the
Landau–Lifshitz path (src/core/lattice_md.cpp):

  molecular_dynamics        Lattice::molecular_dynamics_cpu, method "rk4"
                            (lattice_md.cpp:341-453)
  single_pulse_drive        Lattice::single_pulse_drive (:493-586)
  double_pulse_drive        Lattice::double_pulse_drive (:589-680)
  pump_probe_spectroscopy   Lattice::pump_probe_spectroscopy (:683-926)

with the same argument meaning, the same time grid (Boost integrate_const,
lattice.h:4034-4039) and the same observer gating; every integration runs on
the GPU through libcsl_hip.so.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

from . import _capi
from ._capi import check, dptr, f64, lib, szptr
from .model import Lattice


class GPULattice:
    """Owns a device copy of a Lattice's tables (csl_create_lattice_csr)."""

    def __init__(self, lattice: Lattice, padded: bool = False):
        if lattice.spin_dim != 3:
            raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "spin_dim != 3 needs the mixed-lattice API")
        self.lattice = lattice
        L = lib()
        h = C.c_void_p()
        N, D, A = lattice.lattice_size, lattice.spin_dim, lattice.N_atoms
        field = f64(lattice.field)
        onsite = f64(lattice.onsite)
        if padded:  # reference create_gpu_lattice_data layout
            bil, par, cnt = lattice.padded_tables()
            bil = f64(bil)
            check(L.csl_create_lattice(N, D, A, lattice.num_bi, dptr(field), dptr(onsite), dptr(bil),
                                       szptr(np.ascontiguousarray(par)), szptr(np.ascontiguousarray(cnt)),
                                       C.byref(h)))
        else:
            J = f64(lattice.bi_J)
            check(L.csl_create_lattice_csr(N, D, A, dptr(field), dptr(onsite),
                                           szptr(np.ascontiguousarray(lattice.bi_offset)),
                                           szptr(np.ascontiguousarray(lattice.bi_partner)), dptr(J),
                                           C.byref(h)))
        self.h = h
        self.N, self.D, self.A = N, D, A
        self.set_frames(lattice.frames, lattice.afm_signs)

    # -- lifetime
    def close(self):
        if self.h:
            check(lib().csl_destroy_lattice(self.h))
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

    def info(self):
        vals = [C.c_size_t() for _ in range(5)]
        u = C.c_int()
        check(lib().csl_lattice_info(self.h, *[C.byref(v) for v in vals], C.byref(u)))
        keys = ["N", "D", "n_atoms", "n_padded", "max_bonds"]
        d = {k: v.value for k, v in zip(keys, vals)}
        d["uniform"] = bool(u.value)
        return d

    # -- state and parameters
    def set_frames(self, frames, afm_signs):
        check(lib().csl_set_frames(self.h, dptr(f64(frames)), dptr(f64(afm_signs))))

    def set_spins(self, spins):
        s = f64(spins).reshape(-1)
        check(lib().csl_set_spins(self.h, dptr(s), s.size))

    def get_spins(self) -> np.ndarray:
        out = np.empty(self.N * self.D)
        check(lib().csl_get_spins(self.h, dptr(out), out.size))
        return out.reshape(self.N, self.D)

    def set_pulse(self, field_drive, amp, width, freq, t1, t2):
        fd = f64(field_drive).reshape(-1)
        if fd.size != 2 * self.A * self.D:
            raise ValueError("field_drive must hold 2*N_atoms*spin_dim values")
        check(lib().csl_set_pulse(self.h, dptr(fd), amp, width, freq, t1, t2))

    def set_damping(self, alpha: float):
        check(lib().csl_set_damping(self.h, float(alpha)))

    # -- evaluation
    def energy(self) -> float:
        e = C.c_double()
        check(lib().csl_energy(self.h, C.byref(e)))
        return e.value

    def local_field(self) -> np.ndarray:
        out = np.empty(self.N * 3)
        check(lib().csl_local_field(self.h, dptr(out), out.size))
        return out.reshape(self.N, 3)

    def rhs(self, t: float) -> np.ndarray:
        out = np.empty(self.N * 3)
        check(lib().csl_rhs(self.h, float(t), dptr(out), out.size))
        return out.reshape(self.N, 3)

    def magnetization(self):
        ml, ms = np.empty(3), np.empty(3)
        check(lib().csl_magnetization(self.h, dptr(ml), dptr(ms)))
        return ml, ms

    def normalize(self, spin_length: float = 1.0):
        check(lib().csl_normalize(self.h, float(spin_length)))

    def step(self, t: float, dt: float, method: str = "rk4"):
        check(lib().csl_step(self.h, float(t), float(dt), method.encode()))

    # -- integration
    @staticmethod
    def count_records(T0, T1, dt, save_interval=1, tgrid="gpu"):
        ns, nr = C.c_size_t(), C.c_size_t()
        check(lib().csl_count_records(T0, T1, dt, save_interval, _capi.TGRID[tgrid], C.byref(ns), C.byref(nr)))
        return ns.value, nr.value

    def integrate(self, T0, T1, dt, save_interval=1, method="rk4", tgrid="gpu", engine=0,
                  want_states=False, want_energy=True):
        """gpu::integrate_gpu (lattice_gpu_api.h:117-125) plus M/E records.
        Returns dict(times, mags [n,3,3], energy [n], states [n,N,3] or None)."""
        _, nr = self.count_records(T0, T1, dt, save_interval, tgrid)
        times = np.empty(nr)
        mags = np.empty(nr * 9)
        energy = np.empty(nr) if want_energy else None
        states = np.empty(nr * self.N * 3) if want_states else None
        n = C.c_size_t()
        check(lib().csl_integrate(self.h, T0, T1, dt, save_interval, method.encode(), _capi.TGRID[tgrid],
                                  engine, nr, dptr(times), dptr(states), dptr(mags), dptr(energy),
                                  C.byref(n)))
        return {"times": times, "mags": mags.reshape(nr, 3, 3), "energy": energy,
                "states": None if states is None else states.reshape(nr, self.N, 3)}

    def integrate_batched(self, init, t_pulse, pulse_amp=None, T0=0.0, T1=1.0, dt=0.01,
                          save_interval=1, method="rk4", tgrid="batched", block=0,
                          want_energy=True, want_final=False, want_mags=True):
        """Generalised gpu::integrate_gpu_batched (lattice_gpu_api.h:200-214).
        init: [N,3] shared or [B,N,3]; t_pulse, pulse_amp: [B,2].
        Returns dict(times, mags [n,B,3,3], energy [n,B], final [B,N,3]);
        with want_mags=False the records stay in device memory (mags None)."""
        tp = f64(t_pulse).reshape(-1, 2)
        B = tp.shape[0]
        init = f64(init)
        per = 1 if init.ndim == 3 else 0
        if per and init.shape[0] != B:
            raise ValueError("init has a different batch size than t_pulse")
        amp = None if pulse_amp is None else f64(pulse_amp).reshape(B, 2)
        _, nr = self.count_records(T0, T1, dt, save_interval, tgrid)
        times = np.empty(nr)
        mags = np.empty(nr * B * 9) if want_mags else None
        energy = np.empty(nr * B) if want_energy else None
        final = np.empty(B * self.N * 3) if want_final else None
        n = C.c_size_t()
        check(lib().csl_integrate_batched(self.h, B, dptr(init.reshape(-1)), per, dptr(tp.reshape(-1)),
                                          dptr(None if amp is None else amp.reshape(-1)), T0, T1, dt,
                                          save_interval, method.encode(), _capi.TGRID[tgrid], block, nr,
                                          dptr(times), dptr(mags), dptr(energy), dptr(final), C.byref(n)))
        return {"times": times, "mags": None if mags is None else mags.reshape(nr, B, 3, 3),
                "energy": None if energy is None else energy.reshape(nr, B),
                "final": None if final is None else final.reshape(B, self.N, 3)}

    def quench(self, states, max_sweeps=10000, rel_tol=1e-12, spin_length=1.0):
        """csl_quench: relaxed copies of states ([N,3] or [B,N,3]) and the
        sweep count of each (Lattice::greedy_quench, colour-class order)."""
        st = f64(states).copy()
        single = st.ndim == 2
        st = st.reshape(-1, self.N, 3)
        B = st.shape[0]
        done = np.zeros(B, dtype=np.uintp)
        check(lib().csl_quench(self.h, B, dptr(st.reshape(-1)), max_sweeps, rel_tol, spin_length, szptr(done)))
        return (st[0] if single else st), done.astype(np.int64)

    def structure_factor(self, init, T0, T1, dt, sample_interval, dims, positions, tgrid="boost",
                         want_mags=False):
        """csl_structure_factor: S(q, w) [L1*L2*L3, n_t] of the undriven
        trajectories started from init ([N,3] or [B,N,3]); positions are the
        atoms' fractional coordinates [n_atoms, 3]."""
        init = f64(init)
        B = init.shape[0] if init.ndim == 3 else 1
        per = 1 if init.ndim == 3 else 0
        L1, L2, L3 = (int(v) for v in dims)
        n_steps, _ = self.count_records(T0, T1, dt, 1, tgrid)
        nt_max = n_steps // sample_interval
        sqw = np.empty(L1 * L2 * L3 * max(nt_max, 1))
        mags = np.empty((n_steps + 1) * B * 9) if want_mags else None
        nt = C.c_size_t()
        check(lib().csl_structure_factor(self.h, B, dptr(init.reshape(-1)), per, T0, T1, dt, _capi.TGRID[tgrid],
                                         sample_interval, L1, L2, L3, dptr(f64(positions).reshape(-1)), nt_max,
                                         dptr(sqw), C.byref(nt), dptr(mags)))
        out = {"sqw": sqw[:L1 * L2 * L3 * nt.value].reshape(L1 * L2 * L3, nt.value)}
        if want_mags:
            out["mags"] = mags.reshape(n_steps + 1, B, 3, 3)
        return out

    def autotune(self, B: int, n_probe_steps: int = 64):
        blk, ms = C.c_int(), C.c_double()
        check(lib().csl_autotune_batched(self.h, B, n_probe_steps, C.byref(blk), C.byref(ms)))
        return blk.value, ms.value

    def set_replica_hamiltonians(self, lattices=None, J=None, field=None):
        """Per-trajectory Hamiltonians for integrate_batched: either a list of
        Lattice objects with this lattice's structure, or arrays J [B,nnz,3,3]
        and field [B,N,3].  None clears them."""
        if lattices is None and J is None:
            check(lib().csl_set_replica_hamiltonians(self.h, 0, None, None))
            return
        if lattices is not None:
            J = np.stack([l.bi_J for l in lattices])
            field = np.stack([l.field for l in lattices])
        J = f64(J)
        field = f64(field)
        check(lib().csl_set_replica_hamiltonians(self.h, J.shape[0], dptr(J.reshape(-1)),
                                                 dptr(field.reshape(-1))))

    def set_launch_config(self, B, block, sites_per_thread, flags):
        """flags: bit 0 two LDS state buffers, bit 1 compile-time RK stages."""
        check(lib().csl_set_launch_config(self.h, B, block, sites_per_thread, int(flags)))

    def set_stage_graph(self, enable: bool):
        """csl_set_stage_graph: replay the batched stage-kernel launch
        sequence from a hipGraph (lattices larger than one workgroup)."""
        check(lib().csl_set_stage_graph(self.h, int(bool(enable))))

    def set_partition(self, parts: int):
        """csl_set_partition: 0 = the partitioned kernel for lattices larger
        than one workgroup's LDS (default), P > 0 = always with P parts,
        -1 = never (grid-wide stage kernels)."""
        check(lib().csl_set_partition(self.h, int(parts)))

    def last_partition(self):
        """(parts per trajectory, trajectory groups) of the last batched
        launch; (0, 0) when it did not use the partitioned kernel."""
        p, g = C.c_int(), C.c_int()
        check(lib().csl_last_partition(self.h, C.byref(p), C.byref(g)))
        return p.value, g.value

    def set_batch_split(self, enable: bool):
        """csl_set_batch_split: share the trajectory-steps of a batch larger
        than the co-resident workgroups evenly over them (default on)."""
        check(lib().csl_set_batch_split(self.h, int(bool(enable))))

    def last_grid(self) -> int:
        g = C.c_int()
        check(lib().csl_last_grid(self.h, C.byref(g)))
        return g.value

    def last_launch_config(self):
        b, s, d, v = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib().csl_last_launch_config(self.h, C.byref(b), C.byref(s), C.byref(d)))
        check(lib().csl_last_launch_variant(self.h, C.byref(v)))
        return {"block": b.value, "sites_per_thread": s.value, "double_buffer": bool(d.value & 1),
                "unrolled_stages": bool(d.value & 2), "flags": d.value,
                "field_only": bool(v.value & 1), "atom_major": bool(v.value & 2),
                "partitioned": bool(v.value & 4)}

    def launch_candidates(self):
        """csl_persistent_candidates: every persistent launch shape that fits
        this lattice, as dicts (block, sites_per_thread, flags, field_only,
        atom_major)."""
        n = C.c_size_t()
        m = 64
        blk, spt, fl, var = (np.zeros(m, dtype=np.int32) for _ in range(4))
        ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
        check(lib().csl_persistent_candidates(self.h, m, ip(blk), ip(spt), ip(fl), ip(var), C.byref(n)))
        return [{"block": int(blk[i]), "sites_per_thread": int(spt[i]), "flags": int(fl[i]),
                 "field_only": bool(var[i] & 1), "atom_major": bool(var[i] & 2)} for i in range(n.value)]

    def last_kernel_ms(self):
        ms, n = C.c_double(), C.c_size_t()
        check(lib().csl_last_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def pump_probe(self, ground, field_in, amp, width, freq, tau_start, tau_end, tau_step,
                   T0, T1, T_step, reuse_m0_for_m1=False, stationarity_tol=1e-6, method="rk4",
                   abs_tol=1e-8, rel_tol=1e-8, pulse_window_chunking=True):
        """csl_pump_probe_ex: M0, M1(tau), M01(tau) in one batched launch
        (method "rk4" on the Boost grid, or "dopri5" with step control per
        replica, by pulse-window segments with pulse_window_chunking); with
        reuse_m0_for_m1 and a stationary ground state M1 is the time shift of
        M0 (W1).  The result carries w1 (bool) and max_dsdt."""
        g = f64(ground).reshape(-1)
        fi = f64(field_in).reshape(self.A, 3)
        n_tau_max = int(abs((tau_end - tau_start) / tau_step)) + 2
        ns, _ = self.count_records(T0, T1, T_step, 1, "boost")
        max_rec = ns + 1
        times = np.empty(max_rec)
        M0 = np.empty(max_rec * 9)
        M1 = np.empty(n_tau_max * max_rec * 9)
        M01 = np.empty(n_tau_max * max_rec * 9)
        taus = np.empty(n_tau_max)
        nr, nt = C.c_size_t(), C.c_size_t()
        w1, mx = C.c_int(0), C.c_double(float("nan"))
        check(lib().csl_pump_probe_ex(self.h, dptr(g), dptr(fi.reshape(-1)), amp, width, freq, tau_start,
                                      tau_end, tau_step, T0, T1, T_step, int(bool(reuse_m0_for_m1)),
                                      float(stationarity_tol), int(bool(pulse_window_chunking)), method.encode(),
                                      float(abs_tol), float(rel_tol),
                                      max_rec, n_tau_max, C.byref(nr), C.byref(nt), dptr(times), dptr(M0),
                                      dptr(M1), dptr(M01), dptr(taus), C.byref(w1), C.byref(mx)))
        r, t = nr.value, nt.value
        return {"times": times[:r], "M0": M0.reshape(max_rec, 3, 3)[:r],
                "M1": M1.reshape(n_tau_max, max_rec, 3, 3)[:t, :r],
                "M01": M01.reshape(n_tau_max, max_rec, 3, 3)[:t, :r], "tau": taus[:t],
                "w1": bool(w1.value), "max_dsdt": mx.value}

    def integrate_dopri5(self, init, T0, T1, dt, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=1,
                         tgrid="boost", t_pulse=None, pulse_amp=None, want_energy=True, want_final=True,
                         max_rec=None, max_attempts=0, segments=None):
        """csl_integrate_dopri5: Dormand-Prince 5(4) for B trajectories
        (init [N,3] -> B from t_pulse, or [B,N,3]); control 0 fixed grid
        steps, 1 integrate_adaptive, 2 integrate_const with step control,
        with segments (B sequences of (t0, t1[, dt_init]) rows) one
        integrate_const call per segment.  Returns dict(times, mags, energy:
        lists of per-trajectory arrays, final [B,N,3], n_rec, accepted,
        rejected [B]); a too small record buffer is enlarged and the
        (deterministic) run repeated."""
        init = f64(init)
        per = 1 if init.ndim == 3 else 0
        tp = None if t_pulse is None else f64(t_pulse).reshape(-1, 2)
        B = init.shape[0] if per else (1 if tp is None else tp.shape[0])
        if tp is not None and tp.shape[0] != B:
            raise ValueError("t_pulse has a different batch size than init")
        amp = None if pulse_amp is None else f64(pulse_amp).reshape(B, 2)
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
                ns, nr = self.count_records(T0, T1, dt, 1, tgrid)
                max_rec = nr + 1 if control != 1 else max(64, 2 * nr)
        while True:
            times = np.empty(B * max_rec)
            mags = np.empty(B * max_rec * 9)
            energy = np.empty(B * max_rec) if want_energy else None
            final = np.empty(B * self.N * 3) if want_final else None
            nrec = np.zeros(B, dtype=np.uintp)
            nacc = np.zeros(B, dtype=np.uintp)
            nrej = np.zeros(B, dtype=np.uintp)
            rc = lib().csl_integrate_dopri5(self.h, B, dptr(init.reshape(-1)), per,
                                            dptr(None if tp is None else tp.reshape(-1)),
                                            dptr(None if amp is None else amp.reshape(-1)), T0, T1, dt, control,
                                            abs_tol, rel_tol, save_interval, _capi.TGRID[tgrid],
                                            dptr(None if sg is None else sg.reshape(-1)), szptr(nsg), ms, max_attempts,
                                            max_rec, dptr(times), dptr(mags), dptr(energy), dptr(final),
                                            szptr(nrec), szptr(nacc), szptr(nrej))
            if rc == _capi.CSL_ERR_CAPACITY and int(nrec.max()) > max_rec:
                max_rec = int(nrec.max())
                continue
            check(rc)
            break
        n = [int(v) for v in nrec]
        t2 = times.reshape(B, max_rec)
        m2 = mags.reshape(B, max_rec, 3, 3)
        e2 = None if energy is None else energy.reshape(B, max_rec)
        return {"times": [t2[b, :n[b]] for b in range(B)], "mags": [m2[b, :n[b]] for b in range(B)],
                "energy": None if e2 is None else [e2[b, :n[b]] for b in range(B)],
                "final": None if final is None else final.reshape(B, self.N, 3),
                "n_rec": n, "accepted": [int(v) for v in nacc], "rejected": [int(v) for v in nrej]}


# ---------------------------------------------------------------------------
# Reference-style simulator entry points (CPU semantics, GPU arithmetic)

def integrate_multi(lattices, inits, t_pulse, pulse_amp=None, T0=0.0, T1=1.0, dt=0.01, save_interval=1,
                    tgrid="batched", want_energy=True):
    """csl_integrate_multi: RK4 trajectories on lattices of different sizes
    in one persistent launch.  lattices: GPULattice handles on one device;
    inits[i]: [B_i, N_i, 3] states of lattice i; t_pulse, pulse_amp: [B, 2]
    per trajectory in lattice order (pulse_amp None: each lattice's own
    amplitude).  Returns dict(times, mags [n, B, 3, 3], energy [n, B])."""
    if not lattices or len(lattices) != len(inits):
        raise ValueError("need one [B_i, N_i, 3] array of initial states per lattice")
    ins = [f64(x).reshape(-1, gl.N, 3) for gl, x in zip(lattices, inits)]
    b_of = np.array([x.shape[0] for x in ins], dtype=np.uintp)
    B = int(b_of.sum())
    tp = f64(t_pulse).reshape(B, 2)
    amp = None if pulse_amp is None else f64(pulse_amp).reshape(B, 2)
    init = np.ascontiguousarray(np.concatenate([x.reshape(-1) for x in ins]))
    _, nr = lattices[0].count_records(T0, T1, dt, save_interval, tgrid)
    times = np.empty(nr)
    mags = np.empty(nr * B * 9)
    energy = np.empty(nr * B) if want_energy else None
    hs = (C.c_void_p * len(lattices))(*[gl.h for gl in lattices])
    n = C.c_size_t()
    check(lib().csl_integrate_multi(hs, len(lattices), szptr(b_of), dptr(init), dptr(tp.reshape(-1)),
                                    dptr(None if amp is None else amp.reshape(-1)), T0, T1, dt, save_interval,
                                    _capi.TGRID[tgrid], nr, dptr(times), dptr(mags), dptr(energy), C.byref(n)))
    return {"times": times, "mags": mags.reshape(nr, B, 3, 3),
            "energy": None if energy is None else energy.reshape(nr, B)}


def _drive_dirs(lattice: Lattice, field_in) -> np.ndarray:
    """Lattice::set_pulse frame rotation field_drive = R_atom * B (lattice_md.cpp:215-216)."""
    fi = np.asarray(field_in, dtype=float).reshape(lattice.N_atoms, lattice.spin_dim)
    return np.einsum("aij,aj->ai", lattice.frames, fi)


def _const_calls(t0, t1, dt):
    """Observer calls of Boost integrate_const over [t0, t1] with step dt."""
    k = 0
    while (t0 + k * dt + dt) - t1 <= 2.220446049250313e-16:
        k += 1
    return k + 1


# half-width of a pulse window in pulse widths and the free-segment step-hint
# factor of the reference's W3 chunking (pulse_chunking.h:173,178)
PULSE_WINDOW_WIDTHS = 9.0
FREE_DT_FACTOR = 20.0


def split_schedule(B: int, G: int, n_steps: int) -> np.ndarray:
    """csl_split_schedule: the persistent kernel's even work split of B
    trajectories of n_steps over G workgroups, rows [G, pieces, 5] =
    (trajectory, first step, end step, slot loaded, slot saved), -1 unused."""
    cap = G * 8
    while True:
        rows = np.full(cap * 5, -1, dtype=np.int32)
        pp = C.c_size_t()
        rc = lib().csl_split_schedule(B, G, n_steps, cap, rows.ctypes.data_as(C.POINTER(C.c_int)), C.byref(pp))
        if rc == _capi.CSL_ERR_CAPACITY:
            cap = G * pp.value
            continue
        check(rc)
        return rows[:G * pp.value * 5].reshape(G, pp.value, 5)


def seam_trajectories(B: int, G: int, n_steps: int) -> list:
    """Trajectories the work split cuts between two workgroups (their state
    crosses a seam through global memory)."""
    r = split_schedule(B, G, n_steps)
    return sorted({int(t) for t, _, _, load, _ in r.reshape(-1, 5) if t >= 0 and load >= 0})


def pulse_segments(T_start, T_end, centers, window, T_step, free_dt_factor=FREE_DT_FACTOR):
    """build_pulse_segments (pulse_chunking.h:82-161) through
    csl_pulse_segments: rows (t0, t1, dt_init) of the free and pulse-active
    segments covering [T_start, last grid point <= T_end]."""
    c = f64(np.atleast_1d(np.asarray(centers, dtype=float)))
    cap = 2 * c.size + 2
    out = np.empty(3 * cap)
    n = C.c_size_t()
    check(lib().csl_pulse_segments(float(T_start), float(T_end), dptr(c), c.size, float(window), float(T_step),
                                   float(free_dt_factor), cap, dptr(out), C.byref(n)))
    return out[:3 * n.value].reshape(-1, 3)


def _gate(times, T_start, T_end, step):
    """Pulse-drive observer gate (lattice_md.cpp:532-554)."""
    keep, last = [], T_start
    for r, t in enumerate(times):
        if t - last >= step - 1e-10 or t >= T_end - 1e-10:
            keep.append(r)
            last = t
    return np.array(keep, dtype=np.int64)


def molecular_dynamics(gl: GPULattice, spins, T_start, T_end, dt, save_interval=1, method="rk4",
                       alpha_gilbert=0.0, engine=0, out_dir: Optional[str] = None,
                       abs_tol: float = 1e-6, rel_tol: float = 1e-6):
    """molecular_dynamics_cpu (lattice_md.cpp:341-453): the observer records
    M_antiferro, M_local, M_global and E/N every save_interval observer
    calls.  method "rk4": integrate_const on the Boost grid; "dopri5" (the
    reference's default md_integrator, spin_config.h:93): integrate_adaptive
    with make_controlled(abs_tol, rel_tol), observer after every accepted
    step (lattice_md.cpp:444-453), on the Dormand-Prince kernel.  With
    out_dir the records and the spin snapshots go to out_dir/trajectory.h5 in
    the HDF5MDWriter layout (lattice_md.cpp:363-373, hdf5_io.h:83-300;
    libcsl_h5.so).  Returns dict(times, mags, energy[, states]); dopri5 runs
    add accepted / rejected step counts."""
    gl.set_damping(alpha_gilbert)
    gl.set_pulse(np.zeros(2 * gl.A * 3), 0.0, 1.0, 0.0, 0.0, 0.0)
    if method == "dopri5":
        if out_dir is not None:
            raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "trajectory.h5 snapshots are written for rk4 runs")
        r = gl.integrate_dopri5(spins, T_start, T_end, dt, control=1, abs_tol=abs_tol, rel_tol=rel_tol,
                                save_interval=save_interval)
        gl.set_spins(r["final"][0])
        return {"times": r["times"][0], "mags": r["mags"][0], "energy": r["energy"][0], "states": None,
                "accepted": r["accepted"][0], "rejected": r["rejected"][0]}
    gl.set_spins(spins)
    res = gl.integrate(T_start, T_end, dt, save_interval, method, tgrid="boost", engine=engine,
                       want_states=out_dir is not None)
    if out_dir is not None:
        from . import h5io
        lat = gl.lattice
        os.makedirs(out_dir, exist_ok=True)
        with h5io.MDWriter(os.path.join(out_dir, "trajectory.h5"), lat.lattice_size, lat.spin_dim, lat.N_atoms,
                           (lat.dim1, lat.dim2, lat.dim3), method, dt, T_start, T_end, save_interval,
                           lat.spin_length, positions=lat.positions) as w:
            w.write_steps(res["times"], res["mags"], res["states"])
    return res


def _pulse_run(gl: GPULattice, spins, fd, t1, t2, amp, width, freq, T_start, T_end, step, method="rk4",
               abs_tol=1e-8, rel_tol=1e-8, centers=(), chunking=True):
    if method == "dopri5":
        # integrate_const with the controlled stepper (lattice_md.cpp:562-575),
        # one call per pulse-window segment with chunking (W3)
        segs = None
        if chunking:
            segs = pulse_segments(T_start, T_end, centers, PULSE_WINDOW_WIDTHS * width, step)
            if len(segs) == 0:
                return np.empty(0), np.empty((0, 3, 3))
            segs = [segs]
        gl.set_pulse(fd, amp, width, freq, t1, t2)
        r = gl.integrate_dopri5(spins, T_start, T_end, step, control=2, abs_tol=abs_tol, rel_tol=rel_tol,
                                t_pulse=[[t1, t2]], pulse_amp=[[amp, amp]], want_energy=False, segments=segs)
        times, mags = r["times"][0], r["mags"][0]
    elif method == "rk4":
        # chunked RK4 takes the same steps on the same grid (up to the
        # rounding of the grid times), so the run is not split
        gl.set_pulse(fd, amp, width, freq, t1, t2)
        gl.set_spins(spins)
        res = gl.integrate(T_start, T_end, step, 1, "rk4", tgrid="boost", want_energy=False)
        times, mags = res["times"], res["mags"]
    else:
        raise _capi.CslError(_capi.CSL_ERR_UNSUPPORTED, "methods: 'rk4', 'dopri5'")
    gl.set_pulse(np.zeros_like(fd), 0.0, width, freq, 0.0, 0.0)  # reset (lattice_md.cpp:578-583)
    keep = _gate(times, T_start, T_end, step)
    return times[keep], mags[keep]


def single_pulse_drive(gl: GPULattice, spins, field_in, t_B, pulse_amp, pulse_width, pulse_freq,
                       T_start, T_end, step_size, method="rk4", abs_tol=1e-8, rel_tol=1e-8,
                       pulse_window_chunking=True):
    """Lattice::single_pulse_drive (lattice_md.cpp:493-586), method "rk4" or
    "dopri5" (controlled, tolerances as the reference's pump-probe defaults
    kDefaultPumpProbeAbsTol/RelTol = 1e-8; with pulse_window_chunking one
    integrate_const call per segment of pulse_segments around t_B).
    Returns (times [n], mags [n, 3(M_antiferro, M_local, M_global), 3])."""
    lat = gl.lattice
    fd = np.concatenate([_drive_dirs(lat, field_in).reshape(-1), np.zeros(lat.N_atoms * 3)])
    return _pulse_run(gl, spins, fd, t_B, 0.0, pulse_amp, pulse_width, pulse_freq, T_start, T_end, step_size,
                      method, abs_tol, rel_tol, (t_B,), pulse_window_chunking)


def double_pulse_drive(gl: GPULattice, spins, field_in_1, t_B_1, field_in_2, t_B_2, pulse_amp,
                       pulse_width, pulse_freq, T_start, T_end, step_size, method="rk4", abs_tol=1e-8,
                       rel_tol=1e-8, pulse_window_chunking=True):
    """Lattice::double_pulse_drive (lattice_md.cpp:589-680); the segments
    surround both pulse centres."""
    lat = gl.lattice
    fd = np.concatenate([_drive_dirs(lat, field_in_1).reshape(-1), _drive_dirs(lat, field_in_2).reshape(-1)])
    return _pulse_run(gl, spins, fd, t_B_1, t_B_2, pulse_amp, pulse_width, pulse_freq, T_start, T_end,
                      step_size, method, abs_tol, rel_tol, (t_B_1, t_B_2), pulse_window_chunking)


def pump_probe_spectroscopy(gl: GPULattice, ground, field_in, pulse_amp, pulse_width, pulse_freq,
                            tau_start, tau_end, tau_step, T_start, T_end, T_step,
                            dir_name: Optional[str] = None, method: str = "rk4",
                            Temp_start: float = 0.0, Temp_end: float = 0.0, n_anneal: int = 0,
                            T_zero_quench: bool = False, quench_sweeps: int = 0,
                            reuse_m0_for_m1: bool = False, stationarity_tol: float = 1e-6,
                            abs_tol: float = 1e-8, rel_tol: float = 1e-8, pulse_window_chunking: bool = True):
    """Lattice::pump_probe_spectroscopy (lattice_md.cpp:683-926): M0, M1(tau)
    and M01(tau) for every delay, all integrated in one batched launch.
    With dir_name the result goes to dir_name/pump_probe_spectroscopy.h5 in
    the HDF5PumpProbeWriter layout (hdf5_io.h:729-1000; libcsl_h5.so), with
    the ground state's E/N and |M_local| and the annealing parameters (which
    only describe how the ground state was made) as metadata.
    reuse_m0_for_m1 / stationarity_tol are the reference's W1 option
    (lattice_md.cpp:729-752): M1(tau) synthesised from M0 when max|dS/dt| of
    the undriven ground state is within the tolerance; pulse_window_chunking
    is the reference's W3 option for "dopri5" (lattice_md.cpp:562-571).
    Returns dict(times, M0 [n,3,3], M1 [n_tau,n,3,3], M01, tau, w1, max_dsdt)."""
    res = gl.pump_probe(ground, field_in, pulse_amp, pulse_width, pulse_freq, tau_start, tau_end,
                        tau_step, T_start, T_end, T_step, reuse_m0_for_m1, stationarity_tol, method, abs_tol,
                        rel_tol, pulse_window_chunking)
    if dir_name is not None:
        from . import h5io
        lat = gl.lattice
        os.makedirs(dir_name, exist_ok=True)
        gl.set_spins(ground)
        E_ground = gl.energy() / lat.lattice_size          # energy_density (lattice_md.cpp:716)
        M_ground = np.linalg.norm(gl.magnetization()[0])   # |magnetization_local| (:717)
        h5io.write_pump_probe(
            os.path.join(dir_name, "pump_probe_spectroscopy.h5"), res, lattice_size=lat.lattice_size,
            n_atoms=lat.N_atoms, dims=(lat.dim1, lat.dim2, lat.dim3), spin_dim=lat.spin_dim,
            spin_length=lat.spin_length, pulse_amp=pulse_amp, pulse_width=pulse_width, pulse_freq=pulse_freq,
            T_start=T_start, T_end=T_end, T_step=T_step, method=method, tau_start=tau_start, tau_end=tau_end,
            tau_step=tau_step, ground_state_energy=E_ground, ground_state_magnetization=M_ground,
            Temp_start=Temp_start, Temp_end=Temp_end, n_anneal=n_anneal, T_zero_quench=T_zero_quench,
            quench_sweeps=quench_sweeps, pulse_field_direction=np.asarray(field_in, dtype=float),
            positions=lat.positions)
    return res


def nonlinear_signal(res: dict) -> np.ndarray:
    """M_NL(tau, t) = M01 - M0 - M1 (util/readers_new/reader_pyrochlore.py:760-830)."""
    return res["M01"] - res["M0"][None] - res["M1"]


def nonlinear_spectrum(res: dict, mag_index: int = 2, omega_range=None):
    """2-D spectrum |FFT2(M_NL)| of a pump-probe result on the GPU
    (csl_nonlinear_spectrum; reader_pyrochlore.py:760-830): M_NL over the
    last half of the times, zero-padded to powers of two, fftshifted.
    mag_index selects M_antiferro (0), M_local (1) or M_global (2).  Returns
    dict(M_NL_FF [P_tau, P_t, D], freq_tau, freq_T), frequencies from
    numpy.fft.fftfreq; with omega_range the window |f| <= omega_range."""
    M0 = f64(res["M0"][:, mag_index])
    M1 = f64(res["M1"][:, :, mag_index])
    M01 = f64(res["M01"][:, :, mag_index])
    n_tau, n_t, D = M1.shape
    dom = n_t // 2
    Pt_ = 1 << max(0, (n_tau - 1).bit_length())
    Pd_ = 1 << max(0, (dom - 1).bit_length())
    out = np.empty(Pt_ * Pd_ * D)
    pt, pd = C.c_size_t(), C.c_size_t()
    check(lib().csl_nonlinear_spectrum(n_tau, n_t, D, dptr(M0), dptr(M1), dptr(M01), out.size, dptr(out),
                                       C.byref(pt), C.byref(pd)))
    ff = out.reshape(pt.value, pd.value, D)
    times, taus = np.asarray(res["times"]), np.asarray(res["tau"])
    dT = times[1] - times[0] if len(times) > 1 else 1.0
    dtau = taus[1] - taus[0] if len(taus) > 1 else 1.0
    fq_tau = np.fft.fftshift(np.fft.fftfreq(pt.value, d=dtau))
    fq_T = np.fft.fftshift(np.fft.fftfreq(pd.value, d=dT))
    if omega_range is not None:
        it = (fq_tau >= -omega_range) & (fq_tau <= omega_range)
        iT = (fq_T >= -omega_range) & (fq_T <= omega_range)
        ff, fq_tau, fq_T = ff[np.ix_(it, iT)], fq_tau[it], fq_T[iT]
    return {"M_NL_FF": ff, "freq_tau": fq_tau, "freq_T": fq_T}
