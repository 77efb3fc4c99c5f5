"""ctypes wrapper of oracle/liboracle.so — TEST INFRASTRUCTURE ONLY.

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
checker / CPU baseline.  The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_psz = C.POINTER(C.c_size_t)
_sz, _d, _i, _vp = C.c_size_t, C.c_double, C.c_int, C.c_void_p

_SIG = {
    "orc_lattice_new": (_vp, [_i, _i, _i, _i, _i, _i, _pi, _pi, _pi, _pd, _pd, _pd, _pd, _pd]),
    "orc_lattice_from_csr": (_vp, [_i, _i, _i, _i, _i, _psz, _psz, _pd, _pd, _pd, _pd, _pd]),
    "orc_lattice_free": (None, [_vp]),
    "orc_lattice_size": (_sz, [_vp]),
    "orc_total_bonds": (_sz, [_vp]),
    "orc_flat_tables": (None, [_vp, _psz, _psz, _pd]),
    "orc_set_alpha": (None, [_vp, _d]),
    "orc_set_pulse": (None, [_vp, _pd, _d, _pd, _d, _d, _d, _d]),
    "orc_clear_pulse": (None, [_vp]),
    "orc_local_field": (None, [_vp, _pd, _sz, _pd]),
    "orc_rhs": (None, [_vp, _pd, _pd, _d]),
    "orc_energy": (_d, [_vp, _pd]),
    "orc_mags": (None, [_vp, _pd, _pd]),
    "orc_rk4_step": (None, [_vp, _pd, _d, _d]),
    "orc_boost_num_steps": (_sz, [_d, _d, _d]),
    "orc_integrate_md": (_sz, [_vp, _pd, _d, _d, _d, _sz, _sz, _pd, _pd, _pd]),
    "orc_single_pulse_drive": (_sz, [_vp, _pd, _pd, _d, _d, _d, _d, _d, _d, _d, _i, _sz, _pd, _pd]),
    "orc_double_pulse_drive": (_sz, [_vp, _pd, _pd, _d, _pd, _d, _d, _d, _d, _d, _d, _d, _i, _sz, _pd, _pd]),
    "orc_max_dsdt_no_drive": (_d, [_vp, _pd]),
    "orc_integrate_dopri5": (C.c_long, [_vp, _pd, _d, _d, _d, _i, _d, _d, _sz, _pd, _sz, _sz, _pd, _pd, _pd,
                                        _psz, _psz]),
    "orc_dopri5_scalar": (C.c_long, [_d, _d, _d, _d, _d, _d, _i, _d, _d, _pd, _psz, _psz, _pd]),
    "orc_pump_probe": (_i, [_vp, _pd, _pd, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i, _d, _i, _i, _sz, _sz,
                            _psz, _pd, _pd, _pd, _pd, _pd, _pi]),
    "orc_build_pulse_segments": (_sz, [_d, _d, _pd, _sz, _d, _d, _d, _sz, _pd]),
    "orc_max_threads": (_i, []),
    "orc_set_threads": (None, [_i]),
    "orc_greedy_quench": (_sz, [_vp, _pd, _psz, _d, _d, _sz]),
    "orc_tangent_md": (_sz, [_vp, _pd, _pd, _pd, _pd, _d, _d, _d, _sz, _pd, _pd]),
    "orc_mixed_new": (_vp, [_i, _i, _i, _i, _pd, _pd, _pd, _pd, _i, _pi, _pi, _pi, _pd,
                            _i, _pd, _pd, _pd, _i, _pi, _pi, _pi, _pd, _i, _pi, _pi, _pi, _pd]),
    "orc_mixed_free": (None, [_vp]),
    "orc_mixed_sizes": (None, [_vp, _psz, _psz, _psz, _psz, _psz]),
    "orc_mixed_tables": (None, [_vp, _psz, _psz, _pd, _psz, _psz, _pd, _psz, _psz, _pd, _psz, _psz, _pd]),
    "orc_mixed_set_damping": (None, [_vp, _d, _pd, _pd]),
    "orc_mixed_set_pulse": (None, [_vp, _i, _pd, _d, _pd, _d, _d, _d, _d]),
    "orc_mixed_local_field": (None, [_vp, _pd, _i, _sz, _pd]),
    "orc_mixed_rhs": (None, [_vp, _pd, _pd, _d]),
    "orc_mixed_energy": (_d, [_vp, _pd]),
    "orc_mixed_record": (None, [_vp, _pd, _pd]),
    "orc_mixed_integrate_md": (_sz, [_vp, _pd, _d, _d, _d, _sz, _pd, _pd]),
    "orc_integrate_gpu": (_i, [_vp, _pd, _d, _d, _d, _sz, C.c_char_p, _sz, _psz, _pd, _pd, _pd, _pd]),
    "orc_mixed_integrate_gpu": (_i, [_vp, _pd, _d, _d, _d, _sz, C.c_char_p, _sz, _psz, _pd, _pd, _pd]),
    "orc_mixed_integrate_dopri5": (C.c_long, [_vp, _pd, _d, _d, _d, _i, _d, _d, _sz, _pd, _sz, _sz, _pd, _pd,
                                              _psz, _psz]),
}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        for k, (r, a) in _SIG.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_pd)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class OracleLattice:
    """Oracle lattice built from a model.UnitCell-like object (bonds in
    insertion order) and dimensions."""

    @classmethod
    def from_csr(cls, D, n_atoms, dims, offset, partner, J, field_site, onsite_site, frames, afm):
        self = cls.__new__(cls)
        self.D, self.A = D, n_atoms
        off = np.ascontiguousarray(offset, dtype=np.uintp)
        par = np.ascontiguousarray(partner, dtype=np.uintp)
        Jf, fs, os_, fr, af = (_f(J).reshape(-1), _f(field_site).reshape(-1), _f(onsite_site).reshape(-1),
                               _f(frames).reshape(-1), _f(afm).reshape(-1))
        self._keep = (off, par, Jf, fs, os_, fr, af)
        self.h = lib().orc_lattice_from_csr(D, n_atoms, dims[0], dims[1], dims[2], off.ctypes.data_as(_psz),
                                            par.ctypes.data_as(_psz), _p(Jf), _p(fs), _p(os_), _p(fr), _p(af))
        self.N = lib().orc_lattice_size(self.h)
        return self

    def __init__(self, uc, d1, d2, d3):
        self.D = uc.N
        self.A = uc.N_atoms
        nb = len(uc.bilinear)
        src = np.array([b.source for b in uc.bilinear], dtype=np.int32)
        par = np.array([b.partner for b in uc.bilinear], dtype=np.int32)
        off = np.array([b.offset for b in uc.bilinear], dtype=np.int32).reshape(-1)
        J = _f(np.array([b.J for b in uc.bilinear]).reshape(-1)) if nb else np.zeros(1)
        self._keep = (src, par, off, J)
        ip = lambda a: a.ctypes.data_as(_pi)
        self.h = lib().orc_lattice_new(self.D, self.A, d1, d2, d3, nb, ip(src) if nb else None,
                                       ip(par) if nb else None, ip(off) if nb else None, _p(J),
                                       _p(_f(uc.field)), _p(_f(uc.onsite_interaction)),
                                       _p(_f(uc.sublattice_frames)), _p(_f(uc.afm_sublattice_signs)))
        if not self.h:
            raise RuntimeError("oracle rejected the lattice (bond offset beyond the lattice size)")
        self.N = lib().orc_lattice_size(self.h)

    def __del__(self):
        try:
            lib().orc_lattice_free(self.h)
        except Exception:
            pass

    def tables(self):
        nb = lib().orc_total_bonds(self.h)
        off = np.zeros(self.N + 1, dtype=np.uintp)
        par = np.zeros(max(nb, 1), dtype=np.uintp)
        J = np.zeros(max(nb, 1) * self.D * self.D)
        lib().orc_flat_tables(self.h, off.ctypes.data_as(_psz), par.ctypes.data_as(_psz), _p(J))
        return off, par[:nb], J[: nb * self.D * self.D].reshape(nb, self.D, self.D)

    def set_alpha(self, a):
        lib().orc_set_alpha(self.h, a)

    def set_pulse(self, f1, t1, f2, t2, amp, width, freq):
        lib().orc_set_pulse(self.h, _p(_f(f1)), t1, _p(None if f2 is None else _f(f2)), t2, amp, width, freq)

    def clear_pulse(self):
        lib().orc_clear_pulse(self.h)

    def local_field(self, state, site):
        H = np.zeros(self.D)
        lib().orc_local_field(self.h, _p(_f(state)), site, _p(H))
        return H

    def rhs(self, state, t=0.0):
        s = _f(state)
        out = np.zeros_like(s)
        lib().orc_rhs(self.h, _p(s), _p(out), t)
        return out

    def energy(self, state):
        return lib().orc_energy(self.h, _p(_f(state)))

    def mags(self, state):
        out = np.zeros(3 * self.D)
        lib().orc_mags(self.h, _p(_f(state)), _p(out))
        return out.reshape(3, self.D)

    def rk4_step(self, state, t, dt):
        s = _f(state).copy()
        lib().orc_rk4_step(self.h, _p(s), t, dt)
        return s

    def integrate_md(self, state, T0, T1, dt, save_interval=1):
        n = lib().orc_boost_num_steps(T0, T1, dt)
        mx = n // save_interval + 2
        s = _f(state).copy()
        t = np.zeros(mx)
        m = np.zeros(mx * 3 * self.D)
        e = np.zeros(mx)
        nr = lib().orc_integrate_md(self.h, _p(s), T0, T1, dt, save_interval, mx, _p(t), _p(m), _p(e))
        return {"times": t[:nr], "mags": m.reshape(mx, 3, self.D)[:nr], "energy": e[:nr], "final": s}

    def integrate_dopri5(self, state, T0, T1, dt, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=1,
                         max_rec=None, segments=None):
        """Dormand-Prince 5(4): control 0 fixed grid steps, 1 integrate_adaptive,
        2 integrate_const with the controlled stepper (one call per row
        (t0, t1[, dt_init]) of segments when given)."""
        sg = None if segments is None else _f(np.asarray(segments, dtype=float).reshape(len(segments), -1)[:, :2])
        s = _f(state).copy()
        mx = max_rec or 1000
        while True:
            t = np.zeros(mx)
            m = np.zeros(mx * 3 * self.D)
            e = np.zeros(mx)
            na, nr_ = C.c_size_t(), C.c_size_t()
            s = _f(state).copy()
            nr = lib().orc_integrate_dopri5(self.h, _p(s), T0, T1, dt, control, abs_tol, rel_tol, save_interval,
                                            None if sg is None else _p(sg), 0 if sg is None else len(sg),
                                            mx, _p(t), _p(m), _p(e), C.byref(na), C.byref(nr_))
            if nr < 0:
                raise RuntimeError("step size control failed (500 rejections in a row)")
            if nr <= mx:
                break
            mx = nr
        return {"times": t[:nr], "mags": m.reshape(mx, 3, self.D)[:nr], "energy": e[:nr], "final": s,
                "accepted": na.value, "rejected": nr_.value}

    def integrate_gpu(self, state, T0, T1, dt, method, save_interval=1, want_states=False):
        """The reference GPU integrate_gpu loop (t += dt while t < T1) with any
        gpu::ode::step method; records times, mags, E/N and optionally the
        full states.  'known' is False when the name fell back to ssprk53."""
        n_steps = gpu_num_steps(T0, T1, dt)
        mx = n_steps // save_interval + 2
        s = _f(state).copy()
        t = np.zeros(mx)
        m = np.zeros(mx * 3 * self.D)
        e = np.zeros(mx)
        st = np.zeros(mx * s.size) if want_states else None
        nr = C.c_size_t()
        known = lib().orc_integrate_gpu(self.h, _p(s), T0, T1, dt, save_interval, method.encode(), mx, C.byref(nr),
                                        _p(t), _p(st), _p(m), _p(e))
        k = nr.value
        out = {"times": t[:k], "mags": m.reshape(mx, 3, self.D)[:k], "energy": e[:k], "final": s,
               "known": bool(known)}
        if want_states:
            out["states"] = st.reshape(mx, -1)[:k]
        return out

    def single_pulse_drive(self, init, field_in, t_B, amp, width, freq, T0, T1, step, chunking=False):
        n = lib().orc_boost_num_steps(T0, T1, step) + 4
        t = np.zeros(n)
        m = np.zeros(n * 3 * self.D)
        nr = lib().orc_single_pulse_drive(self.h, _p(_f(init)), _p(_f(field_in)), t_B, amp, width, freq, T0, T1,
                                          step, int(chunking), n, _p(t), _p(m))
        return t[:nr], m.reshape(n, 3, self.D)[:nr]

    def double_pulse_drive(self, init, f1, t1, f2, t2, amp, width, freq, T0, T1, step, chunking=False):
        n = lib().orc_boost_num_steps(T0, T1, step) + 4
        t = np.zeros(n)
        m = np.zeros(n * 3 * self.D)
        nr = lib().orc_double_pulse_drive(self.h, _p(_f(init)), _p(_f(f1)), t1, _p(_f(f2)), t2, amp, width, freq,
                                          T0, T1, step, int(chunking), n, _p(t), _p(m))
        return t[:nr], m.reshape(n, 3, self.D)[:nr]

    def greedy_quench(self, state, order, spin_length=1.0, rel_tol=1e-12, max_sweeps=10000):
        """Lattice::greedy_quench with every sweep visiting the sites in
        `order`; returns (relaxed state, sweeps done)."""
        s = _f(state).copy()
        o = np.ascontiguousarray(order, dtype=np.uintp)
        n = lib().orc_greedy_quench(self.h, _p(s), o.ctypes.data_as(_psz), spin_length, rel_tol, max_sweeps)
        return s, int(n)

    def max_dsdt_no_drive(self, state):
        return lib().orc_max_dsdt_no_drive(self.h, _p(_f(state)))

    def pump_probe(self, ground, field_in, amp, width, freq, tau_start, tau_end, tau_step, T0, T1, T_step,
                   reuse_m0=False, stat_tol=1e-6, n_outer=1, chunking=False):
        mr = lib().orc_boost_num_steps(T0, T1, T_step) + 4
        mt = int(abs((tau_end - tau_start) / tau_step)) + 2
        D3 = 3 * self.D
        times = np.zeros(mr)
        M0 = np.zeros(mr * D3)
        M1 = np.zeros(mt * mr * D3)
        M01 = np.zeros(mt * mr * D3)
        taus = np.zeros(mt)
        nrec = C.c_size_t()
        w1 = C.c_int()
        nt = lib().orc_pump_probe(self.h, _p(_f(ground)), _p(_f(field_in)), amp, width, freq, tau_start, tau_end,
                                  tau_step, T0, T1, T_step, int(reuse_m0), stat_tol, n_outer, int(chunking), mr,
                                  mt, C.byref(nrec), _p(times), _p(M0), _p(M1), _p(M01), _p(taus),
                                  C.byref(w1))
        if nt < 0:
            raise RuntimeError("tau grid larger than buffer")
        r = nrec.value
        return {"times": times[:r], "M0": M0.reshape(mr, 3, self.D)[:r],
                "M1": M1.reshape(mt, mr, 3, self.D)[:nt, :r], "M01": M01.reshape(mt, mr, 3, self.D)[:nt, :r],
                "tau": taus[:nt], "used_w1": bool(w1.value)}

    def tangent_md(self, state, dstate, dJ, dB, T0, T1, dt):
        n = lib().orc_boost_num_steps(T0, T1, dt) + 2
        s = _f(state).copy()
        ds = _f(dstate).copy()
        m = np.zeros(n * 9)
        dm = np.zeros(n * 9)
        nr = lib().orc_tangent_md(self.h, _p(s), _p(ds), _p(None if dJ is None else _f(dJ)),
                                  _p(None if dB is None else _f(dB)), T0, T1, dt, n, _p(m), _p(dm))
        return m.reshape(n, 3, 3)[:nr], dm.reshape(n, 3, 3)[:nr]


def _bonds(cell):
    nb = len(cell.bilinear)
    src = np.array([b.source for b in cell.bilinear], dtype=np.int32)
    par = np.array([b.partner for b in cell.bilinear], dtype=np.int32)
    off = np.array([b.offset for b in cell.bilinear], dtype=np.int32).reshape(-1)
    J = _f(np.array([b.J for b in cell.bilinear]).reshape(-1)) if nb else np.zeros(1)
    return nb, src, par, off, J


class MixedOracle:
    """Oracle mixed SU(2)+SU(3) lattice built from a MixedUnitCell-like
    object (SU2_cell, SU3_cell, bilinear_SU2_SU3) and dimensions."""

    REC = 34

    def __init__(self, muc, d1, d2, d3):
        c2, c3 = muc.SU2_cell, muc.SU3_cell
        if c2.N != 3 or c3.N != 8:
            raise ValueError("mixed oracle handles spin_dim 3 + 8")
        b2, b3 = _bonds(c2), _bonds(c3)
        mb = muc.bilinear_SU2_SU3
        nbm = len(mb)
        srcm = np.array([b.source for b in mb], dtype=np.int32)
        parm = np.array([b.partner for b in mb], dtype=np.int32)
        offm = np.array([b.offset for b in mb], dtype=np.int32).reshape(-1)
        Jm = _f(np.array([b.J for b in mb]).reshape(-1)) if nbm else np.zeros(1)
        arrs = [_f(c2.field), _f(c2.onsite_interaction), _f(c2.sublattice_frames), _f(c2.afm_sublattice_signs),
                _f(c3.field), _f(c3.onsite_interaction), _f(c3.sublattice_frames)]
        self._keep = (b2, b3, srcm, parm, offm, Jm, arrs)
        ip = lambda a, n: a.ctypes.data_as(_pi) if n else None
        self.h = lib().orc_mixed_new(d1, d2, d3, c2.N_atoms, _p(arrs[0]), _p(arrs[1]), _p(arrs[2]), _p(arrs[3]),
                                     b2[0], ip(b2[1], b2[0]), ip(b2[2], b2[0]), ip(b2[3], b2[0]), _p(b2[4]),
                                     c3.N_atoms, _p(arrs[4]), _p(arrs[5]), _p(arrs[6]),
                                     b3[0], ip(b3[1], b3[0]), ip(b3[2], b3[0]), ip(b3[3], b3[0]), _p(b3[4]),
                                     nbm, ip(srcm, nbm), ip(parm, nbm), ip(offm, nbm), _p(Jm))
        if not self.h:
            raise RuntimeError("oracle rejected the mixed lattice (bond offset beyond the lattice size)")
        n2, n3, a, b, c = (C.c_size_t() for _ in range(5))
        lib().orc_mixed_sizes(self.h, C.byref(n2), C.byref(n3), C.byref(a), C.byref(b), C.byref(c))
        self.N2, self.N3 = n2.value, n3.value
        self.nnz = (a.value, b.value, c.value)

    def __del__(self):
        try:
            lib().orc_mixed_free(self.h)
        except Exception:
            pass

    @property
    def n_state(self):
        return self.N2 * 3 + self.N3 * 8

    def tables(self):
        n22, n33, n23 = self.nnz
        o22, o33 = np.zeros(self.N2 + 1, np.uintp), np.zeros(self.N3 + 1, np.uintp)
        o23, o32 = np.zeros(self.N2 + 1, np.uintp), np.zeros(self.N3 + 1, np.uintp)
        p22, p33 = np.zeros(max(n22, 1), np.uintp), np.zeros(max(n33, 1), np.uintp)
        p23, p32 = np.zeros(max(n23, 1), np.uintp), np.zeros(max(n23, 1), np.uintp)
        J22, J33 = np.zeros(max(n22, 1) * 9), np.zeros(max(n33, 1) * 64)
        J23, J32 = np.zeros(max(n23, 1) * 24), np.zeros(max(n23, 1) * 24)
        z = lambda a: a.ctypes.data_as(_psz)
        lib().orc_mixed_tables(self.h, z(o22), z(p22), _p(J22), z(o33), z(p33), _p(J33), z(o23), z(p23), _p(J23),
                               z(o32), z(p32), _p(J32))
        return {"22": (o22, p22[:n22], J22[:n22 * 9].reshape(-1, 3, 3)),
                "33": (o33, p33[:n33], J33[:n33 * 64].reshape(-1, 8, 8)),
                "23": (o23, p23[:n23], J23[:n23 * 24].reshape(-1, 3, 8)),
                "32": (o32, p32[:n23], J32[:n23 * 24].reshape(-1, 8, 3))}

    def set_damping(self, alpha=0.0, gamma=None, eq=None):
        lib().orc_mixed_set_damping(self.h, alpha, _p(None if gamma is None else _f(gamma)),
                                    _p(None if eq is None else _f(eq)))

    def set_pulse(self, species, f1, t1, f2, t2, amp, width, freq):
        lib().orc_mixed_set_pulse(self.h, species, _p(_f(f1)), t1, _p(_f(f2)), t2, amp, width, freq)

    def local_field(self, x, species, site):
        H = np.zeros(8)
        lib().orc_mixed_local_field(self.h, _p(_f(x)), species, site, _p(H))
        return H[:3] if species == 2 else H

    def rhs(self, x, t=0.0):
        out = np.zeros(self.n_state)
        lib().orc_mixed_rhs(self.h, _p(_f(x)), _p(out), t)
        return out

    def energy(self, x):
        return lib().orc_mixed_energy(self.h, _p(_f(x)))

    def record(self, x):
        out = np.zeros(self.REC)
        lib().orc_mixed_record(self.h, _p(_f(x)), _p(out))
        return out

    def integrate_md(self, state, T0, T1, dt):
        n = lib().orc_boost_num_steps(T0, T1, dt) + 2
        s = _f(state).copy()
        t = np.zeros(n)
        r = np.zeros(n * self.REC)
        nr = lib().orc_mixed_integrate_md(self.h, _p(s), T0, T1, dt, n, _p(t), _p(r))
        return {"times": t[:nr], "records": r.reshape(n, self.REC)[:nr], "final": s}

    def integrate_gpu(self, state, T0, T1, dt, method, save_interval=1, want_states=False):
        """integrate_mixed_gpu with any gpu::ode::step method (records of 34
        values, optional full states)."""
        n_steps = gpu_num_steps(T0, T1, dt)
        mx = n_steps // save_interval + 2
        s = _f(state).copy()
        t = np.zeros(mx)
        r = np.zeros(mx * self.REC)
        st = np.zeros(mx * s.size) if want_states else None
        nr = C.c_size_t()
        known = lib().orc_mixed_integrate_gpu(self.h, _p(s), T0, T1, dt, save_interval, method.encode(), mx,
                                              C.byref(nr), _p(t), _p(st), _p(r))
        k = nr.value
        out = {"times": t[:k], "records": r.reshape(mx, self.REC)[:k], "final": s, "known": bool(known)}
        if want_states:
            out["states"] = st.reshape(mx, -1)[:k]
        return out

    def integrate_dopri5(self, state, T0, T1, dt, control=1, abs_tol=1e-6, rel_tol=1e-6, save_interval=1,
                         segments=None):
        """Dormand-Prince 5(4) over the flat state with the controls of
        OracleLattice.integrate_dopri5; records of 34 values."""
        sg = None if segments is None else _f(np.asarray(segments, dtype=float).reshape(len(segments), -1)[:, :2])
        mx = 1000
        while True:
            s = _f(state).copy()
            t = np.zeros(mx)
            r = np.zeros(mx * self.REC)
            na, nr_ = C.c_size_t(), C.c_size_t()
            nr = lib().orc_mixed_integrate_dopri5(self.h, _p(s), T0, T1, dt, control, abs_tol, rel_tol,
                                                  save_interval, None if sg is None else _p(sg),
                                                  0 if sg is None else len(sg), mx, _p(t), _p(r), C.byref(na),
                                                  C.byref(nr_))
            if nr < 0:
                raise RuntimeError("step size control failed (500 rejections in a row)")
            if nr <= mx:
                break
            mx = nr
        return {"times": t[:nr], "records": r.reshape(mx, self.REC)[:nr], "final": s,
                "accepted": na.value, "rejected": nr_.value}


def boost_num_steps(T0, T1, dt):
    return lib().orc_boost_num_steps(T0, T1, dt)


def gpu_num_steps(T0, T1, dt):
    """Steps of the reference GPU loop t += dt while t < T1 (integrator.cuh:477-484)."""
    t, n = T0, 0
    while t < T1:
        t += dt
        n += 1
    return n


def pulse_segments(T0, T1, centers, window, T_step, free_dt_factor):
    c = _f(centers)
    out = np.zeros(3 * 64)
    n = lib().orc_build_pulse_segments(T0, T1, _p(c), c.size, window, T_step, free_dt_factor, 64, _p(out))
    return out.reshape(64, 3)[:n]


def max_threads():
    return lib().orc_max_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def dopri5_scalar(lam, omega, y0, T0, T1, dt, control=1, abs_tol=1e-6, rel_tol=1e-6):
    """The oracle's Dormand-Prince stepper on y' = lam*y + cos(omega*t).
    Returns (y(T1), accepted, rejected, (err estimate, y1) of one step of size dt)."""
    y = C.c_double()
    na, nr = C.c_size_t(), C.c_size_t()
    e = np.zeros(2)
    rc = lib().orc_dopri5_scalar(lam, omega, y0, T0, T1, dt, control, abs_tol, rel_tol, C.byref(y), C.byref(na),
                                 C.byref(nr), _p(e))
    if rc < 0:
        raise RuntimeError("step size control failed")
    return y.value, na.value, nr.value, (e[0], e[1])
