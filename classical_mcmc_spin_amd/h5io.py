"""HDF5 output of the MD and pump-probe drivers (ctypes binding of libcsl_h5.so).

The files have the reference's layout (include/csl_h5.h):

  pump_probe_spectroscopy.h5   include/classical_spin/io/hdf5_io.h:729-1000 (HDF5PumpProbeWriter)
  trajectory.h5                include/classical_spin/io/hdf5_io.h:83-300 (HDF5MDWriter)

The writer is native (csrc/csl_h5.cpp, HDF5 C API).  A missing library raises
:class:`H5Error`; nothing falls back to another writer.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsl_h5.so")

_sz = C.c_size_t
_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_psz = C.POINTER(C.c_size_t)
_pi = C.POINTER(C.c_int)
_cs = C.c_char_p
_vp = C.c_void_p


class PumpProbeMeta(C.Structure):
    """csl_h5_pump_probe_meta: HDF5PumpProbeWriter arguments (hdf5_io.h:729-750)."""
    _fields_ = [("lattice_size", _sz), ("spin_dim", _sz), ("n_atoms", _sz),
                ("dim1", _sz), ("dim2", _sz), ("dim3", _sz), ("spin_length", _d),
                ("pulse_amp", _d), ("pulse_width", _d), ("pulse_freq", _d),
                ("T_start", _d), ("T_end", _d), ("T_step", _d), ("integration_method", _cs),
                ("tau_start", _d), ("tau_end", _d), ("tau_step", _d),
                ("ground_state_energy", _d), ("ground_state_magnetization", _d),
                ("Temp_start", _d), ("Temp_end", _d), ("n_anneal", _sz),
                ("T_zero_quench", _i), ("quench_sweeps", _sz),
                ("pulse_field_direction", _pd), ("positions", _pd)]


SIGNATURES = {
    "csl_h5_last_error": (_cs, []),
    "csl_h5_pump_probe_create": (_i, [_cs, C.POINTER(PumpProbeMeta), _psz, C.POINTER(_vp)]),
    "csl_h5_pump_probe_write_reference": (_i, [_vp, _sz, _pd, _pd]),
    "csl_h5_pump_probe_write_tau": (_i, [_vp, _sz, _d, _sz, _pd, _pd]),
    "csl_h5_md_create": (_i, [_cs, _sz, _sz, _sz, _sz, _sz, _sz, _cs, _d, _d, _d, _sz, _d, _pd, _i,
                              C.POINTER(_vp)]),
    "csl_h5_md_write_steps": (_i, [_vp, _sz, _pd, _pd, _pd]),
    "csl_h5_md_steps": (_i, [_vp, _psz]),
    "csl_h5_close": (_i, [_vp]),
    "csl_h5_dataset_shape": (_i, [_cs, _cs, _pi, _psz]),
    "csl_h5_read_doubles": (_i, [_cs, _cs, _pd, _sz]),
    "csl_h5_read_attr_double": (_i, [_cs, _cs, _cs, _pd]),
    "csl_h5_read_attr_string": (_i, [_cs, _cs, _cs, C.c_char_p, _sz]),
}


class H5Error(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"csl_h5 error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise H5Error(-3, f"{LIB_PATH} is missing: build it with `python -m classical_mcmc_spin_amd.build`")
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:
        raise H5Error(-3, f"cannot load {LIB_PATH}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        msg = lib().csl_h5_last_error()
        raise H5Error(rc, msg.decode() if msg else "")


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(_pd)


def _b(s: str) -> bytes:
    return os.fsencode(s)


# ---------------------------------------------------------------------------
# pump_probe_spectroscopy.h5

def write_pump_probe(path: str, res: dict, *, lattice_size: int, n_atoms: int, dims, spin_dim: int = 3,
                     spin_length: float = 1.0, pulse_amp: float, pulse_width: float, pulse_freq: float,
                     T_start: float, T_end: float, T_step: float, method: str = "rk4",
                     tau_start: float, tau_end: float, tau_step: float,
                     ground_state_energy: float, ground_state_magnetization: float,
                     Temp_start: float = 0.0, Temp_end: float = 0.0, n_anneal: int = 0,
                     T_zero_quench: bool = False, quench_sweeps: int = 0,
                     pulse_field_direction=None, positions=None) -> str:
    """Writes a pump_probe_spectroscopy result (dict with times [n], M0 [n,3,D],
    M1/M01 [n_tau,n,3,D]) in the layout of HDF5PumpProbeWriter
    (hdf5_io.h:729-1000), in the order lattice_md.cpp:771-787 and :893-926
    write it: metadata and tau_values, the reference trajectory, one group
    per delay."""
    times, M0 = _f64(res["times"]), _f64(res["M0"])
    M1, M01 = _f64(res["M1"]), _f64(res["M01"])
    n = times.shape[0]
    d1, d2, d3 = (int(x) for x in dims)
    pfd = None if pulse_field_direction is None else _f64(pulse_field_direction).reshape(n_atoms, spin_dim)
    pos = None if positions is None else _f64(positions).reshape(lattice_size, 3)
    meta = PumpProbeMeta(lattice_size, spin_dim, n_atoms, d1, d2, d3, spin_length, pulse_amp, pulse_width,
                         pulse_freq, T_start, T_end, T_step, method.encode(), tau_start, tau_end, tau_step,
                         ground_state_energy, ground_state_magnetization, Temp_start, Temp_end, n_anneal,
                         int(bool(T_zero_quench)), quench_sweeps, _p(pfd), _p(pos))
    if M0.shape != (n, 3, spin_dim) or M1.ndim != 4 or M1.shape[1:] != (n, 3, spin_dim) or M01.shape != M1.shape:
        raise ValueError(f"pump-probe arrays do not match times [{n}] and spin_dim {spin_dim}")
    h, n_tau = _vp(), C.c_size_t()
    _check(lib().csl_h5_pump_probe_create(_b(path), C.byref(meta), C.byref(n_tau), C.byref(h)))
    try:
        if M1.shape[0] != n_tau.value:
            raise ValueError(f"{M1.shape[0]} delay records for {n_tau.value} delays of the scan")
        taus = read_tau_values(tau_start, tau_end, tau_step)
        _check(lib().csl_h5_pump_probe_write_reference(h, n, _p(times), _p(M0)))
        for k in range(n_tau.value):
            _check(lib().csl_h5_pump_probe_write_tau(h, k, float(taus[k]), n, _p(M1[k]), _p(M01[k])))
    finally:
        _check(lib().csl_h5_close(h))
    return path


def read_tau_values(tau_start: float, tau_end: float, tau_step: float) -> np.ndarray:
    """tau_i = tau_start + i * tau_step, i < int(|(tau_end - tau_start) / tau_step|) + 1
    (hdf5_io.h:769-773, lattice_md.cpp:768)."""
    n = int(abs((tau_end - tau_start) / tau_step)) + 1
    return np.array([tau_start + i * tau_step for i in range(n)])


# ---------------------------------------------------------------------------
# trajectory.h5

class MDWriter:
    """HDF5MDWriter (hdf5_io.h:83-300): extendible /trajectory datasets,
    appended a block of saved steps at a time."""

    def __init__(self, path: str, lattice_size: int, spin_dim: int, n_atoms: int, dims,
                 method: str, dt_initial: float, T_start: float, T_end: float, save_interval: int,
                 spin_length: float = 1.0, positions=None, deflate: int = 6):
        d1, d2, d3 = (int(x) for x in dims)
        pos = None if positions is None else _f64(positions).reshape(lattice_size, 3)
        self.h = _vp()
        self.path = path
        self.lattice_size, self.spin_dim = int(lattice_size), int(spin_dim)
        _check(lib().csl_h5_md_create(_b(path), lattice_size, spin_dim, n_atoms, d1, d2, d3, method.encode(),
                                      dt_initial, T_start, T_end, save_interval, spin_length, _p(pos),
                                      deflate, C.byref(self.h)))

    def write_steps(self, times, mags, spins):
        """times [n], mags [n,3,D] (M_antiferro, M_local, M_global), spins [n,N,D]."""
        t = _f64(times).reshape(-1)
        n = t.shape[0]
        m = _f64(mags).reshape(n, 3, self.spin_dim)
        s = _f64(spins).reshape(n, self.lattice_size, self.spin_dim)
        _check(lib().csl_h5_md_write_steps(self.h, n, _p(t), _p(m), _p(s)))

    @property
    def steps(self) -> int:
        n = C.c_size_t()
        _check(lib().csl_h5_md_steps(self.h, C.byref(n)))
        return n.value

    def close(self):
        if self.h:
            h, self.h = self.h, _vp()
            _check(lib().csl_h5_close(h))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------------------
# read-back

def read_dataset(path: str, name: str) -> np.ndarray:
    nd = C.c_int()
    dims = (C.c_size_t * 4)()
    _check(lib().csl_h5_dataset_shape(_b(path), name.encode(), C.byref(nd), dims))
    shape = tuple(int(dims[i]) for i in range(nd.value))
    out = np.empty(int(np.prod(shape)) if shape else 1)
    _check(lib().csl_h5_read_doubles(_b(path), name.encode(), _p(out), out.size))
    return out.reshape(shape)


def read_attr(path: str, obj: str, name: str) -> float:
    v = C.c_double()
    _check(lib().csl_h5_read_attr_double(_b(path), obj.encode(), name.encode(), C.byref(v)))
    return v.value


def read_attr_str(path: str, obj: str, name: str, cap: int = 256) -> str:
    buf = C.create_string_buffer(cap)
    _check(lib().csl_h5_read_attr_string(_b(path), obj.encode(), name.encode(), buf, cap))
    return buf.value.decode()
