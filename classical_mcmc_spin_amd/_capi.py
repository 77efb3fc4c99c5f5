"""ctypes binding of libcsl_hip.so (the C ABI declared in include/csl_gpu.h).

The library is the only compute path: if it is missing or fails to load, every
call raises :class:`CslError` — there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcsl_hip.so")

CSL_OK = 0
CSL_ERR_NULL = -1
CSL_ERR_ARG = -2
CSL_ERR_UNSUPPORTED = -3
CSL_ERR_NOSTATE = -4
CSL_ERR_HIP = -5
CSL_ERR_CAPACITY = -6
CSL_ERR_NONFINITE = -7
CSL_ERR_STEP = -8

TGRID = {"gpu": 0, "batched": 1, "boost": 2}

# (name, restype, argtypes) for every exported entry point of csl_gpu.h
_vp = C.c_void_p
_sz = C.c_size_t
_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_psz = C.POINTER(C.c_size_t)
_pi = C.POINTER(C.c_int)
_cs = C.c_char_p

SIGNATURES = {
    "csl_last_error": (C.c_char_p, []),
    "csl_device_count": (_i, [_pi]),
    "csl_version": (_i, [_pi, _pi]),
    "csl_create_lattice": (_i, [_sz, _sz, _sz, _sz, _pd, _pd, _pd, _psz, _psz, C.POINTER(_vp)]),
    "csl_create_lattice_csr": (_i, [_sz, _sz, _sz, _pd, _pd, _psz, _psz, _pd, C.POINTER(_vp)]),
    "csl_destroy_lattice": (_i, [_vp]),
    "csl_lattice_info": (_i, [_vp, _psz, _psz, _psz, _psz, _psz, _pi]),
    "csl_set_pulse": (_i, [_vp, _pd, _d, _d, _d, _d, _d]),
    "csl_set_damping": (_i, [_vp, _d]),
    "csl_set_frames": (_i, [_vp, _pd, _pd]),
    "csl_set_spins": (_i, [_vp, _pd, _sz]),
    "csl_get_spins": (_i, [_vp, _pd, _sz]),
    "csl_step": (_i, [_vp, _d, _d, _cs]),
    "csl_integrate": (_i, [_vp, _d, _d, _d, _sz, _cs, _i, _i, _sz, _pd, _pd, _pd, _pd, _psz]),
    "csl_energy": (_i, [_vp, _pd]),
    "csl_normalize": (_i, [_vp, _d]),
    "csl_magnetization": (_i, [_vp, _pd, _pd]),
    "csl_local_field": (_i, [_vp, _pd, _sz]),
    "csl_rhs": (_i, [_vp, _d, _pd, _sz]),
    "csl_integrate_batched": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _sz, _cs, _i, _i, _sz,
                                   _pd, _pd, _pd, _pd, _psz]),
    "csl_count_records": (_i, [_d, _d, _d, _sz, _i, _psz, _psz]),
    "csl_pump_probe": (_i, [_vp, _pd, _pd, _d, _d, _d, _d, _d, _d, _d, _d, _d, _sz, _sz, _psz, _psz,
                            _pd, _pd, _pd, _pd, _pd]),
    "csl_pump_probe_ex": (_i, [_vp, _pd, _pd, _d, _d, _d, _d, _d, _d, _d, _d, _d, _i, _d, _i, _cs, _d, _d, _sz,
                               _sz, _psz, _psz, _pd, _pd, _pd, _pd, _pd, _pi, _pd]),
    "csl_integrate_dopri5": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _i, _d, _d, _sz, _i, _pd, _psz, _sz,
                                  _sz, _sz, _pd, _pd, _pd, _pd, _psz, _psz, _psz]),
    "csl_pulse_segments": (_i, [_d, _d, _pd, _sz, _d, _d, _d, _sz, _pd, _psz]),
    "csl_autotune_batched": (_i, [_vp, _sz, _sz, _pi, _pd]),
    "csl_last_kernel_ms": (_i, [_vp, _pd, _psz]),
    "csl_last_launch_config": (_i, [_vp, _pi, _pi, _pi]),
    "csl_set_launch_config": (_i, [_vp, _sz, _i, _i, _i]),
    "csl_set_batch_split": (_i, [_vp, _i]),
    "csl_set_partition": (_i, [_vp, _i]),
    "csl_last_partition": (_i, [_vp, _pi, _pi]),
    "csl_set_stage_graph": (_i, [_vp, _i]),
    "csl_mixed_set_stage_graph": (_i, [_vp, _i]),
    "csl_integrate_batched_noalloc": (_i, [_vp, _sz, _vp, _i, _vp, _vp, _d, _d, _d, _sz, _i, _vp, _vp, _vp, _psz]),
    "csl_persistent_candidates": (_i, [_vp, _sz, _pi, _pi, _pi, _pi, _psz]),
    "csl_last_launch_variant": (_i, [_vp, _pi]),
    "csl_last_grid": (_i, [_vp, _pi]),
    "csl_split_schedule": (_i, [_sz, _sz, _sz, _sz, _pi, _psz]),
    "csl_integrate_multi": (_i, [C.POINTER(_vp), _sz, _psz, _pd, _pd, _pd, _d, _d, _d, _sz, _i, _sz, _pd, _pd, _pd,
                                 _psz]),
    "csl_set_replica_hamiltonians": (_i, [_vp, _sz, _pd, _pd]),
    "csl_quench": (_i, [_vp, _sz, _pd, _sz, _d, _d, _psz]),
    "csl_nonlinear_spectrum": (_i, [_sz, _sz, _sz, _pd, _pd, _pd, _sz, _pd, _psz, _psz]),
    "csl_structure_factor": (_i, [_vp, _sz, _pd, _i, _d, _d, _d, _i, _sz, _sz, _sz, _sz, _pd, _sz, _pd,
                                  _psz, _pd]),
    "csl_adjoint_gradient": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _i, _pd, _sz, _pd, _sz, _pd, _pd,
                                  _sz, _pd, _pd, _pd, _pd]),
    # mixed SU(2) + SU(3) path (include/csl_mixed.h)
    "csl_mixed_create": (_i, [_sz, _sz, _sz, _pd, _pd, _psz, _psz, _pd, _pd, _pd, _psz, _psz, _pd,
                              _psz, _psz, _pd, _psz, _psz, _pd, C.POINTER(_vp)]),
    "csl_mixed_destroy": (_i, [_vp]),
    "csl_mixed_info": (_i, [_vp, _psz, _psz, _pi, _pi, _pi, _pi]),
    "csl_mixed_set_frames": (_i, [_vp, _pd, _pd, _pd]),
    "csl_mixed_set_pulse": (_i, [_vp, _i, _pd, _d, _d, _d, _d, _d]),
    "csl_mixed_set_damping": (_i, [_vp, _d, _pd, _pd]),
    "csl_mixed_set_spins": (_i, [_vp, _pd]),
    "csl_mixed_get_spins": (_i, [_vp, _pd]),
    "csl_mixed_record": (_i, [_vp, _pd]),
    "csl_mixed_energy": (_i, [_vp, _pd]),
    "csl_mixed_rhs": (_i, [_vp, _d, _pd]),
    "csl_mixed_normalize": (_i, [_vp, _d, _d]),
    "csl_mixed_step": (_i, [_vp, _d, _d, _cs]),
    "csl_mixed_integrate": (_i, [_vp, _d, _d, _d, _sz, _cs, _i, _sz, _pd, _pd, _pd, _psz]),
    "csl_mixed_integrate_batched": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _sz, _cs, _i, _sz, _pd,
                                         _pd, _pd, _psz]),
    "csl_mixed_integrate_dopri5": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _i, _d, _d, _sz, _i, _pd, _psz,
                                        _sz, _sz, _sz, _pd, _pd, _pd, _psz, _psz, _psz]),
    "csl_mixed_scale_gradient": (_i, [_vp, _sz, _pd, _i, _pd, _pd, _d, _d, _d, _sz, _i, _pd, _i, _pd, _pd, _pd,
                                      _sz, _pd, _pd, _psz]),
    "csl_mixed_last_kernel_ms": (_i, [_vp, _pd, _psz]),
    "csl_mixed_set_persistent": (_i, [_vp, _i]),
    "csl_mixed_last_path": (_i, [_vp, _pi, _pi, _psz, _pd]),
}


class CslError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"csl error {code}: {msg}")
        self.code = code


_lib = None
_load_error = None


def lib():
    """Return the loaded library, raising CslError if it cannot be loaded."""
    global _lib, _load_error
    if _lib is not None:
        return _lib
    if _load_error is not None:
        raise CslError(CSL_ERR_HIP, _load_error)
    if not os.path.exists(LIB_PATH):
        _load_error = (f"{LIB_PATH} is missing: build it with "
                       "`python -m classical_mcmc_spin_amd.build` (needs hipcc)")
        raise CslError(CSL_ERR_HIP, _load_error)
    # torch ships its own HIP runtime (torch/lib/libamdhip64.so, soname
    # libamdhip64.so.7) and loads it by path.  Were this library loaded first,
    # it would bind the system runtime, torch would then load a second one,
    # and the second runtime finds no GPU ("No HIP GPUs are available").
    # Importing torch first makes this library bind the runtime torch already
    # holds (same soname), so device memory from torch and these entry points
    # share one runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    try:
        L = C.CDLL(LIB_PATH)
    except OSError as e:  # missing HIP runtime etc.
        _load_error = f"cannot load {LIB_PATH}: {e}"
        raise CslError(CSL_ERR_HIP, _load_error) from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != CSL_OK:
        msg = lib().csl_last_error()
        raise CslError(rc, msg.decode() if msg else "")


def dptr(a):
    """Pointer to a C-contiguous float64 array (None -> NULL)."""
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "need contiguous float64"
    return a.ctypes.data_as(_pd)


def szptr(a):
    if a is None:
        return None
    assert a.dtype == np.uintp and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_psz)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def device_count() -> int:
    n = C.c_int(0)
    check(lib().csl_device_count(C.byref(n)))
    return n.value
