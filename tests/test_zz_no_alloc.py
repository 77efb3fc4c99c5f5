"""csl_integrate_batched_noalloc (the allocation-free variant of
csl_integrate_batched: device inputs and outputs, scratch in a preallocated
per-GPU device region) against csl_integrate_batched on the same inputs:
two calls in a row, the second with a smaller batch; the first large enough
that the persistent kernel's work split (seam states and flags in the
scratch region) is active.  Same kernel and launch shape, so the records and
final states are bit-identical."""
import ctypes as C

import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, _capi
from helpers import kitaev, random_state, unit_x_field

pytestmark = pytest.mark.gpu


def _noalloc(gl, init, tp, am, T0, T1, dt, torch):
    B = tp.shape[0]
    per = 1 if init.ndim == 3 else 0
    dev = torch.device("cuda", 0)
    d_init = torch.from_numpy(np.ascontiguousarray(init.reshape(-1))).to(dev)
    d_tp = torch.from_numpy(np.ascontiguousarray(tp.reshape(-1))).to(dev)
    d_am = torch.from_numpy(np.ascontiguousarray(am.reshape(-1))).to(dev)
    _, n_rec = gl.count_records(T0, T1, dt, 1, "boost")
    d_mags = torch.zeros(n_rec * B * 9, dtype=torch.float64, device=dev)
    d_energy = torch.zeros(n_rec * B, dtype=torch.float64, device=dev)
    d_final = torch.zeros(B * gl.N * 3, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    n = C.c_size_t()
    ptr = lambda t: C.c_void_p(t.data_ptr())
    _capi.check(_capi.lib().csl_integrate_batched_noalloc(
        gl.h, B, ptr(d_init), per, ptr(d_tp), ptr(d_am), T0, T1, dt, 1, _capi.TGRID["boost"], ptr(d_mags),
        ptr(d_energy), ptr(d_final), C.byref(n)))
    assert n.value == n_rec
    return (d_mags.cpu().numpy().reshape(n_rec, B, 3, 3), d_energy.cpu().numpy().reshape(n_rec, B),
            d_final.cpu().numpy().reshape(B, gl.N, 3), gl.last_grid())


def test_noalloc_matches_integrate_batched_twice():
    torch = pytest.importorskip("torch")
    uc, lat = kitaev(12, 12, h=0.1)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    rng = np.random.default_rng(17)
    T0, dt = 0.0, 0.01
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
        for B, n_steps, per in ((3000, 60, True), (40, 25, False)):
            print(f"[noalloc] 12x12 Kitaev, B {B}, {n_steps} steps", flush=True)
            tp = rng.uniform(0.0, 0.4, size=(B, 2))
            am = rng.uniform(0.2, 0.6, size=(B, 2))
            init = (np.stack([random_state(lat, 500 + b) for b in range(B)]) if per
                    else random_state(lat, 77))
            T1 = T0 + (n_steps + 0.5) * dt
            ref = gl.integrate_batched(init, tp, am, T0, T1, dt, tgrid="boost", want_final=True)
            g_ref = gl.last_grid()
            m, e, fin, g = _noalloc(gl, init, tp, am, T0, T1, dt, torch)
            assert g == g_ref
            if B == 3000:
                assert g < B, "the first call must exercise the work split"
            assert np.array_equal(m, ref["mags"])
            assert np.array_equal(e, ref["energy"])
            assert np.array_equal(fin, ref["final"])


def test_noalloc_rejects_oversized_scratch():
    """A launch whose scratch (here the time grid of 3.3 million steps) does
    not fit the region is refused before anything runs; the next call on the
    same handle works."""
    torch = pytest.importorskip("torch")
    uc, lat = kitaev(6, 6, h=0.1)
    dev = torch.device("cuda", 0)
    with GPULattice(lat) as gl:
        init = random_state(lat, 3)
        d_init = torch.from_numpy(init.reshape(-1).copy()).to(dev)
        d_tp = torch.zeros(2, dtype=torch.float64, device=dev)
        d_am = torch.zeros(2, dtype=torch.float64, device=dev)
        si = 1_000_000
        n_steps = 3_300_000
        dt = 0.01
        d_mags = torch.zeros((n_steps // si + 2) * 9, dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        n = C.c_size_t()
        ptr = lambda t: C.c_void_p(t.data_ptr())
        rc = _capi.lib().csl_integrate_batched_noalloc(gl.h, 1, ptr(d_init), 0, ptr(d_tp), ptr(d_am), 0.0,
                                                        (n_steps + 0.5) * dt, dt, si, _capi.TGRID["boost"],
                                                        ptr(d_mags), None, None, C.byref(n))
        assert rc == _capi.CSL_ERR_CAPACITY
        assert b"scratch" in _capi.lib().csl_last_error()
        m, e, fin, g = _noalloc(gl, init, np.zeros((2, 2)), np.zeros((2, 2)), 0.0, 0.105, 0.01, torch)
        ref = gl.integrate_batched(init, np.zeros((2, 2)), np.zeros((2, 2)), 0.0, 0.105, 0.01, tgrid="boost",
                                   want_final=True)
        assert np.array_equal(m, ref["mags"]) and np.array_equal(fin, ref["final"])
