"""The gpu:: C++ drop-in API (lattice_gpu_api.h) driven from C++ the way the
reference's Lattice glue calls it, checked against the oracle."""
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import kitaev, random_state, unit_x_field

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cxx", "test_gpu_api_shim")


def _w(f, arr, kind):
    a = np.ascontiguousarray(arr, dtype=np.float64 if kind == "d" else np.uint64).reshape(-1)
    f.write(struct.pack("<Q", a.size))
    f.write(a.tobytes())


def _r(f):
    (n,) = struct.unpack("<Q", f.read(8))
    return np.frombuffer(f.read(8 * n), dtype=np.float64).copy()


def test_shim_binary_built():
    assert os.path.exists(EXE), "tests/cxx/test_gpu_api_shim is built by __graft_entry__.build()"


@pytest.mark.gpu
def test_shim_against_oracle(tmp_path):
    import pyoracle
    uc, lat = kitaev(6, 6, h=0.1)
    o = pyoracle.OracleLattice(uc, 6, 6, 1)
    s = random_state(lat, 42)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    amp, width, freq, t1, t2 = 0.5, 0.5, 0.0, 0.2, 0.7
    T0, T1, dt = 0.0, 1.0, 0.01
    tau2 = np.array([0.3, 0.5, 0.9])
    bil, par, cnt = lat.padded_tables()
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    with open(inp, "wb") as fh:
        _w(fh, [lat.lattice_size, 3, lat.N_atoms, lat.num_bi], "u")
        _w(fh, lat.field, "d")
        _w(fh, lat.onsite, "d")
        _w(fh, bil, "d")
        _w(fh, par, "u")
        _w(fh, cnt, "u")
        _w(fh, np.concatenate([fd, fd]), "d")
        _w(fh, [amp, width, freq, t1, t2], "d")
        _w(fh, s, "d")
        _w(fh, [T0, T1, dt], "d")
        _w(fh, tau2, "d")
        _w(fh, lat.afm_signs, "d")
        _w(fh, lat.frames, "d")
    print(f"running {EXE}", flush=True)
    subprocess.run(["timeout", "-k", "5", "60", EXE, str(inp), str(out)], check=True)
    with open(out, "rb") as fh:
        meta, t_out, s_out, ml, ms, bmeta, btimes, bmag = (_r(fh) for _ in range(8))
    # energy and the unsupported-method error path
    assert meta[0] == pytest.approx(o.energy(s.reshape(-1)), rel=1e-12)
    assert meta[1] == 1.0
    # integrate_gpu: GPU grid t += dt while t < T1, records every 10 steps + final
    o.set_pulse(f, t1, f, t2, amp, width, freq)
    grid = []
    t = T0
    while t < T1:
        grid.append(t)
        t += dt
    grid.append(t)
    x = s.reshape(-1).copy()
    ref_t, ref_s = [], []
    for k in range(len(grid) - 1):
        if k % 10 == 0:
            ref_t.append(grid[k])
            ref_s.append(x.copy())
        x = o.rk4_step(x, grid[k], dt)
    ref_t.append(grid[-1])
    ref_s.append(x.copy())
    assert np.array_equal(t_out, np.array(ref_t))
    assert np.max(np.abs(s_out - np.concatenate(ref_s))) <= 1e-10
    # compute_magnetization_gpu: host sums with (-1)^i signs (lattice_gpu.cu:729-760)
    fin = ref_s[-1].reshape(-1, 3)
    sg = np.where(np.arange(lat.lattice_size) % 2 == 0, 1.0, -1.0)
    assert np.allclose(ml, fin.mean(axis=0), atol=1e-10)
    assert np.allclose(ms, (fin * sg[:, None]).mean(axis=0), atol=1e-10)
    # integrate_gpu_batched: t_pulse_1 from the handle, t_pulse_2 = tau2[b]
    B = len(tau2)
    assert tuple(bmeta) == (B, 3, len(btimes))
    bmag = bmag.reshape(len(btimes), B, 3, 3)
    for b in range(B):
        o.set_pulse(f, t1, f, tau2[b], amp, width, freq)
        x = s.reshape(-1).copy()
        t = T0
        k = 0
        while t < T1 - 1e-12:
            assert t == btimes[k]
            assert np.max(np.abs(o.mags(x) - bmag[k, b])) <= 1e-10
            x = o.rk4_step(x, t, dt)
            t += dt
            k += 1
        assert np.max(np.abs(o.mags(x) - bmag[k, b])) <= 1e-10
