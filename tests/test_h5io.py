"""HDF5 output (libcsl_h5.so): the pump-probe file of HDF5PumpProbeWriter
(include/classical_spin/io/hdf5_io.h:729-1000) and the MD trajectory file of
HDF5MDWriter (hdf5_io.h:83-300).

Names, HDF5 types, shapes, chunking and values are checked with h5dump, the
HDF5 distribution's own dump tool, which shares no code with the writer; the
read-back functions of the library are checked against the same data."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from classical_mcmc_spin_amd import h5io

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

try:
    h5io.lib()
    _LIB_ERR = None
except h5io.H5Error as e:  # the writer is optional output, not the compute path
    _LIB_ERR = str(e)

pytestmark = pytest.mark.skipif(_LIB_ERR is not None, reason=f"libcsl_h5.so unavailable: {_LIB_ERR}")
needs_h5dump = pytest.mark.skipif(H5DUMP is None, reason="h5dump not installed")


def h5dump(*args) -> str:
    return subprocess.run([H5DUMP, *args], capture_output=True, text=True, check=True).stdout


def dump_values(path, name) -> np.ndarray:
    out = h5dump("-d", name, "-y", "-w", "0", "-m", "%.17g", path)
    body = out[out.index("DATA {") + 6:out.rindex("}", 0, out.rindex("}"))]
    return np.array([float(x) for x in re.findall(r"[-+0-9.eE]+|nan|inf", body)])


def dump_attr(path, obj, name):
    out = h5dump("-a", f"{obj}/{name}", path)
    dtype = re.search(r"DATATYPE\s+(H5T_\w+)", out).group(1)
    data = out[out.index("DATA {"):]
    m = re.search(r'\(0\):\s*"([^"]*)"', data)
    value = m.group(1) if m else float(re.search(r"\(0\):\s*([-+0-9.eE]+)", data).group(1))
    return dtype, value


def dump_header(path, name) -> str:
    return h5dump("-H", "-p", "-d", name, path)


def _pp_case(rng, n=9, tau=(0.0, 1.0, 0.25)):
    n_tau = int(abs((tau[1] - tau[0]) / tau[2])) + 1
    return {"times": np.linspace(0.0, 2.0, n), "M0": rng.normal(size=(n, 3, 3)),
            "M1": rng.normal(size=(n_tau, n, 3, 3)), "M01": rng.normal(size=(n_tau, n, 3, 3))}


def _write_pp(path, res, tau=(0.0, 1.0, 0.25), **kw):
    args = dict(lattice_size=8, n_atoms=2, dims=(2, 2, 1), spin_length=1.0, pulse_amp=0.7, pulse_width=0.38,
                pulse_freq=1.25, T_start=-5.0, T_end=20.0, T_step=0.05, tau_start=tau[0], tau_end=tau[1],
                tau_step=tau[2], ground_state_energy=-0.4375, ground_state_magnetization=0.125,
                Temp_start=1.0, Temp_end=1e-3, n_anneal=1000, T_zero_quench=True, quench_sweeps=50,
                pulse_field_direction=np.arange(6.0).reshape(2, 3), positions=np.arange(24.0).reshape(8, 3))
    args.update(kw)
    return h5io.write_pump_probe(path, res, **args)


def test_pump_probe_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    res = _pp_case(rng)
    p = _write_pp(str(tmp_path / "pump_probe_spectroscopy.h5"), res)
    assert np.array_equal(h5io.read_dataset(p, "/reference/times"), res["times"])
    for m, name in enumerate(("antiferro", "local", "global")):
        assert np.array_equal(h5io.read_dataset(p, f"/reference/M_{name}"), res["M0"][:, m])
        for k in range(res["M1"].shape[0]):
            assert np.array_equal(h5io.read_dataset(p, f"/tau_scan/tau_{k}/M1_{name}"), res["M1"][k, :, m])
            assert np.array_equal(h5io.read_dataset(p, f"/tau_scan/tau_{k}/M01_{name}"), res["M01"][k, :, m])
    taus = h5io.read_dataset(p, "/tau_scan/tau_values")
    assert np.array_equal(taus, [0.0, 0.25, 0.5, 0.75, 1.0])
    for k, t in enumerate(taus):
        assert h5io.read_attr(p, f"/tau_scan/tau_{k}", "tau_value") == t
    assert h5io.read_attr(p, "/metadata", "tau_steps") == 5
    assert h5io.read_attr_str(p, "/metadata", "T_zero_quench") == "true"
    assert h5io.read_attr(p, "/metadata", "quench_sweeps") == 50
    assert np.array_equal(h5io.read_dataset(p, "/metadata/positions"), np.arange(24.0).reshape(8, 3))


@needs_h5dump
def test_pump_probe_layout_matches_reference_writer(tmp_path):
    """Every /metadata attribute of write_metadata (hdf5_io.h:855-915) with
    its HDF5 type, and the dataset values as h5dump prints them."""
    rng = np.random.default_rng(2)
    res = _pp_case(rng)
    p = _write_pp(str(tmp_path / "pp.h5"), res)
    U64, F64, I32, S = "H5T_STD_U64LE", "H5T_IEEE_F64LE", "H5T_STD_I32LE", "H5T_STRING"
    expect = {"experiment_type": (S, "pump_probe_spectroscopy"), "code_version": (S, "ClassicalSpin_Cpp v1.0"),
              "file_format": (S, "HDF5_PumpProbe_v1.0"), "lattice_size": (U64, 8), "spin_dim": (U64, 3),
              "n_atoms": (U64, 2), "dim1": (U64, 2), "dim2": (U64, 2), "dim3": (U64, 1),
              "spin_length": (F64, 1.0), "pulse_amp": (F64, 0.7), "pulse_width": (F64, 0.38),
              "pulse_freq": (F64, 1.25), "T_start": (F64, -5.0), "T_end": (F64, 20.0), "T_step": (F64, 0.05),
              "integration_method": (S, "rk4"), "tau_start": (F64, 0.0), "tau_end": (F64, 1.0),
              "tau_step": (F64, 0.25), "tau_steps": (I32, 5), "ground_state_energy": (F64, -0.4375),
              "ground_state_magnetization": (F64, 0.125), "Temp_start": (F64, 1.0), "Temp_end": (F64, 1e-3),
              "n_anneal": (U64, 1000), "T_zero_quench": (S, "true"), "quench_sweeps": (U64, 50)}
    for name, (dtype, value) in expect.items():
        got_t, got_v = dump_attr(p, "/metadata", name)
        assert got_t == dtype, name
        assert got_v == value, name
    _, stamp = dump_attr(p, "/metadata", "creation_time")
    assert re.fullmatch(r"\d{4}-\d\d-\d\dT\d\d:\d\d:\d\d", stamp)
    assert "SIMPLE { ( 9, 3 ) / ( 9, 3 ) }" in dump_header(p, "/reference/M_local")
    assert "SIMPLE { ( 2, 3 ) / ( 2, 3 ) }" in dump_header(p, "/metadata/pulse_field_direction")
    assert np.array_equal(dump_values(p, "/reference/M_local"), res["M0"][:, 1].ravel())
    assert np.array_equal(dump_values(p, "/tau_scan/tau_4/M01_antiferro"), res["M01"][4, :, 0].ravel())
    assert np.array_equal(dump_values(p, "/tau_scan/tau_values"), [0.0, 0.25, 0.5, 0.75, 1.0])
    assert dump_attr(p, "/tau_scan/tau_2", "tau_value") == (F64, 0.5)


@needs_h5dump
def test_pump_probe_without_quench_or_optional_data(tmp_path):
    """quench_sweeps is written only with T_zero_quench (hdf5_io.h:912-914);
    pulse_field_direction and positions only when given."""
    rng = np.random.default_rng(3)
    res = _pp_case(rng, n=4, tau=(2.0, 0.0, -0.5))
    p = _write_pp(str(tmp_path / "pp.h5"), res, tau=(2.0, 0.0, -0.5), T_zero_quench=False,
                  pulse_field_direction=None, positions=None)
    out = h5dump("-H", p)
    assert 'ATTRIBUTE "quench_sweeps"' not in out
    assert 'DATASET "positions"' not in out and 'DATASET "pulse_field_direction"' not in out
    assert dump_attr(p, "/metadata", "T_zero_quench")[1] == "false"
    assert np.array_equal(h5io.read_dataset(p, "/tau_scan/tau_values"), [2.0, 1.5, 1.0, 0.5, 0.0])


def test_pump_probe_errors(tmp_path):
    rng = np.random.default_rng(4)
    res = _pp_case(rng)
    with pytest.raises(ValueError):   # 5 delays in the scan, 4 records given
        _write_pp(str(tmp_path / "a.h5"), {**res, "M1": res["M1"][:4], "M01": res["M01"][:4]})
    with pytest.raises(h5io.H5Error):
        _write_pp(str(tmp_path / "b.h5"), res, tau_step=0.0)
    with pytest.raises(h5io.H5Error):
        _write_pp(str(tmp_path / "no_such_dir" / "c.h5"), res)
    # one group per delay, once; the reference once
    import ctypes as C
    L = h5io.lib()
    meta = h5io.PumpProbeMeta(8, 3, 2, 2, 2, 1, 1.0, 0, 1, 1, 0, 1, 0.1, b"rk4", 0.0, 1.0, 0.5, 0, 0, 0, 0, 0,
                              0, 0, None, None)
    h, nt = C.c_void_p(), C.c_size_t()
    assert L.csl_h5_pump_probe_create(str(tmp_path / "d.h5").encode(), C.byref(meta), C.byref(nt), C.byref(h)) == 0
    try:
        assert nt.value == 3
        z = np.zeros((2, 3, 3))
        pz = z.ctypes.data_as(C.POINTER(C.c_double))
        assert L.csl_h5_pump_probe_write_tau(h, 1, 0.5, 2, pz, pz) == 0
        assert L.csl_h5_pump_probe_write_tau(h, 1, 0.5, 2, pz, pz) == -4
        assert L.csl_h5_pump_probe_write_tau(h, 3, 1.5, 2, pz, pz) == -2
        t = np.zeros(2)
        assert L.csl_h5_pump_probe_write_reference(h, 2, t.ctypes.data_as(C.POINTER(C.c_double)), pz) == 0
        assert L.csl_h5_pump_probe_write_reference(h, 2, t.ctypes.data_as(C.POINTER(C.c_double)), pz) == -4
        assert L.csl_h5_md_write_steps(h, 1, None, None, None) == -4
    finally:
        assert L.csl_h5_close(h) == 0


def _md_writer(path, N=12, A=3, deflate=6):
    return h5io.MDWriter(path, N, 3, A, (2, 2, 1), "rk4", 0.01, 0.0, 5.0, 10, spin_length=1.0,
                         positions=np.arange(N * 3.0).reshape(N, 3), deflate=deflate)


def test_md_trajectory_appends(tmp_path):
    rng = np.random.default_rng(5)
    N = 12
    p = str(tmp_path / "trajectory.h5")
    blocks = [(rng.normal(size=k), rng.normal(size=(k, 3, 3)), rng.normal(size=(k, N, 3))) for k in (4, 1, 7)]
    with _md_writer(p, N) as w:
        for t, m, s in blocks:
            w.write_steps(t, m, s)
        w.write_steps(np.zeros(0), np.zeros((0, 3, 3)), np.zeros((0, N, 3)))  # empty block: no-op
        assert w.steps == 12
    times = np.concatenate([b[0] for b in blocks])
    mags = np.concatenate([b[1] for b in blocks])
    spins = np.concatenate([b[2] for b in blocks])
    assert np.array_equal(h5io.read_dataset(p, "/trajectory/times"), times)
    for m, name in enumerate(("antiferro", "local", "global")):
        assert np.array_equal(h5io.read_dataset(p, f"/trajectory/magnetization_{name}"), mags[:, m])
    assert np.array_equal(h5io.read_dataset(p, "/trajectory/spins"), spins)
    assert h5io.read_attr(p, "/metadata", "save_interval") == 10
    assert h5io.read_attr_str(p, "/metadata", "file_format") == "HDF5_MD_v1.0"


@needs_h5dump
@pytest.mark.parametrize("deflate", [6, 0])
def test_md_layout_matches_reference_writer(tmp_path, deflate):
    """Extendible datasets with the chunk shapes and deflate level of
    hdf5_io.h:154-192, metadata types of hdf5_io.h:113-137."""
    rng = np.random.default_rng(6)
    N = 12
    p = str(tmp_path / "t.h5")
    s = rng.normal(size=(3, N, 3))
    with _md_writer(p, N, deflate=deflate) as w:
        w.write_steps(np.arange(3.0), rng.normal(size=(3, 3, 3)), s)
    hdr = dump_header(p, "/trajectory/spins")
    assert "SIMPLE { ( 3, 12, 3 ) / ( H5S_UNLIMITED, 12, 3 ) }" in hdr
    assert "CHUNKED ( 1, 12, 3 )" in hdr
    assert ("COMPRESSION DEFLATE { LEVEL 6 }" in hdr) == (deflate == 6)
    hdr = dump_header(p, "/trajectory/magnetization_global")
    assert "SIMPLE { ( 3, 3 ) / ( H5S_UNLIMITED, 3 ) }" in hdr and "CHUNKED ( 100, 3 )" in hdr
    hdr = dump_header(p, "/trajectory/times")
    assert "SIMPLE { ( 3 ) / ( H5S_UNLIMITED ) }" in hdr and "CHUNKED ( 1000 )" in hdr
    assert np.array_equal(dump_values(p, "/trajectory/spins"), s.ravel())
    assert "H5T_STD_U64LE" in dump_header(p, "/metadata/dimensions")
    assert np.array_equal(dump_values(p, "/metadata/dimensions"), [2, 2, 1])
    for name, dtype in (("lattice_size", "H5T_STD_U64LE"), ("n_atoms", "H5T_STD_U64LE"),
                        ("save_interval", "H5T_STD_U64LE"), ("dt_initial", "H5T_IEEE_F64LE"),
                        ("spin_length", "H5T_IEEE_F64LE"), ("integration_method", "H5T_STRING"),
                        ("code_version", "H5T_STRING")):
        assert dump_attr(p, "/metadata", name)[0] == dtype, name


def test_md_errors(tmp_path):
    with pytest.raises(h5io.H5Error):
        _md_writer(str(tmp_path / "x.h5"), deflate=10)
    w = _md_writer(str(tmp_path / "y.h5"))
    with pytest.raises(ValueError):   # spins of the wrong lattice size
        w.write_steps(np.zeros(1), np.zeros((1, 3, 3)), np.zeros((1, 5, 3)))
    w.close()
    w.close()  # idempotent
