"""CPU-only tests: model layer, seeding, golden fixtures, C ABI exports and the
host-side logic of libcsl_hip.so that needs no GPU (time grids, argument
checks)."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

from classical_mcmc_spin_amd import _capi, model, seeding
from helpers import kitaev, random_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden", "kitaev_6x6_seed1234.npz")


# ---------------------------------------------------------------- model layer

def test_offsets_beyond_lattice_rejected(oracle):
    uc = model.build_kitaev_honeycomb(J3=0.02)   # J3 bond offset (1,-2,0)
    with pytest.raises(ValueError):
        model.Lattice(uc, 1, 1, 1)
    with pytest.raises(RuntimeError):
        oracle.OracleLattice(uc, 1, 1, 1)


def test_tables_match_oracle_expansion(oracle):
    uc = model.build_kitaev_honeycomb(J=0.3)
    lat = model.Lattice(uc, 1, 1, 1)                 # every bond wraps onto the same pair
    assert np.array_equal(oracle.OracleLattice(uc, 1, 1, 1).tables()[2], lat.bi_J)
    for dims in [(2, 3, 1), (6, 6, 1), (4, 5, 2)]:
        uc = model.build_kitaev_honeycomb(J=0.3, field_strength=0.2, field_direction=(1, 0, 0),
                                          J2_A=0.1, J2_B=-0.05, J3=0.02, J_perp=0.07, Ka=0.1)
        lat = model.Lattice(uc, *dims)
        o = oracle.OracleLattice(uc, *dims)
        off, par, J = o.tables()
        assert np.array_equal(off, lat.bi_offset)
        assert np.array_equal(par, lat.bi_partner)
        assert np.array_equal(J, lat.bi_J)


def test_honeycomb_geometry():
    """Every site has 3 nearest neighbours of the other sublattice at distance
    1/sqrt(3) (minimum image), and each bond appears with J at one end and J^T
    at the other (lattice.h:790-803)."""
    uc, lat = kitaev(5, 4)
    a1, a2 = uc.lattice_vectors[0], uc.lattice_vectors[1]
    off = lat.bi_offset.astype(np.int64)
    par = lat.bi_partner.astype(np.int64)
    for i in range(lat.lattice_size):
        ks = range(off[i], off[i + 1])
        assert len(ks) == 3
        for k in ks:
            p = par[k]
            assert p % 2 != i % 2
            d = lat.positions[p] - lat.positions[i]
            best = min(np.linalg.norm(d + m * 5 * a1 + n * 4 * a2) for m in (-1, 0, 1) for n in (-1, 0, 1))
            assert best == pytest.approx(1 / math.sqrt(3), abs=1e-12)
            back = [kk for kk in range(off[p], off[p + 1]) if par[kk] == i]
            assert any(np.array_equal(lat.bi_J[kk], lat.bi_J[k].T) for kk in back)


def test_kitaev_bond_matrices_closed_form():
    Jx, Jy, Jz = model.kitaev_bond_matrices(-1.0, 0.25, -0.02, 0.1)
    for M, a in ((Jx, 0), (Jy, 1), (Jz, 2)):
        assert np.allclose(M, M.T)
        assert M[a, a] == pytest.approx(-0.9)
        others = [i for i in range(3) if i != a]
        assert M[others[0], others[1]] == 0.25
        assert all(M[a, o] == -0.02 for o in others)
    R = model.kitaev_frame()
    assert np.allclose(R @ R.T, np.eye(3)) and np.linalg.det(R) == pytest.approx(1.0)


def test_padded_tables_round_trip():
    uc, lat = kitaev(3, 4, J2_A=0.1)
    bil, par, cnt = lat.padded_tables()
    assert bil.shape == (lat.lattice_size, lat.num_bi, 3, 3)
    off = lat.bi_offset.astype(np.int64)
    for i in range(lat.lattice_size):
        n = int(cnt[i])
        assert np.array_equal(bil[i, :n], lat.bi_J[off[i]:off[i] + n])
        assert np.array_equal(par[i, :n], lat.bi_partner[off[i]:off[i] + n])
        assert not bil[i, n:].any()


def test_spin_config_text_round_trip(tmp_path):
    uc, lat = kitaev(2, 2)
    s = random_state(lat, 3)
    p = tmp_path / "initial_spins.txt"
    lat.save_spin_config(str(p), s)
    assert np.array_equal(lat.load_spin_config(str(p)), s)


# ---------------------------------------------------------------- seeding

def test_seeding_deterministic_and_uniform():
    a = seeding.random_spins(4096, 1234)
    b = seeding.random_spins(4096, 1234)
    assert np.array_equal(a, b)
    assert np.allclose(np.linalg.norm(a, axis=1), 1.0, atol=1e-15)
    assert np.all(np.abs(a.mean(axis=0)) < 0.05)        # isotropic
    assert not np.array_equal(a, seeding.random_spins(4096, 1235))
    g = seeding.normal(200000, 7)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01


def test_quench_reaches_stationary_state():
    uc, lat = kitaev(6, 6, h=0.3)
    s0 = random_state(lat, 2)

    def energy(S):
        return 0.5 * np.sum(S * (lat.local_fields_host(S) - lat.field)) / lat.lattice_size

    e = [energy(s0)]
    q = s0
    for _ in range(4):
        q = seeding.quench(lat, q, sweeps=100)
        e.append(energy(q))
    assert all(b <= a + 1e-14 for a, b in zip(e, e[1:]))   # each sweep lowers E
    H = lat.local_fields_host(q)
    torque0 = np.max(np.linalg.norm(np.cross(lat.local_fields_host(s0), s0), axis=1))
    torque = np.max(np.linalg.norm(np.cross(H, q), axis=1))
    assert torque < 1e-3 * torque0 and np.allclose(np.linalg.norm(q, axis=1), 1.0)


def test_golden_fixture_reproduced(oracle):
    g = np.load(GOLDEN)
    uc, lat = kitaev(6, 6, h=0.1)
    s = seeding.random_spins(lat.lattice_size, 1234)
    assert np.array_equal(s, g["init"])
    o = oracle.OracleLattice(uc, 6, 6, 1)
    assert np.array_equal(o.rhs(s.reshape(-1)).reshape(-1, 3), g["dsdt"])
    md = o.integrate_md(s.reshape(-1), 0.0, 2.005, 0.01)
    assert np.array_equal(md["mags"], g["md_mags"]) and np.array_equal(md["energy"], g["md_energy"])


# ---------------------------------------------------------------- C ABI (no GPU needed)

def _declared_symbols(headers=None, exclude=("csl_h5.h",)):
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in headers or os.listdir(inc):
        if fn.endswith(".h") and fn not in (exclude if headers is None else ()):
            txt = open(os.path.join(inc, fn)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(csl_\w+)\s*\(", txt, re.M))
    return sorted(names)


def test_capi_exports_every_declared_symbol():
    L = _capi.lib()
    names = _declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert set(names) <= set(_capi.SIGNATURES)


def test_h5_library_exports_every_declared_symbol():
    """libcsl_h5.so (HDF5 output) exports what include/csl_h5.h declares."""
    from classical_mcmc_spin_amd import h5io
    try:
        L = h5io.lib()
    except h5io.H5Error as e:
        pytest.skip(f"libcsl_h5.so unavailable: {e}")
    names = _declared_symbols(["csl_h5.h"])
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(h5io.SIGNATURES)


def test_capi_host_only_paths():
    L = _capi.lib()
    ns, nr = C.c_size_t(), C.c_size_t()
    _capi.check(L.csl_count_records(0.0, 100.005, 0.01, 1, _capi.TGRID["boost"], C.byref(ns), C.byref(nr)))
    assert (ns.value, nr.value) == (10000, 10001)
    _capi.check(L.csl_count_records(0.0, 1.0, 0.25, 1, _capi.TGRID["gpu"], C.byref(ns), C.byref(nr)))
    assert ns.value == 4
    _capi.check(L.csl_count_records(0.0, 1.0, 0.1, 3, _capi.TGRID["batched"], C.byref(ns), C.byref(nr)))
    assert nr.value == (ns.value + 2) // 3 + 1
    # error paths: return codes plus a message, no GPU touched
    assert L.csl_count_records(0.0, 1.0, 0.0, 1, 2, C.byref(ns), C.byref(nr)) == _capi.CSL_ERR_ARG
    assert L.csl_count_records(0.0, 1.0, 0.1, 0, 2, C.byref(ns), C.byref(nr)) == _capi.CSL_ERR_ARG
    assert L.csl_count_records(0.0, 1.0, 0.1, 1, 9, C.byref(ns), C.byref(nr)) == _capi.CSL_ERR_ARG
    assert L.csl_destroy_lattice(None) == _capi.CSL_ERR_NULL
    assert b"null" in L.csl_last_error()
    assert L.csl_set_spins(None, None, 0) == _capi.CSL_ERR_NULL
    with pytest.raises(_capi.CslError):
        _capi.check(L.csl_energy(None, None))


def test_gpu_grid_policy_matches_reference_loops():
    """CSL_TGRID_GPU / BATCHED reproduce the accumulated-time loops of
    gpu::ode::integrate (integrator.cuh:470-487) and the batched path
    (lattice_gpu.cu:557-565) including their final observation."""
    L = _capi.lib()
    for (T0, T1, dt) in [(0.0, 1.0, 0.1), (-1.0, 2.0, 0.3), (0.0, 81.92, 0.01)]:
        t, n = T0, 0
        while t < T1:
            t += dt
            n += 1
        ns, nr = C.c_size_t(), C.c_size_t()
        _capi.check(L.csl_count_records(T0, T1, dt, 1, 0, C.byref(ns), C.byref(nr)))
        assert ns.value == n and nr.value == n + 1
        t, n = T0, 0
        while t < T1 - 1e-12:
            t += dt
            n += 1
        _capi.check(L.csl_count_records(T0, T1, dt, 1, 1, C.byref(ns), C.byref(nr)))
        assert ns.value == n


def test_pulse_segments_product_equals_oracle():
    """csl_pulse_segments (host code of the product) against the oracle's
    build_pulse_segments restatement, on the cases the oracle is pinned with
    (tests/test_oracle_physics.py) and on random windows: identical rows."""
    import pyoracle as po
    from classical_mcmc_spin_amd import gpu as G
    cases = [(-100.0, 100.0, [0.0, 50.0], 4.5, 0.01), (-10.0, 10.0, [0.0, 1.0], 4.5, 0.01),
             (-10.0, 10.0, [0.0, 5.0], 9 * 0.45, 0.1), (-10.0, 10.0, [0.0, 0.55], 3.6, 0.1),
             (0.0, 10.013, [0.0, 5.0], 3.6, 0.1), (0.0, 10.0, [100.0], 3.6, 0.1), (0.0, 0.0, [0.0], 1.0, 0.1),
             (1.0, 0.0, [0.0], 1.0, 0.1), (0.0, 1.0, [0.0], 0.1, 0.0), (-5.0, 20.0, [], 1.0, 0.05)]
    rng = np.random.default_rng(3)
    for _ in range(200):
        T0 = rng.uniform(-20, 0)
        T1 = T0 + rng.uniform(0.5, 40)
        cases.append((T0, T1, list(rng.uniform(T0 - 5, T1 + 5, size=rng.integers(1, 4))),
                      rng.uniform(0.01, 6), rng.choice([0.01, 0.05, 0.1, 0.013])))
    for T0, T1, c, w, st in cases:
        a = G.pulse_segments(T0, T1, c, w, st, 20.0)
        b = po.pulse_segments(T0, T1, c, w, st, 20.0) if c else po.pulse_segments(T0, T1, [1e300], w, st, 20.0)
        assert a.shape == b.shape and np.array_equal(a, b), (T0, T1, c, w, st)
    with pytest.raises(_capi.CslError):
        G.pulse_segments(0.0, float("nan"), [0.0], 1.0, 0.1)


def test_split_schedule_covers_every_step_once():
    """csl_split_schedule (the persistent kernel's work split): every step of
    every trajectory in exactly one piece, pieces of a trajectory contiguous,
    a cut trajectory begun first on workgroup g (saving slot g) and finished
    last on g + 1 (loading slot g), and no workgroup with more than
    ceil(B * n / G) steps (the launch's length)."""
    L = _capi.lib()
    rng = np.random.default_rng(5)
    cases = [(513, 256, 8192), (300, 256, 300), (257, 256, 1), (1000, 7, 13)]
    cases += [(int(g + rng.integers(1, 3 * g + 5)), int(g), int(rng.integers(1, 400)))
              for g in rng.integers(1, 300, size=60)]
    for B, G, n in cases:
        cap = G * (2 + (B * n + G - 1) // G // n + 2)
        rows = np.empty(cap * 5, dtype=np.int32)
        pp = C.c_size_t()
        _capi.check(L.csl_split_schedule(B, G, n, cap, rows.ctypes.data_as(C.POINTER(C.c_int)), C.byref(pp)))
        P = pp.value
        r = rows[:G * P * 5].reshape(G, P, 5)
        cover = {}
        heads = {}
        work = []
        for g in range(G):
            used = [q for q in r[g] if q[0] >= 0]
            work.append(sum(q[2] - q[1] for q in used))
            for i, (j, s0, s1, si, so) in enumerate(used):
                assert 0 <= s0 < s1 <= n
                cover.setdefault(j, []).append((s0, s1))
                if so >= 0:
                    assert (i, so, s0) == (0, g, 0) and s1 < n
                    heads[g] = (j, s1)
                if si >= 0:
                    assert i == len(used) - 1 and si == g - 1 and s0 > 0
        for g in range(G):
            for (j, s0, s1, si, so) in r[g]:
                if j >= 0 and si >= 0:
                    assert heads[si] == (j, s0)
        assert sorted(cover) == list(range(B))
        for j, iv in cover.items():
            iv.sort()
            assert iv[0][0] == 0 and iv[-1][1] == n and all(a[1] == b[0] for a, b in zip(iv, iv[1:]))
        assert max(work) == -(-B * n // G)
    assert L.csl_split_schedule(10, 10, 5, 100, None, None) == _capi.CSL_ERR_ARG
