"""The partitioned persistent kernel (csl_set_partition; one trajectory
spread over P co-resident workgroups that exchange halo values through global
memory every RK stage) against the oracle.

* Kitaev 12x12 (h = 0.1 along (1,1,1)) with the partition forced to P = 1, 2,
  3, 4, 6 parts: three trajectories with their own pulses, 150 RK4 steps,
  records and final spins within 1e-10 of the oracle.  P = 1 has no halo;
  P >= 2 exchange across the periodic boundary of the cell order.
* Kitaev 24x24 with an anisotropy matrix on every site (onsite path) and a
  damping alpha = 0.05 (Gilbert term), P = 3.
* Pyrochlore 6x6x6 (3456 sites, 6 bonds per site, 16 sublattices) with P = 4
  and P = 9 (parts that end inside a cell row).
* More trajectories than trajectory groups: each trajectory equals the same
  trajectory run alone (bit for bit: same parts, same arithmetic).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, model, workloads
from helpers import kitaev, random_state, unit_x_field

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT = 0.01
TP = np.array([[0.0, 0.0], [0.3, 0.0], [0.1, 0.6]])
AM = np.array([[0.5, 0.0], [0.5, 0.0], [0.5, 0.4]])


def _t1(T0, n, dt):
    return T0 + (n + 0.5) * dt


def _oracle_refs(oracle, uc, lat, s, f, n, alpha=0.0):
    o = oracle.OracleLattice(uc, lat.dim1, lat.dim2, lat.dim3)
    if alpha:
        o.set_alpha(alpha)
    refs = []
    for b in range(len(TP)):
        # per-pulse amplitudes of the batch scale the pulse directions (the
        # reference's pulse has one amplitude, 0.5)
        o.set_pulse(f * (AM[b, 0] / 0.5), TP[b, 0], f * (AM[b, 1] / 0.5), TP[b, 1], 0.5, 0.5, 0.0)
        refs.append(o.integrate_md(s.reshape(-1), 0.0, _t1(0.0, n, DT), DT))
    return refs


def _run(gl, lat, s, f, n, parts, alpha=0.0):
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
    if alpha:
        gl.set_damping(alpha)
    gl.set_partition(parts)
    res = gl.integrate_batched(s, TP, AM, 0.0, _t1(0.0, n, DT), DT, tgrid="boost", want_final=True)
    assert gl.last_launch_config()["partitioned"]
    assert gl.last_partition()[0] == parts
    return res


def _check(res, refs, tag):
    for b, ref in enumerate(refs):
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL, (tag, b)
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL, (tag, b)
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL, (tag, b)


def test_kitaev_forced_parts_vs_oracle(oracle):
    uc, lat = kitaev(12, 12, h=0.1)
    s = random_state(lat, 41)
    f = unit_x_field(lat)
    n = 150
    refs = _oracle_refs(oracle, uc, lat, s, f, n)
    with GPULattice(lat) as gl:
        for parts in (1, 2, 3, 4, 6):
            print(f"[partition] kitaev 12x12, P {parts}", flush=True)
            res = _run(gl, lat, s, f, n, parts)
            _check(res, refs, parts)


def test_kitaev_onsite_damped_vs_oracle(oracle):
    uc = model.build_kitaev_honeycomb(field_strength=0.1, field_direction=(1.0, 1.0, 1.0))
    for a in range(uc.N_atoms):
        uc.set_onsite_interaction(np.diag([0.15, -0.075, 0.03]), a)
    lat = model.Lattice(uc, 24, 24, 1)
    s = random_state(lat, 43)
    f = unit_x_field(lat)
    n = 120
    refs = _oracle_refs(oracle, uc, lat, s, f, n, alpha=0.05)
    with GPULattice(lat) as gl:
        print("[partition] kitaev 24x24 onsite + damping, P 3", flush=True)
        res = _run(gl, lat, s, f, n, 3, alpha=0.05)
    _check(res, refs, "onsite")


@pytest.mark.parametrize("parts", [4, 9])
def test_pyrochlore_forced_parts_vs_oracle(oracle, parts):
    lat = workloads.pyrochlore_lattice(6)
    st, _ = workloads.pyrochlore_tilted_aiao(lat, 2, seed=5)
    n, dt = 120, 0.02
    with GPULattice(lat) as gl:
        print(f"[partition] pyrochlore 6^3, P {parts}", flush=True)
        gl.set_partition(parts)
        res = gl.integrate_batched(st, np.zeros((2, 2)), np.zeros((2, 2)), 0.0, _t1(0.0, n, dt), dt,
                                   tgrid="boost", want_final=True)
        assert gl.last_partition()[0] == parts
    o = oracle.OracleLattice(lat.unit_cell, 6, 6, 6)
    for b in range(2):
        ref = o.integrate_md(st[b].reshape(-1), 0.0, _t1(0.0, n, dt), dt)
        assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL, b
        assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL, b
        assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL, b


def test_trajectory_loop_bit_identical():
    """A batch larger than the launch's trajectory groups: every group runs
    several trajectories in turn; each equals the trajectory run alone."""
    uc, lat = kitaev(12, 12, h=0.1)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    rng = np.random.default_rng(9)
    B, n = 3000, 30
    tp = rng.uniform(0.0, 0.3, size=(B, 2))
    am = rng.uniform(0.2, 0.6, size=(B, 2))
    init = np.stack([random_state(lat, 100 + b) for b in range(B)])
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
        gl.set_partition(4)
        print(f"[partition] trajectory loop, B {B}", flush=True)
        res = gl.integrate_batched(init, tp, am, 0.0, _t1(0.0, n, DT), DT, tgrid="boost", want_final=True)
        parts, groups = gl.last_partition()
        assert parts == 4 and groups < B
        for b in (0, groups, B - 1):
            one = gl.integrate_batched(init[b], tp[b:b + 1], am[b:b + 1], 0.0, _t1(0.0, n, DT), DT,
                                       tgrid="boost", want_final=True)
            assert np.array_equal(one["mags"][:, 0], res["mags"][:, b]), b
            assert np.array_equal(one["energy"][:, 0], res["energy"][:, b]), b
            assert np.array_equal(one["final"][0], res["final"][b]), b
