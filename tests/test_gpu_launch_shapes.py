"""Every persistent launch shape the autotuner can pick, forced through
csl_set_launch_config, against the oracle.

For each lattice below, every candidate of csl_persistent_candidates (sites
per thread 1/2/3/4/6 as the lattice admits x one or two LDS buffers x
runtime or compile-time RK stages; the launcher's FIELD_ONLY and ATOM_MAJOR
specialisations where the lattice qualifies) runs three trajectories with
their own pulses for 200 RK4 steps: records and final spins within 1e-10 of
the oracle, and the launch reports the forced shape.  The batch work split
(more trajectories than co-resident workgroups) is checked per shape as
bit-identity with the one-workgroup-per-trajectory launch.

Lattices: Kitaev 12x12 and 24x24 (h = 0.1: per-sublattice field, full
coordination -> FIELD_ONLY / ATOM_MAJOR shapes), Kitaev 24x24 with a site
anisotropy (onsite matrix -> the general table path), Kitaev 36x36 (the 2DCS
lattice).
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, model
from helpers import random_state, unit_x_field

pytestmark = pytest.mark.gpu

TOL = 1e-10
DT = 0.01
N_STEPS = 200
LATTICES = {
    "kitaev12_h": dict(L=12, h=0.1),
    "kitaev24_h": dict(L=24, h=0.1),
    "kitaev24_onsite": dict(L=24, h=0.1, onsite=0.15),
    "kitaev36_h": dict(L=36, h=0.1),
}
TP = np.array([[0.0, 0.0], [0.3, 0.0], [0.1, 0.6]])
AM = np.array([[0.5, 0.0], [0.5, 0.0], [0.5, 0.4]])


def _lattice(name):
    p = LATTICES[name]
    uc = model.build_kitaev_honeycomb(field_strength=p["h"], field_direction=(1.0, 1.0, 1.0))
    if "onsite" in p:
        for a in range(uc.N_atoms):
            A = np.diag([p["onsite"], -0.5 * p["onsite"], 0.2 * p["onsite"]])
            uc.set_onsite_interaction(A, a)
    lat = model.Lattice(uc, p["L"], p["L"], 1)
    return uc, lat


_REF = {}


def _reference(oracle, name):
    if name not in _REF:
        uc, lat = _lattice(name)
        o = oracle.OracleLattice(uc, lat.dim1, lat.dim2, 1)
        s = random_state(lat, 55)
        f = unit_x_field(lat)
        T1 = (N_STEPS + 0.5) * DT
        refs = []
        for b in range(len(TP)):
            # the reference's pulse has one amplitude (0.5): per-pulse
            # amplitudes of the batch scale the pulse directions instead
            o.set_pulse(f * (AM[b, 0] / 0.5), TP[b, 0], f * (AM[b, 1] / 0.5), TP[b, 1], 0.5, 0.5, 0.0)
            refs.append(o.integrate_md(s.reshape(-1), 0.0, T1, DT))
        _REF[name] = (uc, lat, s, f, refs)
    return _REF[name]


# the candidate shapes of a lattice are enumerated on the device inside the
# test (collection runs without a GPU); each shape is printed before it runs
@pytest.mark.parametrize("name", sorted(LATTICES))
def test_every_candidate_vs_oracle(oracle, name):
    uc, lat, s, f, refs = _reference(oracle, name)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    T1 = (N_STEPS + 0.5) * DT
    seen = set()
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
        cands = gl.launch_candidates()
        assert cands, name
        for c in cands:
            tag = (f"{name} block {c['block']} spt {c['sites_per_thread']} flags {c['flags']} "
                   f"field_only {c['field_only']} atom_major {c['atom_major']}")
            print(f"[shape] {tag}", flush=True)
            gl.set_launch_config(len(TP), c["block"], c["sites_per_thread"], c["flags"])
            res = gl.integrate_batched(s, TP, AM, 0.0, T1, DT, tgrid="boost", want_final=True)
            got = gl.last_launch_config()
            assert (got["block"], got["sites_per_thread"], got["flags"]) == \
                (c["block"], c["sites_per_thread"], c["flags"]), tag
            assert (got["field_only"], got["atom_major"]) == (c["field_only"], c["atom_major"]), tag
            for b, ref in enumerate(refs):
                assert np.max(np.abs(res["mags"][:, b] - ref["mags"])) <= TOL, (tag, b)
                assert np.max(np.abs(res["energy"][:, b] - ref["energy"])) <= TOL, (tag, b)
                assert np.max(np.abs(res["final"][b] - ref["final"].reshape(-1, 3))) <= TOL, (tag, b)
            seen.add((c["field_only"], c["atom_major"], c["sites_per_thread"]))
    if name != "kitaev24_onsite":
        assert any(fo for fo, _, _ in seen), "no FIELD_ONLY shape among the candidates"
    else:
        assert not any(fo for fo, _, _ in seen), "an onsite matrix must take the general path"
    if name in ("kitaev12_h", "kitaev24_h", "kitaev36_h"):
        assert any(am for _, am, _ in seen), "no ATOM_MAJOR shape among the candidates"


@pytest.mark.parametrize("name", ["kitaev12_h", "kitaev24_onsite"])
def test_split_bit_identical_every_candidate(name):
    """More trajectories than co-resident workgroups: the even work split
    (seam states handed over through global memory) gives the same bits as
    one workgroup per trajectory, for every launch shape."""
    uc, lat = _lattice(name)
    rng = np.random.default_rng(3)
    B = 9000 if lat.dim1 <= 12 else 3000
    n = 40
    tp = rng.uniform(0.0, 0.3, size=(B, 2))
    am = np.full((B, 2), 0.5)
    s = random_state(lat, 8)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.5, 0.0, 0.0, 0.0)
        for c in gl.launch_candidates():
            tag = f"{name} block {c['block']} spt {c['sites_per_thread']} flags {c['flags']}"
            print(f"[split] {tag} B {B}", flush=True)
            gl.set_launch_config(B, c["block"], c["sites_per_thread"], c["flags"])
            gl.set_batch_split(True)
            a = gl.integrate_batched(s, tp, am, 0.0, (n + 0.5) * DT, DT, tgrid="boost", want_final=True)
            G = gl.last_grid()
            assert G < B, f"{tag}: batch not split (grid {G})"
            gl.set_batch_split(False)
            b = gl.integrate_batched(s, tp, am, 0.0, (n + 0.5) * DT, DT, tgrid="boost", want_final=True)
            assert gl.last_grid() == B
            assert np.array_equal(a["mags"], b["mags"]), tag
            assert np.array_equal(a["energy"], b["energy"]), tag
            assert np.array_equal(a["final"], b["final"]), tag


def test_default_shape_balances_simds(oracle):
    """Without a tuned entry the launcher takes the shape whose busiest SIMD
    carries the fewest 64-site units: for one 1152-site trajectory the
    atom-major 640 x 2 shape (9 units per sublattice dealt 2, 2, 2, 2, 1 over
    its 5 waves: SIMD shares 5, 5, 4, 4; the exact tilings carry 6), and its
    records equal the oracle's."""
    uc, lat = _lattice("kitaev24_h")
    s = random_state(lat, 56)
    T1 = (N_STEPS + 0.5) * DT
    o = oracle.OracleLattice(uc, lat.dim1, lat.dim2, 1)
    ref = o.integrate_md(s.reshape(-1), 0.0, T1, DT)
    with GPULattice(lat) as gl:
        gl.set_spins(s)
        print("[shape] kitaev24_h default plain persistent", flush=True)
        res = gl.integrate(0.0, T1, DT, 1, tgrid="boost", engine=2)
        got = gl.last_launch_config()
    assert (got["block"], got["sites_per_thread"], got["atom_major"]) == (640, 2, True), got
    assert np.max(np.abs(res["mags"] - ref["mags"])) <= TOL
    assert np.max(np.abs(res["energy"] - ref["energy"])) <= TOL
