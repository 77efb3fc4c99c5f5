"""Mixed SU(2) + SU(3) oracle and model tables (CPU).

Pins of the mixed oracle (none retypes its formulas):
  * tables: the package's vectorised MixedLattice expansion and the oracle's
    site-by-site expansion (build_sublattice + build_mixed_interactions
    order) agree bit for bit; every SU(2)->SU(3) bond has its transposed
    SU(3)->SU(2) twin;
  * local fields: H = dE/dx by central finite differences of the oracle
    energy, for both species, with every term switched on (fields, onsite,
    SU(2)-SU(2), SU(3)-SU(3) and mixed bonds);
  * SU(3) right-hand side: f_abc H_b lambda_c with f_abc computed here from
    the Gell-Mann matrices, f_abc = -(i/4) Tr(l_a [l_b, l_c]);
  * Gilbert damping: the mixed form (S (S.H) - H |S|^2) alpha/|S| equals the
    single-species oracle's S x (H x S) form on the same SU(2) lattice;
  * drive: a pulse enters exactly as the field -f1 d1 - f2 d2;
  * conservation: S.dS/dt = 0, lambda.dlambda/dt = 0 and sum H.dx/dt = 0
    without damping and drive; RK4 runs keep E and |lambda| to RK4 accuracy;
  * records: magnetisations recomputed here with numpy from the state.
"""
import numpy as np
import pytest

from classical_mcmc_spin_amd import mixed_model as mm
from classical_mcmc_spin_amd import model, seeding

RNG = np.random.default_rng(77)


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def rich_cell(seed=3):
    """TmFeO3-like cell with every term of the mixed Hamiltonian switched on."""
    rng = np.random.default_rng(seed)
    muc = mm.build_tmfeo3_mixed(field_strength=0.3, field_direction=(0.2, -0.5, 1.0), seed=seed)
    fe, tm = muc.SU2_cell, muc.SU3_cell
    for a in range(4):
        A = 0.05 * rng.normal(size=(8, 8))
        tm.set_onsite_interaction(A + A.T, a)
        tm.set_field(tm.field[a] + 0.1 * rng.normal(size=8), a)
    Jt = 0.05 * rng.normal(size=(8, 8))
    for o in ((0, 0, 0), (0, 1, 0)):
        tm.set_bilinear_interaction(Jt, 0, 2, o)
        tm.set_bilinear_interaction(Jt.T, 1, 3, o)
    R = model.kitaev_frame()
    for a in range(4):
        fe.set_sublattice_frame(R if a % 2 else np.eye(3), a)
    return muc


def random_flat(ml, seed):
    s2 = seeding.random_spins(ml.N2, seed)
    s3 = mm.random_su3_states(ml.N3, seed + 1)
    return ml.flat_state(s2, s3)


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 2)])
def test_tables_match_oracle_expansion(po, dims):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, *dims)
    o = po.MixedOracle(muc, *dims)
    t = o.tables()
    ref = {"22": (ml.su2.bi_offset, ml.su2.bi_partner, ml.su2.bi_J),
           "33": (ml.su3.bi_offset, ml.su3.bi_partner, ml.su3.bi_J),
           "23": (ml.m2_offset, ml.m2_partner, ml.m2_J),
           "32": (ml.m3_offset, ml.m3_partner, ml.m3_J)}
    for k, (off, par, J) in ref.items():
        assert np.array_equal(t[k][0], off), k
        assert np.array_equal(t[k][1], par), k
        assert np.array_equal(t[k][2], J), k
    # every mixed bond i->q (3x8) has the transposed twin q->i (8x3)
    fwd = {}
    for i in range(ml.N2):
        for k in range(int(ml.m2_offset[i]), int(ml.m2_offset[i + 1])):
            fwd.setdefault((i, int(ml.m2_partner[k])), []).append(ml.m2_J[k])
    n = 0
    for q in range(ml.N3):
        for k in range(int(ml.m3_offset[q]), int(ml.m3_offset[q + 1])):
            i = int(ml.m3_partner[k])
            assert any(np.array_equal(ml.m3_J[k], J.T) for J in fwd[(i, q)])
            n += 1
    assert n == int(ml.m2_offset[-1]) == 32 * np.prod(dims)


def test_local_field_is_energy_gradient(po):
    muc = rich_cell()
    o = po.MixedOracle(muc, 2, 2, 2)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    x = random_flat(ml, 5)
    h = 1e-6
    for species, sites in ((2, (0, 5, 17, 31)), (3, (0, 6, 19, 30))):
        D = 3 if species == 2 else 8
        base = 0 if species == 2 else ml.N2 * 3
        for s in sites:
            H = o.local_field(x, species, s)
            for d in range(D):
                xp, xm = x.copy(), x.copy()
                xp[base + s * D + d] += h
                xm[base + s * D + d] -= h
                g = (o.energy(xp) - o.energy(xm)) / (2 * h)
                assert abs(g - H[d]) < 1e-7 * max(1.0, abs(H[d])), (species, s, d)


def _fabc():
    lm = mm.gell_mann()
    f = np.zeros((8, 8, 8))
    for a in range(8):
        for b in range(8):
            for c in range(8):
                comm = lm[b] @ lm[c] - lm[c] @ lm[b]
                f[a, b, c] = np.real(-0.25j * np.trace(lm[a] @ comm))
    return f


def test_su3_rhs_and_bloch_damping(po):
    muc = rich_cell()
    o = po.MixedOracle(muc, 2, 2, 2)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    x = random_flat(ml, 9)
    gamma = np.array([0.0, 0.1, 0.0, 0.2, 0.05, 0.0, 0.0, 0.3])
    eq = RNG.normal(size=(ml.N3, 8))
    o.set_damping(0.0, gamma, eq)
    d = o.rhs(x)[ml.N2 * 3:].reshape(ml.N3, 8)
    f = _fabc()
    lam = x[ml.N2 * 3:].reshape(ml.N3, 8)
    for s in range(ml.N3):
        H = o.local_field(x, 3, s)
        ref = np.einsum("abc,b,c->a", f, H, lam[s]) - gamma * (lam[s] - eq[s])
        assert np.max(np.abs(d[s] - ref)) < 1e-13


def test_gilbert_form_matches_single_species(po):
    muc = rich_cell()
    fe_only = mm.MixedUnitCell(muc.SU2_cell, muc.SU3_cell)   # no mixed bonds
    o = po.MixedOracle(fe_only, 2, 2, 2)
    single = po.OracleLattice(muc.SU2_cell, 2, 2, 2)
    ml = mm.MixedLattice(fe_only, 2, 2, 2)
    x = random_flat(ml, 11)
    o.set_damping(0.07)
    single.set_alpha(0.07)
    d_mixed = o.rhs(x)[:ml.N2 * 3]
    d_single = single.rhs(x[:ml.N2 * 3])
    assert np.max(np.abs(d_mixed - d_single)) < 1e-14


def test_drive_enters_as_field(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    x = random_flat(ml, 13)
    o = po.MixedOracle(muc, 2, 2, 2)
    d2a, d2b = RNG.normal(size=(4, 3)), RNG.normal(size=(4, 3))
    d3a, d3b = RNG.normal(size=(4, 8)), RNG.normal(size=(4, 8))
    amp, w, fr, t = 0.4, 0.7, 1.3, 0.35
    o.set_pulse(2, d2a, 0.1, d2b, 0.9, amp, w, fr)
    o.set_pulse(3, d3a, 0.1, d3b, 0.9, amp, w, fr)
    d = o.rhs(x, t)

    def env(tc):
        return amp * np.exp(-((t - tc) / (2 * w)) ** 2) * np.cos(fr * (t - tc))
    f1, f2 = env(0.1), env(0.9)
    # same lattice with the drive folded into the fields (frames: F^T B)
    c2, c3 = muc.SU2_cell, muc.SU3_cell
    fe = model.UnitCell(3, 4, c2.lattice_pos, c2.lattice_vectors)
    tm = model.UnitCell(8, 4, c3.lattice_pos, c3.lattice_vectors)
    fe.bilinear, tm.bilinear = c2.bilinear, c3.bilinear
    fe.onsite_interaction, tm.onsite_interaction = c2.onsite_interaction, c3.onsite_interaction
    fe.sublattice_frames, tm.sublattice_frames = c2.sublattice_frames, c3.sublattice_frames
    for a in range(4):
        Fa = c2.sublattice_frames[a]
        fe.set_field(c2.field[a] + Fa.T @ (f1 * d2a[a] + f2 * d2b[a]), a)
        tm.set_field(c3.field[a] + f1 * d3a[a] + f2 * d3b[a], a)
    shifted = mm.MixedUnitCell(fe, tm)
    shifted.bilinear_SU2_SU3 = muc.bilinear_SU2_SU3
    ref = po.MixedOracle(shifted, 2, 2, 2).rhs(x, t)
    assert np.max(np.abs(d - ref)) < 1e-12


def test_conservation_laws(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = po.MixedOracle(muc, 2, 2, 2)
    x = random_flat(ml, 17)
    d = o.rhs(x)
    s2, s3 = ml.split_state(x)
    d2, d3 = ml.split_state(d)
    assert np.max(np.abs(np.sum(s2 * d2, axis=1))) < 1e-13
    assert np.max(np.abs(np.sum(s3 * d3, axis=1))) < 1e-13
    H = np.concatenate([np.concatenate([o.local_field(x, 2, i) for i in range(ml.N2)]),
                        np.concatenate([o.local_field(x, 3, i) for i in range(ml.N3)])])
    assert abs(H @ d) < 1e-11
    r = o.integrate_md(x, 0.0, 200.5 * 0.005, 0.005)
    e = r["records"][:, 33]
    assert np.max(np.abs(e - e[0])) < 1e-9
    f2, f3 = ml.split_state(r["final"])
    assert np.max(np.abs(np.linalg.norm(f3, axis=1) - 1.0)) < 1e-9
    assert np.max(np.abs(np.linalg.norm(f2, axis=1) - 1.0)) < 1e-9


def test_records_recomputed(po):
    muc = rich_cell()
    ml = mm.MixedLattice(muc, 2, 2, 2)
    o = po.MixedOracle(muc, 2, 2, 2)
    x = random_flat(ml, 19)
    r = o.record(x)
    s2, s3 = ml.split_state(x)
    atom2 = np.arange(ml.N2) % 4
    F2 = ml.su2.frames[atom2]
    g2 = np.einsum("nij,nj->ni", F2, s2)
    assert np.allclose(r[0:3], (ml.su2.afm_signs[atom2][:, None] * g2).mean(0), atol=1e-15)
    assert np.allclose(r[3:6], s2.mean(0), atol=1e-15)
    assert np.allclose(r[6:9], g2.mean(0), atol=1e-15)
    sg = np.where(np.arange(ml.N3) % 2 == 0, 1.0, -1.0)
    assert np.allclose(r[9:17], (sg[:, None] * s3).mean(0), atol=1e-15)
    assert np.allclose(r[17:25], s3.mean(0), atol=1e-15)
    assert np.allclose(r[25:33], s3.mean(0), atol=1e-15)     # identity SU(3) frames
    assert abs(r[33] * (ml.N2 + ml.N3) - o.energy(x)) < 1e-12


def test_random_su3_states():
    v = mm.random_su3_states(50, 4)
    assert v.shape == (50, 8)
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0)
    assert np.array_equal(v, mm.random_su3_states(50, 4))


def test_mixed_handle_marshalling_without_device():
    """The host wrapper marshals every table into csl_mixed_create; without a
    GPU the library reports a HIP error (never a Python-side failure)."""
    from classical_mcmc_spin_amd import _capi
    from classical_mcmc_spin_amd.gpu_mixed import GPUMixedLattice
    try:
        has_gpu = _capi.device_count() > 0
    except _capi.CslError:
        has_gpu = False
    if has_gpu:
        pytest.skip("a GPU is present")
    ml = mm.MixedLattice(rich_cell(), 2, 2, 2)
    with pytest.raises(_capi.CslError):
        GPUMixedLattice(ml)
