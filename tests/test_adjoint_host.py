"""Host side of the adjoint path (no GPU): the parameter tangents are the
exact derivatives of the tables, the oracle tangent-linear model matches
central differences of the oracle trajectory, and the C ABI rejects bad
adjoint requests before touching the device."""
import numpy as np
import pytest

from classical_mcmc_spin_amd import adjoint as AD
from classical_mcmc_spin_amd import model, seeding


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def test_kitaev_tangents_are_linear_derivatives():
    theta = np.array([-1.1, 0.27, -0.03, 0.02, 0.07])
    lat = AD.kitaev_lattice(theta, 4, 3)
    dJ, dB = AD.kitaev_tangents(4, 3)
    assert dJ.shape == (5,) + lat.bi_J.shape and dB.shape == (5,) + lat.field.shape
    assert np.allclose(np.einsum("p,pkij->kij", theta, dJ), lat.bi_J, atol=1e-15)
    assert np.allclose(np.einsum("p,pnd->nd", theta, dB), lat.field, atol=1e-15)
    # the same sparsity as every parameter start
    lat2 = AD.kitaev_lattice(theta * 0.9, 4, 3)
    assert np.array_equal(lat.bi_partner, lat2.bi_partner)


def test_fit_tables_equal_the_builders():
    """KitaevFit forms the per-start tables as theta . (dJ, dB); for the
    Kitaev parametrisation that is bit for bit what the builder gives."""
    dJ, dB = AD.kitaev_tangents(6, 5)
    rng = np.random.default_rng(12)
    star = np.array([-1.0, 0.25, -0.02, 0.0, 0.05])
    for theta in star * (1.0 + 0.2 * rng.uniform(-1, 1, size=(40, 5))):
        lat = AD.kitaev_lattice(theta, 6, 5)
        assert np.array_equal(np.tensordot(theta, dJ, 1), lat.bi_J)
        assert np.array_equal(np.tensordot(theta, dB, 1), lat.field)


def test_oracle_tangent_matches_finite_differences(po):
    """dm/dK of the damped, driven trajectory from the oracle tangent-linear
    model vs central differences of the oracle trajectory (step 1e-6)."""
    L1, L2, n, dt = 3, 3, 120, 0.01
    lat = model.Lattice(model.build_kitaev_honeycomb(), L1, L2, 1)
    x0 = seeding.random_spins(lat.lattice_size, 3).reshape(-1)
    f = np.zeros((2, 3))
    f[:, 0] = 1.0

    def oracle(theta):
        uc = model.build_kitaev_honeycomb(K=theta[0], Gamma=theta[1], Gammap=theta[2], J=theta[3],
                                          field_strength=theta[4], field_direction=(1.0, 1.0, 1.0))
        o = po.OracleLattice(uc, L1, L2, 1)
        o.set_alpha(0.02)
        o.set_pulse(f, 0.3, f * 0.0, 0.0, 0.5, 0.5, 0.0)
        return o
    theta = np.array([-1.0, 0.25, -0.02, 0.0, 0.05])
    e = np.zeros(5)
    e[0] = 1.0
    uc = model.build_kitaev_honeycomb(K=1.0, Gamma=0.0, Gammap=0.0)
    dJ = np.array([b.J for b in uc.bilinear]).reshape(-1)
    _, dm = oracle(theta).tangent_md(x0, np.zeros_like(x0), dJ, np.zeros(6), 0.0, (n + 0.5) * dt, dt)
    h = 1e-6
    mp = oracle(theta + h * e).integrate_md(x0, 0.0, (n + 0.5) * dt, dt)["mags"]
    mm = oracle(theta - h * e).integrate_md(x0, 0.0, (n + 0.5) * dt, dt)["mags"]
    fd = (mp - mm) / (2 * h)
    assert np.max(np.abs(fd - dm)) <= 1e-7 * max(1.0, np.max(np.abs(dm)))


def test_adjoint_argument_errors():
    from classical_mcmc_spin_amd import _capi
    L = _capi.lib()
    rc = L.csl_adjoint_gradient(None, 1, None, 0, None, None, 0.0, 1.0, 0.01, 2, None, 0, None, 0, None,
                                None, 0, None, None, None, None)
    assert rc == _capi.CSL_ERR_NULL
