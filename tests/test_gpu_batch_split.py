"""Work split of the persistent kernel (csl_set_batch_split): a batch larger
than the co-resident workgroups (here 300 trajectories of the 36x36 Kitaev
lattice, whose 130 KB LDS state allows one workgroup per CU) is shared evenly
by the workgroups that fit, trajectories cut between two workgroups
continuing from a saved state.  The arithmetic is the same, so the records
and final states equal those of the one-workgroup-per-trajectory launch bit
for bit, and two trajectories are checked against the oracle."""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice, seeding
from helpers import kitaev, unit_x_field

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def po():
    import pyoracle
    pyoracle.lib()
    return pyoracle


def test_split_batch_equals_rounds_and_oracle(po):
    L = 36
    uc, lat = kitaev(L, L, h=0.1)
    g = seeding.random_spins(lat.lattice_size, 5)
    B, dt, n = 300, 0.01, 300
    T1 = (n + 0.5) * dt
    tp = np.zeros((B, 2))
    tp[:, 1] = 0.01 * np.arange(B)
    am = np.full((B, 2), 0.5)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.3, 1.0, 0.0, 0.0)
        split = gl.integrate_batched(g, tp, am, 0.0, T1, dt, tgrid="boost", want_final=True)
        grid = gl.last_grid()
        gl.set_batch_split(False)
        rounds = gl.integrate_batched(g, tp, am, 0.0, T1, dt, tgrid="boost", want_final=True)
        assert gl.last_grid() == B
    if grid >= B:
        pytest.skip(f"{grid} workgroups fit at once: no split for {B} trajectories")
    print(f"split over {grid} workgroups", flush=True)
    assert np.array_equal(split["mags"], rounds["mags"])
    assert np.array_equal(split["energy"], rounds["energy"])
    assert np.array_equal(split["final"], rounds["final"])
    o = po.OracleLattice(uc, L, L, 1)
    for b in (7, B - 1):
        o.set_pulse(f * am[b, 0], tp[b, 0], f * am[b, 1], tp[b, 1], 1.0, 0.3, 1.0)
        ref = o.integrate_md(g.reshape(-1), 0.0, T1, dt)
        assert np.max(np.abs(split["mags"][:, b] - ref["mags"])) <= 1e-10, b
    o.clear_pulse()
