"""The batched stage-kernel launch sequences replayed from a hipGraph
(csl_set_stage_graph, csl_mixed_set_stage_graph) against the same sequences
launched eagerly: every replay bit-identical to the eager run on the same
inputs, with eager work and host synchronisation between the replays and
new inputs (initial states, pulse times) refreshed in place before them."""
import numpy as np
import pytest

from classical_mcmc_spin_amd import GPULattice
from classical_mcmc_spin_amd import gpu_mixed as GM
from classical_mcmc_spin_amd import mixed_model as mm
from helpers import kitaev, random_state, unit_x_field
from test_mixed_oracle import random_flat

pytestmark = pytest.mark.gpu


def test_plain_stage_graph_replays_bit_identical():
    uc, lat = kitaev(8, 8, h=0.1)
    f = unit_x_field(lat)
    fd = np.einsum("aij,aj->ai", lat.frames, f).reshape(-1)
    B, n, dt = 6, 120, 0.01
    T1 = (n + 0.5) * dt
    rng = np.random.default_rng(4)
    with GPULattice(lat) as gl:
        gl.set_pulse(np.concatenate([fd, fd]), 0.5, 0.4, 0.0, 0.0, 0.0)
        for rep in range(3):
            print(f"[graph] plain replay {rep}", flush=True)
            init = np.stack([random_state(lat, 10 * rep + b) for b in range(B)])
            tp = rng.uniform(0.0, 0.8, size=(B, 2))
            am = np.full((B, 2), 0.5)
            gl.set_stage_graph(False)
            eager = gl.integrate_batched(init, tp, am, 0.0, T1, dt, tgrid="boost", block=-1, want_final=True)
            gl.set_stage_graph(True)
            graph = gl.integrate_batched(init, tp, am, 0.0, T1, dt, tgrid="boost", block=-1, want_final=True)
            assert gl.last_launch_config()["block"] == -1
            for k in ("mags", "energy", "final"):
                assert np.array_equal(graph[k], eager[k]), (rep, k)
            # eager work between replays
            gl.set_stage_graph(False)
            gl.integrate_batched(init[:2], tp[:2], am[:2], 0.0, 0.205, dt, tgrid="boost", block=-1)


def test_mixed_stage_graph_replays_bit_identical():
    muc = mm.build_tmfeo3_mixed(seed=7)
    ml = mm.MixedLattice(muc, 2, 2, 2)
    f2 = np.zeros((4, 3))
    f2[:, 0] = 1.0
    f3 = np.zeros((4, 8))
    f3[:, 1] = 1.0
    B, n, dt = 4, 100, 0.005
    T0 = -0.2
    T1 = T0 + (n + 0.5) * dt
    rng = np.random.default_rng(9)
    with GM.GPUMixedLattice(ml) as gm:
        GM._set_pulses(gm, f2, f2, f3, f3, 0.0, 0.0, 0.5, 0.3, 0.0, 0.4, 0.3, 1.0)
        gm.set_persistent(-1)  # the eager run issues the same stage-kernel sequence
        for rep in range(3):
            print(f"[graph] mixed replay {rep}", flush=True)
            inits = np.stack([random_flat(ml, 100 * rep + b) for b in range(B)])
            tp = rng.uniform(-0.1, 0.2, size=(B, 2))
            sc = np.ones((B, 2))
            gm.set_stage_graph(False)
            eager = gm.integrate_batched(inits, tp, sc, T0, T1, dt, want_final=True)
            gm.set_stage_graph(True)
            graph = gm.integrate_batched(inits, tp, sc, T0, T1, dt, want_final=True)
            assert np.array_equal(graph["records"], eager["records"]), rep
            assert np.array_equal(graph["final"], eager["final"]), rep
            gm.set_stage_graph(False)
            gm.integrate_batched(inits[:1], tp[:1], sc[:1], T0, T0 + 0.0525, dt)
