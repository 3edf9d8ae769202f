"""LDS-staged neighbour cells against the FP32 oracle (uniform fluids; other
compositions and batches: test_gpu_staging_mixed.py).

Uniform single-scene fluids run their neighbour build, lambda, solve and
XSPH kernels from neighbour positions copied into LDS per 256-particle unit
(the table then holds LDS indices, which ps_read_neighbors maps back).  The
oracle knows nothing of this, so every comparison below checks that staging
changes no result bit, on scenes where the units do stage: grid rows of 256
particles (one row per unit) and of 248 (units across two rows, and a last
partial unit)."""
import numpy as np
import pytest

from particlesolver_amd import System, scenes
from oracle import load_scene
from test_gpu_parity import assert_bitexact, staged_compare

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dims", [(64, 16, 32), (62, 16, 30)])
def test_staged_tables_match_oracle(gpu_available, dims):
    sc = scenes.dam_break(*dims, max_neighbors=96)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for s in (g, ref):
            s.predict(sc.dt)
            s.build_grid()
            s.find_neighbors()
        st = g.status()
        assert st.units == (sc.n + 255) // 256
        assert st.staged_units == st.units          # every unit spans <= 3 rows here
        gn, gt = g.neighbors()
        rn, rt = ref.neighbors()
        assert_bitexact(gn, rn, "neighbour counts")
        assert_bitexact(gt, rt, "neighbour table (LDS indices mapped back)")


def test_staged_iterations_bitexact(gpu_available):
    staged_compare(scenes.dam_break(62, 16, 30, max_neighbors=96), 2)


def test_staged_trajectory_bitexact(gpu_available):
    sc = scenes.dam_break(64, 16, 32, max_neighbors=96)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for block in range(3):
            g.step(sc.dt, 10)
            ref.step(sc.dt, 10)
            assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % (10 * block + 10))
            assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d substeps" % (10 * block + 10))
        st = g.status()
        assert st.staged_units > 0 and st.flags == ref.status().flags


def test_staged_graph_replays(gpu_available):
    sc = scenes.dam_break(64, 16, 32, max_neighbors=96)
    with System(sc.params) as ge, System(sc.params) as gg:
        sc.load_into(ge)
        sc.load_into(gg)
        for rep in range(2):
            ge.step(sc.dt, 4)
            gg.step_graph(sc.dt, 4)
            assert_bitexact(gg.positions(), ge.positions(), "graph replay %d" % rep)
