"""Edge cases of the substep on the GPU against the FP32 oracle: a grid
that covers only part of the box (particles on both sides live in the
clamped edge cells, whose rows then hold far more than 32 candidates), a
crowded grid whose rows run to several hundred candidates, and inputs the
C-ABI must refuse without disturbing the system."""
import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes
from oracle import load_scene
from test_gpu_parity import assert_bitexact, staged_compare

pytestmark = pytest.mark.gpu


def partial_grid_scene():
    # a 12x8x12 block in [0, 1.2] x [0, 0.8] x [0, 1.2] of a 2.4x1.6x2.4 box;
    # the grid's 4x3x4 cells of 0.2 cover x, z in [0.4, 1.2) and y in
    # [0.2, 0.8) only, so keys are clamped on both sides of every axis
    sc = scenes.dam_break(12, 8, 12, max_neighbors=192)
    sc.params.origin[:] = [0.4, 0.2, 0.4]
    sc.params.grid[:] = [4, 3, 4]
    return sc


def test_partial_grid_full_substeps_bitexact(gpu_available):
    sc = partial_grid_scene()
    staged_compare(sc, 3)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 6)
        ref.step(sc.dt, 6)
        assert_bitexact(g.positions(), ref.positions(), "positions after 6 substeps")
        assert_bitexact(g.velocities(), ref.velocities(), "velocities after 6 substeps")
        gs, rs = g.status(), ref.status()
        assert gs.flags == rs.flags
        assert gs.flags & capi.PS_STATUS_OUTSIDE_GRID
        assert gs.max_neighbors == rs.max_neighbors
        # the clamped edge cells are crowded: some stencil row spans several
        # 32-candidate windows of the scan
        occ = np.diff(g.grid()[2].astype(np.int64))
        assert occ.max() > 32


def test_crowded_cells_bitexact(gpu_available):
    # cells of 4 h with the search radius h: 27 cells of ~500 particles
    # each, so every stencil row runs to ~1500 candidates and most of them
    # fail the distance test
    sc = scenes.dam_break(16, 8, 16, max_neighbors=96)
    sc.params.cell = 4.0 * sc.params.h
    sc.params.grid[:] = [int(np.ceil((sc.params.box_max[a] - sc.params.box_min[a]) / sc.params.cell))
                         for a in range(3)]
    staged_compare(sc, 2)


@pytest.mark.parametrize("bad", [np.nan, np.inf, -np.inf])
def test_non_finite_input_refused_and_system_intact(gpu_available, bad):
    sc = scenes.dam_break(6, 4, 6, max_neighbors=64)
    sc.params.max_particles = sc.n + 8       # room for the refused particle
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        n = g.n
        pos = np.array([[0.3, 0.3, 0.3]], np.float32)
        one = (np.ones(1, np.float32), np.zeros(1, np.int32), np.full(1, 100.0, np.float32))
        for p, v in ((pos.copy(), None), (pos.copy(), np.zeros((1, 3), np.float32))):
            if v is None:
                p[0, 1] = bad
            else:
                v[0, 2] = bad
            with pytest.raises(capi.PsError) as e:
                g.add_particles(p, v, *one)
            assert e.value.code == capi.PS_ERR_NOT_FINITE
        assert g.n == n
        g.step(sc.dt, 2)
        ref.step(sc.dt, 2)
        assert_bitexact(g.positions(), ref.positions(), "positions after a refused add")
        assert g.status().flags & capi.PS_STATUS_NONFINITE == 0
