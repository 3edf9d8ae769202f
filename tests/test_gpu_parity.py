"""HIP path vs the FP32 oracle, through the C-ABI.

Integer/index outputs (cell keys, sort order, cell offsets, neighbour
tables) and FP32 state are compared bit for bit: the kernels perform the
oracle's IEEE operations in the oracle's order.
"""
import numpy as np
import pytest

from particlesolver_amd import System, scenes
from oracle import Oracle, load_scene

pytestmark = pytest.mark.gpu


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_bitexact(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    if a.dtype.kind == "f" or b.dtype.kind == "f":
        ok = np.array_equal(bits(a), bits(b))
        if not ok:
            bad = np.argwhere(bits(a) != bits(b))
            i = tuple(bad[0])
            raise AssertionError("%s differs at %d places, first %s: gpu %r oracle %r (max |d| %g)"
                                 % (what, len(bad), i, a[i], b[i], float(np.abs(a - b).max())))
    else:
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            i = tuple(bad[0])
            raise AssertionError("%s differs at %d places, first %s: gpu %r oracle %r"
                                 % (what, len(bad), i, a[i], b[i]))


def grid_params(n, box=(4.0, 3.0, 5.0), h=0.25, kmax=96):
    return scenes.pbf_params((0.0, 0.0, 0.0), box, h, h / 4.0, 2, n, kmax)


def random_positions(n, box, seed, edge_frac=0.3, cell=0.25):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.3, 1.0, size=(n, 3)) * np.asarray(box)[None, :]
    # put a fraction exactly on cell boundaries (the floor() edge cases)
    m = rng.random(n) < edge_frac
    k = rng.integers(0, 12, size=(n, 3)).astype(np.float64)
    pos[m] = k[m] * cell
    return pos.astype(np.float32)


@pytest.mark.parametrize("seed,n", [(1, 1), (2, 7), (3, 1000), (4, 20000)])
def test_grid_and_neighbors_bitexact(gpu_available, seed, n):
    box = (4.0, 3.0, 5.0)
    p = grid_params(n, box)
    pos = random_positions(n, box, seed)
    w = np.ones(n, np.float32)
    ph = np.zeros(n, np.int32)
    r0 = np.full(n, 100.0, np.float32)
    ref = Oracle(p, "f32")
    ref.add_particles(pos, None, w, ph, r0)
    with System(p) as g:
        g.add_particles(pos, None, w, ph, r0)
        for sysx in (g, ref):
            sysx.predict(0.0)
            sysx.build_grid()
            sysx.find_neighbors()
        gi, gc, go = g.grid()
        ri, rc, ro = ref.grid()
        assert_bitexact(gi, ri, "sorted insertion index")
        assert_bitexact(gc, rc, "cell key")
        assert_bitexact(go, ro, "cell offsets")
        gn, gt = g.neighbors()
        rn, rt = ref.neighbors()
        assert_bitexact(gn, rn, "neighbour counts")
        assert_bitexact(gt, rt, "neighbour table")
        assert g.status().max_neighbors == ref.status().max_neighbors
        assert g.status().flags == ref.status().flags


@pytest.mark.parametrize("kmax", [4, 8, 12])
def test_neighbor_overflow_and_partial_chunks(gpu_available, kmax):
    n = 3000
    box = (2.0, 2.0, 2.0)
    p = grid_params(n, box, kmax=kmax)
    pos = random_positions(n, box, 11, edge_frac=0.1)
    args = (pos, None, np.ones(n, np.float32), np.zeros(n, np.int32), np.full(n, 100.0, np.float32))
    ref = Oracle(p, "f32")
    ref.add_particles(*args)
    with System(p) as g:
        g.add_particles(*args)
        for sysx in (g, ref):
            sysx.predict(0.0)
            sysx.build_grid()
            sysx.find_neighbors()
        gn, gt = g.neighbors()
        rn, rt = ref.neighbors()
        assert_bitexact(gn, rn, "neighbour counts")
        assert_bitexact(gt, rt, "neighbour table")
        gs, rs = g.status(), ref.status()
        assert gs.max_neighbors == rs.max_neighbors > kmax
        assert gs.flags == rs.flags and (gs.flags & 1)
        # the truncated tables still drive identical iterations
        for sysx in (g, ref):
            sysx.iterate()
        assert_bitexact(g.sorted_state()[0], ref.sorted_state()[0], "iterate after overflow")


def test_empty_system(gpu_available):
    p = grid_params(16)
    with System(p) as g:
        g.step(1.0 / 60.0, 2)
        assert g.n == 0
        assert g.positions().shape == (0, 3)


def staged_compare(sc, substeps):
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for s in range(substeps):
            g.begin(sc.dt)
            ref.begin(sc.dt)
            assert_bitexact(g.grid()[0], ref.grid()[0], "substep %d sort order" % s)
            assert_bitexact(g.neighbors()[1], ref.neighbors()[1], "substep %d neighbours" % s)
            for it in range(sc.params.iterations):
                g.iterate()
                ref.iterate()
                assert_bitexact(g.sorted_state()[0], ref.sorted_state()[0],
                                "substep %d iteration %d iterate" % (s, it))
            g.end()
            ref.end()
            assert_bitexact(g.positions(), ref.positions(), "substep %d positions" % s)
            assert_bitexact(g.velocities(), ref.velocities(), "substep %d velocities" % s)


def test_dam_break_stages_bitexact(gpu_available):
    staged_compare(scenes.dam_break(12, 8, 10, max_neighbors=64), 3)


def test_dam_break_trajectory_bitexact(gpu_available):
    sc = scenes.dam_break(24, 12, 20, max_neighbors=96)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for block in range(4):
            g.step(sc.dt, 10)
            ref.step(sc.dt, 10)
            assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % (10 * block + 10))
            assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d substeps" % (10 * block + 10))
        assert g.status().flags == ref.status().flags


def test_determinism(gpu_available):
    sc = scenes.dam_break(16, 8, 16, max_neighbors=64)
    out = []
    for _ in range(2):
        with System(sc.params) as g:
            sc.load_into(g)
            g.step(sc.dt, 8)
            out.append((g.positions(), g.velocities()))
    assert_bitexact(out[0][0], out[1][0], "positions run-to-run")
    assert_bitexact(out[0][1], out[1][1], "velocities run-to-run")


def test_rebuild_every_iteration_mode(gpu_available):
    from particlesolver_amd import capi
    sc = scenes.dam_break(10, 6, 8, max_neighbors=64)
    sc.params.flags = capi.PS_FLAG_REBUILD_EVERY_ITERATION
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 3)
        ref.step(sc.dt, 3)
        assert_bitexact(g.positions(), ref.positions(), "positions")
        assert_bitexact(g.velocities(), ref.velocities(), "velocities")
