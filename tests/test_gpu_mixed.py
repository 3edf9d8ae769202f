"""Mixed-phase, 2D, sleeping and C++-facade runs on the GPU, checked against
the FP32 oracle bit for bit (contacts with friction and mass scaling,
distance constraints with per-family stiffness, pins, shape matching,
two-density fluid)."""
import os
import subprocess

import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes
from oracle import load_scene
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_both(sc, substeps, check_every=1, staged=False):
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        done = 0
        while done < substeps:
            if staged:
                g.begin(sc.dt)
                ref.begin(sc.dt)
                for it in range(sc.params.iterations):
                    g.iterate()
                    ref.iterate()
                    assert_bitexact(g.sorted_state()[0], ref.sorted_state()[0],
                                    "substep %d iteration %d" % (done, it))
                g.end()
                ref.end()
                k = 1
            else:
                k = min(check_every, substeps - done)
                g.step(sc.dt, k)
                ref.step(sc.dt, k)
            done += k
            assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % done)
            assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d substeps" % done)
        gs, rs = g.status(), ref.status()
        assert gs.flags == rs.flags and gs.max_neighbors == rs.max_neighbors
        return g.positions()


def small_mixed():
    return scenes.mixed(cloth_n=24, cubes=(3, 1, 3), fluid=(10, 6, 8), box=(6.0, 8.0, 6.0))


def test_mixed_stages_bitexact(gpu_available):
    run_both(small_mixed(), 3, staged=True)


def test_mixed_trajectory_bitexact(gpu_available):
    x = run_both(small_mixed(), 60, check_every=20)
    assert np.isfinite(x).all()


def test_mixed_full_size_bitexact(gpu_available):
    # 765,536 particles, 390,658 links, 3,125 bodies: 30 substeps bit-exact,
    # and the scene's neighbour table (256) holds every particle's candidates
    sc = scenes.mixed()
    assert sc.n == 765536 and sc.params.max_neighbors == 256
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for done in (10, 20, 30):
            g.step(sc.dt, 10)
            ref.step(sc.dt, 10)
            assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % done)
            assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d substeps" % done)
        gs, rs = g.status(), ref.status()
        assert gs.flags == rs.flags and gs.max_neighbors == rs.max_neighbors
        assert not gs.flags & capi.PS_STATUS_NEIGHBOR_OVERFLOW, gs.max_neighbors


def test_two_d_fluid_bitexact(gpu_available):
    sc = scenes.two_d_fluid()
    x = run_both(sc, 40, check_every=10)
    assert np.all(x[:, 2] == np.float32(0.5 * sc.params.h))   # z held fixed


def test_sleeping_flag_bitexact(gpu_available):
    # unjittered block away from the walls under weak gravity: some interior
    # particles move less than 1e-4 in a substep and sleep (v = 0, x kept)
    n = 8
    p = scenes.pbf_params((0.0, 0.0, 0.0), (3.0, 3.0, 3.0), 0.2, 0.05, 4, n ** 3, 64)
    p.gravity[:] = [0.0, -0.02, 0.0]
    p.flags = capi.PS_FLAG_SLEEPING
    pos = scenes.fluid_block(n, n, n, 0.1, (1.0, 1.0, 1.0), 0.0, 1)
    S = scenes.lattice_rest_density(0.2, 0.1)
    sc = scenes.Scene(p, pos, np.ones(n ** 3, np.float32), np.zeros(n ** 3, np.int32),
                      np.full(n ** 3, S, np.float32), 1.0 / 60.0)
    ref = load_scene(sc, "f32")
    slept = 0
    with System(sc.params) as g:
        sc.load_into(g)
        for s in range(6):
            g.step(sc.dt, 1)
            ref.step(sc.dt, 1)
            assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % (s + 1))
            v = g.velocities()
            assert_bitexact(v, ref.velocities(), "velocities after %d substeps" % (s + 1))
            slept += int((np.abs(v).sum(1) == 0).sum())
    assert slept > 0


def test_cpp_facade_demo_runs():
    exe = os.path.join(ROOT, "particlesolver_amd", "ps_demo")
    assert os.path.exists(exe), "build the demo with __graft_entry__.build()"
    out = subprocess.run([exe, "90"], capture_output=True, text=True, timeout=100)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = [l for l in out.stdout.splitlines() if l.startswith(("combo", "ropes_on_sphere"))]
    assert len(lines) == 2
    for l in lines:
        assert "bad 0" in l
