"""Batched scenes and captured substep graphs on the GPU.

A batch must evolve every scene exactly as that scene alone would, so the
checks are bit-exact: batch vs per-scene systems vs the FP32 oracle, and
graph replays vs eager stepping."""
import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes
from oracle import load_scene
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu


def small_batch(nscenes=5, nmin=300, nmax=2500):
    return scenes.ensemble(nscenes=nscenes, nmin=nmin, nmax=nmax)


def split(x, b):
    out, first = [], 0
    for s in b.scenes:
        out.append(x[first:first + s.n])
        first += s.n
    return out


def test_batch_equals_each_scene_alone(gpu_available):
    b = small_batch()
    with System(b.params, b.descs) as g:
        b.load_into(g)
        assert g.lib.ps_num_scenes(g.h) == len(b.scenes)
        g.step(b.dt, 8)
        xb, vb = split(g.positions(), b), split(g.velocities(), b)
        stb = g.status()
    for q in range(len(b.scenes)):
        sc = b.plain(q)
        with System(sc.params) as p:
            sc.load_into(p)
            p.step(sc.dt, 8)
            assert_bitexact(xb[q], p.positions(), "scene %d positions batch vs alone" % q)
            assert_bitexact(vb[q], p.velocities(), "scene %d velocities batch vs alone" % q)
        ref = load_scene(sc, "f32")
        ref.step(sc.dt, 8)
        assert_bitexact(xb[q], ref.positions(), "scene %d positions batch vs oracle" % q)
    assert stb.flags == 0


def test_graph_replays_equal_eager_stepping(gpu_available):
    b = small_batch(3)
    with System(b.params, b.descs) as ge, System(b.params, b.descs) as gg:
        b.load_into(ge)
        b.load_into(gg)
        for rep in range(3):
            ge.step(b.dt, 8)
            gg.step_graph(b.dt, 8)
            # eager work and a synchronisation between replays
            xe, xg = ge.positions(), gg.positions()
            assert_bitexact(xg, xe, "graph replay %d positions" % rep)
            assert_bitexact(gg.velocities(), ge.velocities(), "graph replay %d velocities" % rep)
        assert gg.status().iterations == ge.status().iterations == 3 * 8 * b.params.iterations
        # eager steps after replays continue the same trajectory
        ge.step(b.dt, 2)
        gg.step(b.dt, 2)
        assert_bitexact(gg.positions(), ge.positions(), "eager after graph")


def test_graph_single_scene_dam_break(gpu_available):
    sc = scenes.dam_break(16, 8, 16, max_neighbors=64)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        for rep in range(2):
            g.step_graph(sc.dt, 4)
            ref.step(sc.dt, 4)
            assert_bitexact(g.positions(), ref.positions(), "graph replay %d" % rep)


def test_graph_rejects_partial_rotation(gpu_available):
    sc = scenes.dam_break(6, 4, 6, max_neighbors=64)
    with System(sc.params) as g:
        sc.load_into(g)
        with pytest.raises(capi.PsError) as e:
            g.step_graph(sc.dt, 3)
        assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT
        g.step(sc.dt, 1)                # the handle is still usable
        assert np.isfinite(g.positions()).all()


def test_batch_rejects_out_of_order_scenes(gpu_available):
    b = small_batch(2)
    with System(b.params, b.descs) as g:
        s1 = b.scenes[1]
        g.add_particles(s1.pos, None, s1.inv_mass, s1.phase, s1.rest_density, scene=1)
        s0 = b.scenes[0]
        with pytest.raises(capi.PsError):
            g.add_particles(s0.pos, None, s0.inv_mass, s0.phase, s0.rest_density, scene=0)
        with pytest.raises(capi.PsError):
            g.add_particles(s0.pos, None, s0.inv_mass, s0.phase, s0.rest_density, scene=7)


TUNABLE = ("predict_hash", "reorder", "neighbors", "lambda", "solve", "finalize")


def test_autotune_restores_state_and_block_size_never_changes_results(gpu_available):
    sc = scenes.dam_break(16, 8, 16, max_neighbors=64)
    ref = load_scene(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 2)
        ref.step(sc.dt, 2)
        x0, v0, it0 = g.positions(), g.velocities(), g.status().iterations
        blocks = g.autotune(sc.dt, 1)
        assert all(blocks[k] in (64, 128, 256, 512) for k in TUNABLE)
        assert_bitexact(g.positions(), x0, "positions restored after tuning")
        assert_bitexact(g.velocities(), v0, "velocities restored after tuning")
        assert g.status().iterations == it0
        for bs in (blocks, {k: 64 for k in TUNABLE}, {k: 512 for k in TUNABLE}):
            g.set_tuning(bs)
            g.step(sc.dt, 2)
            ref.step(sc.dt, 2)
            assert_bitexact(g.positions(), ref.positions(), "positions with blocks %s" % bs)
            assert_bitexact(g.velocities(), ref.velocities(), "velocities with blocks %s" % bs)
            # the tuner's staging choice: every unit staged (as far as the
            # slots allow) or none
            st = g.status()
            assert st.units > 0 and (st.staged_units == 0 or st.staged_units > st.units // 2), \
                (st.units, st.staged_units)
        with pytest.raises(capi.PsError):
            g.set_tuning({k: 96 for k in TUNABLE})
