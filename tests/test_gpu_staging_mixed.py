"""LDS staging of scenes that are not a uniform fluid, and of batches.

The staged kernels read a unit's neighbour cells from LDS: lambda stages
(x, y, z, mj) with a particle of w = 0 at infinity, the solve (x, y, z,
lambda) the same way, and in a unit holding a contact-walking particle also
each slot's phase, so the counting walk runs from LDS (a partner with w = 0,
staged at infinity, is read from memory: the rope below lands on a static
slab) while the corrections gather through the unit's chunk table; XSPH
stages the neighbours as well, with each slot's mass where fluid masses
differ (the scene below has two).  Where contact-walking particles are at
most 1 in 32 (the rope scene below: 12 of 3,532) the solve is two launches,
the units without such a particle through the variant without contacts.  A
batch stages the units that lie in one scene.  None of this may change a
bit: every run below is compared with the FP32 oracle and with the same
system built with staging off (PS_FLAG_NO_LDS_STAGING), and checks that
units did stage."""
import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes
from oracle import load_scene
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu


def fluid_with_obstacles_and_rope(seed=0x5EED0C01):
    """A collapsing fluid block (two masses) under a slab of static
    particles (w = 0, SOLID, just above the fluid's top) with a RIGID rope
    pinned above the slab: fluid sums that leave out w = 0 neighbours,
    contacts with static and moving partners, distance constraints and a
    pin, all in staged units."""
    spacing = 0.1
    sc = scenes.dam_break(20, 10, 16, max_neighbors=96, seed=seed)
    p = sc.params.copy()
    nf = sc.n
    w = np.ones(nf, np.float32)
    w[nf // 2:] = 0.5                                  # two fluid masses
    gx = np.arange(0, 20) * spacing + 0.05
    gz = np.arange(0, 16) * spacing + 0.05
    zz, xx = np.meshgrid(gz, gx, indexing="ij")
    slab = np.stack([xx.ravel() + 0.4, np.full(xx.size, 1.1), zz.ravel() + 0.2], 1).astype(np.float32)
    rope = np.stack([0.5 + np.arange(12) * spacing, np.full(12, 1.6), np.full(12, 0.8)], 1).astype(np.float32)
    pos = np.concatenate([sc.pos, slab, rope]).astype(np.float32)
    n = pos.shape[0]
    inv = np.concatenate([w, np.zeros(slab.shape[0], np.float32), np.ones(12, np.float32)])
    ph = np.concatenate([np.zeros(nf, np.int32), np.full(slab.shape[0], capi.PHASE_SOLID, np.int32),
                         np.full(12, capi.PHASE_RIGID, np.int32)])
    r0 = np.concatenate([np.asarray(sc.rest_density, np.float32), np.ones(n - nf, np.float32)])
    p.max_particles = n
    p.mu_s, p.mu_k = 0.6, 0.4
    out = scenes.Scene(p, pos, inv, ph, r0, sc.dt, name="fluid_obstacles_rope")
    pairs = np.stack([np.arange(n - 12, n - 1), np.arange(n - 11, n)], 1).astype(np.uint32)
    out.distance = (pairs, np.full(11, spacing, np.float32), np.full(11, 0.8, np.float32))
    out.pins = (np.array([n - 12], np.uint32), rope[:1].copy())
    return out


def unstaged(params):
    p = params.copy()
    p.flags = p.flags | capi.PS_FLAG_NO_LDS_STAGING
    return p


def run_three(sc, blocks, per, descs=None):
    """the HIP path staged, the HIP path unstaged, and the oracle (single
    scenes) in lockstep; returns the staged system's last status"""
    ref = load_scene(sc, "f32") if descs is None else None
    with System(sc.params, descs) as g, System(unstaged(sc.params), descs) as u:
        sc.load_into(g)
        sc.load_into(u)
        for blk in range(blocks):
            g.step(sc.dt, per)
            u.step(sc.dt, per)
            assert_bitexact(g.positions(), u.positions(), "staged vs unstaged after %d substeps" % (per * (blk + 1)))
            assert_bitexact(g.velocities(), u.velocities(), "velocities, staged vs unstaged")
            if ref is not None:
                ref.step(sc.dt, per)
                assert_bitexact(g.positions(), ref.positions(), "positions after %d substeps" % (per * (blk + 1)))
                assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d substeps" % (per * (blk + 1)))
        st, su = g.status(), u.status()
        assert su.staged_units == 0
        assert st.flags == su.flags and st.max_neighbors == su.max_neighbors
        return st


def test_static_obstacles_rope_and_two_fluid_masses_staged(gpu_available):
    st = run_three(fluid_with_obstacles_and_rope(), 3, 8)
    assert st.staged_units > 0


def test_mixed_scene_staged(gpu_available):
    # fluid rows of 40 particles: 256-particle fluid units span two grid rows
    # and stage (the sparse cloth and cube units keep global slots)
    sc = scenes.mixed(cloth_n=24, cubes=(3, 1, 3), fluid=(40, 6, 8), box=(6.0, 8.0, 6.0))
    st = run_three(sc, 2, 10)
    assert st.staged_units > 0


def test_batch_stages_units_inside_one_scene(gpu_available):
    # scenes of 2,000-6,000 particles: most 256-particle units lie in one
    # scene, those across a boundary keep global slots
    b = scenes.ensemble(nscenes=6, nmin=2000, nmax=6000)
    st = run_three(b, 2, 6, descs=b.descs)
    assert 0 < st.staged_units < st.units
    # and every scene equals the oracle run on that scene alone
    with System(b.params, b.descs) as g:
        b.load_into(g)
        g.step(b.dt, 6)
        x = g.positions()
        first = 0
        for q, s in enumerate(b.scenes):
            ref = load_scene(b.plain(q), "f32")
            ref.step(b.dt, 6)
            assert_bitexact(x[first:first + s.n], ref.positions(), "scene %d" % q)
            first += s.n
