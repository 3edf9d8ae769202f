"""The reference's ParticleSystem host class (gpu/src/particlesystem.h:22-118)
in Python, over the native library: the constructor, the scene builders
(gpu/src/particlesystem.cpp:357-698), the deferred single-particle and fluid
additions (:249-354) and update (:144-253).

It restates the C++ class of csrc/particle_system.hip one FP32 operation at a
time and draws the builders' jitter from the same SplitMix64 stream (seeded
by params.seed), so a scene built here and the same calls made on the C++
class give identical particle, pin and link arrays in identical order.

    ps = ParticleSystem(0.25, (64, 64, 64), 15000, (-50, 0, -50), (50, 200, 50), 5)
    ps.addRope((-17, 20, -17), (0, -0.5, 0.001), 0.4, 30, 1.0, True)
    ps.setParticleToAdd((0, 30, 0), (0, -10, 0), 2.0)
    for _ in range(60):
        ps.update(1 / 60)
    x = ps.readPositions()          # (n, 3) float32, insertion order

`backend` builds the system the calls go to from a ps_params; the default is
capi.System (the HIP path).  Any object with add_particles,
add_distance_constraints, add_point_constraints, step and n serves.
"""
import math

import numpy as np

from . import capi

f32 = np.float32
MASK64 = (1 << 64) - 1


def facade_params(radius, max_particles, min_bounds, max_bounds, iterations, h=2.0):
    """ps_params of ParticleSystem(radius, gridSize, maxParticles, minBounds,
    maxBounds, iterations): contacts at 2.001 r (integration_kernel.cuh:319),
    a dense grid of cell max(h, 2.001 r) over the bounds, 500 neighbours
    (MAX_FLUID_NEIGHBORS, integration_kernel.cuh:23)."""
    p = capi.default_params()
    p.radius = f32(radius)
    p.h = f32(h)
    p.collide_dist = f32(radius) * f32(2.001)
    p.cell = max(f32(p.h), f32(p.collide_dist))
    for a in range(3):
        p.box_min[a] = float(min_bounds[a])
        p.box_max[a] = float(max_bounds[a])
        p.origin[a] = float(min_bounds[a])
        ext = float(max_bounds[a]) - float(min_bounds[a])
        p.grid[a] = int(max(1.0, math.ceil(ext / float(f32(p.cell)))))
    p.iterations = max(1, int(iterations))
    p.max_particles = int(max_particles)
    p.max_neighbors = 500
    return p


def _count(lo, hi, step):
    """(int)((float)(int)ceil(hi - lo) / step) of the builders"""
    return int(f32(hi - lo) / f32(step))


class ParticleSystem:
    def __init__(self, particle_radius, grid_size, max_particles, min_bounds, max_bounds,
                 iterations, h=2.0, backend=None):
        """particlesystem.h:25; grid_size is accepted for the signature only:
        the dense grid follows the bounds and h."""
        params = facade_params(particle_radius, max_particles, min_bounds, max_bounds,
                               iterations, h)
        self._init(params, backend)
        self.min_bounds = tuple(int(v) for v in min_bounds)
        self.max_bounds = tuple(int(v) for v in max_bounds)

    @classmethod
    def from_params(cls, params, backend=None):
        self = cls.__new__(cls)
        self._init(params, backend)
        return self

    def _init(self, params, backend):
        self.params = params.copy()
        self.system = (backend or capi.System)(self.params)
        self.min_bounds = tuple(int(v) for v in self.params.box_min)
        self.max_bounds = tuple(int(v) for v in self.params.box_max)
        self.rigid_index = 0
        self.rng = int(self.params.seed) & MASK64
        self.particles_to_add = []   # (pos, vel, mass)
        self.fluids_to_add = []      # (pos, mass, density)

    def close(self):
        if hasattr(self.system, "close"):
            self.system.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- queries
    def getNumParticles(self):
        return int(self.system.n)

    def getParticleRadius(self):
        return float(self.params.radius)

    def getMinBounds(self):
        return self.min_bounds

    def getMaxBounds(self):
        return self.max_bounds

    def readPositions(self):
        return self.system.positions()

    def readVelocities(self):
        return self.system.velocities()

    # -- jitter
    def frand(self):
        """SplitMix64 (Steele, Lea, Flood 2014), top 24 bits -> [0, 1)"""
        self.rng = (self.rng + 0x9E3779B97F4A7C15) & MASK64
        z = self.rng
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        z ^= z >> 31
        return f32((z >> 40) / 16777216.0)

    def _jitter(self, amount):
        return (self.frand() * f32(2.0) - f32(1.0)) * amount

    # -- stepping
    def update(self, delta_time):
        """particlesystem.cpp:144-253: one substep of min(dt, 0.05), then the
        queued single particles (SOLID, rest density 1.5) and queued fluids
        (velocity (0, -1, 0))."""
        dt = min(f32(delta_time), f32(0.05))
        if self.getNumParticles() > 0 and dt > 0:
            self.system.step(dt, 1)
        queued, self.particles_to_add = self.particles_to_add, []
        for pos, vel, mass in queued:
            if self.getNumParticles() == self.params.max_particles:
                continue                                  # particlesystem.cpp:303-304
            self._add(np.array([pos], f32), np.array([vel], f32), np.array([f32(1.0) / mass], f32),
                      np.array([1.5], f32), np.array([capi.PHASE_SOLID], np.int32), single=True)
        if self.fluids_to_add:
            queued, self.fluids_to_add = self.fluids_to_add, []
            pos = np.array([q[0] for q in queued], f32)
            vel = np.tile(np.array([0.0, -1.0, 0.0], f32), (len(queued), 1))
            w = np.array([f32(1.0) / q[1] for q in queued], f32)
            ro = np.array([q[2] for q in queued], f32)
            self._add(pos, vel, w, ro, np.full(len(queued), capi.PHASE_FLUID, np.int32))

    def _add(self, pos, vel, w, ro, phase, single=False):
        """addParticleMultiple (particlesystem.cpp:333-347): a block that
        would reach max_particles is dropped whole, as there; False then."""
        n = pos.shape[0]
        if n == 0:
            return True
        if not single and self.getNumParticles() + n >= self.params.max_particles:
            return False
        self.system.add_particles(pos, vel, w, phase, ro)
        return True

    # -- builders
    def _lattice(self, ll, ur, distance, jitter):
        r = self.params
        cx = _count(ll[0], ur[0], distance)
        cy = _count(ll[1], ur[1], distance)
        cz = 1 if r.dims == 2 else _count(ll[2], ur[2], distance)
        pos = []
        for z in range(cz):
            for y in range(cy):
                for x in range(cx):
                    jx = self._jitter(jitter)
                    jy = self._jitter(jitter)
                    jz = self._jitter(jitter)
                    pos.append((f32(ll[0]) + f32(x) * distance + jx,
                                f32(ll[1]) + f32(y) * distance + jy,
                                f32(r.box_min[2]) if r.dims == 2 else f32(ll[2]) + f32(z) * distance + jz))
        return np.array(pos, f32).reshape(-1, 3)

    def addFluid(self, ll, ur, mass, density, color=(0.0, 0.0, 1.0)):
        """particlesystem.cpp:357-405: lattice spacing 2.5 r, jitter r * 0.01"""
        r = f32(self.params.radius)
        pos = self._lattice(ll, ur, r * f32(2.5), r * f32(0.01))
        n = pos.shape[0]
        self._add(pos, np.zeros((n, 3), f32), np.full(n, f32(1.0) / f32(mass), f32),
                  np.full(n, f32(density), f32), np.full(n, capi.PHASE_FLUID, np.int32))

    def addParticleGrid(self, ll, ur, mass, add_jitter):
        """particlesystem.cpp:408-458: spacing 2.002 r, SOLID phase"""
        r = f32(self.params.radius)
        pos = self._lattice(ll, ur, r * f32(2.002), r * f32(0.01) if add_jitter else f32(0.0))
        n = pos.shape[0]
        self._add(pos, np.zeros((n, 3), f32), np.full(n, f32(1.0) / f32(mass), f32),
                  np.ones(n, f32), np.full(n, capi.PHASE_SOLID, np.int32))

    def addHorizCloth(self, ll, ur, spacing, dist, mass, hold_edges):
        """particlesystem.cpp:461-568: a horizontal sheet at height spacing.y,
        links dist.x along x and dist.y along z; the x == 0 column is pinned,
        and with hold_edges the other three edges too (corners twice)."""
        start = self.getNumParticles()
        sx, sy, sz = (f32(v) for v in spacing)
        cx = _count(ll[0], ur[0], sx)
        cz = _count(ll[1], ur[1], sz)
        pos, pins, pin_xyz, pairs, rest = [], [], [], [], []
        for z in range(cz):
            for x in range(cx):
                p = (f32(ll[0]) + f32(x) * sx, sy, f32(ll[1]) + f32(z) * sz)
                pos.append(p)
                idx = start + z * cx + x
                if x > 0:
                    pairs.append((idx - 1, idx))
                    rest.append(dist[0])
                else:
                    pins.append(idx)
                    pin_xyz.append(p)
                if z > 0:
                    pairs.append((idx - cx, idx))
                    rest.append(dist[1])
                elif hold_edges:
                    pins.append(idx)
                    pin_xyz.append(p)
                if x == cx - 1 and hold_edges:
                    pins.append(idx)
                    pin_xyz.append(p)
                if z == cz - 1 and hold_edges:
                    pins.append(idx)
                    pin_xyz.append(p)
        n = len(pos)
        if self._add(np.array(pos, f32).reshape(-1, 3), np.zeros((n, 3), f32),
                     np.full(n, f32(1.0) / f32(mass), f32), np.ones(n, f32),
                     np.full(n, capi.PHASE_RIGID + self.rigid_index, np.int32)) and n:
            if pins:
                self.system.add_point_constraints(np.array(pins, np.uint32), np.array(pin_xyz, f32))
            if rest:
                self.system.add_distance_constraints(np.array(pairs, np.uint32), np.array(rest, f32))
        self.rigid_index += 1

    def addRope(self, start, spacing, dist, num_links, mass, constrain_start):
        """particlesystem.cpp:570-616: num_links + 1 particles at
        start + i * spacing, linked in order; the first pinned at start."""
        s0 = self.getNumParticles()
        st = [f32(v) for v in start]
        sp = [f32(v) for v in spacing]
        pos = [(st[0] + f32(i) * sp[0], st[1] + f32(i) * sp[1], st[2] + f32(i) * sp[2])
               for i in range(num_links + 1)]
        n = len(pos)
        if self._add(np.array(pos, f32), np.zeros((n, 3), f32), np.full(n, f32(1.0) / f32(mass), f32),
                     np.ones(n, f32), np.full(n, capi.PHASE_RIGID + self.rigid_index, np.int32)):
            if num_links > 0:
                pairs = np.stack([np.arange(s0, s0 + num_links), np.arange(s0 + 1, s0 + num_links + 1)], 1)
                self.system.add_distance_constraints(pairs.astype(np.uint32),
                                                     np.full(num_links, f32(dist), f32))
            if constrain_start:
                self.system.add_point_constraints(np.array([s0], np.uint32), np.array([st], f32))
        self.rigid_index += 1

    def addStaticSphere(self, ll, ur, spacing):
        """particlesystem.cpp:618-687: the lattice points strictly inside the
        inscribed sphere, inverse mass 0.01, each pinned where it starts."""
        s0 = self.getNumParticles()
        r = self.params
        spacing = f32(spacing)
        cx = _count(ll[0], ur[0], spacing)
        cy = _count(ll[1], ur[1], spacing)
        cz = 1 if r.dims == 2 else _count(ll[2], ur[2], spacing)
        radius = f32(ur[0] - ll[0]) * f32(0.5)
        c = [f32(ll[a]) + radius for a in range(3)]
        pos = []
        for z in range(cz):
            for y in range(cy):
                for x in range(cx):
                    p = (f32(ll[0]) + f32(x) * spacing, f32(ll[1]) + f32(y) * spacing,
                         f32(r.box_min[2]) if r.dims == 2 else f32(ll[2]) + f32(z) * spacing)
                    d = [p[a] - c[a] for a in range(3)]
                    if np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < radius:
                        pos.append(p)
        n = len(pos)
        pos = np.array(pos, f32).reshape(-1, 3)
        if self._add(pos, np.zeros((n, 3), f32), np.full(n, f32(0.01), f32), np.ones(n, f32),
                     np.full(n, capi.PHASE_RIGID + self.rigid_index, np.int32)) and n:
            self.system.add_point_constraints(np.arange(s0, s0 + n, dtype=np.uint32), pos)
        self.rigid_index += 1

    # -- deferred additions
    def setParticleToAdd(self, pos, vel, mass):
        """particlesystem.cpp:256-266: jitter r * 0.01 on x and y"""
        j = f32(self.params.radius) * f32(0.01)
        p = [f32(v) for v in pos]
        p[0] = p[0] + self._jitter(j)
        p[1] = p[1] + self._jitter(j)
        self.particles_to_add.append((tuple(p), tuple(f32(v) for v in vel), f32(mass)))

    def setFluidToAdd(self, pos, color, mass, density):
        """particlesystem.cpp:349-354"""
        self.fluids_to_add.append((tuple(f32(v) for v in pos), f32(mass), f32(density)))

    # -- constraints
    def makePointConstraint(self, index, point):
        """particlesystem.cpp:690-693"""
        self.system.add_point_constraints(np.array([index], np.uint32), np.array([point], f32))

    def makeDistanceConstraint(self, index, distance):
        """particlesystem.cpp:695-698"""
        self.system.add_distance_constraints(np.array([index], np.uint32), np.array([distance], f32))
