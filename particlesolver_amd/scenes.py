"""Synthetic scene generators for the benchmark and parity configurations.

All random draws come from SplitMix64 (Steele, Lea, Flood 2014) with
24-bit uniforms: u_k = (mix64(seed + (k+1) * golden) >> 40) / 2^24, formed
in FP64 and rounded once to FP32 where they become positions.  Particle
order follows the reference's fluid builder, x fastest, then y, then z
(gpu/src/particlesystem.cpp:374-392).
"""
import math

import numpy as np

from . import capi

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, count, offset=0):
    """count 64-bit outputs of SplitMix64 starting at draw `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform24(seed, count, offset=0):
    """FP64 uniforms in [0, 1) with 24 random bits."""
    return (splitmix64(seed, count, offset) >> np.uint64(40)).astype(np.float64) / 16777216.0


def poly6_coef(h):
    return 315.0 / (64.0 * math.pi * h ** 9)


def lattice_rest_density(h, spacing, dims=3, mass=1.0):
    """FP64 SPH density at an interior point of an unjittered cubic (square)
    lattice: sum over lattice offsets with r < h of m W_poly6(r), self
    included."""
    R = int(math.ceil(h / spacing)) + 1
    rng = np.arange(-R, R + 1)
    if dims == 3:
        o = np.stack(np.meshgrid(rng, rng, rng, indexing="ij"), -1).reshape(-1, 3)
    else:
        o = np.stack(np.meshgrid(rng, rng, indexing="ij"), -1).reshape(-1, 2)
    r2 = (o.astype(np.float64) * spacing) ** 2
    r2 = r2.sum(1)
    r2 = r2[r2 < h * h]
    return float(mass * poly6_coef(h) * np.sum((h * h - r2) ** 3))


def fluid_block(nx, ny, nz, spacing, origin, jitter, seed):
    """Lattice block, x fastest; jitter uniform in [-jitter, jitter) per axis."""
    i = np.arange(nx, dtype=np.float64)
    j = np.arange(ny, dtype=np.float64)
    k = np.arange(nz, dtype=np.float64)
    z, y, x = np.meshgrid(k, j, i, indexing="ij")
    pos = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    pos = (pos + 0.5) * spacing + np.asarray(origin, dtype=np.float64)
    n = pos.shape[0]
    if jitter:
        u = uniform24(seed, 3 * n).reshape(n, 3)
        pos = pos + (2.0 * u - 1.0) * jitter
    return pos.astype(np.float32)


def pbf_params(box_min, box_max, h, radius, iterations, max_particles, max_neighbors=96,
               relaxation=600.0, xsph_c=0.01, seed=0x5EED0002, dims=3):
    """Parameters of a PBF fluid scene with cell = h and a grid covering the box."""
    p = capi.PsParams()
    p.abi_version = capi.PS_ABI_VERSION
    p.dims = dims
    p.radius = radius
    p.h = h
    p.cell = h
    p.collide_dist = radius * 2.001
    p.gravity[:] = [0.0, -9.8, 0.0]
    p.box_min[:] = list(box_min)
    p.box_max[:] = list(box_max)
    p.origin[:] = list(box_min)
    g = [max(1, int(math.ceil((box_max[a] - box_min[a]) / h - 1e-9))) for a in range(3)]
    if dims == 2:
        g[2] = 1
    p.grid[:] = g
    p.iterations = iterations
    p.relaxation = relaxation
    p.k_corr = 0.1
    p.n_corr = 4
    p.dq_corr = 0.2
    p.xsph_c = xsph_c
    p.s_solid = 1.0
    p.mu_s = 0.005
    p.mu_k = 0.0002
    p.wall_mu_s = 0.005
    p.wall_mu_k = 0.0002
    p.wall_jitter = 0.01
    p.shape_stiffness = 1.0
    p.max_particles = max_particles
    p.max_neighbors = max_neighbors
    p.seed = seed
    p.flags = 0
    return p


class Scene:
    """A scene: parameters plus the arrays ps_add_particles takes."""

    def __init__(self, params, pos, inv_mass, phase, rest_density, dt, vel=None, name=""):
        self.params = params
        self.pos = pos
        self.vel = vel
        self.inv_mass = inv_mass
        self.phase = phase
        self.rest_density = rest_density
        self.dt = dt
        self.name = name
        self.distance = None   # (pairs, rest, stiffness)
        self.pins = None       # (idx, xyz)
        self.bodies = []       # list of (idx, stiffness)

    @property
    def n(self):
        return self.pos.shape[0]

    def load_into(self, sys_):
        """Add this scene to a capi.System (or an oracle system with the same methods)."""
        sys_.add_particles(self.pos, self.vel, self.inv_mass, self.phase, self.rest_density)
        if self.distance is not None:
            sys_.add_distance_constraints(*self.distance)
        if self.pins is not None:
            sys_.add_point_constraints(*self.pins)
        for idx, st in self.bodies:
            sys_.add_rigid_body(idx, st)


def dam_break(nx=128, ny=64, nz=128, spacing=0.1, box_scale=(2.0, 2.0, 2.0), jitter=1e-3,
              h=None, iterations=4, max_neighbors=96, seed=0x5EED0002, dt=1.0 / 60.0):
    """3D single-phase PBF dam break: an nx*ny*nz lattice (spacing = diameter)
    in one corner of a box box_scale times the block.  With the defaults this
    is the 1,048,576-particle dam break: 128x64x128 at spacing 0.1 in a
    25.6 x 12.8 x 25.6 box, h = 0.2 = cell, k=0.1, n=4, dq=0.2h, XSPH 0.01,
    eps=600, 4 iterations, dt 1/60."""
    h = 2.0 * spacing if h is None else h
    radius = 0.5 * spacing
    ext = [nx * spacing * box_scale[0], ny * spacing * box_scale[1], nz * spacing * box_scale[2]]
    n = nx * ny * nz
    p = pbf_params((0.0, 0.0, 0.0), ext, h, radius, iterations, n, max_neighbors, seed=seed)
    pos = fluid_block(nx, ny, nz, spacing, (0.0, 0.0, 0.0), jitter, seed)
    rho0 = lattice_rest_density(h, spacing)
    s = Scene(p, pos, np.ones(n, np.float32), np.zeros(n, np.int32),
              np.full(n, np.float32(rho0), np.float32), dt, name="dam_break_%dx%dx%d" % (nx, ny, nz))
    s.rho0 = rho0
    return s
