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

PHASE_RIGID0 = capi.PHASE_RIGID

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
        if self.bodies:
            off = np.zeros(len(self.bodies) + 1, np.uint32)
            off[1:] = np.cumsum([len(i) for i, _ in self.bodies])
            idx = np.concatenate([i for i, _ in self.bodies])
            st = np.array([s for _, s in self.bodies], np.float32)
            sys_.add_rigid_bodies(off, idx, st)


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


def cloth_constraints(nx, nz, first, spacing, stiffness=(1.0, 0.5, 0.1)):
    """Structural (x and z neighbours), shear (both diagonals) and bend
    (two apart) distance constraints of an nx*nz cloth, rest length = the
    initial distance, with the given per-family stiffness."""
    idx = np.arange(nx * nz, dtype=np.int64).reshape(nz, nx) + first
    fams = []
    a, b = idx[:, :-1].ravel(), idx[:, 1:].ravel()
    c, d = idx[:-1, :].ravel(), idx[1:, :].ravel()
    fams.append((np.concatenate([a, c]), np.concatenate([b, d]), spacing, stiffness[0]))
    e, f = idx[:-1, :-1].ravel(), idx[1:, 1:].ravel()
    g, hh = idx[1:, :-1].ravel(), idx[:-1, 1:].ravel()
    fams.append((np.concatenate([e, g]), np.concatenate([f, hh]), spacing * math.sqrt(2.0), stiffness[1]))
    i2, j2 = idx[:, :-2].ravel(), idx[:, 2:].ravel()
    k2, l2 = idx[:-2, :].ravel(), idx[2:, :].ravel()
    fams.append((np.concatenate([i2, k2]), np.concatenate([j2, l2]), 2.0 * spacing, stiffness[2]))
    pairs = np.concatenate([np.stack([p, q], 1) for p, q, _, _ in fams]).astype(np.uint32)
    rest = np.concatenate([np.full(len(p), r, np.float32) for p, _, r, _ in fams])
    stiff = np.concatenate([np.full(len(p), s, np.float32) for p, _, _, s in fams])
    return pairs, rest, stiff


def mixed(cloth_n=256, cubes=(25, 5, 25), fluid=(100, 50, 50), spacing=0.1, box=(30.0, 20.0, 30.0),
          iterations=4, max_neighbors=256, seed=0x5EED0003, dt=1.0 / 60.0):
    """3D mixed-phase scene: a cloth_n^2 cloth (structural/shear/bend links,
    stiffness 1.0/0.5/0.1, corners pinned) above prod(cubes) rigid 4x4x4
    cubes (shape matching, mu_s 0.6 / mu_k 0.4) dropped from a seeded
    lattice, above two fluid blocks of rest density 1.0 and 2.0 (in units of
    the lattice kernel sum: particle masses 1 and 2).  Defaults: 65,536 +
    200,000 + 500,000 = 765,536 particles."""
    rng_off = 0
    h = 2.0 * spacing
    radius = 0.5 * spacing
    S = lattice_rest_density(h, spacing)
    parts = []
    # two fluid blocks side by side on the floor
    fx, fy, fz = fluid
    f1 = fluid_block(fx, fy, fz, spacing, (1.0, 0.0, 1.0), 1e-3, seed)
    f2 = fluid_block(fx, fy, fz, spacing, (1.0, 0.0, 1.0 + fz * spacing), 1e-3, seed + 1)
    nf = f1.shape[0]
    parts.append((f1, np.ones(nf, np.float32), np.zeros(nf, np.int32), np.full(nf, S, np.float32)))
    parts.append((f2, np.full(nf, 0.5, np.float32), np.zeros(nf, np.int32), np.full(nf, 2.0 * S, np.float32)))
    # rigid cubes on a jittered lattice above the fluid
    cx, cy, cz = cubes
    ncube = cx * cy * cz
    u = uniform24(seed + 2, 3 * ncube).reshape(ncube, 3)
    k = np.arange(4, dtype=np.float64)
    cz_, cy_, cx_ = np.meshgrid(k, k, k, indexing="ij")
    local = np.stack([cx_.ravel(), cy_.ravel(), cz_.ravel()], 1) * spacing
    pitch = 1.0
    cube_pos = []
    for q in range(ncube):
        i, rem = q % cx, q // cx
        j, l = rem % cy, rem // cy
        base = np.array([3.0 + i * pitch, fy * spacing + 3.0 + j * pitch, 3.0 + l * pitch])
        base = base + (u[q] - 0.5) * 0.3
        cube_pos.append(local + base)
    cube_pos = np.concatenate(cube_pos).astype(np.float32) if ncube else np.zeros((0, 3), np.float32)
    nc = cube_pos.shape[0]
    cube_phase = (PHASE_RIGID0 + 1 + np.repeat(np.arange(ncube), 64)).astype(np.int32)
    parts.append((cube_pos, np.ones(nc, np.float32), cube_phase, np.ones(nc, np.float32)))
    # cloth above everything
    ny_top = fy * spacing + 3.0 + cy * pitch + 2.0
    gx = np.arange(cloth_n, dtype=np.float64) * spacing + 1.0
    zz, xx = np.meshgrid(gx, gx, indexing="ij")
    cloth = np.stack([xx.ravel(), np.full(cloth_n * cloth_n, ny_top), zz.ravel()], 1).astype(np.float32)
    ncl = cloth.shape[0]
    parts.append((cloth, np.ones(ncl, np.float32), np.full(ncl, PHASE_RIGID0, np.int32),
                  np.ones(ncl, np.float32)))
    pos = np.concatenate([p[0] for p in parts])
    w = np.concatenate([p[1] for p in parts])
    ph = np.concatenate([p[2] for p in parts])
    r0 = np.concatenate([p[3] for p in parts])
    n = pos.shape[0]
    p = pbf_params((0.0, 0.0, 0.0), box, h, radius, iterations, n, max_neighbors, seed=seed)
    p.mu_s, p.mu_k = 0.6, 0.4
    s = Scene(p, pos, w, ph, r0, dt, name="mixed_%d" % n)
    first_cloth = n - ncl
    s.distance = cloth_constraints(cloth_n, cloth_n, first_cloth, spacing)
    corners = np.array([0, cloth_n - 1, cloth_n * (cloth_n - 1), cloth_n * cloth_n - 1]) + first_cloth
    s.pins = (corners.astype(np.uint32), pos[corners])
    first_cube = 2 * nf
    s.bodies = [(np.arange(first_cube + 64 * q, first_cube + 64 * (q + 1), dtype=np.uint32), 1.0)
                for q in range(ncube)]
    s.rho0 = S
    del rng_off
    return s


def rope_over_fluid(fluid=(5, 3, 5), rope=10, spacing=0.1, iterations=3, seed=0x5EED0006,
                    dt=1.0 / 60.0, stiffness=(0.35, 0.9)):
    """A fluid block on the floor with a GAS-phase rope (phase 1: no
    particle contacts, counted in the fluid's density with s_solid) hanging
    just above it: distance constraints between rope neighbours, rest =
    spacing, stiffness ramped over the given range, first rope particle
    pinned where it starts.  The scene the adjoint is checked on (fluid,
    distance constraints, pins, walls and XSPH together)."""
    nx, ny, nz = fluid
    h = 2.0 * spacing
    ext = [max(nx * 2.0, rope + 3.0) * spacing, ny * spacing * 4.0, nz * spacing * 2.0]
    pf = fluid_block(nx, ny, nz, spacing, (0.0, 0.0, 0.0), 1e-3, seed)
    x = (np.arange(rope, dtype=np.float64) + 0.5) * spacing
    pr = np.stack([x, np.full(rope, (ny + 0.3) * spacing), np.full(rope, 0.5 * nz * spacing)], 1)
    pos = np.concatenate([pf, pr.astype(np.float32)])
    n, nf = pos.shape[0], pf.shape[0]
    p = pbf_params((0.0, 0.0, 0.0), ext, h, 0.5 * spacing, iterations, n, 64, seed=seed)
    rho0 = lattice_rest_density(h, spacing)
    phase = np.zeros(n, np.int32)
    phase[nf:] = 1
    s = Scene(p, pos, np.ones(n, np.float32), phase, np.full(n, np.float32(rho0), np.float32), dt,
              name="rope_over_fluid")
    pairs = np.stack([np.arange(nf, n - 1), np.arange(nf + 1, n)], 1).astype(np.uint32)
    s.distance = (pairs, np.full(rope - 1, spacing, np.float32),
                  np.linspace(stiffness[0], stiffness[1], rope - 1).astype(np.float32))
    s.pins = (np.array([nf], np.uint32), pos[nf:nf + 1].copy())
    s.rho0 = rho0
    return s


def cloth_and_fluid(cloth=(6, 6), fluid=(4, 3, 4), spacing=0.1, cloth_spacing=0.08, iterations=3,
                    seed=0x5EED0007, dt=1.0 / 60.0, velocity=(1.2, 0.0, 0.3),
                    stiffness=(0.9, 0.5, 0.2)):
    """A fluid block in the box corner and a cloth (structural, shear and
    bend distance constraints, first particle pinned) lying on the floor
    beside it, sliding away with the given initial velocity: cloth-cloth and
    cloth-fluid particle contacts, the floor's push-out and kinetic wall
    friction every iteration, fluid, XSPH and a pin in one small scene.  The
    cloth spacing is below the contact distance, so adjacent cloth particles
    stay strictly inside their contact range (a contact projects its pair
    onto the range's edge, where the contact set switches)."""
    fx, fy, fz = fluid
    cx, cz = cloth
    h = 2.0 * spacing
    radius = 0.5 * spacing
    pf = fluid_block(fx, fy, fz, spacing, (0.0, 0.0, 0.0), 1e-3, seed)
    i = np.arange(cx, dtype=np.float64)
    k = np.arange(cz, dtype=np.float64)
    zz, xx = np.meshgrid(k, i, indexing="ij")
    pc = np.stack([(fx + 0.2) * spacing + xx.ravel() * cloth_spacing,
                   np.full(cx * cz, 0.9 * radius),
                   spacing + zz.ravel() * cloth_spacing], 1).astype(np.float32)
    pos = np.concatenate([pf, pc])
    n, nf = pos.shape[0], pf.shape[0]
    ext = (max(2.0, (fx + cx + 8) * spacing), 1.0, max(1.0, (cz + 4) * spacing))
    p = pbf_params((0.0, 0.0, 0.0), ext, h, radius, iterations, n, 64, seed=seed)
    rho0 = lattice_rest_density(h, spacing)
    phase = np.zeros(n, np.int32)
    phase[nf:] = capi.PHASE_CLOTH
    vel = np.zeros((n, 3), np.float32)
    vel[nf:] = velocity
    s = Scene(p, pos, np.ones(n, np.float32), phase, np.full(n, np.float32(rho0), np.float32), dt,
              vel=vel, name="cloth_and_fluid")
    s.distance = cloth_constraints(cx, cz, nf, cloth_spacing, stiffness)
    s.pins = (np.array([nf], np.uint32), pos[nf:nf + 1].copy())
    s.rho0 = rho0
    return s


def cloth_over_pool(cloth=(128, 128), fluid=(64, 16, 64), spacing=0.1, cloth_spacing=0.1,
                    height=2.5, iterations=4, seed=0x5EED0005, dt=1.0 / 60.0,
                    stiffness=(0.9, 0.5, 0.2)):
    """The adjoint workload: a cloth (structural, shear and bend distance
    constraints, the four corners pinned) held `height` above a fluid block
    standing on the floor of the box, falling onto it as the block spreads.
    With the defaults, 16,384 cloth + 65,536 fluid particles.  The cloth
    spacing is 2r as in the reference's cloth: adjacent cloth particles sit
    at the edge of their contact range.  (A spacing well inside the range
    sets contacts against the distance constraints; the linearised Jacobi
    iteration then amplifies perturbations and the gradients over 60
    substeps overflow FP32.)"""
    fx, fy, fz = fluid
    cx, cz = cloth
    h = 2.0 * spacing
    radius = 0.5 * spacing
    wx = max(fx * spacing, cx * cloth_spacing) + 4.0 * spacing
    wz = max(fz * spacing, cz * cloth_spacing) + 4.0 * spacing
    ext = (wx, height + 1.5, wz)
    pf = fluid_block(fx, fy, fz, spacing, (0.5 * (wx - fx * spacing), 0.0, 0.5 * (wz - fz * spacing)), 1e-3,
                     seed)
    i = np.arange(cx, dtype=np.float64)
    k = np.arange(cz, dtype=np.float64)
    zz, xx = np.meshgrid(k, i, indexing="ij")
    x0 = 0.5 * (wx - (cx - 1) * cloth_spacing)
    z0 = 0.5 * (wz - (cz - 1) * cloth_spacing)
    pc = np.stack([x0 + xx.ravel() * cloth_spacing, np.full(cx * cz, height),
                   z0 + zz.ravel() * cloth_spacing], 1).astype(np.float32)
    pos = np.concatenate([pf, pc])
    n, nf = pos.shape[0], pf.shape[0]
    p = pbf_params((0.0, 0.0, 0.0), ext, h, radius, iterations, n, 96, seed=seed)
    rho0 = lattice_rest_density(h, spacing)
    phase = np.zeros(n, np.int32)
    phase[nf:] = capi.PHASE_CLOTH
    s = Scene(p, pos, np.ones(n, np.float32), phase, np.full(n, np.float32(rho0), np.float32), dt,
              name="cloth_over_pool_%dx%d_%d" % (cx, cz, nf))
    s.distance = cloth_constraints(cx, cz, nf, cloth_spacing, stiffness)
    corners = np.array([nf, nf + cx - 1, nf + (cz - 1) * cx, nf + cz * cx - 1], np.uint32)
    s.pins = (corners, pos[corners].copy())
    s.rho0 = rho0
    return s


def two_d_fluid(nx=40, ny=20, spacing=1.0, h=2.0, iterations=3, seed=0x5EED0001, dt=1.0 / 60.0):
    """2D PBF block (dims = 2, z fixed) with rest density from the square
    lattice kernel sum."""
    radius = 0.5 * spacing
    ext = (nx * spacing * 1.5, ny * spacing * 3.0, h)
    p = pbf_params((0.0, 0.0, 0.0), ext, h, radius, iterations, nx * ny, 32, seed=seed, dims=2,
                   relaxation=600.0)
    pos = fluid_block(nx, ny, 1, spacing, (0.0, 0.0, 0.0), 0.01, seed)
    pos[:, 2] = 0.5 * h
    n = nx * ny
    S = lattice_rest_density(h, spacing, dims=2)
    return Scene(p, pos, np.ones(n, np.float32), np.zeros(n, np.int32),
                 np.full(n, S, np.float32), dt, name="fluid2d")


class FluidObjectsScene:
    """Parameters plus fluid objects (positions, rest density), added with
    ps_add_fluid_object in creation order (Simulation::createFluid,
    cpu/src/simulation.cpp:431-449), and gas objects (positions, rest
    density, "gas" or "open gas"), added with ps_add_gas_object
    (Simulation::createGas, :408-428).  An object is (pos, rho0) or
    (pos, rho0, kind); kind "static" adds immovable solids (new Particle(pos,
    0), appended to the particle list as the volcano scene does, :1182-1185)
    and takes no object index.  A fluid object may carry an inverse mass as
    a fourth entry.  Smoke emitters (Simulation::createSmokeEmitter,
    :454-457) follow the objects, each (pos, particles_per_sec, object,
    max_markers) with object the index of the gas object it feeds or -1;
    then fluid emitters (createFluidEmitter, :459-461), each (pos,
    particles_per_sec, object)."""

    def __init__(self, params, objects, dt, name="", emitters=(), fluid_emitters=()):
        self.params = params
        self.objects = objects
        self.dt = dt
        self.name = name
        self.emitters = list(emitters)
        self.fluid_emitters = list(fluid_emitters)

    @property
    def n(self):
        return sum(o[0].shape[0] for o in self.objects)

    def load_into(self, sys_):
        for o in self.objects:
            pos, rho0 = o[0], o[1]
            kind = o[2] if len(o) > 2 else "fluid"
            if kind == "fluid":
                sys_.add_fluid_object(pos, None, o[3] if len(o) > 3 else 1.0, rho0)
            elif kind == "static":
                sys_.add_static_particles(pos)
            else:
                sys_.add_gas_object(pos, None, 1.0, rho0, open=(kind == "open gas"))
        for e in self.emitters:
            sys_.add_smoke_emitter(*e)
        for e in self.fluid_emitters:
            sys_.add_fluid_emitter(*e)


def cpu_reference_params(box_min, box_max, radius, h, iterations, max_particles, dims=2,
                         seed=0x5EED0003):
    """PS_FLAG_CPU_REFERENCE parameters with the reference CPU solver's
    constants (cpu/src/constraint/totalfluidconstraint.h:22-30:
    RELAXATION .01, K_P .1, E_P 4, DQ_P .2, S_SOLID 0; boundary jitter
    frand() * .003, boundaryconstraint.cpp:19; fluids have no wall friction,
    particle.h:49-50).  The grid is unused in this mode."""
    p = capi.PsParams()
    p.abi_version = capi.PS_ABI_VERSION
    p.dims = dims
    p.radius = radius
    p.h = h
    p.cell = h
    p.collide_dist = h
    p.gravity[:] = [0.0, -9.8, 0.0]
    p.box_min[:] = list(box_min)
    p.box_max[:] = list(box_max)
    p.origin[:] = list(box_min)
    p.grid[:] = [1, 1, 1]
    p.iterations = iterations
    p.relaxation = 0.01
    p.k_corr = 0.1
    p.n_corr = 4
    p.dq_corr = 0.2
    p.xsph_c = 0.0
    p.s_solid = 0.0
    p.mu_s = p.mu_k = 0.0
    p.wall_mu_s = p.wall_mu_k = 0.0
    p.wall_jitter = 0.003 / radius
    p.shape_stiffness = 1.0
    p.max_particles = max_particles
    p.max_neighbors = 4
    p.seed = seed
    p.flags = capi.PS_FLAG_CPU_REFERENCE | capi.PS_FLAG_SLEEPING
    return p


def rayleigh_taylor_2d(nx=40, ny=20, spacing=1.0, jitter=0.01, iterations=3, seed=0x5EED0004,
                       dt=1.0 / 60.0):
    """Two-fluid Rayleigh-Taylor scene in 2D: a light fluid (rest density 1)
    block of nx x ny particles on the floor and a heavy one (rest density 3)
    on top of it, spacing 1.0 with seeded jitter, radius spacing/2, h = 2r,
    box x in [-25, 25], y in [0, 1e6) (open top, as the reference's
    m_yBoundaries, simulation.cpp:661-662).  The light object is created
    first, as initFluid creates its objects in increasing density
    (simulation.cpp:894-913)."""
    radius = 0.5 * spacing
    h = 2.0 * radius
    half = 0.5 * nx * spacing
    light = fluid_block(nx, ny, 1, spacing, (-half, 0.0, 0.0), jitter, seed)
    heavy = fluid_block(nx, ny, 1, spacing, (-half, ny * spacing, 0.0), jitter, seed + 1)
    light[:, 2] = 0.0
    heavy[:, 2] = 0.0
    p = cpu_reference_params((-25.0, 0.0, -1.0), (25.0, 1.0e6, 1.0), radius, h, iterations,
                             2 * nx * ny)
    return FluidObjectsScene(p, [(light, 1.0), (heavy, 3.0)], dt, name="rayleigh_taylor_2d")


def reference_two_fluids(seed=0x5EED0005, dt=0.01):
    """The reference CPU app's fluid scene, Simulation::initFluid
    (cpu/src/simulation.cpp:894-913): two fluid objects side by side, rest
    densities 1 and 1.75, lattice step .7 from x = -8 (and 0) and y = -8,
    jitter .2 (u - .5) per axis (seeded SplitMix64 in place of frand()),
    H = 2, radius .25, box x in [-8, 8], y in [-8, 40], 3 iterations, dt =
    .01 (the app's fixed tick, cpu/src/view.cpp:197)."""
    scale, delta = 4.0, 0.7
    objects = []
    for d in range(2):
        start = -2.0 * scale + 4.0 * scale * (d / 2.0)
        pts = []
        x = start
        while x < start + 4.0 * scale / 2.0:
            y = -2.0 * scale
            while y < scale:
                pts.append((x, y))
                y += delta
            x += delta
        pts = np.array(pts, np.float64)
        u = uniform24(seed + d, 2 * len(pts)).reshape(-1, 2)
        pos = np.zeros((len(pts), 3), np.float64)
        pos[:, :2] = pts + 0.2 * (u - 0.5)
        objects.append((pos.astype(np.float32), 1.0 + 0.75 * d))
    n = sum(o[0].shape[0] for o in objects)
    p = cpu_reference_params((-2.0 * scale, -2.0 * scale, -1.0), (2.0 * scale, 10.0 * scale, 1.0),
                             0.25, 2.0, 3, n)
    return FluidObjectsScene(p, objects, dt, name="reference_two_fluids")


def _reference_lattice(x0, x1, y0, y1, delta, seed):
    """The CPU app's scene lattice: x from x0 while x < x1, y from y0 while
    y < y1, step delta, plus .2 (u - .5) per axis (seeded SplitMix64 in
    place of frand())."""
    pts = []
    x = x0
    while x < x1:
        y = y0
        while y < y1:
            pts.append((x, y))
            y += delta
        x += delta
    pts = np.array(pts, np.float64)
    u = uniform24(seed, 2 * len(pts)).reshape(-1, 2)
    pos = np.zeros((len(pts), 3), np.float64)
    pos[:, :2] = pts + 0.2 * (u - 0.5)
    return pos.astype(np.float32)


def reference_gas(seed=0x5EED0008, dt=0.01):
    """The reference CPU app's gas scene, Simulation::initGas
    (cpu/src/simulation.cpp:985-1016): two closed gas objects side by side
    (rest densities .75 and 3.75, lattice step .7 over x in [-4, 4), y in
    [-4, 4)) and above them two fluid objects (rest densities 4.75 and 5.5,
    x in [-6, 6), y in [4, 16)), created in that order; box x in [-4, 4],
    y in [-4, 20] (set before the fluid lattice's scale changes, so the
    fluid starts partly outside the box, as there); H = 2, radius .25,
    3 iterations, dt .01."""
    delta = 0.7
    objects = []
    scale = 2.0
    for d in range(2):
        start = -2.0 * scale + 4.0 * scale * (d / 2.0)
        pos = _reference_lattice(start, start + 4.0 * scale / 2.0, -2.0 * scale, 2.0 * scale, delta,
                                 seed + d)
        objects.append((pos, 0.75 + 3.0 * d, "gas"))
    scale = 3.0
    for d in range(2):
        start = -2.0 * scale + 4.0 * scale * (d / 2.0)
        pos = _reference_lattice(start, start + 4.0 * scale / 2.0, -2.0 * scale + 10.0,
                                 2.0 * scale + 10.0, delta, seed + 2 + d)
        objects.append((pos, 4.0 + 0.75 * (d + 1)))
    n = sum(o[0].shape[0] for o in objects)
    p = cpu_reference_params((-4.0, -4.0, -1.0), (4.0, 20.0, 1.0), 0.25, 2.0, 3, n)
    return FluidObjectsScene(p, objects, dt, name="reference_gas")


def reference_smoke(seed=0x5EED0009, dt=0.01, open_gas=True, emitter=False, max_emitted=512):
    """The CPU app's smoke scenes, Simulation::initSmokeOpen and
    initSmokeClosed (cpu/src/simulation.cpp:1096-1134): one gas object of
    rest density 1.5, lattice step .63 over x in [-4, 4), y in [-4, 4); open:
    box x in [-6, 6], y in [-4, 200] and the open-gas drag; closed: box
    [-4, 4] x [-4, 4].  With `emitter`, their smoke emitter at (0, -3), 15
    per second: feeding the gas object (open) or markers only (closed), with
    room for `max_emitted` emissions."""
    scale, delta = 2.0, 0.63
    pos = _reference_lattice(-2.0 * scale, 2.0 * scale, -2.0 * scale, 2.0 * scale, delta, seed)
    cap = pos.shape[0] + (max_emitted if emitter and open_gas else 0)
    if open_gas:
        p = cpu_reference_params((-3.0 * scale, -2.0 * scale, -1.0), (3.0 * scale, 100.0 * scale, 1.0),
                                 0.25, 2.0, 3, cap)
    else:
        p = cpu_reference_params((-2.0 * scale, -2.0 * scale, -1.0), (2.0 * scale, 2.0 * scale, 1.0),
                                 0.25, 2.0, 3, cap)
    kind = "open gas" if open_gas else "gas"
    emitters = []
    if emitter:
        emitters.append(((0.0, -2.0 * scale + 1.0, 0.0), 15.0, 0 if open_gas else -1, max_emitted))
    return FluidObjectsScene(p, [(pos, 1.5, kind)], dt, emitters=emitters,
                             name="reference_smoke_%s%s" % ("open" if open_gas else "closed",
                                                            "_emitter" if emitter else ""))


def reference_volcano(seed=0x5EED000A, dt=0.01, emitter=True):
    """The CPU app's volcano, Simulation::initVolcano
    (cpu/src/simulation.cpp:1178-1203): immovable solids along the cone
    y = 10 - |x| (x from 1 while x <= 10, step .2, the -x one first), then
    one fluid object of rest density 1 and mass 1.1 per particle (lattice
    step .8: y from 0 while y < 9, x from 0 while x < 9 - y, the particles at
    x and -x, each with .2 (frand() - .5) per axis), and a fluid emitter at
    the origin, 40 per second for its first 5 s, that also turns slow fluid
    into solid.  Box x in [-20, 20], y in [0, 100]; H = 2, radius .25,
    3 iterations, dt .01."""
    scale = 10.0
    walls = []
    x = 1.0
    while x <= scale:
        walls += [(-x, scale - x, 0.0), (x, scale - x, 0.0)]
        x += 0.2
    pts = []
    y = 0.0
    while y < scale - 1.0:
        x = 0.0
        while x < scale - y - 1.0:
            pts += [(x, y), (-x, y)]
            x += 0.8
        y += 0.8
    pts = np.array(pts, np.float64)
    u = uniform24(seed, 2 * len(pts)).reshape(-1, 2)
    fluid = np.zeros((len(pts), 3), np.float64)
    fluid[:, :2] = pts + 0.2 * (u - 0.5)
    walls = np.array(walls, np.float32)
    n = walls.shape[0] + fluid.shape[0]
    p = cpu_reference_params((-2.0 * scale, 0.0, -1.0), (2.0 * scale, 10.0 * scale, 1.0), 0.25, 2.0,
                             3, n + (200 if emitter else 0))
    objects = [(walls, 0.0, "static"),
               (fluid.astype(np.float32), 1.0, "fluid", np.float32(1.0 / 1.1))]
    fe = [((0.0, 0.0, 0.0), 4.0 * scale, 0)] if emitter else []
    return FluidObjectsScene(p, objects, dt, fluid_emitters=fe, name="reference_volcano")


def fluid_object_metrics(params, objects, pos):
    """FP64 densities of every particle at `pos` (all-pairs poly6 sum with the
    3D normalisation the reference's CPU solver uses in 2D,
    totalfluidconstraint.cpp:123-129) and the density constraint value
    C_i = rho_i / rho0_i - 1 of each particle's object."""
    x = np.asarray(pos, np.float64)
    h = float(params.h)
    c = poly6_coef(h)
    rho0 = np.concatenate([np.full(o[0].shape[0], o[1]) for o in objects])
    rho = np.zeros(x.shape[0])
    for a in range(0, x.shape[0], 512):
        d2 = ((x[a:a + 512, None, :] - x[None, :, :]) ** 2).sum(-1)
        w = np.where(d2 < h * h, c * (h * h - d2) ** 3, 0.0)
        rho[a:a + 512] = w.sum(1)
    return rho, rho / rho0 - 1.0


def pbf_density_metrics(sc, pos):
    """FP64 densities of a 3D fluid scene at `pos`, rho_i = m_i W(0) +
    sum_{j != i, r < h} m_j W_poly6(r) with m = 1/inverse mass (the density of
    findLambdasD, integration_kernel.cuh:552-577), and the density constraint
    C_i = rho_i / rho0_i - 1.  Pairs come from a k-d tree (scipy)."""
    from scipy.spatial import cKDTree
    ph = np.asarray(sc.phase)
    w = np.asarray(sc.inv_mass, np.float64)
    if np.any(ph != capi.PHASE_FLUID) or np.any(w <= 0.0):
        raise ValueError("pbf_density_metrics takes movable fluid particles only")
    x = np.asarray(pos, np.float64)
    h = float(sc.params.h)
    c = poly6_coef(h)
    m = 1.0 / w
    pairs = cKDTree(x, balanced_tree=False, compact_nodes=False).query_pairs(h, output_type="ndarray")
    i, j = pairs[:, 0], pairs[:, 1]
    d2 = ((x[i] - x[j]) ** 2).sum(1)
    wij = np.where(d2 < h * h, c * (h * h - d2) ** 3, 0.0)
    rho = m * c * h ** 6
    rho += np.bincount(i, wij * m[j], minlength=x.shape[0])
    rho += np.bincount(j, wij * m[i], minlength=x.shape[0])
    return rho, rho / np.asarray(sc.rest_density, np.float64) - 1.0


class Batch:
    """Scenes for ps_create_batch: a shared parameter set, one PsSceneDesc per
    scene, and each scene's particles/constraints (global indices are
    assigned when the batch is loaded, scenes filled in order)."""

    def __init__(self, params, descs, scenes, dt):
        self.params = params
        self.descs = descs
        self.scenes = scenes       # list of Scene (local indices)
        self.dt = dt

    @property
    def n(self):
        return sum(s.n for s in self.scenes)

    def load_into(self, sys_):
        first = 0
        for q, s in enumerate(self.scenes):
            sys_.add_particles(s.pos, s.vel, s.inv_mass, s.phase, s.rest_density, scene=q)
            if s.distance is not None:
                pairs, rest, st = s.distance
                sys_.add_distance_constraints(pairs + np.uint32(first), rest, st)
            if s.pins is not None:
                idx, xyz = s.pins
                sys_.add_point_constraints(idx + np.uint32(first), xyz)
            first += s.n

    def plain(self, q):
        """Scene q as a single-scene system (its own params and box)."""
        s = self.scenes[q]
        d = self.descs[q]
        p = self.params.copy()
        p.max_particles = s.n
        p.box_min[:] = list(d.box_min)
        p.box_max[:] = list(d.box_max)
        p.origin[:] = list(d.box_min)
        for a in range(3):
            ext = float(np.float32(d.box_max[a])) - float(np.float32(d.box_min[a]))
            p.grid[a] = max(1, int(math.ceil(ext / float(np.float32(p.cell)) - 1e-9)))
        g = np.float32(self.params.gravity[:]) * np.float32(d.gravity_scale)
        p.gravity[:] = [float(v) for v in g]
        p.seed = d.seed
        return Scene(p, s.pos, s.inv_mass, s.phase, s.rest_density, self.dt, vel=s.vel,
                     name="%s_plain_%d" % (s.name, q)) if s.distance is None else \
            _with_constraints(Scene(p, s.pos, s.inv_mass, s.phase, s.rest_density, self.dt,
                                    vel=s.vel), s)


def _with_constraints(dst, src):
    dst.distance = src.distance
    dst.pins = src.pins
    return dst


def ensemble(nscenes=512, nmin=2048, nmax=65536, links=64, spacing=0.1, iterations=4,
             max_neighbors=96, seed=0x5EED0004, dt=1.0 / 60.0):
    """Batched ensemble: nscenes independent scenes, particle counts
    log-uniform in [nmin, nmax] (rounded down), each a fluid block (rest
    density U[0.8, 1.2] times the lattice kernel sum) plus one `links`-link
    rope pinned at one end (stiffness U[0.3, 1.0]); gravity scale U[0.5, 1.5]."""
    h = 2.0 * spacing
    radius = 0.5 * spacing
    S = lattice_rest_density(h, spacing)
    u = uniform24(seed, 4 * nscenes).reshape(nscenes, 4)
    counts = np.floor(np.exp(np.log(nmin) + u[:, 0] * (np.log(nmax) - np.log(nmin)))).astype(np.int64)
    scenes, descs = [], []
    nrope = links + 1
    for q in range(nscenes):
        nf = int(counts[q]) - nrope
        nx = max(2, int(round((2.0 * nf) ** (1.0 / 3.0))))
        ny = max(1, int(math.ceil(nf / float(nx * nx))))
        blk = fluid_block(nx, ny, nx, spacing, (0.0, 0.0, 0.0), 1e-3, seed + 17 * (q + 1))[:nf]
        bx, by = nx * spacing, ny * spacing
        ext = (max(2.0 * bx, links * spacing + 1.0), max(2.0 * by, 2.0) + 1.0, 2.0 * bx)
        top = ext[1] - 0.5
        rope = np.stack([0.5 + np.arange(nrope) * spacing, np.full(nrope, top),
                         np.full(nrope, 0.5 * ext[2])], 1).astype(np.float32)
        pos = np.concatenate([blk, rope]).astype(np.float32)
        n = pos.shape[0]
        rho = np.float32((0.8 + 0.4 * u[q, 1]) * S)
        r0 = np.concatenate([np.full(nf, rho, np.float32), np.ones(nrope, np.float32)])
        ph = np.concatenate([np.zeros(nf, np.int32), np.full(nrope, PHASE_RIGID0, np.int32)])
        sc = Scene(None, pos, np.ones(n, np.float32), ph, r0, dt, name="scene%d" % q)
        pairs = np.stack([np.arange(nf, nf + links), np.arange(nf + 1, nf + nrope)], 1).astype(np.uint32)
        sc.distance = (pairs, np.full(links, spacing, np.float32),
                       np.full(links, 0.3 + 0.7 * u[q, 2], np.float32))
        sc.pins = (np.array([nf], np.uint32), rope[:1].copy())
        scenes.append(sc)
        d = capi.PsSceneDesc()
        d.box_min[:] = [0.0, 0.0, 0.0]
        d.box_max[:] = list(ext)
        d.gravity_scale = 0.5 + u[q, 3]
        d.seed = seed + q
        descs.append(d)
    total = int(sum(s.n for s in scenes))
    p = pbf_params((0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h, radius, iterations, total, max_neighbors,
                   seed=seed)
    return Batch(p, descs, scenes, dt)
