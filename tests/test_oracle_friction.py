"""Pins the oracle's contact friction, wall friction and height-scaled contact
mass (no GPU needed).

Each case runs one Jacobi iteration of the FP64 oracle on a one- or
two-particle scene whose predicted positions are set through the
velocities (dt = 1, no gravity), and checks what the projection must do
physically, not the oracle's own arithmetic:

- static friction (|t| < mu_s d) leaves the pair with no relative
  tangential displacement since the substep start;
- kinetic friction shortens the relative tangential displacement by exactly
  mu_k d (the CPU solver's RigidContactConstraint weights both branches by
  the scaled-mass share, rigidcontactconstraint.cpp:80-95; the GPU
  reference's collideD weights only the static one,
  integration_kernel.cuh:453-457), and removes it when mu_k d >= |t|;
- the corrections of the two particles conserve momentum under the
  height-scaled masses m* = e^{-y}/w (Particle::scaleMass, particle.h:67-73;
  collideD's sW, integration_kernel.cuh:401-403,423-425), each particle with
  its own height;
- a wall stops a particle whose tangential displacement is below
  sqrt(mu_s) r and otherwise shortens it by sqrt(mu_k) r
  (collide_world_functor, integration_kernel.cuh:131-151).

test_mutations_fail_the_pins rebuilds the oracle with one deliberate
mistake at a time (mu_s and mu_k swapped, the kinetic share dropped, e^{-y}
for e^{y}, one height for both particles, the friction clamp dropped, the
square roots of the wall coefficients dropped) and checks that some pin
here fails for each: the pins see those mistakes.
"""
import math
import os
import subprocess

import numpy as np
import pytest

from particlesolver_amd import scenes
from oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SOLID, CLOTH = 3, 2


def base_params(mu_s, mu_k, wall_mu_s=0.005, wall_mu_k=0.0002, collide=0.1):
    p = scenes.pbf_params((0.0, 0.0, 0.0), (8.0, 8.0, 8.0), 0.2, 0.05, 1, 4, 64)
    p.gravity[:] = [0.0, 0.0, 0.0]
    p.k_corr = 0.0
    p.collide_dist = collide
    p.mu_s, p.mu_k = mu_s, mu_k
    p.wall_mu_s, p.wall_mu_k = wall_mu_s, wall_mu_k
    return p


def one_iteration(make, p, x, v, w, phase):
    """Positions (insertion order, FP64) after predict + one iteration."""
    o = make(p)
    o.add_particles(np.asarray(x, np.float32), np.asarray(v, np.float32), np.asarray(w, np.float32),
                    np.asarray(phase, np.int32), np.ones(len(x), np.float32))
    o.begin(1.0)
    o.iterate()
    it, _ = o.sorted_state()
    order = o.grid()[0]
    out = np.empty((len(x), 3))
    out[order] = it[:, :3]
    return out


def pair_geometry(x, v):
    """Predicted positions, unit normal from j to i and the tangential part
    of the relative displacement (all FP64 from the FP32 inputs)."""
    x = np.asarray(x, np.float32).astype(np.float64)
    v = np.asarray(v, np.float32).astype(np.float64)
    p = x + v
    d = p[0] - p[1]
    n = d / np.linalg.norm(d)
    r = v[0] - v[1]
    t = r - np.dot(r, n) * n
    return x, p, n, t, np.linalg.norm(d)


def rel_tangent_after(x, out, n):
    """Relative tangential displacement since the substep start, projected
    on the plane normal to the pre-projection contact normal n."""
    r = (out[0] - x[0]) - (out[1] - x[1])
    return r - np.dot(r, n) * n


# ------------------------------------------------------------ particle pairs
# Pair in the x-z plane, 0.093 apart after prediction (contact distance 0.1),
# particle 0 sliding in +z relative to particle 1; equal heights, so the
# scaled masses equal the inverse masses' reciprocals.
X_PAIR = [[3.0, 3.0, 3.0], [3.09, 3.0, 3.025]]
V_SLIDE = [[0.0, 0.0, 0.05], [0.0, 0.0, 0.0]]


def static_pair(make):
    p = base_params(0.6, 0.4)
    w = [1.0, 0.5]
    out = one_iteration(make, p, X_PAIR, V_SLIDE, w, [SOLID, SOLID])
    x, pp, n, t, dist = pair_geometry(X_PAIR, V_SLIDE)
    assert np.linalg.norm(t) < 0.6 * dist                       # the static branch applies
    # no relative tangential displacement is left
    assert np.linalg.norm(rel_tangent_after(x, out, n)) < 1e-12
    # momentum of the corrections under m = 1/w (heights equal)
    dm = (out[0] - pp[0]) / w[0] + (out[1] - pp[1]) / w[1]
    assert np.abs(dm).max() < 1e-12


def kinetic_pair(make):
    mu_s, mu_k = 0.1, 0.05
    p = base_params(mu_s, mu_k)
    w = [1.0, 0.5]
    out = one_iteration(make, p, X_PAIR, V_SLIDE, w, [SOLID, SOLID])
    x, pp, n, t, dist = pair_geometry(X_PAIR, V_SLIDE)
    lt = np.linalg.norm(t)
    mud = float(np.float32(mu_k)) * dist
    assert lt > mu_s * dist and mud < lt                        # kinetic branch, f < 1
    ta = rel_tangent_after(x, out, n)
    # the relative tangential displacement keeps its direction and loses mu_k d
    assert abs(np.linalg.norm(ta) - (lt - mud)) < 1e-9
    assert np.linalg.norm(np.cross(ta, t)) < 1e-9 * lt
    dm = (out[0] - pp[0]) / w[0] + (out[1] - pp[1]) / w[1]
    assert np.abs(dm).max() < 1e-12


def clamped_pair(make):
    # mu_k d > |t| >= mu_s d: the kinetic factor clamps at 1
    p = base_params(0.1, 2.0)
    w = [1.0, 1.0]
    out = one_iteration(make, p, X_PAIR, V_SLIDE, w, [SOLID, SOLID])
    x, pp, n, t, dist = pair_geometry(X_PAIR, V_SLIDE)
    assert 0.1 * dist < np.linalg.norm(t) < 2.0 * dist
    assert np.linalg.norm(rel_tangent_after(x, out, n)) < 1e-12


# Vertical pair: particle 0 above particle 1 by 0.08, so the scaled masses
# differ by e^{0.08}; no tangential motion.
X_STACK = [[3.0, 3.08, 3.0], [3.03, 3.0, 3.0]]
V_REST = [[0.0, 0.0, 0.0], [0.0, 0.0, 0.0]]


def height_scaled_pair(make):
    p = base_params(0.6, 0.4)
    w = [1.0, 1.0]
    out = one_iteration(make, p, X_STACK, V_REST, w, [SOLID, SOLID])
    x, pp, n, t, dist = pair_geometry(X_STACK, V_REST)
    d0, d1 = out[0] - pp[0], out[1] - pp[1]
    # penetration resolved exactly along the normal
    assert abs(np.linalg.norm(out[0] - out[1]) - float(np.float32(0.1))) < 1e-12
    # the moves are in the ratio of the scaled inverse masses w e^{y}
    ratio = np.linalg.norm(d0) / np.linalg.norm(d1)
    assert abs(ratio - math.exp(x[0][1] - x[1][1])) < 1e-9
    # momentum under m* = e^{-y} / w
    m0, m1 = math.exp(-x[0][1]) / w[0], math.exp(-x[1][1]) / w[1]
    assert np.abs(m0 * d0 + m1 * d1).max() < 1e-12


def cloth_pair_is_not_height_scaled(make):
    # cloth-solid contacts use the plain inverse masses and no friction
    p = base_params(0.6, 0.4)
    w = [1.0, 1.0]
    out = one_iteration(make, p, X_STACK, V_REST, w, [CLOTH, SOLID])
    x, pp, n, t, dist = pair_geometry(X_STACK, V_REST)
    d0, d1 = out[0] - pp[0], out[1] - pp[1]
    assert abs(np.linalg.norm(d0) / np.linalg.norm(d1) - 1.0) < 1e-12


# ------------------------------------------------------------ walls
# A cloth particle predicted 0.04 below the floor's push-out height while
# sliding 0.03 in x since the substep start.
X_WALL = [[1.0, 0.06, 1.0]]
V_WALL = [[0.03, -0.05, 0.0]]


def wall_case(make, wall_mu_s, wall_mu_k):
    p = base_params(0.6, 0.4, wall_mu_s, wall_mu_k)
    out = one_iteration(make, p, X_WALL, V_WALL, [1.0], [CLOTH])
    return p, out[0]


def wall_static(make):
    # sqrt(0.5) r = 0.0354 > 0.03: the particle sticks
    p, y = wall_case(make, 0.5, 0.2)
    assert y[0] == X_WALL[0][0] and y[2] == np.float32(X_WALL[0][2])
    assert y[1] >= p.radius  # pushed out of the floor


def wall_kinetic(make):
    # sqrt(0.1) r = 0.0158 < 0.03: it slides, 0.03 shortened by sqrt(0.04) r = 0.01
    p, y = wall_case(make, 0.1, 0.04)
    slide = float(np.float32(V_WALL[0][0]))  # the FP64 prediction moves x by v dt exactly
    r = float(np.float32(p.radius))
    want = slide - math.sqrt(float(np.float32(0.04))) * r
    assert abs((y[0] - float(np.float32(X_WALL[0][0]))) - want) < 1e-9
    assert y[2] == np.float32(X_WALL[0][2])


def wall_clamped(make):
    # sqrt(mu_k) r = 0.0354 > 0.03 >= sqrt(mu_s) r: the kinetic step clamps at the whole slide
    p, y = wall_case(make, 0.1, 0.5)
    assert abs(y[0] - float(np.float32(X_WALL[0][0]))) < 1e-12


PINS = {
    "static_pair": static_pair,
    "kinetic_pair": kinetic_pair,
    "clamped_pair": clamped_pair,
    "height_scaled_pair": height_scaled_pair,
    "cloth_pair_is_not_height_scaled": cloth_pair_is_not_height_scaled,
    "wall_static": wall_static,
    "wall_kinetic": wall_kinetic,
    "wall_clamped": wall_clamped,
}


def oracle64(p):
    return Oracle(p, "f64")


@pytest.mark.parametrize("name", sorted(PINS))
def test_friction_pin(name):
    PINS[name](oracle64)


def test_fp32_oracle_agrees_with_fp64_on_the_pins():
    # the FP32 build takes the same branches and lands within FP32 rounding
    for (x, v, mu, ph) in ((X_PAIR, V_SLIDE, (0.1, 0.05), [SOLID, SOLID]),
                           (X_STACK, V_REST, (0.6, 0.4), [SOLID, SOLID])):
        p = base_params(*mu)
        a = one_iteration(lambda q: Oracle(q, "f32"), p, x, v, [1.0, 0.5], ph)
        b = one_iteration(oracle64, p, x, v, [1.0, 0.5], ph)
        assert np.abs(a - b).max() < 2e-6


# ------------------------------------------------------------ mutations
# (name, original text, replacement, pins expected to fail)
MUTATIONS = [
    ("mu_s/mu_k swapped in particle friction",
     "if (lt < (d->mu_s * dist)) {", "if (lt < (d->mu_k * dist)) {", ["static_pair"]),
    ("kinetic particle friction without the mass share",
     "c3[0] = c3[0] - ((tx * f) * share);", "c3[0] = c3[0] - (tx * f);", ["kinetic_pair"]),
    ("e^{-y} mass scaling",
     "real ewi = wi * EXP(py);", "real ewi = wi * EXP(-py);", ["height_scaled_pair"]),
    ("particle i's height for both (the GPU reference's form)",
     "cw2 = wj * EXP(qy);", "cw2 = wj * EXP(py);", ["height_scaled_pair"]),
    ("no clamp of the kinetic factor",
     "if (f > R(1)) f = R(1);\n                        c3[0]", "c3[0]", ["clamped_pair"]),
    ("height scaling also for cloth",
     "int solid = ph >= PS_PHASE_SOLID && pj >= PS_PHASE_SOLID;",
     "int solid = ph >= PS_PHASE_CLOTH && pj >= PS_PHASE_CLOTH;", ["cloth_pair_is_not_height_scaled"]),
    ("wall coefficients without the square root",
     "d->wall_k = R(sqrt((double)p->wall_mu_k) * (double)p->radius);",
     "d->wall_k = R((double)p->wall_mu_k * (double)p->radius);", ["wall_kinetic"]),
    ("wall mu_s/mu_k swapped",
     "if (l < d->wall_s) {", "if (l < d->wall_k) {", ["wall_static"]),
]


def build_mutant(tmp, name, old, new):
    src = open(os.path.join(ROOT, "oracle", "ps_oracle.c")).read()
    assert src.count(old) >= 1, "mutation site not found: " + name
    src = src.replace(old, new)
    path = os.path.join(tmp, "ps_oracle_mut.c")
    with open(path, "w") as f:
        f.write(src)
    lib = os.path.join(tmp, "liboracle_mut_f64.so")
    inc = os.path.join(ROOT, "oracle")
    subprocess.run(["gcc", "-std=c99", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-mfma",
                    "-DORC_DOUBLE=1", "-I" + inc, path, os.path.join(ROOT, "oracle", "ref2d.c"),
                    "-o", lib, "-lm"], check=True, cwd=inc)
    return lib


@pytest.mark.parametrize("mut", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_mutations_fail_the_pins(mut, tmp_path):
    name, old, new, expect = mut
    lib = build_mutant(str(tmp_path), name, old, new)
    make = lambda p: Oracle(p, "f64", lib=lib)
    for pin in expect:
        with pytest.raises(AssertionError):
            PINS[pin](make)


def test_fp32_exp_accuracy():
    """e^y of the height-scaled contact mass: the FP32 oracle's Cody-Waite +
    degree-7 Taylor form (the GPU performs the same steps) is within 4 ulp
    of e^x over the heights scenes reach, and clamps at |x| = 80; the FP64
    build calls libm."""
    import ctypes
    from oracle import _lib
    f32, f64 = _lib("f32").orc_exp_f32, _lib("f64").orc_exp_f64
    f32.restype, f32.argtypes = ctypes.c_float, [ctypes.c_float]
    f64.restype, f64.argtypes = ctypes.c_double, [ctypes.c_double]
    xs = np.float32(np.linspace(-60.0, 60.0, 24001))
    xs = np.concatenate([xs, np.float32([0.0, 1.0, -1.0, 0.5, 0.34657359, -0.34657359, 88.0])])
    got = np.array([f32(float(x)) for x in xs], np.float64)
    ref = np.exp(np.minimum(xs.astype(np.float64), 80.0))
    ulp = np.spacing(ref.astype(np.float32)).astype(np.float64)
    assert (np.abs(got - ref) / ulp).max() <= 4.0
    assert f32(0.0) == 1.0
    for x in (-3.0, 0.08, 2.5):
        assert f64(x) == math.exp(x)
