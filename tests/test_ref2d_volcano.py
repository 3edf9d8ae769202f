"""Pinning of ref2d's immovable solids, ContactConstraint and FluidEmitter
(the reference's volcano scene, Simulation::initVolcano,
cpu/src/simulation.cpp:1178-1203) against closed forms and the rules as the
reference states them.

Each check catches one plausible slip:
- the contact test (closer than a diameter less EPSILON, a solid in the
  pair, not both immovable) and its projection (overlap removed along the
  line of centres, divided by the particle's constraint count, contacts in
  creation order, immovable partners unmoved);
- the emitter's schedule (emissions while the total time, this tick
  included, stays below 5 s) and the emitted state (mass 1, v = (u, 1));
- solidification: t counts down 4 -> 0 under lambda <= 0, up to 3
  otherwise, the solid has imass 0 and v = 0 and leaves the object, and
  `lambdas` is looked up with the position in the object's list, so a
  particle reads the lambda of the particle whose index equals its position
  (0 when that one is outside the object);
- the stabilisation pre-pass (PS_FLAG_STABILIZATION): the start position
  moves with the prediction, so the push-out leaves no velocity.

The height-scaled contact mass tmass = m e^{-y} (particle.h:67-73) cancels
from every contact with an immovable partner, the only contacts here, so
only its rounding is visible; the GPU's bit-exact parity covers it and the
closed forms below do not (parity of the formula itself: unpinned)."""
import math

import numpy as np
import pytest

from particlesolver_amd import scenes
from oracle import Ref2D, load_ref2d

H = 2.0
C6 = 315.0 / (64.0 * math.pi * H ** 9)


def W(r2):
    return C6 * (H * H - r2) ** 3 if r2 < H * H else 0.0


def params(box=50.0, g=0.0, max_particles=512, iterations=1):
    p = scenes.cpu_reference_params((-box, -box, -1.0), (box, box, 1.0), 0.25, H, iterations,
                                    max_particles)
    p.gravity[:] = [0.0, g, 0.0]
    p.flags = p.flags & ~2          # no sleeping
    return p


def f32(v):
    return np.asarray(v, np.float32).astype(np.float64)


def test_contacts_push_out_in_creation_order():
    r = Ref2D(params(), "f64")
    a, b, f = f32([0.0, 0.0, 0.0]), f32([0.35, 0.3, 0.0]), f32([0.0, 0.3, 0.0])
    r.add_static_particles(np.array([a, b]))
    r.add_fluid_object(f[None], rest_density=1.0)
    r.step(1.0, 1)
    # contact (A, F) first, then (B, F); F counts two constraints, so each
    # removes half of its overlap, along the line of centres
    x = f.copy()
    for s in (a, b):
        d = x - s
        x = x + d / np.linalg.norm(d) * (0.5 - np.linalg.norm(d)) / 2.0
    got = r.positions()
    np.testing.assert_allclose(got[2], x, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(got[:2], np.array([a, b]))     # immovable
    np.testing.assert_allclose(r.velocities()[2], x - f, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("gap,touch", [(0.4998, True), (0.49995, False)])
def test_contact_threshold_is_a_diameter_less_epsilon(gap, touch):
    r = Ref2D(params(), "f64")
    r.add_static_particles(np.array([[0.0, 0.0, 0.0]]))
    r.add_fluid_object(np.array([[gap, 0.0, 0.0]]), rest_density=1.0)
    r.step(1.0, 1)
    x = r.positions()[1, 0]
    if touch:
        assert x == pytest.approx(0.5, rel=1e-12)
    else:
        assert x == np.float32(gap)


def test_no_contact_without_a_solid_or_between_immovables():
    d, rho0 = float(np.float32(0.3)), 0.5
    r = Ref2D(params(), "f64")
    r.add_static_particles(np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]]))   # overlapping solids
    r.add_fluid_object(np.array([[0.0, 10.0, 0.0], [d, 10.0, 0.0]]), rest_density=rho0)
    x0 = r.positions()
    r.step(1.0, 1)
    x = r.positions()
    np.testing.assert_array_equal(x[:2], x0[:2])
    # the two fluid particles overlap, yet only their fluid constraint acts:
    # the two-particle projection with divisor 2 (themselves, no contact)
    # (relaxation, K_P and DQ_P enter as floats)
    relax, kp, dq = (float(np.float32(v)) for v in (0.01, 0.1, 0.2))
    CS = 45.0 / (math.pi * H ** 6)
    rho = W(0.0) + W(d * d)
    lam = -(rho / rho0 - 1.0) / (2.0 * (CS * (H - d) ** 2 / rho0) ** 2 + relax)
    corr = -kp * (W(d * d) / W((dq * H) ** 2)) ** 4
    move = -(2.0 * lam + corr) * CS * (H - d) ** 2 / rho0 / 2.0
    assert x[3, 0] - x0[3, 0] == pytest.approx(move, rel=1e-9)
    assert x0[2, 0] - x[2, 0] == pytest.approx(move, rel=1e-9)


def test_emission_schedule_and_state():
    r = Ref2D(params(max_particles=400), "f64")
    r.add_fluid_object(np.array([[40.0, 40.0, 0.0]]), rest_density=1.0)
    r.add_fluid_emitter((0.0, 0.0, 0.0), 40.0, 0)
    counts = []
    for k in range(25):
        r.step(0.25, 1)
        counts.append(r.n - 1)
        if k == 0:
            v = r.velocities()[1:]
            assert v.shape == (10, 3)
            assert np.all(v[:, 1] == 1.0) and np.all(v[:, 2] == 0.0)
            assert np.all((v[:, 0] >= 0.0) & (v[:, 0] < 1.0)) and len(set(v[:, 0])) == 10
            np.testing.assert_array_equal(r.positions()[1:], 0.0)
            ph, w, t = r.attrs()
            assert np.all(ph[1:] == 0) and np.all(w[1:] == 1.0) and np.all(t[1:] == 4.0)
    # 10 per tick while the total (this tick included) is below 5 s: the
    # 20th tick brings it to 5.0 exactly and emits nothing
    assert counts == [10 * min(k + 1, 19) for k in range(25)]


def test_solidification_and_the_lambda_lookup_by_list_position():
    # particle 0 is a solid outside the object; the object is [A, B, C]:
    # A alone (density W(0) < rho0: lambda > 0), B and C a pair 1.5 apart
    # whose density exceeds rho0 by 1e-5 (lambda < 0), so slightly that their
    # correction stays below the sleeping threshold and nothing moves.
    # Positions in the list are A 0, B 1, C 2, so A reads lambdas[0]
    # (particle 0: no entry, 0), B reads lambdas[1] (A's, > 0) and C reads
    # lambdas[2] (B's, < 0).
    rho0 = (W(0.0) + W(2.25)) / (1.0 + 1e-5)
    assert W(0.0) < rho0
    p = params()
    p.flags |= 2                                     # sleeping, as the reference
    r = Ref2D(p, "f64")
    r.add_static_particles(np.array([[-30.0, 0.0, 0.0]]))
    r.add_fluid_object(np.array([[20.0, 0.0, 0.0], [-10.0, 0.0, 0.0], [-8.5, 0.0, 0.0]]),
                       rest_density=rho0)
    r.add_fluid_emitter((0.0, 30.0, 0.0), 1.0, 0)    # dt = 10: never emits (total >= 5)
    dt = 10.0
    # (A, B, C): A reads 0 and C reads B's lambda, so both count down and
    # turn solid together; B reads A's lambda and stays at 3; once A is a
    # solid, B (now first in the list) reads particle 0's: 0
    expect = [(3.0, 3.0, 3.0), (2.0, 3.0, 2.0), (1.0, 3.0, 1.0), (0.0, 3.0, 0.0), (0.0, 2.0, 0.0)]
    for tick, e in enumerate(expect):
        r.step(dt, 1)
        v = np.linalg.norm(r.velocities(), axis=1)
        assert np.all(v == 0.0), v                      # asleep: slow and low
        ph, w, t = r.attrs()
        assert tuple(t[1:]) == e, (tick, t)
    assert r.n == 4
    ph, w, t = r.attrs()
    assert list(ph) == [3, 3, 0, 3] and w[1] == 0.0 and w[3] == 0.0
    assert np.all(r.velocities()[[1, 3]] == 0.0)
    # as a solid, A no longer moves
    x = r.positions()[1].copy()
    r.step(dt, 1)
    np.testing.assert_array_equal(r.positions()[1], x)


def test_fp32_tracks_fp64_on_the_volcano():
    sc = scenes.reference_volcano()
    a, b = load_ref2d(sc, "f32"), load_ref2d(sc, "f64")
    a.step(np.float32(sc.dt), 10)
    b.step(np.float32(sc.dt), 10)
    assert a.n == b.n
    assert np.abs(a.positions().astype(np.float64) - b.positions()).max() < 1e-3


def test_stabilization_moves_the_start_position_with_the_prediction():
    # a particle predicted 0.1 below the floor: the pre-pass puts p and ep at
    # the push-out height (r plus its jitter), so the substep ends at rest
    # there; without the pre-pass the push-out shows up as velocity
    dt = 0.01
    out = {}
    for stab in (False, True):
        p = params(box=10.0)
        p.box_min[1] = 0.0
        if stab:
            p.flags |= 0x10
        r = Ref2D(p, "f64")
        r.add_fluid_object(np.array([[0.0, 0.1, 0.0]]), np.array([[0.0, -20.0, 0.0]]),
                           rest_density=1.0)
        r.step(dt, 1)
        out[stab] = (r.positions()[0], r.velocities()[0])
    jit = float(np.float32(0.003 / 0.25)) * 0.25
    y, vy = out[True][0][1], out[True][1][1]
    assert 0.25 <= y < 0.25 + jit and vy == 0.0
    y0, vy0 = out[False][0][1], out[False][1][1]
    y_start = float(np.float32(0.1))
    assert 0.25 <= y0 < 0.25 + jit and vy0 == pytest.approx((y0 - y_start) / dt, rel=1e-12)
    assert y != y0                                # its own jitter draw
