"""CPU-reference mode (PS_FLAG_CPU_REFERENCE) on the GPU against ref2d, the
oracle's restatement of the reference's CPU solver (Simulation::tick,
cpu/src/simulation.cpp:115-369).

FP32 GPU vs FP32 ref2d is bit-exact at every checked step.  Against the FP64
ref2d the FP32 error is stated per result over a short horizon: positions
(max and RMS), densities and the density-constraint residual."""
import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes
from oracle import load_ref2d
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu


def run_parity(sc, checkpoints):
    ref = load_ref2d(sc, "f32")
    with System(sc.params) as g:
        sc.load_into(g)
        done = 0
        for k in checkpoints:
            g.step(sc.dt, k - done)
            ref.step(sc.dt, k - done)
            done = k
            assert_bitexact(g.positions(), ref.positions(), "positions after %d steps" % k)
            assert_bitexact(g.velocities(), ref.velocities(), "velocities after %d steps" % k)
            for e in range(len(getattr(sc, "emitters", ()))):
                assert_bitexact(g.markers(e), ref.markers(e), "markers of %d after %d steps" % (e, k))
            if getattr(sc, "fluid_emitters", None):
                assert np.array_equal(g.phases(), ref.attrs()[0]), "phases after %d steps" % k
        st = g.status()
        assert st.flags == 0
        assert st.iterations == done * sc.params.iterations
        return g.positions()


def test_rayleigh_taylor_200_steps_bitexact(gpu_available):
    sc = scenes.rayleigh_taylor_2d()            # 1,600 particles, two fluid objects
    assert sc.n == 1600
    x = run_parity(sc, [1, 2, 10, 50, 200])
    assert np.all(x[:, 2] == 0.0)               # 2D: z never moves
    assert np.all(x[:, 1] >= 0.0)               # floor


def test_reference_two_fluids_bitexact(gpu_available):
    run_parity(scenes.reference_two_fluids(), [1, 5, 60, 300])


def test_staged_and_graph_equal_step(gpu_available):
    sc = scenes.reference_two_fluids()
    with System(sc.params) as a, System(sc.params) as b, System(sc.params) as c:
        for s in (a, b, c):
            sc.load_into(s)
        for rep in range(3):
            a.step(sc.dt, 4)
            for _ in range(4):
                b.begin(sc.dt)
                for _ in range(sc.params.iterations):
                    b.iterate()
                b.end()
            c.step_graph(sc.dt, 4)
            xa = a.positions()
            assert_bitexact(b.positions(), xa, "staged, round %d" % rep)
            assert_bitexact(c.positions(), xa, "graph replay %d" % rep)
            assert_bitexact(c.velocities(), a.velocities(), "graph replay %d velocities" % rep)


def test_extra_staged_iterations_keep_boundaries(gpu_available):
    # an iteration beyond params.iterations still projects its boundaries
    # first: compare with a system whose iteration count is one higher
    sc = scenes.reference_two_fluids()
    p4 = sc.params.copy()
    p4.iterations = sc.params.iterations + 1
    with System(sc.params) as a, System(p4) as b:
        sc.load_into(a)
        sc.load_into(b)
        for _ in range(3):
            a.begin(sc.dt)
            for _ in range(p4.iterations):
                a.iterate()
            a.end()
            b.step(sc.dt, 1)
        assert_bitexact(a.positions(), b.positions(), "extra staged iteration")


def test_fp32_error_vs_fp64_short_horizon(gpu_available):
    sc = scenes.reference_two_fluids()
    ref64 = load_ref2d(sc, "f64")
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 10)
        ref64.step(sc.dt, 10)
        x32, x64 = g.positions().astype(np.float64), ref64.positions()
    err = np.abs(x32 - x64).max(1)
    rho32, c32 = scenes.fluid_object_metrics(sc.params, sc.objects, x32)
    rho64, c64 = scenes.fluid_object_metrics(sc.params, sc.objects, x64)
    # tolerances for 10 steps of 3 iterations in FP32 (measured max 2.9e-5)
    assert err.max() < 2e-4, err.max()
    assert np.sqrt((err ** 2).mean()) < 5e-5
    assert np.abs(rho32 - rho64).max() < 1e-3
    assert abs(np.abs(c32).mean() - np.abs(c64).mean()) < 1e-4


def test_unsupported_calls_in_cpu_reference_mode(gpu_available):
    sc = scenes.reference_two_fluids()
    with System(sc.params) as g:
        sc.load_into(g)
        for call in (lambda: g.add_particles(np.zeros((1, 3), np.float32)),
                     lambda: g.add_distance_constraints(np.array([[0, 1]]), np.array([1.0])),
                     lambda: g.add_point_constraints(np.array([0]), np.zeros((1, 3))),
                     lambda: g.predict(sc.dt),
                     lambda: g.grid(),
                     lambda: g.autotune(sc.dt, 1)):
            with pytest.raises(capi.PsError) as e:
                call()
            assert e.value.code == capi.PS_ERR_UNSUPPORTED
        p = sc.params.copy()
        p.flags = 0
        with pytest.raises(capi.PsError) as e:
            g.set_params(p)
        assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT
        g.step(sc.dt, 1)                     # the handle still works
        assert np.isfinite(g.positions()).all()


def test_reference_gas_bitexact(gpu_available):
    # Simulation::initGas: two closed gas objects and two fluid objects
    sc = scenes.reference_gas()
    assert [o[0].shape[0] for o in sc.objects] == [72, 72, 162, 162]
    run_parity(sc, [1, 5, 50, 200])


@pytest.mark.parametrize("open_gas", [True, False])
def test_smoke_gas_bitexact(gpu_available, open_gas):
    x = run_parity(scenes.reference_smoke(open_gas=open_gas), [1, 10, 100])
    assert np.all(np.isfinite(x))


@pytest.mark.parametrize("open_gas", [True, False])
def test_smoke_emitter_bitexact(gpu_available, open_gas):
    # initSmokeOpen: the emitter feeds the open gas; initSmokeClosed: markers only
    sc = scenes.reference_smoke(open_gas=open_gas, emitter=True)
    n0 = sc.n
    run_parity(sc, [1, 7, 10, 100, 300])
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 300)
        assert g.markers(0).shape[0] == 44          # 300 x 0.01f s at 15/s
        assert g.n == n0 + (44 if open_gas else 0)


def test_smoke_emitter_rules(gpu_available):
    sc = scenes.reference_smoke(open_gas=True, emitter=False)
    p = sc.params.copy()
    p.max_particles = sc.n + 2
    with System(p) as g:
        sc.load_into(g)
        with pytest.raises(capi.PsError) as e:       # a fluid object cannot be fed
            g.add_fluid_object(np.array([[5.0, 5.0, 0.0]], np.float32), None, 1.0, 1.0)
            g.add_smoke_emitter((0.0, 0.0, 0.0), 15.0, 1, 8)
        assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT
    with System(p) as g:
        sc.load_into(g)
        g.add_fluid_object(np.array([[5.0, 5.0, 0.0]], np.float32), None, 1.0, 1.0)
        with pytest.raises(capi.PsError) as e:       # only the last object
            g.add_smoke_emitter((0.0, 0.0, 0.0), 15.0, 0, 8)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
    with System(p) as g:
        sc.load_into(g)
        g.add_smoke_emitter((0.0, -3.0, 0.0), 15.0, 0, 8)
        with pytest.raises(capi.PsError) as e:       # nothing may follow it
            g.add_fluid_object(np.array([[5.0, 5.0, 0.0]], np.float32), None, 1.0, 1.0)
        assert e.value.code == capi.PS_ERR_STATE
        with pytest.raises(capi.PsError) as e:
            g.step_graph(sc.dt, 12)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
        # room for 2 emissions: 14 substeps emit 2, 21 would emit 3 and fail
        # before the first substep, leaving the handle as it was
        x0 = g.positions()
        with pytest.raises(capi.PsError) as e:
            g.step(sc.dt, 21)
        assert e.value.code == capi.PS_ERR_CAPACITY
        assert np.array_equal(g.positions(), x0) and g.status().iterations == 0
        g.step(sc.dt, 14)
        assert g.n == sc.n + 2 and g.markers(0).shape[0] == 2
    q = sc.params.copy()
    q.dims = 3
    q.box_min[2], q.box_max[2] = -4.0, 4.0
    with System(q) as g:
        with pytest.raises(capi.PsError) as e:
            g.add_smoke_emitter((0.0, 0.0, 0.0), 15.0, -1, 8)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED


def test_volcano_bitexact(gpu_available):
    # Simulation::initVolcano: immovable cone, fluid of mass 1.1, a fluid
    # emitter at 40/s for 5 s that turns slow fluid into solid
    sc = scenes.reference_volcano()
    assert [o[0].shape[0] for o in sc.objects] == [92, 156]
    run_parity(sc, [1, 7, 50, 200, 520, 700])
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 700)
        ph = g.phases()
        assert g.n == sc.n + 199                      # 5 s at 40/s, the 500th substep excluded
        assert (ph == 3).sum() > 92                   # some fluid turned solid


def test_static_obstacles_bitexact(gpu_available):
    # fluid poured onto a slope of immovable solids: contacts with many
    # partners per particle, and solids inside the boundary band
    p = scenes.cpu_reference_params((-6.0, 0.0, -1.0), (6.0, 12.0, 1.0), 0.25, 2.0, 3, 600)
    xs = np.arange(-6.0, 6.01, 0.3)
    slope = np.stack([xs, 0.15 * (xs + 6.0), np.zeros_like(xs)], 1).astype(np.float32)
    from particlesolver_amd.scenes import _reference_lattice, FluidObjectsScene
    fluid = _reference_lattice(-3.0, 3.0, 4.0, 8.0, 0.5, 0x5EED0011)
    sc = FluidObjectsScene(p, [(slope, 0.0, "static"), (fluid, 1.0, "fluid")], 0.01,
                           name="slope")
    run_parity(sc, [1, 10, 100, 300])


@pytest.mark.parametrize("make", [scenes.reference_two_fluids, scenes.reference_volcano,
                                  lambda: scenes.reference_smoke(open_gas=True, emitter=True)])
def test_stabilization_bitexact(gpu_available, make):
    # the USE_STABILIZATION pre-pass (PS_FLAG_STABILIZATION) on three scenes
    sc = make()
    sc.params.flags |= capi.PS_FLAG_STABILIZATION
    run_parity(sc, [1, 10, 100, 250])


def test_stabilization_needs_the_cpu_reference_mode(gpu_available):
    p = scenes.dam_break(4, 4, 4).params.copy()
    p.flags |= capi.PS_FLAG_STABILIZATION
    with pytest.raises(capi.PsError) as e:
        System(p)
    assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT


def test_fluid_emitter_rules(gpu_available):
    sc = scenes.reference_volcano(emitter=False)
    roomy = sc.params.copy()
    roomy.max_particles += 8
    with System(roomy) as g:
        sc.load_into(g)
        with pytest.raises(capi.PsError) as e:           # a movable solid
            g.add_particles(np.zeros((1, 3), np.float32), None, np.ones(1, np.float32),
                            np.full(1, 3, np.int32), np.zeros(1, np.float32))
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
        g.add_gas_object(np.array([[0.0, 50.0, 0.0]], np.float32))
        with pytest.raises(capi.PsError) as e:           # a gas object
            g.add_fluid_emitter((0.0, 0.0, 0.0), 40.0, 1)
        assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT
        with pytest.raises(capi.PsError) as e:           # not the last object
            g.add_fluid_emitter((0.0, 0.0, 0.0), 40.0, 0)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
    with System(sc.params) as g:
        sc.load_into(g)
        g.add_fluid_emitter((0.0, 0.0, 0.0), 40.0, 0)
        with pytest.raises(capi.PsError) as e:           # nothing may follow it
            g.add_static_particles(np.array([[5.0, 5.0, 0.0]]))
        assert e.value.code == capi.PS_ERR_STATE
        with pytest.raises(capi.PsError) as e:
            g.step_graph(sc.dt, 12)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
        with pytest.raises(capi.PsError) as e:           # max_particles = the scene: no room
            g.step(sc.dt, 3)
        assert e.value.code == capi.PS_ERR_CAPACITY
        assert g.status().iterations == 0


def test_gas_objects_need_the_2d_cpu_reference_mode(gpu_available):
    pos = np.zeros((1, 3), np.float32)
    sc = scenes.dam_break(4, 4, 4)
    with System(sc.params) as g:
        with pytest.raises(capi.PsError) as e:
            g.add_gas_object(pos)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
    p = scenes.reference_smoke().params.copy()
    p.dims = 3
    p.box_min[2], p.box_max[2] = -4.0, 4.0
    with System(p) as g:
        with pytest.raises(capi.PsError) as e:
            g.add_gas_object(pos)
        assert e.value.code == capi.PS_ERR_UNSUPPORTED
