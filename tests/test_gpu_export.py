"""ps_export_device (insertion-order state written to device memory, the
reference's position-VBO layout) and trajectory files written by the
library, against the synchronous readbacks ps_read_positions /
ps_read_velocities and the scene's own phase array."""
import numpy as np
import pytest

from particlesolver_amd import System, capi, scenes, trajectory
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu


def export(g, field, first=0, n=None):
    n = g.n - first if n is None else n
    shape = (n,) if field == capi.PS_FIELD_PHASES else (n, 4)
    dtype = np.int32 if field == capi.PS_FIELD_PHASES else np.float32
    with capi.DeviceBuffer(4 * int(np.prod(shape))) as buf:
        buf.fill(0xFF)                    # NaN / -1 where nothing is written
        g.export_device(field, buf.ptr, first, n)
        g.synchronize()
        return buf.numpy(dtype, shape)


@pytest.mark.parametrize("make", [lambda: scenes.dam_break(12, 10, 12),
                                  lambda: scenes.mixed(cloth_n=8, cubes=(2, 1, 2), fluid=(8, 6, 8),
                                                       box=(5.0, 5.0, 5.0))])
def test_export_matches_readback(gpu_available, make):
    sc = make()
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 5)                      # storage is cell-sorted by now
        x, v = g.positions(), g.velocities()
        e = export(g, capi.PS_FIELD_POSITIONS)
        assert_bitexact(e[:, :3], x, "exported positions")
        assert np.all(e[:, 3] == 1.0)
        e = export(g, capi.PS_FIELD_VELOCITIES)
        assert_bitexact(e[:, :3], v, "exported velocities")
        assert np.all(e[:, 3] == 0.0)
        assert np.array_equal(export(g, capi.PS_FIELD_PHASES), np.asarray(sc.phase, np.int32))
        a, k = sc.n // 3, sc.n // 4           # a sub-range
        e = export(g, capi.PS_FIELD_POSITIONS, a, k)
        assert_bitexact(e[:, :3], x[a:a + k], "exported sub-range")


def test_export_argument_errors(gpu_available):
    sc = scenes.dam_break(4, 4, 4)
    with System(sc.params) as g:
        sc.load_into(g)
        with capi.DeviceBuffer(16 * sc.n) as buf:
            for field, first, n in ((3, 0, sc.n), (8, 0, sc.n), (1, 1, sc.n), (1, sc.n + 1, 0)):
                with pytest.raises(capi.PsError) as e:
                    g.export_device(field, buf.ptr, first, n)
                assert e.value.code == capi.PS_ERR_INVALID_ARGUMENT
            g.export_device(1, buf.ptr, sc.n, 0)      # empty range at the end


def test_trajectory_round_trip(gpu_available, tmp_path):
    sc = scenes.dam_break(10, 8, 10)
    path = tmp_path / "dam.pstraj"
    expect = []
    with System(sc.params) as g, trajectory.Writer(path, 7, 3) as w:
        sc.load_into(g)
        for _ in range(4):
            w.write(g)
            expect.append((g.status().substeps, g.positions(), g.velocities()))
            g.step(sc.dt, 3)
    got = trajectory.read(path)
    assert len(got) == 4
    for f, (sub, x, v) in zip(got, expect):
        assert f.substeps == sub
        assert_bitexact(f.positions, x, "trajectory positions")
        assert_bitexact(f.velocities, v, "trajectory velocities")
        assert np.array_equal(f.phases, np.asarray(sc.phase, np.int32))


def test_trajectory_follows_a_growing_system(gpu_available, tmp_path):
    # the open smoke scene's emitter adds a gas particle every 1/15 s
    sc = scenes.reference_smoke(open_gas=True, emitter=True)
    path = tmp_path / "smoke.pstraj"
    xs = []
    with System(sc.params) as g, trajectory.Writer(path, capi.PS_FIELD_POSITIONS | capi.PS_FIELD_PHASES,
                                                  2) as w:
        sc.load_into(g)
        for _ in range(5):
            g.step(sc.dt, 10)
            w.write(g)
            xs.append(g.positions())
    got = trajectory.read(path)
    assert [f.n for f in got] == [x.shape[0] for x in xs] == [sc.n + k for k in (1, 2, 4, 5, 7)]
    for f, x in zip(got, xs):
        assert_bitexact(f.positions, x, "smoke trajectory")
        assert np.all(f.phases == 1)            # PS_PHASE_GAS
