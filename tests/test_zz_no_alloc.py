"""ps_grid_sort_noalloc (its workspace is the library's static device array)
against ps_grid_sort (workspace allocated per call), bit for bit, on two
calls in a row on one GPU, the second with fewer inputs on fewer cells, so
that it runs over what the first left in the array; and its refusal of
sizes beyond its limits."""
import ctypes

import numpy as np
import pytest

from particlesolver_amd import capi

pytestmark = pytest.mark.gpu


def cloud(seed, n, origin, cell, grid):
    rng = np.random.default_rng(seed)
    lo = np.array(origin, np.float64)
    hi = lo + cell * np.array(grid)
    pos = rng.uniform(lo - 0.5, hi + 0.5, (n, 3))
    pos[: n // 10] = lo + cell * rng.integers(0, 4, (n // 10, 3))   # crowded corner cells
    w = rng.uniform(0.5, 2.0, (n, 1))
    return np.concatenate([pos, w], axis=1).astype(np.float32)


def grid_sort(name, pos4, origin, cell, grid):
    n = pos4.shape[0]
    cells = int(np.prod(grid))
    g = capi.PsGridDesc()
    g.origin[:] = origin
    g.cell = cell
    g.grid[:] = grid
    lib = capi.load()
    bufs = [capi.DeviceBuffer(max(16 * n, 4)), capi.DeviceBuffer(4 * max(n, 1)), capi.DeviceBuffer(4 * max(n, 1)),
            capi.DeviceBuffer(4 * cells), capi.DeviceBuffer(4 * cells), capi.DeviceBuffer(max(16 * n, 4))]
    try:
        dpos, dkey, didx, dstart, dend, dsorted = bufs
        if n:
            dpos.upload(pos4)
        for b in (dkey, didx, dstart, dend, dsorted):
            b.fill(0xAB)
        rc = getattr(lib, name)(ctypes.byref(g), n, dpos.ptr, dkey.ptr, didx.ptr, dstart.ptr, dend.ptr,
                                dsorted.ptr, None)
        assert rc == 0, (name, lib.ps_last_error(None))
        return [dkey.numpy(np.uint32, (n,)), didx.numpy(np.uint32, (n,)), dstart.numpy(np.uint32, (cells,)),
                dend.numpy(np.uint32, (cells,)), dsorted.numpy(np.float32, (n, 4)).view(np.uint32)]
    finally:
        for b in bufs:
            b.free()


CASES = [
    (0x5EED0701, 300000, (-1.0, -1.0, -1.0), 0.125, (64, 32, 64)),   # 131,072 cells
    (0x5EED0702, 5000, (0.0, 0.0, 0.0), 0.5, (8, 4, 8)),             # then far fewer of both
]


def test_noalloc_equals_allocating_variant_on_two_calls_in_a_row():
    got = []
    for seed, n, origin, cell, grid in CASES:
        print("noalloc case: seed %#x n %d grid %s" % (seed, n, grid), flush=True)
        got.append(grid_sort("ps_grid_sort_noalloc", cloud(seed, n, origin, cell, grid), origin, cell, grid))
    for (seed, n, origin, cell, grid), mine in zip(CASES, got):
        ref = grid_sort("ps_grid_sort", cloud(seed, n, origin, cell, grid), origin, cell, grid)
        for a, b in zip(mine, ref):
            assert np.array_equal(a, b), (seed, n)
        # and the order is the stable (cell, input index) one
        key, idx = ref[0], ref[1]
        assert np.all((key[1:] > key[:-1]) | ((key[1:] == key[:-1]) & (idx[1:] > idx[:-1])))


def test_noalloc_refuses_sizes_beyond_its_limits():
    lib = capi.load()
    g = capi.PsGridDesc()
    g.origin[:] = (0.0, 0.0, 0.0)
    g.cell = 1.0
    g.grid[:] = (16, 16, 16)
    dummy = 0x1000  # never dereferenced: both calls fail their checks first
    rc = lib.ps_grid_sort_noalloc(ctypes.byref(g), capi.PS_GRID_NOALLOC_MAX_PARTICLES + 1, dummy, dummy, dummy,
                                  dummy, dummy, None, None)
    assert rc == capi.PS_ERR_CAPACITY
    assert b"PS_GRID_NOALLOC_MAX_PARTICLES" in lib.ps_last_error(None)
    g.grid[:] = (256, 128, 128)   # 4,194,304 cells
    rc = lib.ps_grid_sort_noalloc(ctypes.byref(g), 1, dummy, dummy, dummy, dummy, dummy, None, None)
    assert rc == capi.PS_ERR_CAPACITY
    assert b"PS_GRID_NOALLOC_MAX_CELLS" in lib.ps_last_error(None)
