"""ctypes binding of the C-ABI declared in include/ps.h.

The library (libparticlesolver.so, built in-tree by __graft_entry__.build())
is the product: every call here goes straight into the HIP implementation.
There is no CPU fallback; if the library is missing, loading raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libparticlesolver.so")

PS_ABI_VERSION = 1

PHASE_FLUID = 0
PHASE_GAS = 1
PHASE_CLOTH = 2
PHASE_SOLID = 3
PHASE_RIGID = 4

PS_OK = 0
PS_ERR_INVALID_ARGUMENT = -1
PS_ERR_CAPACITY = -2
PS_GRID_NOALLOC_MAX_PARTICLES = 1 << 20
PS_GRID_NOALLOC_MAX_CELLS = 1 << 21
PS_ERR_DEVICE = -3
PS_ERR_NO_DEVICE = -4
PS_ERR_STATE = -5
PS_ERR_OUT_OF_MEMORY = -6
PS_ERR_NOT_FINITE = -7
PS_ERR_UNSUPPORTED = -8
PS_ERR_IO = -9

PS_FIELD_POSITIONS = 1
PS_FIELD_VELOCITIES = 2
PS_FIELD_PHASES = 4

PS_FLAG_REBUILD_EVERY_ITERATION = 0x1
PS_FLAG_SLEEPING = 0x2
PS_FLAG_CPU_REFERENCE = 0x4
PS_FLAG_STABILIZATION = 0x10
PS_FLAG_NO_LDS_STAGING = 0x8
PS_FLAG_UNSCALED_TENSILE = 0x20

PS_STATUS_NEIGHBOR_OVERFLOW = 0x1
PS_STATUS_NONFINITE = 0x2
PS_STATUS_OUTSIDE_GRID = 0x4
PS_STATUS_INTERNAL = 0x80000000

KERNELS = ["predict_hash", "scan", "scatter", "reorder", "neighbors", "lambda", "solve",
           "shape", "finalize", "adj_xsph", "adj_finalize", "adj_pre", "adj_lambda",
           "adj_position", "adj_stiffness", "adj_predict", "export"]
K_PREDICT_HASH, K_SCAN, K_SCATTER, K_REORDER, K_NEIGHBORS, K_LAMBDA, K_SOLVE, K_SHAPE, \
    K_FINALIZE = range(9)


class PsParams(ctypes.Structure):
    _fields_ = [
        ("abi_version", ctypes.c_uint32),
        ("dims", ctypes.c_uint32),
        ("radius", ctypes.c_float),
        ("h", ctypes.c_float),
        ("cell", ctypes.c_float),
        ("collide_dist", ctypes.c_float),
        ("gravity", ctypes.c_float * 3),
        ("box_min", ctypes.c_float * 3),
        ("box_max", ctypes.c_float * 3),
        ("origin", ctypes.c_float * 3),
        ("grid", ctypes.c_uint32 * 3),
        ("iterations", ctypes.c_uint32),
        ("relaxation", ctypes.c_float),
        ("k_corr", ctypes.c_float),
        ("n_corr", ctypes.c_uint32),
        ("dq_corr", ctypes.c_float),
        ("xsph_c", ctypes.c_float),
        ("s_solid", ctypes.c_float),
        ("mu_s", ctypes.c_float),
        ("mu_k", ctypes.c_float),
        ("wall_mu_s", ctypes.c_float),
        ("wall_mu_k", ctypes.c_float),
        ("wall_jitter", ctypes.c_float),
        ("shape_stiffness", ctypes.c_float),
        ("max_particles", ctypes.c_uint32),
        ("max_neighbors", ctypes.c_uint32),
        ("seed", ctypes.c_uint64),
        ("flags", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32 * 7),
    ]

    def copy(self):
        out = PsParams()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(PsParams))
        return out

    def as_dict(self):
        d = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            d[name] = list(v) if hasattr(v, "__len__") else v
        return d


class PsDerived(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in
                ("h2", "poly6", "spiky", "w_zero", "w_dq", "inv_w_dq", "tensile_k", "search_r2",
                 "collide_r2", "wall_s", "wall_k", "jitter")] + [
        ("num_cells", ctypes.c_uint32), ("w_scale", ctypes.c_float), ("w_scale_h2", ctypes.c_float),
        ("reserved", ctypes.c_uint32 * 1)]


class PsSceneDesc(ctypes.Structure):
    _fields_ = [("box_min", ctypes.c_float * 3), ("box_max", ctypes.c_float * 3),
                ("gravity_scale", ctypes.c_float), ("reserved0", ctypes.c_uint32),
                ("seed", ctypes.c_uint64), ("reserved", ctypes.c_uint32 * 4)]


class PsGridDesc(ctypes.Structure):
    _fields_ = [("origin", ctypes.c_float * 3), ("cell", ctypes.c_float),
                ("grid", ctypes.c_uint32 * 3), ("reserved", ctypes.c_uint32)]


class PsStatus(ctypes.Structure):
    _fields_ = [("flags", ctypes.c_uint32), ("max_neighbors", ctypes.c_uint32),
                ("num_particles", ctypes.c_uint32), ("num_cells", ctypes.c_uint32),
                ("substeps", ctypes.c_uint64), ("iterations", ctypes.c_uint64),
                ("units", ctypes.c_uint32), ("staged_units", ctypes.c_uint32)]


_P = ctypes.c_void_p
_u32 = ctypes.c_uint32
_f32 = ctypes.c_float
_int = ctypes.c_int

# name -> (restype, argtypes); the full exported surface of include/ps.h
PROTOTYPES = {
    "ps_params_default": (_int, [ctypes.POINTER(PsParams)]),
    "ps_params_validate": (_int, [ctypes.POINTER(PsParams)]),
    "ps_derive": (_int, [ctypes.POINTER(PsParams), ctypes.POINTER(PsDerived)]),
    "ps_create": (_int, [ctypes.POINTER(PsParams), ctypes.POINTER(_P)]),
    "ps_create_batch": (_int, [ctypes.POINTER(PsParams), _u32, _P, ctypes.POINTER(_P)]),
    "ps_num_scenes": (_u32, [_P]),
    "ps_add_scene_particles": (_int, [_P, _u32, _u32, _P, _P, _P, _P, _P, ctypes.POINTER(_u32)]),
    "ps_step_graph": (_int, [_P, _f32, _u32]),
    "ps_step_backward": (_int, [_P, _f32, _u32, _u32, _P, _P, _P, _P, _P, _P, _P]),
    "ps_destroy": (_int, [_P]),
    "ps_set_params": (_int, [_P, ctypes.POINTER(PsParams)]),
    "ps_get_params": (_int, [_P, ctypes.POINTER(PsParams)]),
    "ps_last_error": (ctypes.c_char_p, [_P]),
    "ps_add_particles": (_int, [_P, _u32, _P, _P, _P, _P, _P, ctypes.POINTER(_u32)]),
    "ps_add_fluid_object": (_int, [_P, _u32, _P, _P, _P, _f32, ctypes.POINTER(_u32)]),
    "ps_add_gas_object": (_int, [_P, _u32, _P, _P, _P, _f32, _int, ctypes.POINTER(_u32)]),
    "ps_add_smoke_emitter": (_int, [_P, _P, ctypes.c_double, ctypes.c_int32, _u32,
                                    ctypes.POINTER(_u32)]),
    "ps_read_markers": (_int, [_P, _u32, ctypes.POINTER(_u32), _P]),
    "ps_add_fluid_emitter": (_int, [_P, _P, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(_u32)]),
    "ps_export_device": (_int, [_P, _u32, _u32, _u32, _P]),
    "ps_timing_sample": (_int, [_P, _u32]),
    "ps_trajectory_open": (_int, [ctypes.c_char_p, _u32, _u32, ctypes.POINTER(_P)]),
    "ps_trajectory_write": (_int, [_P, _P]),
    "ps_trajectory_close": (_int, [_P]),
    "ps_add_distance_constraints": (_int, [_P, _u32, _P, _P, _P]),
    "ps_num_distance_constraints": (_u32, [_P]),
    "ps_add_point_constraints": (_int, [_P, _u32, _P, _P]),
    "ps_add_rigid_body": (_int, [_P, _u32, _P, _f32, ctypes.POINTER(_u32)]),
    "ps_add_rigid_bodies": (_int, [_P, _u32, _P, _P, _P, ctypes.POINTER(_u32)]),
    "ps_set_velocities": (_int, [_P, _u32, _u32, _P]),
    "ps_step": (_int, [_P, _f32, _u32]),
    "ps_substep_begin": (_int, [_P, _f32]),
    "ps_substep_iterate": (_int, [_P]),
    "ps_substep_end": (_int, [_P]),
    "ps_predict": (_int, [_P, _f32]),
    "ps_build_grid": (_int, [_P]),
    "ps_find_neighbors": (_int, [_P]),
    "ps_synchronize": (_int, [_P]),
    "ps_num_particles": (_u32, [_P]),
    "ps_read_positions": (_int, [_P, _u32, _u32, _P]),
    "ps_read_velocities": (_int, [_P, _u32, _u32, _P]),
    "ps_read_grid": (_int, [_P, _P, _P, _P]),
    "ps_read_neighbors": (_int, [_P, _P, _P]),
    "ps_read_sorted_state": (_int, [_P, _P, _P]),
    "ps_get_status": (_int, [_P, ctypes.POINTER(PsStatus)]),
    "ps_clear_status": (_int, [_P]),
    "ps_timing_enable": (_int, [_P, _u32]),
    "ps_timing_reset": (_int, [_P]),
    "ps_timing_read": (_int, [_P, _int, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_uint64)]),
    "ps_kernel_name": (ctypes.c_char_p, [_int]),
    "ps_autotune": (_int, [_P, _f32, _u32, _P]),
    "ps_get_tuning": (_int, [_P, _P]),
    "ps_set_tuning": (_int, [_P, _P]),
    "ps_grid_sort": (_int, [ctypes.POINTER(PsGridDesc), _u32, _P, _P, _P, _P, _P, _P, _P]),
    "ps_grid_sort_noalloc": (_int, [ctypes.POINTER(PsGridDesc), _u32, _P, _P, _P, _P, _P, _P, _P]),
}

_lib = None


def load(path=None):
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError("native library %s is missing: run __graft_entry__.build()" % p)
    lib = ctypes.CDLL(p)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself links
    (libamdhip64.so.7), for ps_export_device in tests and tools: no second
    runtime (such as the one bundled with torch) enters the process."""

    def __init__(self, nbytes):
        load()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(max(self.nbytes, 1)))
        if rc:
            raise PsError(PS_ERR_OUT_OF_MEMORY, "hipMalloc(%d) failed: %d" % (self.nbytes, rc))
        self.ptr = p.value

    def fill(self, byte):
        """set every byte; waits for the fill, which runs on the null stream
        and so is not ordered with a system's (non-blocking) stream"""
        rc = self.hip.hipMemset(ctypes.c_void_p(self.ptr), int(byte), ctypes.c_size_t(self.nbytes))
        if not rc:
            rc = self.hip.hipDeviceSynchronize()
        if rc:
            raise PsError(PS_ERR_DEVICE, "hipMemset failed: %d" % rc)

    def upload(self, a):
        """synchronous copy of a host array into the start of the buffer"""
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        rc = self.hip.hipMemcpy(ctypes.c_void_p(self.ptr), a.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_size_t(a.nbytes), 1)       # hipMemcpyHostToDevice
        if rc:
            raise PsError(PS_ERR_DEVICE, "hipMemcpy failed: %d" % rc)

    def numpy(self, dtype, shape):
        """synchronous copy to the host (hipMemcpy waits for prior work)"""
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        rc = self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.ptr),
                                ctypes.c_size_t(out.nbytes), 2)     # hipMemcpyDeviceToHost
        if rc:
            raise PsError(PS_ERR_DEVICE, "hipMemcpy failed: %d" % rc)
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.free()


class PsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (code %d)" % (msg, code))
        self.code = code


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def default_params():
    lib = load()
    p = PsParams()
    rc = lib.ps_params_default(ctypes.byref(p))
    assert rc == 0
    return p


def derive(params):
    lib = load()
    d = PsDerived()
    rc = lib.ps_derive(ctypes.byref(params), ctypes.byref(d))
    if rc:
        raise PsError(rc, lib.ps_last_error(None).decode())
    return d


class System:
    """One particle system on the GPU (wraps a ps_system handle)."""

    def __init__(self, params, scenes=None):
        """scenes: None for one scene, or a list of PsSceneDesc (a batch)."""
        self.lib = load()
        self.params = params.copy()
        h = ctypes.c_void_p()
        if scenes is None:
            rc = self.lib.ps_create(ctypes.byref(self.params), ctypes.byref(h))
        else:
            arr = (PsSceneDesc * len(scenes))(*scenes)
            rc = self.lib.ps_create_batch(ctypes.byref(self.params), len(scenes),
                                          ctypes.cast(arr, ctypes.c_void_p), ctypes.byref(h))
        if rc:
            raise PsError(rc, self.lib.ps_last_error(None).decode())
        self.h = h

    # -- plumbing
    def _check(self, rc):
        if rc:
            raise PsError(rc, self.lib.ps_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ps_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scene
    def add_particles(self, pos, vel=None, inv_mass=None, phase=None, rest_density=None,
                      scene=0):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        ph = np.ascontiguousarray(np.broadcast_to(PHASE_FLUID if phase is None else phase, (n,)),
                                  dtype=np.int32)
        r0 = np.ascontiguousarray(np.broadcast_to(1.0 if rest_density is None else rest_density,
                                                  (n,)), dtype=np.float32)
        first = ctypes.c_uint32()
        self._check(self.lib.ps_add_scene_particles(self.h, scene, n, _ptr(pos), _ptr(vel), _ptr(w),
                                                    _ptr(ph), _ptr(r0), ctypes.byref(first)))
        return first.value

    def add_fluid_object(self, pos, vel=None, inv_mass=None, rest_density=1.0):
        """One fluid object (Simulation::createFluid): in CPU-reference mode its
        particles form one density constraint solved as a unit."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        first = ctypes.c_uint32()
        self._check(self.lib.ps_add_fluid_object(self.h, n, _ptr(pos), _ptr(vel), _ptr(w),
                                                 float(rest_density), ctypes.byref(first)))
        return first.value

    def add_gas_object(self, pos, vel=None, inv_mass=None, rest_density=1.0, open=False):
        """ps_add_gas_object: one GasConstraint object (CPU-reference mode, 2D)."""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        first = ctypes.c_uint32()
        self._check(self.lib.ps_add_gas_object(self.h, n, _ptr(pos), _ptr(vel), _ptr(w),
                                               float(rest_density), int(bool(open)),
                                               ctypes.byref(first)))
        return first.value

    def add_smoke_emitter(self, pos, particles_per_sec, object=-1, max_markers=4096):
        """ps_add_smoke_emitter: an OpenSmokeEmitter at pos feeding gas object
        `object` (-1: markers only); returns the emitter index."""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        e = ctypes.c_uint32()
        self._check(self.lib.ps_add_smoke_emitter(self.h, _ptr(p), float(particles_per_sec),
                                                  int(object), int(max_markers), ctypes.byref(e)))
        return e.value

    def add_fluid_emitter(self, pos, particles_per_sec, object):
        """ps_add_fluid_emitter: a FluidEmitter at pos feeding fluid object
        `object`; returns the emitter index."""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        e = ctypes.c_uint32()
        self._check(self.lib.ps_add_fluid_emitter(self.h, _ptr(p), float(particles_per_sec),
                                                  int(object), ctypes.byref(e)))
        return e.value

    def add_static_particles(self, pos):
        """immovable solids (phase SOLID, inverse mass 0) through
        ps_add_particles"""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        return self.add_particles(pos, None, np.zeros(n, np.float32),
                                  np.full(n, PHASE_SOLID, np.int32), np.zeros(n, np.float32))

    def phases(self):
        """phase of every particle, insertion order (through ps_export_device)"""
        with DeviceBuffer(4 * max(self.n, 1)) as buf:
            if self.n:
                self.export_device(PS_FIELD_PHASES, buf.ptr)
                self.synchronize()
            return buf.numpy(np.int32, (self.n,))

    def markers(self, emitter):
        """ps_read_markers: the emitter's marker positions, (count, 3)"""
        c = ctypes.c_uint32()
        self._check(self.lib.ps_read_markers(self.h, emitter, ctypes.byref(c), None))
        out = np.empty((c.value, 3), dtype=np.float32)
        if c.value:
            self._check(self.lib.ps_read_markers(self.h, emitter, ctypes.byref(c), _ptr(out)))
        return out

    def export_device(self, field, dst_ptr, first=0, n=None):
        """ps_export_device: field PS_FIELD_* of [first, first+n) into the
        device pointer dst_ptr (an int), enqueued on the system's stream"""
        n = self.n - first if n is None else n
        self._check(self.lib.ps_export_device(self.h, field, first, n, ctypes.c_void_p(dst_ptr)))

    def add_distance_constraints(self, pairs, rest, stiffness=None):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        rest = np.ascontiguousarray(rest, dtype=np.float32).reshape(-1)
        st = None if stiffness is None else np.ascontiguousarray(stiffness, dtype=np.float32)
        self._check(self.lib.ps_add_distance_constraints(self.h, pairs.shape[0], _ptr(pairs),
                                                         _ptr(rest), _ptr(st)))

    @property
    def n_dist(self):
        """distance constraints held by the system (ps_num_distance_constraints)"""
        return self.lib.ps_num_distance_constraints(self.h)

    def add_point_constraints(self, idx, xyz):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self._check(self.lib.ps_add_point_constraints(self.h, idx.shape[0], _ptr(idx), _ptr(xyz)))

    def add_rigid_body(self, idx, stiffness=0.0):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        b = ctypes.c_uint32()
        self._check(self.lib.ps_add_rigid_body(self.h, idx.shape[0], _ptr(idx), stiffness,
                                               ctypes.byref(b)))
        return b.value

    def add_rigid_bodies(self, offsets, idx, stiffness=None):
        offsets = np.ascontiguousarray(offsets, dtype=np.uint32).reshape(-1)
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        st = None if stiffness is None else np.ascontiguousarray(stiffness, dtype=np.float32)
        b = ctypes.c_uint32()
        self._check(self.lib.ps_add_rigid_bodies(self.h, offsets.shape[0] - 1, _ptr(offsets),
                                                 _ptr(idx), _ptr(st), ctypes.byref(b)))
        return b.value

    def set_velocities(self, vel, first=0):
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        self._check(self.lib.ps_set_velocities(self.h, first, vel.shape[0], _ptr(vel)))

    def set_params(self, params):
        self._check(self.lib.ps_set_params(self.h, ctypes.byref(params)))
        self.params = params.copy()

    # -- stepping
    def step(self, dt, substeps=1):
        self._check(self.lib.ps_step(self.h, dt, substeps))

    def step_graph(self, dt, substeps):
        self._check(self.lib.ps_step_graph(self.h, dt, substeps))

    def step_backward(self, dt, substeps, dl_dx=None, dl_dv=None, checkpoint_every=0):
        """ps_step_backward: runs `substeps` substeps and returns
        (dL/dx0, dL/dv0, dL/drho0, dL/dstiffness, dL/dwall_mu_k) for the
        loss gradients dl_dx, dl_dv of the final state (insertion order)."""
        n = self.n
        gx = None if dl_dx is None else np.ascontiguousarray(dl_dx, dtype=np.float32).reshape(n, 3)
        gv = None if dl_dv is None else np.ascontiguousarray(dl_dv, dtype=np.float32).reshape(n, 3)
        dx = np.empty((n, 3), np.float32)
        dv = np.empty((n, 3), np.float32)
        dr = np.empty(n, np.float32)
        nd = self.lib.ps_num_distance_constraints(self.h)
        ds = np.empty(max(nd, 1), np.float32)
        # dL/dwall_mu_k is None where wall_mu_k = 0 (no finite derivative there)
        dm = np.zeros(1, np.float32) if self.params.wall_mu_k > 0.0 else None
        self._check(self.lib.ps_step_backward(self.h, dt, substeps, checkpoint_every, _ptr(gx),
                                              _ptr(gv), _ptr(dx), _ptr(dv), _ptr(dr), _ptr(ds),
                                              _ptr(dm)))
        return dx, dv, dr, ds[:nd], None if dm is None else float(dm[0])

    def begin(self, dt):
        self._check(self.lib.ps_substep_begin(self.h, dt))

    def iterate(self):
        self._check(self.lib.ps_substep_iterate(self.h))

    def end(self):
        self._check(self.lib.ps_substep_end(self.h))

    def predict(self, dt):
        self._check(self.lib.ps_predict(self.h, dt))

    def build_grid(self):
        self._check(self.lib.ps_build_grid(self.h))

    def find_neighbors(self):
        self._check(self.lib.ps_find_neighbors(self.h))

    def synchronize(self):
        self._check(self.lib.ps_synchronize(self.h))

    # -- readback
    @property
    def n(self):
        return self.lib.ps_num_particles(self.h)

    def positions(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        self._check(self.lib.ps_read_positions(self.h, 0, self.n, _ptr(out)))
        return out

    def velocities(self):
        out = np.empty((self.n, 3), dtype=np.float32)
        self._check(self.lib.ps_read_velocities(self.h, 0, self.n, _ptr(out)))
        return out

    def grid(self):
        n = self.n
        nc = self.params.grid[0] * self.params.grid[1] * self.params.grid[2]
        idx = np.empty(n, np.uint32)
        cell = np.empty(n, np.uint32)
        off = np.empty(nc + 1, np.uint32)
        self._check(self.lib.ps_read_grid(self.h, _ptr(idx), _ptr(cell), _ptr(off)))
        return idx, cell, off

    def neighbors(self):
        n, k = self.n, self.params.max_neighbors
        cnt = np.empty(n, np.uint32)
        tab = np.empty((n, k), np.uint32)
        self._check(self.lib.ps_read_neighbors(self.h, _ptr(cnt), _ptr(tab)))
        return cnt, tab

    def sorted_state(self):
        n = self.n
        it = np.empty((n, 4), np.float32)
        st = np.empty((n, 3), np.float32)
        self._check(self.lib.ps_read_sorted_state(self.h, _ptr(it), _ptr(st)))
        return it, st

    def status(self):
        s = PsStatus()
        self._check(self.lib.ps_get_status(self.h, ctypes.byref(s)))
        return s

    def clear_status(self):
        self._check(self.lib.ps_clear_status(self.h))

    # -- tuning
    def autotune(self, dt, substeps=2):
        out = np.zeros(len(KERNELS), np.uint32)
        self._check(self.lib.ps_autotune(self.h, dt, substeps, _ptr(out)))
        return dict(zip(KERNELS, out.tolist()))

    def tuning(self):
        out = np.zeros(len(KERNELS), np.uint32)
        self._check(self.lib.ps_get_tuning(self.h, _ptr(out)))
        return dict(zip(KERNELS, out.tolist()))

    def set_tuning(self, blocks):
        arr = np.array([blocks.get(k, 256) for k in KERNELS], np.uint32)
        self._check(self.lib.ps_set_tuning(self.h, _ptr(arr)))

    # -- timing
    def timing(self, kernels=True):
        """True: every kernel; False: none; or an iterable of kernel names."""
        if kernels is True:
            mask = 0xFFFFFFFF
        elif not kernels:
            mask = 0
        else:
            mask = 0
            for name in kernels:
                mask |= 1 << KERNELS.index(name)
        self._check(self.lib.ps_timing_enable(self.h, mask))

    def timing_reset(self):
        self._check(self.lib.ps_timing_reset(self.h))

    def timing_sample(self, period):
        """ps_timing_sample: events around one launch in `period`"""
        self._check(self.lib.ps_timing_sample(self.h, period))

    def timing_read(self):
        out = {}
        for k, name in enumerate(KERNELS):
            ms = ctypes.c_double()
            cnt = ctypes.c_uint64()
            self._check(self.lib.ps_timing_read(self.h, k, ctypes.byref(ms), ctypes.byref(cnt)))
            out[name] = (ms.value, cnt.value)
        return out
