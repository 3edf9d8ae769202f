"""ctypes binding of the CPU oracle (liboracle_f32.so / liboracle_f64.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke entry point and
by the CPU-baseline leg of bench.py.  The product library never loads it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_libs = {}


def _lib(prec, path=None):
    """The oracle library of precision prec; `path` names another build of
    the same sources (the mutation tests load deliberately broken copies).
    PS_ORACLE_DIR, when set, names the folder of the default builds (the
    sanitizer builds of `oracle/build.py --sanitize` live in oracle/_san)."""
    key = (prec, path)
    if key not in _libs:
        if path is None:
            path = os.path.join(os.environ.get("PS_ORACLE_DIR") or HERE, "liboracle_%s.so" % prec)
        if not os.path.exists(path):
            raise RuntimeError("oracle library %s missing: run oracle/build.py" % path)
        _libs[key] = ctypes.CDLL(path)
    return _libs[key]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Oracle:
    """CPU restatement of the substep, FP32 ('f32') or FP64 ('f64')."""

    def __init__(self, params, prec="f32", lib=None):
        self.prec = prec
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.real = ctypes.c_float if prec == "f32" else ctypes.c_double
        self.L = _lib(prec, lib)
        self.params = params.copy()
        s = self.prec
        self._f = lambda name: getattr(self.L, "%s_%s" % (name, s))
        h = ctypes.c_void_p()
        rc = self._f("orc_create")(ctypes.byref(self.params), ctypes.byref(h))
        if rc:
            raise RuntimeError("orc_create failed: %d" % rc)
        self.h = h
        self._f("orc_num_particles").restype = ctypes.c_uint32
        for name in ("orc_predict", "orc_begin", "orc_step"):
            fn = self._f(name)
            fn.argtypes = [ctypes.c_void_p, self.real] + ([ctypes.c_uint32] if name == "orc_step" else [])
        self._f("orc_wall_uniform").restype = ctypes.c_double
        self._f("orc_wall_uniform").argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32,
                                                ctypes.c_uint32]

    def close(self):
        if getattr(self, "h", None):
            self._f("orc_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc:
            raise RuntimeError("oracle call failed: %d" % rc)

    @property
    def n(self):
        return self._f("orc_num_particles")(self.h)

    def add_particles(self, pos, vel=None, inv_mass=None, phase=None, rest_density=None):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        ph = np.ascontiguousarray(np.broadcast_to(0 if phase is None else phase, (n,)),
                                  dtype=np.int32)
        r0 = np.ascontiguousarray(np.broadcast_to(1.0 if rest_density is None else rest_density,
                                                  (n,)), dtype=np.float32)
        self._check(self._f("orc_add_particles")(self.h, ctypes.c_uint32(n), _ptr(pos), _ptr(vel),
                                                 _ptr(w), _ptr(ph), _ptr(r0)))

    def add_distance_constraints(self, pairs, rest, stiffness=None):
        pairs = np.ascontiguousarray(pairs, dtype=np.uint32).reshape(-1, 2)
        rest = np.ascontiguousarray(rest, dtype=np.float32).reshape(-1)
        st = None if stiffness is None else np.ascontiguousarray(stiffness, dtype=np.float32)
        self._check(self._f("orc_add_distance_constraints")(self.h, ctypes.c_uint32(pairs.shape[0]),
                                                            _ptr(pairs), _ptr(rest), _ptr(st)))

    def add_point_constraints(self, idx, xyz):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        xyz = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
        self._check(self._f("orc_add_point_constraints")(self.h, ctypes.c_uint32(idx.shape[0]),
                                                         _ptr(idx), _ptr(xyz)))

    def add_rigid_body(self, idx, stiffness=0.0):
        idx = np.ascontiguousarray(idx, dtype=np.uint32).reshape(-1)
        b = ctypes.c_uint32()
        self._check(self._f("orc_add_rigid_body")(self.h, ctypes.c_uint32(idx.shape[0]), _ptr(idx),
                                                  ctypes.c_float(stiffness), ctypes.byref(b)))
        return b.value

    def add_rigid_bodies(self, offsets, idx, stiffness=None):
        offsets = np.asarray(offsets).reshape(-1)
        idx = np.asarray(idx).reshape(-1)
        first = None
        for b in range(len(offsets) - 1):
            st = 0.0 if stiffness is None else float(stiffness[b])
            bid = self.add_rigid_body(idx[offsets[b]:offsets[b + 1]], st)
            first = bid if first is None else first
        return first

    def set_params(self, params):
        self.params = params.copy()
        self._check(self._f("orc_set_params")(self.h, ctypes.byref(self.params)))

    def step(self, dt, substeps=1):
        self._check(self._f("orc_step")(self.h, dt, substeps))

    def begin(self, dt):
        self._check(self._f("orc_begin")(self.h, dt))

    def iterate(self):
        self._check(self._f("orc_iterate")(self.h))

    def end(self):
        self._check(self._f("orc_end")(self.h))

    def compute_lambda(self):
        """lambda pass only; returns lambda per sorted slot"""
        self._check(self._f("orc_compute_lambda")(self.h))
        return self.sorted_state()[0][:, 3].copy()

    def predict(self, dt):
        self._check(self._f("orc_predict")(self.h, dt))

    def build_grid(self):
        self._check(self._f("orc_build_grid")(self.h))

    def find_neighbors(self):
        self._check(self._f("orc_find_neighbors")(self.h))

    def positions(self):
        out = np.empty((self.n, 3), self.dtype)
        self._check(self._f("orc_read_positions")(self.h, _ptr(out)))
        return out

    def velocities(self):
        out = np.empty((self.n, 3), self.dtype)
        self._check(self._f("orc_read_velocities")(self.h, _ptr(out)))
        return out

    def grid(self):
        n = self.n
        nc = self.params.grid[0] * self.params.grid[1] * self.params.grid[2]
        idx = np.empty(n, np.uint32)
        cell = np.empty(n, np.uint32)
        off = np.empty(nc + 1, np.uint32)
        self._check(self._f("orc_read_grid")(self.h, _ptr(idx), _ptr(cell), _ptr(off)))
        return idx, cell, off

    def neighbors(self):
        n, k = self.n, self.params.max_neighbors
        cnt = np.empty(n, np.uint32)
        tab = np.empty((n, k), np.uint32)
        self._check(self._f("orc_read_neighbors")(self.h, _ptr(cnt), _ptr(tab)))
        # entries past the count are unspecified; zero them like the library does
        tab[np.arange(k)[None, :] >= cnt[:, None]] = 0
        return cnt, tab

    def sorted_state(self):
        n = self.n
        it = np.empty((n, 4), self.dtype)
        st = np.empty((n, 3), self.dtype)
        self._check(self._f("orc_read_sorted_state")(self.h, _ptr(it), _ptr(st)))
        return it, st

    def status(self):
        from particlesolver_amd.capi import PsStatus
        s = PsStatus()
        self._check(self._f("orc_get_status")(self.h, ctypes.byref(s)))
        return s

    def derived(self):
        out = np.empty(14, self.dtype)
        self._f("orc_derive")(ctypes.byref(self.params), _ptr(out))
        return dict(zip(["h2", "poly6", "spiky", "w_zero", "w_dq", "inv_w_dq", "tensile_k",
                         "search_r2", "collide_r2", "wall_s", "wall_k", "jitter", "w_scale", "w_scale_h2"],
                        out))

    def set_state(self, x=None, v=None):
        """positions / velocities in insertion order at the oracle's precision"""
        x = None if x is None else np.ascontiguousarray(x, dtype=self.dtype).reshape(-1, 3)
        v = None if v is None else np.ascontiguousarray(v, dtype=self.dtype).reshape(-1, 3)
        self._check(self._f("orc_set_state")(self.h, _ptr(x), _ptr(v)))

    def set_coefficients(self, rest_density=None, stiffness=None):
        """per-particle rest densities (insertion order) and per-constraint
        distance stiffnesses at the oracle's precision"""
        r0 = None if rest_density is None else np.ascontiguousarray(rest_density, dtype=self.dtype)
        st = None if stiffness is None else np.ascontiguousarray(stiffness, dtype=self.dtype)
        self._check(self._f("orc_set_coefficients")(self.h, _ptr(r0), _ptr(st)))

    def set_wall_k(self, wall_k):
        """kinetic wall-friction distance sqrt(mu_k) radius, full precision"""
        fn = self._f("orc_set_wall_k")
        fn.argtypes = [ctypes.c_void_p, self.real]
        self._check(fn(self.h, wall_k))

    def step_backward(self, dt, substeps, gx=None, gv=None, n_dist=0, with_mu=False):
        """Reverse-mode gradient of L(x_T, v_T) after `substeps` substeps from
        the current state, given gx = dL/dx_T and gv = dL/dv_T.  Returns
        (dL/dx_0, dL/dv_0, dL/drho0, dL/dstiffness) in insertion order, plus
        dL/dwall_mu_k when with_mu, and leaves the system in the final
        state."""
        n = self.n
        gx = None if gx is None else np.ascontiguousarray(gx, dtype=self.dtype).reshape(n, 3)
        gv = None if gv is None else np.ascontiguousarray(gv, dtype=self.dtype).reshape(n, 3)
        dx = np.empty((n, 3), self.dtype)
        dv = np.empty((n, 3), self.dtype)
        dr = np.empty(n, self.dtype)
        ds = np.empty(max(n_dist, 1), self.dtype)
        dm = np.zeros(1, self.dtype)
        fn = self._f("orc_step_backward")
        fn.argtypes = [ctypes.c_void_p, self.real, ctypes.c_uint32] + [ctypes.c_void_p] * 7
        rc = fn(self.h, dt, substeps, _ptr(gx), _ptr(gv), _ptr(dx), _ptr(dv), _ptr(dr), _ptr(ds),
                _ptr(dm))
        if rc:
            raise RuntimeError("orc_step_backward failed: %d" % rc)
        out = (dx, dv, dr, ds[:n_dist])
        return out + (float(dm[0]),) if with_mu else out

    def wall_uniform(self, seed, it, orig, wall):
        return self._f("orc_wall_uniform")(seed, it, orig, wall)

    def extract_rotation(self, A, q0=(1.0, 0.0, 0.0, 0.0), sweeps=10):
        A = np.ascontiguousarray(A, dtype=self.dtype).reshape(9)
        q = np.ascontiguousarray(q0, dtype=self.dtype)
        self._f("orc_extract_rotation")(_ptr(A), _ptr(q), ctypes.c_int(sweeps))
        return q


def load_scene(scene, prec="f32"):
    """An Oracle holding `scene` (a particlesolver_amd.scenes.Scene)."""
    o = Oracle(scene.params, prec)
    scene.load_into(o)
    return o


class Ref2D:
    """The reference's own CPU solver, Simulation::tick (ref2d.c), FP32 ('f32')
    or FP64 ('f64').  Particles enter as fluid objects (createFluid)."""

    def __init__(self, params, prec="f64"):
        self.prec = prec
        self.dtype = np.float32 if prec == "f32" else np.float64
        self.real = ctypes.c_float if prec == "f32" else ctypes.c_double
        self.L = _lib(prec)
        self.params = params.copy()
        self._f = lambda name: getattr(self.L, "%s_%s" % (name, prec))
        h = ctypes.c_void_p()
        rc = self._f("r2d_create")(ctypes.byref(self.params), ctypes.byref(h))
        if rc:
            raise RuntimeError("r2d_create failed: %d" % rc)
        self.h = h
        self._f("r2d_step").argtypes = [ctypes.c_void_p, self.real, ctypes.c_uint32]
        self._f("r2d_add_fluid_group").argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                                   ctypes.c_uint32, ctypes.c_float]
        self._f("r2d_add_gas_group").argtypes = [ctypes.c_void_p, ctypes.c_uint32,
                                                 ctypes.c_uint32, ctypes.c_float, ctypes.c_int]
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            self._f("r2d_destroy")(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_fluid_object(self, pos, vel=None, inv_mass=None, rest_density=1.0):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        ph = np.zeros(n, np.int32)
        first = self.n
        rc = self._f("r2d_add_particles")(self.h, n, _ptr(pos), _ptr(vel), _ptr(w), _ptr(ph))
        if rc:
            raise RuntimeError("r2d_add_particles failed: %d" % rc)
        rc = self._f("r2d_add_fluid_group")(self.h, first, n, float(rest_density))
        if rc:
            raise RuntimeError("r2d_add_fluid_group failed: %d" % rc)
        self.n += n
        return first

    def add_gas_object(self, pos, vel=None, inv_mass=None, rest_density=1.0, open=False):
        """createGas (simulation.cpp:408-428): a GasConstraint object, 2D"""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float32).reshape(n, 3)
        w = np.ascontiguousarray(np.broadcast_to(1.0 if inv_mass is None else inv_mass, (n,)),
                                 dtype=np.float32)
        ph = np.full(n, 1, np.int32)
        first = self.n
        rc = self._f("r2d_add_particles")(self.h, n, _ptr(pos), _ptr(vel), _ptr(w), _ptr(ph))
        if rc:
            raise RuntimeError("r2d_add_particles failed: %d" % rc)
        rc = self._f("r2d_add_gas_group")(self.h, first, n, float(rest_density), int(bool(open)))
        if rc:
            raise RuntimeError("r2d_add_gas_group failed: %d" % rc)
        self.n += n
        return first

    def add_static_particles(self, pos):
        """immovable solids, new Particle(pos, 0) (particle.h:31-52): mass
        0, phase SOLID; a contact partner of every other particle"""
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        n = pos.shape[0]
        w = np.zeros(n, np.float32)
        ph = np.full(n, 3, np.int32)
        first = self.n
        rc = self._f("r2d_add_particles")(self.h, n, _ptr(pos), None, _ptr(w), _ptr(ph))
        if rc:
            raise RuntimeError("r2d_add_particles failed: %d" % rc)
        self.n += n
        return first

    def add_fluid_emitter(self, pos, particles_per_sec, object):
        """createFluidEmitter (simulation.cpp:459-461) feeding fluid object
        `object` (index in creation order)"""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        f = self._f("r2d_add_fluid_emitter")
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32]
        rc = f(self.h, _ptr(p), float(particles_per_sec), int(object))
        if rc:
            raise RuntimeError("r2d_add_fluid_emitter failed: %d" % rc)

    def attrs(self):
        """(phase, inverse mass, solidification timer t) per particle"""
        ph = np.zeros(self.n, np.int32)
        w = np.zeros(self.n, self.dtype)
        t = np.zeros(self.n, self.dtype)
        self._f("r2d_read_attrs")(self.h, _ptr(ph), _ptr(w), _ptr(t))
        return ph, w, t

    def add_smoke_emitter(self, pos, particles_per_sec, object=-1, max_markers=4096):
        """createSmokeEmitter (simulation.cpp:454-457) feeding gas object
        `object` (an index in creation order, -1: markers only)"""
        p = np.ascontiguousarray(pos, dtype=np.float32).reshape(3)
        f = self._f("r2d_add_smoke_emitter")
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                      ctypes.c_uint32]
        rc = f(self.h, _ptr(p), float(particles_per_sec), int(object), int(max_markers))
        if rc:
            raise RuntimeError("r2d_add_smoke_emitter failed: %d" % rc)
        self.nem = getattr(self, "nem", 0) + 1
        return self.nem - 1

    def markers(self, emitter):
        f = self._f("r2d_read_markers")
        f.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_void_p]
        f.restype = ctypes.c_uint32
        m = f(self.h, emitter, None)
        out = np.zeros((m, 3), self.dtype)
        if m:
            f(self.h, emitter, _ptr(out))
        return out

    def forces(self):
        """the force each particle carries into the next prediction"""
        f = np.zeros((self.n, 3), self.dtype)
        self._f("r2d_read_force")(self.h, _ptr(f))
        return f

    def step(self, dt, steps=1):
        rc = self._f("r2d_step")(self.h, dt, steps)
        if rc:
            raise RuntimeError("r2d_step failed: %d" % rc)
        f = self._f("r2d_num_particles")
        f.restype = ctypes.c_uint32
        self.n = f(self.h)

    def _read(self):
        x = np.zeros((self.n, 3), self.dtype)
        v = np.zeros((self.n, 3), self.dtype)
        self._f("r2d_read")(self.h, _ptr(x), _ptr(v))
        return x, v

    def positions(self):
        return self._read()[0]

    def velocities(self):
        return self._read()[1]

    def constants(self):
        out = np.zeros(5, np.float64)
        self._f("r2d_constants")(self.h, _ptr(out))
        return dict(zip(("poly6", "spiky", "w_zero", "w_dq", "jitter"), out))

    def set_threads(self, n):
        self._f("r2d_set_threads")(ctypes.c_int(n))


def load_ref2d(scene, prec="f64"):
    """A Ref2D holding `scene` (a particlesolver_amd.scenes.FluidObjectsScene)."""
    r = Ref2D(scene.params, prec)
    scene.load_into(r)
    return r
