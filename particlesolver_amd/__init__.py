"""MI355X (gfx950) unified particle solver: the per-substep hot path of
ebirenbaum/ParticleSolver (predict, uniform-grid neighbour build, Jacobi
constraint projections, velocity + XSPH) as hand-written HIP kernels behind
the C-ABI in include/ps.h.

Python here is plumbing: `capi` binds the native library with ctypes,
`scenes` generates the benchmark/parity configurations and
`particle_system.ParticleSystem` is the reference's host class (scene
builders, deferred additions, update) over the library.
"""
from . import capi, scenes  # noqa: F401
from .capi import System, PsParams, PsError, default_params, derive, load  # noqa: F401
from .particle_system import ParticleSystem  # noqa: F401

__all__ = ["capi", "scenes", "System", "ParticleSystem", "PsParams", "PsError", "default_params", "derive",
           "load"]
