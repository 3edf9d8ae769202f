"""torch.autograd.Function around ps_step / ps_step_backward (test harness
only; the library itself has no torch dependency).

SubstepsFn.apply(x0, v0, rho0, stiffness, runner) runs `runner.T` substeps
of the runner's scene from positions x0, velocities v0, per-particle rest
densities rho0 and per-constraint stiffnesses, and returns (x_T, v_T).  Its
backward asks the runner for the reverse pass.  Two runners:

  GpuRunner    the HIP library through the C-ABI (System.step /
               System.step_backward), FP32 host tensors;
  OracleRunner the CPU oracle in FP64 (Oracle.step / Oracle.step_backward),
               for torch.autograd.gradcheck."""
import numpy as np
import torch

from particlesolver_amd import System, scenes
from oracle import Oracle


def _scene_with(sc, x0, v0, rho0, stiffness):
    out = scenes.Scene(sc.params, np.asarray(x0, np.float32), sc.inv_mass, sc.phase,
                       np.asarray(rho0, np.float32), sc.dt, vel=np.asarray(v0, np.float32),
                       name=sc.name)
    if sc.distance is not None:
        out.distance = (sc.distance[0], sc.distance[1], np.asarray(stiffness, np.float32))
    out.pins = sc.pins
    out.bodies = sc.bodies
    return out


class GpuRunner:
    def __init__(self, scene, T, checkpoint_every=0):
        self.sc, self.T, self.every = scene, T, checkpoint_every

    def _system(self, x0, v0, rho0, stiffness):
        g = System(self.sc.params)
        _scene_with(self.sc, x0, v0, rho0, stiffness).load_into(g)
        return g

    def forward(self, x0, v0, rho0, stiffness):
        with self._system(x0, v0, rho0, stiffness) as g:
            g.step(self.sc.dt, self.T)
            return g.positions(), g.velocities()

    def backward(self, x0, v0, rho0, stiffness, gx, gv):
        with self._system(x0, v0, rho0, stiffness) as g:
            dx, dv, dr, ds, _ = g.step_backward(self.sc.dt, self.T, gx, gv,
                                                checkpoint_every=self.every)
        return dx, dv, dr, ds


class OracleRunner:
    """Full-precision inputs go in through the oracle's FP64 setters."""

    def __init__(self, scene, T):
        self.sc, self.T = scene, T
        self.nd = 0 if scene.distance is None else scene.distance[0].shape[0]

    def _system(self, x0, v0, rho0, stiffness):
        o = Oracle(self.sc.params, "f64")
        self.sc.load_into(o)
        o.set_state(x0, v0)
        o.set_coefficients(rho0, stiffness if self.nd else None)
        return o

    def forward(self, x0, v0, rho0, stiffness):
        o = self._system(x0, v0, rho0, stiffness)
        o.step(self.sc.dt, self.T)
        return o.positions(), o.velocities()

    def backward(self, x0, v0, rho0, stiffness, gx, gv):
        o = self._system(x0, v0, rho0, stiffness)
        return o.step_backward(self.sc.dt, self.T, gx, gv, self.nd)


class SubstepsFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x0, v0, rho0, stiffness, runner):
        args = [t.detach().cpu().numpy() for t in (x0, v0, rho0, stiffness)]
        ctx.runner, ctx.args, ctx.dtype = runner, args, x0.dtype
        x, v = runner.forward(*args)
        return (torch.as_tensor(np.asarray(x), dtype=x0.dtype),
                torch.as_tensor(np.asarray(v), dtype=x0.dtype))

    @staticmethod
    def backward(ctx, gx, gv):
        np_dtype = np.float64 if ctx.dtype == torch.float64 else np.float32
        g = [np.ascontiguousarray(t.detach().cpu().numpy(), dtype=np_dtype) for t in (gx, gv)]
        dx, dv, dr, ds = ctx.runner.backward(*ctx.args, *g)
        t = lambda a: torch.as_tensor(np.asarray(a), dtype=ctx.dtype)
        return t(dx), t(dv), t(dr), t(ds), None
