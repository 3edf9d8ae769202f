"""The C++ Simulation facade (the reference CPU app's Simulation surface,
particlesolver_amd/csrc/simulation.hpp) against the same scenes built by
particlesolver_amd/scenes.py and stepped through the C-ABI: its scene
builders must produce the identical scene, so the states agree bit for bit
(the Python-built scenes are the ones checked against the oracle in
test_gpu_cpuref.py).  The facade runs in its own process (sim_demo), which
writes its final state as a trajectory file."""
import os
import subprocess

import numpy as np
import pytest

from particlesolver_amd import System, scenes, trajectory
from test_gpu_parity import assert_bitexact

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "particlesolver_amd", "sim_demo")

CASES = {
    "fluid": (scenes.reference_two_fluids, 50),
    "gas": (scenes.reference_gas, 50),
    "smoke_open": (lambda: scenes.reference_smoke(open_gas=True, emitter=True), 100),
    "smoke_closed": (lambda: scenes.reference_smoke(open_gas=False, emitter=True), 100),
    "volcano": (scenes.reference_volcano, 200),
}


def inverse_masses(sc, n):
    """per particle: the objects' inverse masses, then 1 for emitted ones"""
    w = []
    for o in sc.objects:
        kind = o[2] if len(o) > 2 else "fluid"
        m = 0.0 if kind == "static" else (float(o[3]) if len(o) > 3 else 1.0)
        w += [m] * o[0].shape[0]
    return np.array(w + [1.0] * (n - len(w)))


@pytest.mark.parametrize("name", sorted(CASES))
def test_facade_matches_the_scene_builders(gpu_available, tmp_path, name):
    make, ticks = CASES[name]
    out = tmp_path / ("%s.pstraj" % name)
    print("sim_demo", name, ticks, flush=True)
    r = subprocess.run([DEMO, name, str(ticks), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    frame, = trajectory.read(out)
    sc = make()
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, ticks)
        x, v, ph = g.positions(), g.velocities(), g.phases()
    assert frame.substeps == ticks and frame.n == x.shape[0]
    assert_bitexact(frame.positions, x, "%s positions" % name)
    assert np.array_equal(frame.phases, ph)
    # Simulation::getKineticEnergy: sum of .5 v.v / imass over movable particles
    w = inverse_masses(sc, x.shape[0])
    ke = 0.0
    for i in range(x.shape[0]):
        if ph[i] != 3 and w[i] != 0.0:
            vx, vy = float(v[i, 0]), float(v[i, 1])
            ke += (0.5 * (vx * vx + vy * vy)) / float(np.float32(w[i]))
    printed = float(r.stdout.split("kinetic_energy")[1])
    assert printed == pytest.approx(ke, rel=1e-7), (printed, ke)
