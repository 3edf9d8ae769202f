"""Benchmark: 1,048,576-particle 3D PBF dam break on one MI355X.

A step is one substep of the solver (predict, grid + neighbour build,
`iterations` Jacobi iterations of lambda + solve, velocity + XSPH) over the
whole particle set, with every input resident in HBM.  The metric is
particle-iterations/s = N * iterations * substeps / wall seconds of the
timed region (synchronised on both sides).

Per-kernel durations come from HIP events recorded around every launch on
the solver's own stream during the timed region; the roofline entry uses the
dominant kernel's algorithmic bytes per launch (per-particle counts below,
K = mean neighbour-table length) over its mean launch time, against the
MI355X HBM peak of 8 TB/s.

The CPU baseline is the FP32 CPU oracle (OpenMP over the host cores given
to this process) on the same scene for a bounded number of substeps.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
ROOF_KERNELS = ("lambda", "solve")

# Algorithmic bytes per particle per launch, K = neighbour-table length.
# Each array read once and written once; neighbour positions are counted
# inside the own-position read.
ALGO_BYTES = {
    # own x,y,z 12 + rho0 4 + count 4 + K indices + lambda out 4
    "lambda": lambda K: 24.0 + 4.0 * K,
    # own x,y,z,lambda 16 + rho0 4 + insertion index 4 + count 4 + K indices + x,y,z out 12
    "solve": lambda K: 40.0 + 4.0 * K,
    # x,y,z 12 + cell key 4 + K indices out + count out 4 (+ cell offsets, ~C/N)
    "neighbors": lambda K: 20.0 + 4.0 * K,
    # p, x*: 24 + rho0 4 + count 4 + K indices + v out 12
    "finalize": lambda K: 44.0 + 4.0 * K,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed substeps")
    ap.add_argument("--warmup", type=int, default=20, help="untimed substeps")
    ap.add_argument("--scale", type=float, default=1.0,
                    help="lattice scale (1.0 = 128x64x128); smaller only for debugging")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-timing", action="store_true",
                    help="no per-kernel events in the timed region")
    ap.add_argument("--workload", default="dam_break", choices=["dam_break", "ensemble"],
                    help="dam_break: the 1M-particle PBF metric; ensemble: 512 batched "
                         "scenes (one graph launch per 8 substeps) vs 512 per-scene systems")
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (the runtime's copy "
                         "kernel), to scale HBM counters against a known byte count")
    ap.add_argument("--breakdown", type=int, default=20,
                    help="substeps of an extra, separately timed run with events on every kernel")
    return ap.parse_args()


def cpu_baseline(sc, budget_s):
    """FP32 CPU oracle on the same scene: as many substeps as fit the budget
    (at least one), timed after one warm-up substep."""
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
        os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import load_scene
    ref = load_scene(sc, "f32")
    ref.step(sc.dt, 1)
    done = 0
    t0 = time.perf_counter()
    while True:
        ref.step(sc.dt, 1)
        done += 1
        el = time.perf_counter() - t0
        if el >= budget_s or el / done * (done + 1) > 2.0 * budget_s:
            break
    rate = sc.n * sc.params.iterations * done / el
    return {"value": rate, "unit": "particle-iterations/s", "cores": threads, "kind": "port",
            "sample": "CPU oracle (C, FP32, OpenMP %d threads) on the same %d-particle dam break: "
                      "%d substeps after 1 warm-up, %.1f s" % (threads, sc.n, done, el)}


def copy_calibration(mib):
    """Three hipMemcpy device-to-device copies of `mib` MiB: each is one
    dispatch of the runtime's copy kernel reading and writing exactly that
    many bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    for _ in range(3):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print("copy calibration: 3 x %d bytes" % n, file=sys.stderr)


def ensemble_main(args):
    """512-scene batched ensemble: one handle, every launch covers all scenes,
    8 substeps per captured graph, vs the same scenes as 512 separate
    handles stepped one after another (each its own launches)."""
    from particlesolver_amd import System, scenes
    b = scenes.ensemble(nscenes=args.scenes)
    N, I, K = b.n, b.params.iterations, 8
    g = System(b.params, b.descs)
    b.load_into(g)
    g.step_graph(b.dt, K)           # capture + first replay (warm-up)
    g.synchronize()
    reps = max(1, args.steps // K)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.step_graph(b.dt, K)
    g.synchronize()
    el_b = time.perf_counter() - t0
    st = g.status()
    g.close()
    batched = N * I * K * reps / el_b
    print("batched: %d scenes %d particles %.2f ms per %d substeps" % (len(b.scenes), N, 1e3 * el_b / reps, K),
          file=sys.stderr, flush=True)
    plains = []
    for q in range(len(b.scenes)):
        sc = b.plain(q)
        h = System(sc.params)
        sc.load_into(h)
        plains.append(h)
    for h in plains:                 # warm-up
        h.step(b.dt, K)
    for h in plains:
        h.synchronize()
    t0 = time.perf_counter()
    for h in plains:
        h.step(b.dt, K)
    for h in plains:
        h.synchronize()
    el_p = time.perf_counter() - t0
    for h in plains:
        h.close()
    plain = N * I * K / el_p
    line = {
        "metric": "particle-iterations/s, 512-scene batched ensemble (8-substep graph) vs per-scene launches",
        "value": round(batched, 1), "unit": "particle-iterations/s", "n_gpus": 1,
        "steps": K * reps, "warmup": K, "ms_per_step": round(1e3 * el_b / (K * reps), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene sizes, densities, stiffness, gravity)",
        "config": {"workload": "%d scenes, %d particles (log-uniform 2k-64k each: fluid block + "
                               "64-link rope), 4 iterations/substep, dt=1/60" % (len(b.scenes), N),
                   "plain_value": round(plain, 1), "speedup_vs_per_scene": round(batched / plain, 2),
                   "plain_ms_per_8_substeps_all_scenes": round(1e3 * el_p, 2),
                   "status_flags": st.flags, "max_neighbors_seen": st.max_neighbors},
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return ensemble_main(args)
    from particlesolver_amd import System, scenes
    import numpy as np

    nx, ny, nz = (max(2, int(round(d * args.scale))) for d in (128, 64, 128))
    sc = scenes.dam_break(nx, ny, nz, max_neighbors=192)
    N, I = sc.n, sc.params.iterations

    g = System(sc.params)
    sc.load_into(g)
    g.step(sc.dt, args.warmup)
    g.synchronize()
    k_start = float(g.neighbors()[0].mean()) if args.steps else 0.0

    # timed region: events only around the solve kernels whose roofline is
    # reported (each event pair costs stream time)
    if not args.no_timing:
        g.timing(ROOF_KERNELS)
        g.timing_reset()
    t0 = time.perf_counter()
    g.step(sc.dt, args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    kt = g.timing_read() if not args.no_timing else {}
    cnt = g.neighbors()[0]
    k_end = float(cnt.mean())
    K = 0.5 * (k_start + k_end)
    st = g.status()
    # diagnostic breakdown of every kernel (not part of the timed region)
    breakdown = {}
    if not args.no_timing and args.breakdown:
        g.timing(True)
        g.timing_reset()
        g.step(sc.dt, args.breakdown)
        g.synchronize()
        for name, (ms, launches) in g.timing_read().items():
            if launches:
                breakdown[name] = round(1000.0 * ms / launches, 2)
        g.timing(False)
    g.close()

    value = N * I * args.steps / el
    kernels = {}
    for name, (ms, launches) in kt.items():
        if launches:
            kernels[name] = {"ms_total": round(ms, 4), "launches": launches,
                             "us_per_launch": round(1000.0 * ms / launches, 3)}
    roof = None
    dom = None
    if kernels:
        timed = {k: v for k, v in kernels.items() if k in ALGO_BYTES}
        dom = max(timed, key=lambda k: timed[k]["ms_total"])
        per_launch_s = timed[dom]["us_per_launch"] * 1e-6
        bytes_per_launch = ALGO_BYTES[dom](K) * N
        achieved = bytes_per_launch / per_launch_s / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": dom, "algo_bytes_per_launch": int(bytes_per_launch),
                "K": round(K, 2)}
        for k in kernels:
            if k in ALGO_BYTES:
                b = ALGO_BYTES[k](K) * N
                kernels[k]["algo_GBs"] = round(b / (kernels[k]["us_per_launch"] * 1e-6) / 1e9, 1)

    cpu = None
    if not args.no_cpu:
        csc = scenes.dam_break(nx, ny, nz, max_neighbors=192)
        cpu = cpu_baseline(csc, args.cpu_seconds)

    line = {
        "metric": "particle-iterations/s, 1M-particle 3D PBF dam break (4 iterations/substep)",
        "value": round(value, 1),
        "unit": "particle-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1000.0 * el / max(1, args.steps), 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 lattice jitter)",
        "config": {"workload": "3D PBF dam break %dx%dx%d = %d particles, h=0.2=cell, "
                               "grid %dx%dx%d, %d iterations/substep, dt=1/60, eps=600, "
                               "k=0.1, n=4, dq=0.2h, XSPH c=0.01"
                               % (nx, ny, nz, N, sc.params.grid[0], sc.params.grid[1],
                                  sc.params.grid[2], I),
                   "particles": N, "iterations": I, "substeps_per_s": round(args.steps / el, 2),
                   "mean_neighbors": round(K, 2), "max_neighbors_seen": st.max_neighbors,
                   "status_flags": st.flags},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "breakdown_us_per_launch_all_events": breakdown,
    }
    print(json.dumps(line))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)


if __name__ == "__main__":
    main()
