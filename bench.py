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

HBM traffic per launch comes from two rocprofv3 counter runs of this script
(FETCH_SIZE, then WRITE_SIZE, with --copy-calibration so the same runs hold a
copy of known size).  `--traffic-report` turns the two counter CSVs into the
summary kept under profiles/; the bench line then quotes the traffic of its
dominant kernel from that summary when the workload matches.
"""
import argparse
import csv
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "traffic_dam_break_1M.md")
ROOF_KERNELS = ("lambda", "solve")
TIMING_PERIOD = 7

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
    ap.add_argument("--workload", default="dam_break",
                    choices=["dam_break", "ensemble", "rayleigh_taylor_2d", "adjoint_cloth_pool", "mixed"],
                    help="dam_break: the 1M-particle PBF metric; ensemble: 512 batched "
                         "scenes (one graph launch per 8 substeps) vs 512 per-scene systems; "
                         "rayleigh_taylor_2d: the 1,600-particle 2D scene in the reference's "
                         "CPU-solver semantics; adjoint_cloth_pool: the reverse pass "
                         "(ps_step_backward) on 16,384 cloth + 65,536 fluid particles; mixed: the "
                         "765,536-particle cloth + rigid cubes + fluid scene")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="dam_break: skip the reverse-pass summary")
    ap.add_argument("--scenes", type=int, default=512)
    ap.add_argument("--no-batched", action="store_true",
                    help="leave the batched-ensemble object out of the default line")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="also run 3 device-to-device copies of MIB MiB (the runtime's copy "
                         "kernel), to scale HBM counters against a known byte count")
    ap.add_argument("--breakdown", type=int, default=20,
                    help="substeps of an extra, separately timed run with events on every kernel")
    ap.add_argument("--no-staging", action="store_true",
                    help="neighbour reads from global memory instead of LDS-staged cells")
    ap.add_argument("--trace-report", nargs=3, metavar=("STATS_CSV", "TRACE_CSV", "BENCH_LOG"),
                    help="summarise a rocprofv3 --kernel-trace --stats run of this script "
                         "(its kernel_stats and kernel_trace CSVs); BENCH_LOG holds that run's "
                         "JSON line")
    ap.add_argument("--counter-report", nargs=3, metavar=("PMCA_CSV", "PMCB_CSV", "BENCH_LOG"),
                    help="summarise two rocprofv3 --pmc runs of SQ counters of this script")
    ap.add_argument("--traffic-report", nargs=3, metavar=("FETCH_CSV", "WRITE_CSV", "BENCH_LOG"),
                    help="summarise two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this "
                         "script with --copy-calibration; BENCH_LOG holds that run's JSON line")
    return ap.parse_args()


KERNEL_OF = {"predict_hash_k": "predict_hash", "scan_partials_k": "scan", "scan_final_k": "scan",
             "scatter_k": "scatter", "reorder_k": "reorder", "neighbors_k": "neighbors",
             "neighbors_staged_k": "neighbors",
             "lambda_k": "lambda", "solve_k": "solve", "shape_k": "shape",
             "finalize_k": "finalize"}


def _counter_means(path):
    """Mean counter value (KiB) per dispatch for each kernel of a rocprofv3
    counter_collection.csv, and the values of the largest runtime copy."""
    import csv
    per, copies = {}, []
    for row in csv.DictReader(open(path)):
        name = row["Kernel_Name"]
        v = float(row["Counter_Value"])
        if "copyBuffer" in name:
            copies.append((int(row["Grid_Size"]), v))
            continue
        for key in KERNEL_OF:
            if "ps::" + key + "(" in name or "ps::" + key + "<" in name:
                per.setdefault(key, []).append(v)
                break
    big = max(g for g, _ in copies)
    out = {}
    for key, v in per.items():          # kernels of one stage add up per launch
        out[KERNEL_OF[key]] = out.get(KERNEL_OF[key], 0.0) + sum(v) / len(v)
    return out, [v for g, v in copies if g == big]


def traffic_report(fetch_csv, write_csv, log):
    """Markdown summary of the HBM traffic per launch (printed to stdout)."""
    line = None
    for l in open(log):
        l = l.strip()
        if l.startswith("{") and '"metric"' in l:
            line = json.loads(l)
    N, K = line["config"]["particles"], line["config"]["mean_neighbors"]
    if "adjoint" in line or "batched" in line:
        raise SystemExit("traffic report: the counted run must use --no-adjoint --no-batched (the "
                         "reverse-pass and ensemble workloads launch the same kernels on other scenes)")
    cal = None
    for l in open(log):
        if l.startswith("copy calibration:"):
            cal = int(l.split("x")[-1].split()[0])
    fetch, fcopy = _counter_means(fetch_csv)
    write, wcopy = _counter_means(write_csv)
    ref_kib = cal / 1024.0
    f_scale = ref_kib / (sum(fcopy) / len(fcopy))
    w_scale = ref_kib / (sum(wcopy) / len(wcopy))
    out = ["# HBM traffic per launch, 1M-particle dam break", "",
           "Commands (repository root on the GPU host, R = its absolute path):", "",
           "    cd /tmp && export TMPDIR=/tmp",
           "    rocprofv3 --pmc FETCH_SIZE --output-format csv -d $R/gpurun_out/pmc_fetch -o fetch "
           "-- python3 $R/bench.py --steps 5 --warmup 2 --no-cpu --no-adjoint --no-batched --breakdown 0 "
           "--copy-calibration 512 > $R/gpurun_out/pmc_fetch.log 2>&1",
           "    rocprofv3 --pmc WRITE_SIZE --output-format csv -d $R/gpurun_out/pmc_write -o write "
           "-- python3 $R/bench.py --steps 5 --warmup 2 --no-cpu --no-adjoint --no-batched --breakdown 0 "
           "--copy-calibration 512 > $R/gpurun_out/pmc_write.log 2>&1",
           "    cd $R && python3 bench.py --traffic-report "
           "gpurun_out/pmc_fetch/fetch_counter_collection.csv "
           "gpurun_out/pmc_write/write_counter_collection.csv gpurun_out/pmc_fetch.log", "",
           "Counters: rocprofv3 FETCH_SIZE and WRITE_SIZE (KiB per dispatch), one run each.",
           "Calibration: the runtime copy kernel of a %d-byte hipMemcpy (device to device) in the "
           "same runs reads FETCH_SIZE %.1f KiB and WRITE_SIZE %.1f KiB for %.1f KiB moved each "
           "way, so FETCH_SIZE is scaled by %.4f and WRITE_SIZE by %.4f."
           % (cal, sum(fcopy) / len(fcopy), sum(wcopy) / len(wcopy), ref_kib, f_scale, w_scale),
           "N = %d particles, mean neighbour-table length K = %.2f." % (N, K), "",
           "| kernel | FETCH KiB | WRITE KiB | traffic MB/launch | algorithmic MB/launch | traffic / algorithmic |",
           "|---|---:|---:|---:|---:|---:|"]
    table = {}
    for k in sorted(set(fetch) | set(write), key=lambda k: -(fetch.get(k, 0) * f_scale + write.get(k, 0))):
        t = (fetch.get(k, 0.0) * f_scale + write.get(k, 0.0) * w_scale) * 1024.0
        table[k] = int(t)
        algo = ALGO_BYTES[k](K) * N if k in ALGO_BYTES else None
        out.append("| %s | %.1f | %.1f | %.1f | %s | %s |" % (
            k, fetch.get(k, 0.0), write.get(k, 0.0), t / 1e6,
            "%.1f" % (algo / 1e6) if algo else "-", "%.2f" % (t / algo) if algo else "-"))
    out += ["", "traffic_bytes_per_launch: " + json.dumps({"particles": N, "kernels": table})]
    print("\n".join(out))


# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz; a wave64 FP32 VALU instruction (no
# packed FP32) occupies a SIMD for 4 cycles
SIMDS, CLOCK_GHZ, VALU_CYCLES = 1024, 2.4, 4.0


def counter_report(a_csv, b_csv, log):
    """Markdown summary of two rocprofv3 --pmc passes of SQ counters over
    this script (printed to stdout): per kernel, the counts per wavefront and
    the VALU issue floor, INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz),
    against the dispatch's own duration in the same runs."""
    rows = {}
    for f in (a_csv, b_csv):
        for r in csv.DictReader(open(f)):
            k = r["Kernel_Name"]
            if "rocclr" in k:
                continue
            k = k.replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            d = rows.setdefault(k, {"dur": [], "c": {}})
            d["c"].setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
            d["dur"].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    line = None
    for l in open(log):
        l = l.strip()
        if l.startswith("{") and '"metric"' in l:
            line = json.loads(l)
    K = line["config"]["mean_neighbors"]
    cmd = ("timeout -s KILL 150 rocprofv3 --pmc %s --output-format csv -d $R/gpurun_out/pmc%s -o pmc%s -- "
           "python3 $R/bench.py --steps 10 --warmup 3 --no-cpu --no-adjoint --no-batched --breakdown 0 "
           "> $R/gpurun_out/pmc%s.log 2>&1")
    out = ["# SQ counters per wavefront: 1,048,576-particle PBF dam break", "",
           "Commands (repository root on the GPU host as R, MI355X):", "", "```",
           "cd /tmp && export TMPDIR=/tmp",
           cmd % ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_LDS_BANK_CONFLICT SQ_WAIT_INST_LDS "
                  "SQ_ACTIVE_INST_VALU SQ_WAVE_CYCLES", "A", "A", "A"),
           cmd % ("SQ_BUSY_CYCLES SQ_WAIT_INST_ANY SQ_ACTIVE_INST_LDS SQ_ACTIVE_INST_ANY SQ_ACTIVE_INST_SCA "
                  "SQ_WAIT_ANY SQ_INSTS_SMEM SQ_INSTS_VMEM_RD", "B", "B", "B"),
           "cd $R && python3 bench.py --counter-report gpurun_out/pmcA/pmcA_counter_collection.csv "
           "gpurun_out/pmcB/pmcB_counter_collection.csv gpurun_out/pmcA.log", "```", "",
           "13 substeps from rest (K = %.2f neighbours per particle). Counts are means per dispatch divided "
           "by SQ_WAVES. The VALU floor is the time to issue the dispatch's VALU instructions at one "
           "wave64 instruction per SIMD every %d cycles on all %d SIMDs at %.1f GHz (FP32 without packed "
           "math); a kernel whose duration is close to it is VALU-issue-bound. Durations are the "
           "dispatches' own under the counters (kernels run one at a time)." % (K, VALU_CYCLES, SIMDS, CLOCK_GHZ),
           "", "| kernel | waves | VALU / wave | SALU / wave | LDS / wave | VMEM rd / wave | "
           "LDS bank conflict cycles / wave | mean us | VALU floor us | floor / duration |",
           "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    for k, d in sorted(rows.items(), key=lambda kv: -sum(kv[1]["dur"])):
        c = {n: sum(v) / len(v) for n, v in d["c"].items()}
        w = c.get("SQ_WAVES", 0.0)
        if w <= 0:
            continue
        dur = sum(d["dur"]) / len(d["dur"])
        floor = c.get("SQ_INSTS_VALU", 0.0) * VALU_CYCLES / (SIMDS * CLOCK_GHZ * 1e3)
        out.append("| `%s` | %.0f | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f | %.1f | %.2f |" % (
            k, w, c.get("SQ_INSTS_VALU", 0) / w, c.get("SQ_INSTS_SALU", 0) / w, c.get("SQ_INSTS_LDS", 0) / w,
            c.get("SQ_INSTS_VMEM_RD", 0) / w, c.get("SQ_LDS_BANK_CONFLICT", 0) / w, dur, floor, floor / dur))
    print("\n".join(out))


def trace_report(stats_csv, trace_csv, log):
    """Markdown summary of a kernel trace of this script (printed to stdout):
    every kernel's statistics over the whole run, and the per-dispatch means
    of the dominant kernels over the bench's timed window."""
    line = None
    for l in open(log):
        l = l.strip()
        if l.startswith("{") and '"metric"' in l:
            line = json.loads(l)
    rows = list(csv.DictReader(open(stats_csv)))
    tot = sum(float(r["TotalDurationNs"]) for r in rows)
    if line["metric"].startswith("reverse-pass"):
        return _trace_report_adjoint(rows, tot, line)
    if "batched ensemble" in line["metric"]:
        return _trace_report_ensemble(rows, tot, line, trace_csv)
    mixed = "mixed-phase" in line["metric"]
    if mixed:
        title = "765,536-particle mixed scene: cloth, rigid cubes, fluid"
        run = "--workload mixed --no-cpu --breakdown 0 --steps %d --warmup %d" % (line["steps"], line["warmup"])
        what = ("%d substeps (%d warm-up + %d timed) of the mixed scene from its initial state "
                "(cloth and cubes falling onto the fluid)." % (line["warmup"] + line["steps"],
                                                                line["warmup"], line["steps"]))
    else:
        title = "1,048,576-particle PBF dam break"
        run = "--no-cpu --no-adjoint --no-batched --breakdown 0"
        what = ("%d substeps (%d warm-up + %d timed) of the dam break from rest, then the exports of "
                "the positions. Durations vary over the run: the collapse phase has the most "
                "neighbours per particle." % (line["warmup"] + line["steps"], line["warmup"], line["steps"]))
    d = "ktm" if mixed else "kt"
    out = ["# Kernel durations: %s (rocprofv3 kernel trace)" % title, "",
           "Commands (repository root on the GPU host as R, MI355X):", "", "```",
           "cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
           "-d /tmp/%s -o kt -- python3 $R/bench.py %s > $R/gpurun_out/%s.log 2>&1" % (d, run, d),
           "cd $R && python3 bench.py --trace-report /tmp/%s/kt_kernel_stats.csv "
           "/tmp/%s/kt_kernel_trace.csv gpurun_out/%s.log" % (d, d, d), "```", "", what,
           "", "| kernel | calls | avg us | min us | max us | share |", "|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append("| `%s` | %s | %.1f | %.1f | %.1f | %.1f %% |" % (
            r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", ""), r["Calls"], float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, 100 * float(r["TotalDurationNs"]) / tot))
    trace = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    warm, steps = line["warmup"], line["steps"]
    out += ["", "Per-dispatch means over the bench's timed window (substeps %d-%d) from the same "
            "trace, and the bench line this run printed:" % (warm + 1, warm + steps), ""]
    for name, per in (("solve_k", 4), ("lambda_k", 4), ("shape_k", 4), ("neighbors_", 1),
                      ("finalize_k", 1), ("reorder_k", 1), ("predict_hash_k", 1)):
        ds = [(int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
              for r in trace if name in r["Kernel_Name"]]
        win = ds[warm * per:(warm + steps) * per]
        if win:
            out.append("    %s: %d dispatches in the timed window, mean %.2f us (all %d: %.2f us)"
                       % (name, len(win), sum(win) / len(win), len(ds), sum(ds) / len(ds)))
    k = line["kernels"]
    out.append("    bench line of this run: value %.4g, ms_per_step %.4f, solve %.2f us (%d sampled "
               "launches), lambda %.2f us, roofline frac %.4f"
               % (line["value"], line["ms_per_step"], k["solve"]["us_per_launch"],
                  k["solve"]["launches"], k["lambda"]["us_per_launch"], line["roofline"]["frac"]))
    print("\n".join(out))


def _trace_report_ensemble(rows, tot, line, trace_csv):
    """The kernel-trace summary of a --workload ensemble run: the kernels of
    the batched handle's graph replays (its stream is the one of the first
    dispatch of predict_hash_k), their busy time against the replays' span,
    and the idle gaps between consecutive kernels -- what one persistent
    launch could at most recover."""
    c = line["config"]
    out = ["# Kernel durations: 512-scene batched ensemble (rocprofv3 kernel trace)", "",
           "Commands (repository root on the GPU host as R, MI355X):", "", "```",
           "cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
           "-d /tmp/kte -o kte -- python3 $R/bench.py --workload ensemble --steps %d "
           "> $R/gpurun_out/kte.log 2>&1" % c["timed_substeps"],
           "cd $R && python3 bench.py --trace-report /tmp/kte/kte_kernel_stats.csv "
           "/tmp/kte/kte_kernel_trace.csv gpurun_out/kte.log", "```", "",
           "%s. One handle steps all scenes with one launch per stage, %d substeps per captured "
           "graph: one warm-up replay, then %d timed replays. Then the same scenes as %d separate "
           "handles, twice issued back to back and once one after another (the rows below "
           "include those runs)." % (c["workload"], c["graph_substeps"],
                                    c["timed_substeps"] // c["graph_substeps"], c["scenes"]),
           "", "| kernel | calls | avg us | min us | max us | share |", "|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append("| `%s` | %s | %.1f | %.1f | %.1f | %.1f %% |" % (
            r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", ""), r["Calls"],
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3,
            100 * float(r["TotalDurationNs"]) / tot))
    trace = sorted(csv.DictReader(open(trace_csv)), key=lambda r: int(r["Start_Timestamp"]))
    first = next(r for r in trace if "predict_hash_k" in r["Kernel_Name"])
    # the batched phase ends where a predict_hash_k of another grid size (a
    # per-scene handle) appears
    stop = next((int(r["Start_Timestamp"]) for r in trace if "predict_hash_k" in r["Kernel_Name"]
                 and r["Grid_Size_X"] != first["Grid_Size_X"]), None)
    mine = [r for r in trace if r["Stream_Id"] == first["Stream_Id"]
            and int(first["Start_Timestamp"]) <= int(r["Start_Timestamp"]) < (stop or 1 << 63)]
    starts = [i for i, r in enumerate(mine) if "predict_hash_k" in r["Kernel_Name"]]
    K = c["graph_substeps"]
    if len(starts) >= 2 * K:
        timed = mine[starts[K]:]                       # after the warm-up replay
        t0 = int(timed[0]["Start_Timestamp"])
        t1 = max(int(r["End_Timestamp"]) for r in timed)
        busy, gaps, end = 0, [], t0
        for r in timed:
            a, b = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
            if a > end:
                gaps.append(a - end)
            busy += max(0, b - max(a, end))
            end = max(end, b)
        subs = len(starts) - K
        out += ["", "The batched handle's timed replays in the same trace: %d substeps, %d dispatches "
                "(%.1f per substep) over %.2f ms, %.2f ms per substep. Kernels are busy %.2f%% of that "
                "span; the %d gaps between consecutive kernels sum to %.3f ms (mean %.2f us, max %.2f us), "
                "%.3f ms per substep." % (subs, len(timed), len(timed) / subs, (t1 - t0) / 1e6,
                                          (t1 - t0) / 1e6 / subs, 100.0 * busy / (t1 - t0), len(gaps),
                                          sum(gaps) / 1e6, sum(gaps) / max(len(gaps), 1) / 1e3,
                                          max(gaps or [0]) / 1e3, sum(gaps) / 1e6 / subs)]
        per = {}
        for r in timed:
            nm = r["Kernel_Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", "")
            per.setdefault(nm, []).append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
        out += ["", "The same replays by kernel (dispatches per substep, mean duration, time per substep):",
                "", "| kernel | per substep | mean us | us per substep |", "|---|---|---|---|"]
        for nm, ds in sorted(per.items(), key=lambda kv: -sum(kv[1])):
            out.append("| `%s` | %.0f | %.1f | %.1f |" % (nm, len(ds) / subs, sum(ds) / len(ds),
                                                      sum(ds) / subs))
    out += ["", "The bench line of the same command: batched %.4g particle-iterations/s (%.3f ms per "
            "substep); per-scene handles issued back to back %.4g (%.2fx), one after another %.4g (%.2fx). "
            "The tracer adds its cost to each of the per-scene runs' many small launches, so their rates "
            "here are lower than without it (the `batched` object of the default bench line)."
            % (line["value"], line["ms_per_step"], c["plain_value"], c["speedup_vs_per_scene"],
               c["sequential_value"], c["speedup_vs_sequential_scenes"])]
    print("\n".join(out))


def _trace_report_adjoint(rows, tot, line):
    """The kernel-trace summary of a --workload adjoint_cloth_pool run."""
    c = line["config"]
    out = ["# Kernel durations: reverse pass on cloth over a fluid pool (rocprofv3 kernel trace)", "",
           "Commands (repository root on the GPU host as R, MI355X):", "", "```",
           "cd /tmp && export TMPDIR=/tmp && rocprofv3 --kernel-trace --stats --output-format csv "
           "-d /tmp/kta -o kta -- python3 $R/bench.py --workload adjoint_cloth_pool --no-cpu "
           "> $R/gpurun_out/kta.log 2>&1",
           "cd $R && python3 bench.py --trace-report /tmp/kta/kta_kernel_stats.csv "
           "/tmp/kta/kta_kernel_trace.csv gpurun_out/kta.log", "```", "",
           "%s. The run holds one forward pass for the target (perturbed parameters), warm-up, "
           "one timed forward pass, four reverse passes (two with checkpoints every %d "
           "substeps, two with the whole horizon on tape) and one more with events on every "
           "reverse kernel. Forward kernels appear in every pass (plain and taping variants); "
           "the `adj_*` kernels only in the reverse passes; `__amd_rocclr_copyBuffer` is the "
           "tape and checkpoint copies." % (c["workload"], c["checkpoint_every"]), "",
           "| kernel | calls | avg us | min us | max us | share |", "|---|---|---|---|---|---|"]
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        out.append("| `%s` | %s | %.1f | %.1f | %.1f | %.1f %% |" % (
            r["Name"].replace("(anonymous namespace)::", "").split("(")[0].replace("void ", ""), r["Calls"], float(r["AverageNs"]) / 1e3,
            float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, 100 * float(r["TotalDurationNs"]) / tot))
    out += ["", "The bench line of the same command (HIP events on the handle's stream): forward "
            "%.2f ms per %d substeps; reverse pass %.2f ms with checkpoints every %d (%.3fx) and "
            "%.2f ms with the whole horizon on tape (%.3fx)."
            % (c["forward_ms"], c["substeps"], c["backward_ms"], c["checkpoint_every"],
               c["backward_over_forward"], c["backward_ms_full_tape"],
               c["backward_over_forward_full_tape"])]
    print("\n".join(out))


def quoted_traffic(kernel, n):
    """Traffic per launch of `kernel` from the committed summary, if it was
    measured on the same particle count."""
    try:
        for l in open(TRAFFIC_SUMMARY):
            if l.startswith("traffic_bytes_per_launch:"):
                d = json.loads(l.split(":", 1)[1])
                if d["particles"] == n and kernel in d["kernels"]:
                    return d["kernels"][kernel]
    except OSError:
        pass
    return None


def cpu_baseline(sc, budget_s, what="dam break"):
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
            "sample": "CPU oracle (C, FP32, OpenMP %d threads) on the same %d-particle %s: "
                      "%d substeps after 1 warm-up, %.1f s" % (threads, sc.n, what, done, el)}


# Algorithmic FP32 flops per accepted neighbour pair (distance, kernel
# values and gradient, 1/r by the seeded Newton form, the tensile power as
# two squarings), and the MI355X FP32 vector peak (packed FMA; the kernels
# are built without packed FP32, so their own ceiling is half of it)
ALGO_FLOPS_PER_PAIR = {"lambda": 25.0, "solve": 30.0}
FP32_VECTOR_PEAK_TFLOPS = 157.3


def copy_bandwidth(mib=512, reps=3):
    """Device-to-device hipMemcpy of `mib` MiB (the runtime's copy kernel)
    timed with HIP events: GB/s moved (read + write), best of `reps`."""
    import ctypes
    from particlesolver_amd import capi
    capi.load()
    hip = ctypes.CDLL("libamdhip64.so.7")      # the runtime the library links
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0 and hip.hipEventCreate(ctypes.byref(e1)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    best = 0.0
    for _ in range(reps + 1):                    # the first is a warm-up
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = ctypes.c_float()
        assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
        best = max(best, 2.0 * n / (ms.value * 1e-3) / 1e9)
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return round(best, 1)


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


def mixed_solve_bytes(K, n, n_nonfluid, n_cloth, n_links, n_rigid):
    """Algorithmic bytes of one solve launch on the mixed scene: every
    particle's iterate 16 + attributes 16 + count 4 + K indices + result 16;
    the contact phases' start position 16 (friction and walls); per cloth
    particle its constraint range 8 and incidence count 4; per link its
    pair 8 and (rest, stiffness) 8, gathered from both ends with the id 4
    and the partner's slot 4 (48 per link); per body member its body, member
    index and goal 24.  Neighbour records count inside the own reads."""
    return ((52.0 + 4.0 * K) * n + 16.0 * n_nonfluid + 12.0 * n_cloth + 48.0 * n_links
            + 24.0 * n_rigid)


MIXED_TIMED = ("lambda", "solve", "shape", "neighbors", "finalize")


def mixed_main(args):
    """The 765,536-particle mixed-phase scene: a 256x256 cloth (structural,
    shear and bend links, corners pinned), 3,125 rigid 4x4x4 cubes with
    friction and 500,000 fluid particles of two rest densities, 4 iterations
    per substep.  Timed like the dam break (warm-up, then `steps` substeps
    between synchronisations, events around one launch in TIMING_PERIOD of
    the main kernels); then an all-events breakdown and the CPU oracle on
    the same scene."""
    import numpy as np
    from particlesolver_amd import System, scenes
    # the settling cubes and cloth pack up to ~210 neighbours within h: a
    # table of 256 keeps every pair (no overflow)
    sc = scenes.mixed(max_neighbors=256)
    N, I = sc.n, sc.params.iterations
    g = System(sc.params)
    sc.load_into(g)
    g.step(sc.dt, args.warmup)
    g.synchronize()
    k_start = float(g.neighbors()[0].mean())
    if not args.no_timing:
        g.timing(MIXED_TIMED)
        g.timing_sample(TIMING_PERIOD)
        g.timing_reset()
    t0 = time.perf_counter()
    g.step(sc.dt, args.steps)
    g.synchronize()
    el = time.perf_counter() - t0
    kt = g.timing_read() if not args.no_timing else {}
    st = g.status()
    K = 0.5 * (k_start + float(g.neighbors()[0].mean()))
    breakdown = {}
    if not args.no_timing and args.breakdown:
        g.timing(True)
        g.timing_sample(1)
        g.timing_reset()
        g.step(sc.dt, args.breakdown)
        g.synchronize()
        for name, (ms, launches) in g.timing_read().items():
            if launches:
                breakdown[name] = round(1000.0 * ms / launches, 2)
        g.timing(False)
    g.close()
    n_fluid = int((np.asarray(sc.phase) == 0).sum())
    n_rigid = int(sum(len(i) for i, _ in sc.bodies))
    n_links = int(sc.distance[0].shape[0]) if sc.distance is not None else 0
    n_cloth = int(np.unique(sc.distance[0]).size) if n_links else 0  # particles with links
    kernels, roof = {}, None
    for name, (ms, launches) in kt.items():
        if launches:
            kernels[name] = {"ms_total": round(ms, 4), "launches": launches,
                             "us_per_launch": round(1000.0 * ms / launches, 3)}
    if "solve" in kernels:
        b = mixed_solve_bytes(K, N, N - n_fluid, n_cloth, n_links, n_rigid)
        ach = b / (kernels["solve"]["us_per_launch"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "solve",
                "algo_bytes_per_launch": int(b), "K": round(K, 2)}
        kernels["solve"]["algo_GBs"] = round(ach, 1)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scenes.mixed(max_neighbors=256), args.cpu_seconds, "mixed-phase scene")
    line = {
        "metric": "particle-iterations/s, 3D mixed-phase scene (cloth, rigid cubes, fluid; 4 iterations/substep)",
        "value": round(N * I * args.steps / el, 1), "unit": "particle-iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(1000.0 * el / max(1, args.steps), 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice jitter, seeded cube lattice and fluid densities)",
        "config": {"workload": "mixed: 256x256 cloth (structural/shear/bend 1.0/0.5/0.1, corners "
                               "pinned) + 3125 rigid 4x4x4 cubes (mu_s 0.6, mu_k 0.4) + 500000 "
                               "fluid in two blocks of rest density 1.0 and 2.0 lattice sums = %d particles" % N,
                   "particles": N, "iterations": I, "substeps_per_s": round(args.steps / el, 2),
                   "mean_neighbors": round(K, 2), "max_neighbors_seen": st.max_neighbors,
                   "status_flags": st.flags, "links": n_links, "bodies": len(sc.bodies)},
        "roofline": roof, "cpu_baseline": cpu, "kernels": kernels,
        "breakdown_us_per_launch_all_events": breakdown,
    }
    print(json.dumps(line))


def ensemble_measure(nscenes=512, K=8, reps=None, steps=64):
    """512-scene batched ensemble: one handle, every launch covers all scenes,
    K substeps per captured graph, against the same scenes as nscenes
    separate handles, each with its own launches: issued back to back (their
    streams overlap on the GPU) and one scene after another (each finished
    before the next starts, as the reference's single legacy stream orders
    its calls).  Returns the rates and the wall times of the phases."""
    from particlesolver_amd import System, scenes
    t_all = time.perf_counter()
    b = scenes.ensemble(nscenes=nscenes)
    N, I = b.n, b.params.iterations
    g = System(b.params, b.descs)
    b.load_into(g)
    g.step_graph(b.dt, K)           # capture + first replay (warm-up)
    g.synchronize()
    reps = reps or max(1, steps // K)
    t0 = time.perf_counter()
    for _ in range(reps):
        g.step_graph(b.dt, K)
    g.synchronize()
    el_b = time.perf_counter() - t0
    st = g.status()
    g.close()
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
    t0 = time.perf_counter()
    for h in plains:
        h.step(b.dt, K)
        h.synchronize()
    el_s = time.perf_counter() - t0
    for h in plains:
        h.close()
    batched = N * I * K * reps / el_b
    plain = N * I * K / el_p
    sequential = N * I * K / el_s
    return {"scenes": len(b.scenes), "particles": N, "iterations": I, "graph_substeps": K,
            "timed_substeps": K * reps, "value": round(batched, 1),
            "ms_per_substep": round(1e3 * el_b / (K * reps), 4),
            "plain_value": round(plain, 1), "speedup_vs_per_scene": round(batched / plain, 2),
            "plain_ms_per_8_substeps_all_scenes": round(1e3 * el_p, 2),
            "sequential_value": round(sequential, 1),
            "speedup_vs_sequential_scenes": round(batched / sequential, 2),
            "sequential_ms_per_8_substeps_all_scenes": round(1e3 * el_s, 2),
            "status_flags": st.flags, "max_neighbors_seen": st.max_neighbors,
            "wall_s": round(time.perf_counter() - t_all, 2)}


def ensemble_main(args):
    r = ensemble_measure(args.scenes, steps=max(args.steps, 8))
    line = {
        "metric": "particle-iterations/s, 512-scene batched ensemble (8-substep graph) vs per-scene launches",
        "value": r["value"], "unit": "particle-iterations/s", "n_gpus": 1,
        "steps": r["timed_substeps"], "warmup": r["graph_substeps"], "ms_per_step": r["ms_per_substep"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 scene sizes, densities, stiffness, gravity)",
        "config": dict({"workload": "%d scenes, %d particles (log-uniform 2k-64k each: fluid block + "
                                    "64-link rope), 4 iterations/substep, dt=1/60" % (r["scenes"], r["particles"])},
                       **{k: v for k, v in r.items() if k != "value"}),
    }
    print(json.dumps(line))


def rt2d_main(args):
    """1,600-particle two-fluid 2D scene with the reference CPU solver's
    semantics: GPU FP32 (graph-captured steps) against ref2d in FP64 on one
    core (the reference's CPU solver is single-threaded); the FP32 error
    against FP64 is stated per result after 1, 10 and 200 steps."""
    import numpy as np
    from particlesolver_amd import System, scenes
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sc = scenes.rayleigh_taylor_2d()
    N, I = sc.n, sc.params.iterations
    steps = args.steps if args.steps % 10 == 0 else 10 * (args.steps // 10 + 1)
    g = System(sc.params)
    sc.load_into(g)
    g.step_graph(sc.dt, 10)              # capture + warm-up
    g.close()
    g = System(sc.params)
    sc.load_into(g)
    g.step_graph(sc.dt, 10)
    g.synchronize()
    g.close()
    # timed: a fresh system, `steps` steps from the initial state
    g = System(sc.params)
    sc.load_into(g)
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps // 10):
        g.step_graph(sc.dt, 10)
    g.synchronize()
    el = time.perf_counter() - t0
    x_gpu = g.positions()
    st = g.status()
    g.close()
    errors = {}
    cpu = None
    if not args.no_cpu:
        from oracle import load_ref2d
        ref = load_ref2d(sc, "f64")
        ref.set_threads(1)
        gpu_at = {}
        with System(sc.params) as h:
            sc.load_into(h)
            done = 0
            for k in (1, 10):
                h.step(sc.dt, k - done)
                done = k
                gpu_at[k] = h.positions()
        gpu_at[steps] = x_gpu
        done, t_cpu = 0, 0.0
        for k in sorted(gpu_at):
            t1 = time.perf_counter()
            ref.step(sc.dt, k - done)
            t_cpu += time.perf_counter() - t1
            done = k
            x64 = ref.positions()
            x32 = gpu_at[k].astype(np.float64)
            e = np.abs(x32 - x64).max(1)
            rho32, c32 = scenes.fluid_object_metrics(sc.params, sc.objects, x32)
            rho64, c64 = scenes.fluid_object_metrics(sc.params, sc.objects, x64)
            errors[str(k)] = {"pos_max": float("%.3e" % e.max()),
                              "pos_rms": float("%.3e" % np.sqrt((e ** 2).mean())),
                              "density_max_abs": float("%.3e" % np.abs(rho32 - rho64).max()),
                              "residual_mean_abs_fp32": float("%.4e" % np.abs(c32).mean()),
                              "residual_mean_abs_fp64": float("%.4e" % np.abs(c64).mean())}
        ref.set_threads(0)
        cpu = {"value": round(N * I * steps / t_cpu, 1), "unit": "particle-iterations/s",
               "cores": 1, "kind": "port",
               "sample": "ref2d (the reference CPU solver restated in C, FP64, 1 thread) on the "
                         "same scene, all %d steps, %.2f s" % (steps, t_cpu)}
    line = {
        "metric": "particle-iterations/s, 2D two-fluid scene in the reference CPU solver's semantics",
        "value": round(N * I * steps / el, 1), "unit": "particle-iterations/s", "n_gpus": 1,
        "steps": steps, "warmup": 10, "ms_per_step": round(1e3 * el / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded SplitMix64 lattice jitter)",
        "config": {"workload": "2D Rayleigh-Taylor, light (rho0 1) and heavy (rho0 3) 40x20 "
                               "blocks = %d particles, r=0.5, h=1, 3 iterations, dt=1/60, "
                               "all-pairs neighbour scan" % N,
                   "status_flags": st.flags, "fp32_vs_fp64_by_step": errors},
        "roofline": None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


# reverse-pass kernels: algorithmic bytes per particle per launch, each
# array read or written once (neighbour data counted inside the own reads)
ADJ_ALGO_BYTES = {
    # elementwise: dL/dp in 16 + clamp 1 + attributes 16 + recorded f and
    # contact count 16 + rho0 term r/w 8 + dL/dy out 16 + mu,1/rho0 out 16
    # + attribute record out 16
    "adj_pre": lambda K: 105.0,
    # attributes 16 + iterate 16 + mu 16 + recorded S,D 16 + recorded rho 4
    # + count 4 + K indices + rho0 term r/w 8 + the pair factors out
    # (mu,Cr 16 and V,u 16)
    "adj_lambda": lambda K: 112.0 + 4.0 * K,
    # iterate 16 + attributes 16 + dL/dy 16 + pair factors 32 + count 4
    # + contact count 4 + K indices + dL/dp out 16
    "adj_position": lambda K: 104.0 + 4.0 * K,
}


def adjoint_measure(args, cloth=(128, 128), fluid=(64, 16, 64), T=60, every=10, cpu=True):
    """The adjoint workload: cloth (128x128, corners pinned) over a fluid pool
    (64x16x64), 60 substeps x 4 iterations, loss L = |x_T - x_T^target|^2
    with the target from seeded perturbed parameters (stiffness x U[0.9,1.1]
    per constraint, rho0 x U[0.95,1.05] per particle).  Times the forward
    pass (ps_step) and the reverse pass (ps_step_backward, checkpoints every
    `every` substeps, its own forward pass included)."""
    import numpy as np
    from particlesolver_amd import System, scenes
    sc = scenes.cloth_over_pool(cloth, fluid)
    N, I = sc.n, sc.params.iterations
    rng = np.random.default_rng(0x5EED0005)
    tgt = scenes.cloth_over_pool(cloth, fluid)
    pairs, rest, stiff = tgt.distance
    tgt.distance = (pairs, rest, (stiff * rng.uniform(0.9, 1.1, stiff.shape)).astype(np.float32))
    tgt.rest_density = (tgt.rest_density * rng.uniform(0.95, 1.05, N)).astype(np.float32)
    with System(tgt.params) as g:
        tgt.load_into(g)
        g.step(sc.dt, T)
        x_target = g.positions().astype(np.float64)
    # warm-up (allocations, code objects) on a throwaway handle
    with System(sc.params) as g:
        sc.load_into(g)
        g.step(sc.dt, 2)
        g.step_backward(sc.dt, 2, np.zeros((N, 3), np.float32), None, checkpoint_every=every)
    with System(sc.params) as g:
        sc.load_into(g)
        g.synchronize()
        t0 = time.perf_counter()
        g.step(sc.dt, T)
        g.synchronize()
        t_fwd = time.perf_counter() - t0
        x_T = g.positions().astype(np.float64)
        K = float(g.neighbors()[0].mean())
        flags = g.status().flags
    loss = float(((x_T - x_target) ** 2).sum())
    gx = (2.0 * (x_T - x_target)).astype(np.float32)
    # Each reverse pass is timed on its second call on the handle (the first
    # allocates the workspace; the second continues from the state the first
    # left, T substeps on, with the same cotangent: the same work).
    def reverse(k):
        with System(sc.params) as g:
            sc.load_into(g)
            out = g.step_backward(sc.dt, T, gx, None, checkpoint_every=k)
            g.synchronize()
            t0 = time.perf_counter()
            g.step_backward(sc.dt, T, gx, None, checkpoint_every=k)
            return out, time.perf_counter() - t0
    (dx, dv, dr, ds, dm), t_bwd = reverse(every)
    full, t_tape = reverse(0)          # the whole horizon on tape, no recomputation
    same = all(np.array_equal(u, v) for u, v in zip(full[:4], (dx, dv, dr, ds))) and full[4] == dm
    # per-kernel breakdown, a separate run with events on every reverse kernel
    kernels = {}
    with System(sc.params) as g:
        sc.load_into(g)
        g.timing([k for k in ("lambda", "solve") + tuple(ADJ_ALGO_BYTES) +
                  ("adj_xsph", "adj_finalize", "adj_predict")])
        g.timing_reset()
        g.step_backward(sc.dt, T, gx, None, checkpoint_every=every)
        for name, (ms, launches) in g.timing_read().items():
            if launches:
                kernels[name] = {"ms_total": round(ms, 3), "launches": launches,
                                 "us_per_launch": round(1000.0 * ms / launches, 2)}
    roof = None
    timed = {k: v for k, v in kernels.items() if k in ADJ_ALGO_BYTES}
    if timed:
        dom = max(timed, key=lambda k: timed[k]["ms_total"])
        b = ADJ_ALGO_BYTES[dom](K) * N
        ach = b / (timed[dom]["us_per_launch"] * 1e-6) / 1e9
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": dom,
                "algo_bytes_per_launch": int(b), "K": round(K, 2)}
        for k in timed:
            kernels[k]["algo_GBs"] = round(ADJ_ALGO_BYTES[k](K) * N / (kernels[k]["us_per_launch"] * 1e-6)
                                           / 1e9, 1)
    cpu_line = None
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import load_scene
        o = load_scene(sc, "f32")
        Tc = 2
        t0 = time.perf_counter()
        o.step_backward(sc.dt, Tc, gx, None, pairs.shape[0])
        tc = time.perf_counter() - t0
        cpu_line = {"value": round(N * I * Tc / tc, 1), "unit": "particle-iterations/s (reverse pass)",
                    "cores": os.cpu_count(), "kind": "port",
                    "sample": "FP32 oracle orc_step_backward, the first %d substeps of the same scene "
                              "(its forward passes use OpenMP, the adjoint loops one thread), %.2f s"
                              % (Tc, tc)}
    return {
        "workload": "cloth %dx%d (corners pinned) over a %dx%dx%d fluid pool = %d particles, "
                    "%d substeps x %d iterations, checkpoints every %d substeps"
                    % (cloth[0], cloth[1], fluid[0], fluid[1], fluid[2], N, T, I, every),
        "particles": N, "substeps": T, "iterations": I, "checkpoint_every": every,
        "forward_ms": round(1e3 * t_fwd, 2), "backward_ms": round(1e3 * t_bwd, 2),
        "backward_over_forward": round(t_bwd / t_fwd, 3),
        "backward_ms_full_tape": round(1e3 * t_tape, 2),
        "backward_over_forward_full_tape": round(t_tape / t_fwd, 3),
        "full_tape_gradients_bitwise_equal": bool(same),
        "value": round(N * I * T / t_bwd, 1), "mean_neighbors": round(K, 2), "status_flags": flags,
        "loss": loss,
        "grad_norms": {"x0": float(np.linalg.norm(dx)), "v0": float(np.linalg.norm(dv)),
                       "rho0": float(np.linalg.norm(dr)), "stiffness": float(np.linalg.norm(ds)),
                       "wall_mu_k": dm},
        "roofline": roof, "kernels": kernels, "cpu_baseline": cpu_line,
    }


def adjoint_main(args):
    r = adjoint_measure(args, cpu=not args.no_cpu)
    line = {
        "metric": "reverse-pass particle-iterations/s (ps_step_backward incl. its forward pass)",
        "value": r.pop("value"), "unit": "particle-iterations/s", "n_gpus": 1,
        "steps": r["substeps"], "warmup": 2,
        "ms_per_step": round(r["backward_ms"] / r["substeps"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded lattice jitter, seeded parameter perturbation for the target)",
        "config": {k: v for k, v in r.items() if k not in ("roofline", "cpu_baseline", "kernels")},
        "roofline": r["roofline"], "cpu_baseline": r["cpu_baseline"], "kernels": r["kernels"],
    }
    print(json.dumps(line))


def main():
    args = parse()
    if args.trace_report:
        return trace_report(*args.trace_report)
    if args.traffic_report:
        return traffic_report(*args.traffic_report)
    if args.counter_report:
        return counter_report(*args.counter_report)
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.workload == "ensemble":
        return ensemble_main(args)
    if args.workload == "rayleigh_taylor_2d":
        return rt2d_main(args)
    if args.workload == "adjoint_cloth_pool":
        return adjoint_main(args)
    if args.workload == "mixed":
        return mixed_main(args)
    from particlesolver_amd import System, scenes
    import numpy as np

    nx, ny, nz = (max(2, int(round(d * args.scale))) for d in (128, 64, 128))
    sc = scenes.dam_break(nx, ny, nz, max_neighbors=192)
    if args.no_staging:
        from particlesolver_amd import capi
        sc.params.flags |= capi.PS_FLAG_NO_LDS_STAGING
    N, I = sc.n, sc.params.iterations

    g = System(sc.params)
    sc.load_into(g)
    g.step(sc.dt, args.warmup)
    g.synchronize()
    k_start = float(g.neighbors()[0].mean()) if args.steps else 0.0

    # timed region: events only around the solve kernels whose roofline is
    # reported, and around one launch in TIMING_PERIOD of each (an event pair
    # costs stream time: on every launch they slowed the substep by 6%); the
    # period is coprime with the 4 iterations, so every iteration is sampled
    if not args.no_timing:
        g.timing(ROOF_KERNELS)
        g.timing_sample(TIMING_PERIOD)
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
    occ = np.diff(g.grid()[2].astype(np.int64))
    occupancy = {"max": int(occ.max()), "mean_nonempty": round(float(occ[occ > 0].mean()), 2),
                 "particles_in_cells_over_16": int(occ[occ > 16].sum())}
    # diagnostic breakdown of every kernel (not part of the timed region)
    breakdown = {}
    if not args.no_timing and args.breakdown:
        g.timing(True)
        g.timing_sample(1)
        g.timing_reset()
        g.step(sc.dt, args.breakdown)
        g.synchronize()
        for name, (ms, launches) in g.timing_read().items():
            if launches:
                breakdown[name] = round(1000.0 * ms / launches, 2)
        g.timing(False)
    # insertion-order export of the positions (what a renderer copies into
    # its vertex buffer each frame): reads X and A, writes one float4
    export = None
    if not args.no_timing:
        from particlesolver_amd import capi
        with capi.DeviceBuffer(16 * N) as buf:
            g.timing({"export"})
            g.timing_sample(1)
            g.timing_reset()
            for _ in range(20):
                g.export_device(capi.PS_FIELD_POSITIONS, buf.ptr)
            g.synchronize()
            ms, launches = g.timing_read()["export"]
            g.timing(False)
        us = 1000.0 * ms / launches
        export = {"kernel": "export_vec_k", "us_per_launch": round(us, 2),
                  "algo_bytes_per_particle": 48,
                  "algo_GBs": round(48.0 * N / (us * 1e-6) / 1e9, 1)}
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
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": quoted_traffic(dom, N),
                "kernel": dom, "algo_bytes_per_launch": int(bytes_per_launch),
                "K": round(K, 2)}
        # the same launches against the FP32 vector peak (algorithmic flops)
        flops = ALGO_FLOPS_PER_PAIR[dom] * K * N
        tf = flops / per_launch_s / 1e12
        roof["valu"] = {"algo_flops_per_pair": ALGO_FLOPS_PER_PAIR[dom],
                        "achieved_TFLOPs": round(tf, 2), "peak_TFLOPs": FP32_VECTOR_PEAK_TFLOPS,
                        "frac": round(tf / FP32_VECTOR_PEAK_TFLOPS, 4)}
        for k in kernels:
            if k in ALGO_BYTES:
                b = ALGO_BYTES[k](K) * N
                kernels[k]["algo_GBs"] = round(b / (kernels[k]["us_per_launch"] * 1e-6) / 1e9, 1)

    if roof is not None:
        # what a plain copy kernel reaches on this GPU, for scale
        roof["copy_kernel_GBs"] = copy_bandwidth()

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
                   "cell_occupancy_end": occupancy, "status_flags": st.flags},
        "roofline": roof,
        "cpu_baseline": cpu,
        "kernels": kernels,
        "breakdown_us_per_launch_all_events": breakdown,
        "export": export,
    }
    if not args.no_batched:
        # the batched ensemble: one handle for all scenes against the same
        # scenes as separate handles
        r = ensemble_measure(args.scenes)
        line["batched"] = dict({"workload": "%d scenes, %d particles (log-uniform 2k-64k each: fluid "
                                            "block + 64-link rope), 4 iterations/substep, one captured "
                                            "8-substep graph" % (r["scenes"], r["particles"])}, **r)
    if not args.no_adjoint:
        r = adjoint_measure(args, cpu=False)
        line["adjoint"] = {k: r[k] for k in ("workload", "forward_ms", "backward_ms",
                                             "backward_over_forward", "backward_ms_full_tape",
                                             "backward_over_forward_full_tape", "value",
                                             "roofline")}
    print(json.dumps(line))
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)


if __name__ == "__main__":
    main()
