"""Build recipe of the native library (libparticlesolver.so, gfx950).

Every HIP source is compiled for gfx950 with IEEE division/sqrt and no
floating-point contraction (the kernels reproduce the CPU oracle's float
operations one by one), and with packed-FP32 instruction selection off.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libparticlesolver.so")

SOURCES = ["ps_kernels.hip", "ps_adjoint.hip", "ps_cpuref.hip", "ps_export.hip", "ps_grid.hip", "ps_api.hip",
           "ps_trajectory.hip", "particle_system.hip", "simulation.hip"]
DEMO_SRC = os.path.join(CSRC, "ps_demo.cpp")
DEMO_BIN = os.path.join(HERE, "ps_demo")
# programs over the C++ facades: (source, binary)
DEMOS = [(DEMO_SRC, DEMO_BIN),
         (os.path.join(CSRC, "sim_demo.cpp"), os.path.join(HERE, "sim_demo"))]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    return "hipcc"


def build_library(verbose=True):
    srcs = [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = LIB + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + srcs + ["-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


def build_cpp_demo(verbose=True):
    """C++ programs that drive the ParticleSystem and Simulation facades
    (the reference's method names) through the library; a program whose
    source is absent is skipped."""
    built = []
    for src, out in DEMOS:
        if not os.path.exists(src):
            continue
        cmd = [hipcc(), "-O2", "-std=c++17", "-ffp-contract=off", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
               src, "-o", out + ".tmp", "-L" + HERE, "-lparticlesolver", "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)
        built.append(out)
    return built


if __name__ == "__main__":
    build_library()
    build_cpp_demo()
    sys.exit(0)
