"""Build recipe of the CPU oracle (test infrastructure only).

ps_oracle.c (the substep of the 3D path) and ref2d.c (the reference's own
CPU solver, Simulation::tick) are compiled together twice with gcc: FP32
(liboracle_f32.so) and FP64 (liboracle_f64.so).  -ffp-contract=off and no -march/-ffast-math keep the
FP32 build to exactly the IEEE operations written in the source.  OpenMP
parallelises the per-particle loops (each particle is a pure gather, so the
results do not depend on the thread count).
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = [os.path.join(HERE, "ps_oracle.c"), os.path.join(HERE, "ref2d.c")]

# -mfma turns the explicit fmaf/fma calls into the FMA instruction (the same
# correctly rounded result as the libm call); -ffp-contract=off still keeps
# every other product and sum separately rounded.
CFLAGS = ["-std=c99", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-mfma", "-fopenmp", "-Wall", "-Wno-unknown-pragmas"]


def build_oracle(verbose=True):
    outs = []
    for dbl, name in ((0, "liboracle_f32.so"), (1, "liboracle_f64.so")):
        out = os.path.join(HERE, name)
        tmp = out + ".tmp"
        cmd = ["gcc"] + CFLAGS + ["-DORC_DOUBLE=%d" % dbl] + SRCS + ["-o", tmp, "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
        outs.append(out)
    return outs


if __name__ == "__main__":
    build_oracle()
