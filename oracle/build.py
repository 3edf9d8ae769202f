"""Build recipe of the CPU oracle (test infrastructure only).

ps_oracle.c (the substep of the 3D path) and ref2d.c (the reference's own
CPU solver, Simulation::tick) are compiled together twice with gcc: FP32
(liboracle_f32.so) and FP64 (liboracle_f64.so).  -ffp-contract=off and no -march/-ffast-math keep the
FP32 build to exactly the IEEE operations written in the source.  OpenMP
parallelises the per-particle loops (each particle is a pure gather, so the
results do not depend on the thread count).

    python oracle/build.py              # the two libraries, next to this file
    python oracle/build.py --sanitize   # ASan+UBSan builds in oracle/_san, then
                                        # the oracle's CPU tests against them
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN_DIR = os.path.join(HERE, "_san")
SRCS = [os.path.join(HERE, "ps_oracle.c"), os.path.join(HERE, "ref2d.c")]

# -mfma turns the explicit fmaf/fma calls into the FMA instruction (the same
# correctly rounded result as the libm call); -ffp-contract=off still keeps
# every other product and sum separately rounded.
CFLAGS = ["-std=c99", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-mfma", "-fopenmp", "-Wall", "-Wno-unknown-pragmas"]


# AddressSanitizer + UndefinedBehaviorSanitizer build (host code only: the
# oracle is plain C), every UB report fatal
SAN_FLAGS = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=all"]

# the CPU tests that exercise the oracle, run against the sanitizer build
SAN_TESTS = ["tests/test_oracle_pinning.py", "tests/test_oracle_adjoint.py", "tests/test_oracle_friction.py",
             "tests/test_ref2d_pinning.py", "tests/test_ref2d_emitter.py", "tests/test_ref2d_gas.py",
             "tests/test_ref2d_volcano.py"]


def build_oracle(verbose=True, sanitize=False):
    outs = []
    outdir = SAN_DIR if sanitize else HERE
    os.makedirs(outdir, exist_ok=True)
    flags = CFLAGS + SAN_FLAGS if sanitize else CFLAGS
    for dbl, name in ((0, "liboracle_f32.so"), (1, "liboracle_f64.so")):
        out = os.path.join(outdir, name)
        tmp = out + ".tmp"
        cmd = ["gcc"] + flags + ["-DORC_DOUBLE=%d" % dbl] + SRCS + ["-o", tmp, "-lm"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        os.replace(tmp, out)
        outs.append(out)
    return outs


def run_sanitized_tests(extra=()):
    """The oracle's CPU tests with the sanitizer build loaded (libasan
    preloaded into the interpreter, leak checks off: CPython's own
    allocations would dominate them)."""
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], check=True, capture_output=True,
                          text=True).stdout.strip()
    env = dict(os.environ, PS_ORACLE_DIR=SAN_DIR, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    cmd = [sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider"] + SAN_TESTS + list(extra)
    print(" ".join(cmd), flush=True)
    return subprocess.run(cmd, cwd=ROOT, env=env).returncode


if __name__ == "__main__":
    if "--sanitize" in sys.argv[1:]:
        build_oracle(sanitize=True)
        sys.exit(run_sanitized_tests())
    build_oracle()
