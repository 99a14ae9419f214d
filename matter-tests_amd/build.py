"""Build recipe for the engine (libmpm_hip.so, gfx950) and, when asked, the CPU oracle used by the tests.

    python matter-tests_amd/build.py            # engine + oracle
    python matter-tests_amd/build.py engine     # engine only

The engine is one shared library compiled by hipcc for gfx950 only. Packed FP32 instructions are
disabled for every HIP translation unit, and floating-point contraction is off so that the FP64 form
rounds like the reference's x86-64 build.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
ENGINE_SO = os.path.join(HERE, "libmpm_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
# The C ABI and host logic, then one translation unit per (dimension, spline degree, transfer type) of the
# step, compiled in parallel and linked into one library.
ENGINE_SOURCES = ["mpm_engine.hip", "mpm_server.hip"] + [f"mpm_step_{d}d_s{s}_{t}.hip" for d in (3, 2) for s in (2, 3) for t in ("f64", "split")]
ENGINE_HEADERS = ["mpm_context.hpp", "mpm_step.hpp", "mpm_batch_kernels.hpp", "mpm_kernels.hpp", "mpm_math.hpp", "mpm_types.hpp"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_engine():
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    inc = ["-I", os.path.join(REPO, "include")]
    objs = [os.path.join(objdir, os.path.splitext(s)[0] + ".o") for s in ENGINE_SOURCES]
    jobs = max(1, min(len(ENGINE_SOURCES), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
    with ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(run, [HIPCC] + HIP_FLAGS + inc + ["-c", "-o", o, os.path.join(CSRC, s)]) for s, o in zip(ENGINE_SOURCES, objs)]
        for f in futs:
            f.result()
    tmp = ENGINE_SO + ".tmp"
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, ENGINE_SO)
    return ENGINE_SO


REFTESTS_SRC = os.path.join(REPO, "tests", "cpp", "reference_tests.cpp")
REFTESTS_BIN = os.path.join(REPO, "tests", "cpp", "reference_tests")


def build_reference_tests():
    """The reference's physics tests against the C++ host API (links libmpm_hip.so)."""
    tmp = REFTESTS_BIN + ".tmp"
    run(["g++", "-O2", "-std=c++17", "-Wall", "-Wno-unused-function", REFTESTS_SRC, "-I", os.path.join(REPO, "include"),
         "-L", HERE, "-lmpm_hip", "-Wl,-rpath,$ORIGIN/../../matter-tests_amd", "-o", tmp])
    os.replace(tmp, REFTESTS_BIN)
    build_sampler()
    return REFTESTS_BIN


SAMPLER_SRC = os.path.join(REPO, "tests", "cpp", "sampler_check.cpp")
SAMPLER_BIN = os.path.join(REPO, "tests", "cpp", "sampler_check_3d")


def build_sampler():
    """The host Poisson-disk sampler as a program (3D), so that a test can regenerate a recorded sample."""
    tmp = SAMPLER_BIN + ".tmp"
    run(["g++", "-O2", "-std=c++17", SAMPLER_SRC, "-o", tmp])
    os.replace(tmp, SAMPLER_BIN)
    return SAMPLER_BIN


def build_oracle():
    run(["make", "-s", "-C", os.path.join(REPO, "oracle")])
    return os.path.join(REPO, "oracle", "liboracle.so")


def main(argv):
    what = argv[1:] or ["engine", "oracle", "reftests"]
    if "engine" in what:
        build_engine()
    if "reftests" in what:
        build_reference_tests()
    if "oracle" in what:
        build_oracle()


if __name__ == "__main__":
    main(sys.argv)
