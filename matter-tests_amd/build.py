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
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result"]
ENGINE_SOURCES = ["mpm_engine.hip"]
ENGINE_HEADERS = ["mpm_kernels.hpp", "mpm_math.hpp", "mpm_types.hpp"]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def build_engine():
    srcs = [os.path.join(CSRC, s) for s in ENGINE_SOURCES]
    tmp = ENGINE_SO + ".tmp"
    run([HIPCC] + HIP_FLAGS + ["-I", os.path.join(REPO, "include"), "-o", tmp] + srcs)
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
    return REFTESTS_BIN


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
