"""Writes tests/golden/reference_cases.json: the CPU oracle's outcome of the two reference cases whose pass/fail at
the reference's tolerance depends on the particle sample, on one recorded sample of each.

    python tests/golden/make_reference_cases.py [--threads N]

* ElasticityTest.BulkModulus (tests/tests.cpp:476-541): the measured bulk modulus K, lambda + 2 mu / 3 and the
  J average after the 100 frames;
* BoundaryTest.MIBF (tests/tests.cpp:308-474): the maximum x after 0.5 s of the three runs (plate friction,
  MIBF grid friction, von Mises q).

The samples come from the host Poisson-disk sampler (tests/cpp/sampler_check_3d, built by
matter-tests_amd/build.py) with the arguments stored beside the results; the particle count and the exactly
rounded sum of the coordinates identify the sample, so that a test can regenerate it and check that it is the
same one. With 6 threads the oracle takes about 9 minutes for the bulk-modulus case and 15 for each MIBF
run. Its parallel sums depend on the thread schedule, so repeated runs differ in the last digits (K in the
15th): the test compares at 1e-12, not bit for bit.
"""
import argparse
import json
import math
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.join(REPO, "matter-tests_amd"))

import mpm_abi as A  # noqa: E402
import physics_cases as PC  # noqa: E402

SAMPLER = os.path.join(REPO, "tests", "cpp", "sampler_check_3d")
OUT = os.path.join(HERE, "reference_cases.json")

# sampler arguments (radius, box, particles per cell, density, seed) of the reference's set-ups: the
# bulk-modulus slab (sampleParticles(sim, 0.02, 8), 1 x 1 x 0.2, rho 10000) and the MIBF block
# (sampleParticles(sim, 0.001), 0.1 x 0.05 x 0.05, rho 1000)
SAMPLES = {
    "bulk_modulus": dict(r=0.02, L=[1.0, 1.0, 0.2], ppc=8.0, rho=10000.0, seed=42),
    "mibf": dict(r=0.001, L=[0.1, 0.05, 0.05], ppc=8.0, rho=1000.0, seed=42),
}


def regenerate(spec):
    """The sample of one set-up and its identity (count, exactly rounded coordinate sum)."""
    s = PC.host_sample(SAMPLER, spec["r"], spec["L"], spec["ppc"], spec["rho"], spec["seed"])
    return s, {"n": int(s[0].shape[0]), "coord_fsum": math.fsum(s[0].ravel().tolist())}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    a = ap.parse_args()
    lib = A.oracle_library()
    out = {}
    s, ident = regenerate(SAMPLES["bulk_modulus"])
    t = time.time()
    K, Kt, Javg = PC.bulk_modulus_on_sample(lib, s, n_threads=a.threads)
    out["bulk_modulus"] = dict(sampler=SAMPLES["bulk_modulus"], **ident, K=float(K), K_true=float(Kt),
                               J_avg=float(Javg), rel=float(abs(K - Kt) / Kt), reference_tol=0.03)
    print("bulk modulus", out["bulk_modulus"], f"{time.time() - t:.0f} s", flush=True)
    s, ident = regenerate(SAMPLES["mibf"])
    r = dict(sampler=SAMPLES["mibf"], **ident, reference_tol=1e-13)
    for label, mibf, mises in (("plate", False, False), ("mibf", True, False), ("mises", False, True)):
        t = time.time()
        r[f"max_x_{label}"] = PC.mibf_slope(lib, s, mibf, mises, n_threads=a.threads)
        print(label, repr(r[f"max_x_{label}"]), f"{time.time() - t:.0f} s", flush=True)
    r["d_mibf"] = abs(r["max_x_plate"] - r["max_x_mibf"])
    r["d_mises"] = abs(r["max_x_plate"] - r["max_x_mises"])
    out["mibf"] = r
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", OUT)


if __name__ == "__main__":
    main()
