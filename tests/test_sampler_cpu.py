"""The host sampler (matter.hpp: sampleParticles / poissonDiskSampling), which stands in for the reference's
tph_poisson call (src/sampling/sampling_particles.hpp:12-123; tph_poisson is fetched at configure time and
is absent here, so positions cannot match the reference's and sample parity is unpinned). What fixes it:

* the reference's set-up formulas (sampling_particles.hpp:28-31, 90-93): dx^dim = ppc * volume / N,
  particle_volume = dx^dim / ppc, particle_mass = rho * particle_volume;
* Poisson-disk properties: every sample in the box, no two samples closer than the radius, every point of
  the box within twice the radius of a sample (Bridson's construction stops only when the shell [r, 2r)
  around each sample is full), and a packing fraction of a random maximal disk set (the jamming limit of
  random sequential adsorption is 0.547 for disks and 0.384 for spheres of diameter r; Bridson's
  construction reaches close to it). Sampling with radius r/2 instead of r, or without the distance test,
  fails these;
* determinism per seed.
"""
import os
import subprocess

import numpy as np
import pytest
from scipy.spatial import cKDTree

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "sampler_check.cpp")


@pytest.fixture(scope="module")
def progs(tmp_path_factory):
    d = tmp_path_factory.mktemp("sampler")
    out = {}
    for dim, flags in ((3, []), (2, ["-DMATTER_TWODIM"])):
        exe = str(d / f"sampler_check_{dim}d")
        subprocess.run(["g++", "-O2", "-std=c++17"] + flags + [SRC, "-o", exe], check=True)
        out[dim] = exe
    return out


def sample(progs, dim, r, L, ppc, rho=1550.0, seed=42):
    res = subprocess.run([progs[dim], str(r), str(L[0]), str(L[1]), str(L[2] if dim == 3 else 0.0), str(ppc), str(rho),
                          str(seed)], check=True, capture_output=True, text=True).stdout.split("\n")
    n, dx, vol, mass = res[0].split()
    x = np.array([[float(v) for v in ln.split()] for ln in res[1:] if ln.strip()])
    return int(n), float(dx), float(vol), float(mass), x


CASES = [(3, 0.02, (0.2, 0.2, 0.2), 8), (3, 0.015, (0.3, 0.1, 0.12), 8), (2, 0.002, (0.2, 0.1, 0.0), 6),
         (2, 0.004, (0.1, 0.3, 0.0), 6)]


@pytest.mark.parametrize("dim,r,L,ppc", CASES)
def test_poisson_disk_properties(progs, dim, r, L, ppc):
    n, dx, vol, mass, x = sample(progs, dim, r, L, ppc)
    assert x.shape == (n, dim) and n > 100
    V = float(np.prod(L[:dim]))
    assert abs(dx ** dim * n / (ppc * V) - 1) < 1e-12
    assert abs(vol - dx ** dim / ppc) <= 1e-15 * vol
    assert abs(mass - 1550.0 * vol) <= 1e-13 * mass
    hi = np.array(L[:dim])
    assert np.all(x >= 0) and np.all(x <= hi)
    tree = cKDTree(x)
    assert not tree.query_pairs(r * (1 - 1e-12)), "two samples closer than the radius"
    probes = np.random.default_rng(3).uniform(0, 1, (20000, dim)) * hi
    gap = float(tree.query(probes)[0].max())
    assert gap < 2 * r, f"a point of the box is {gap / r:.2f} r from every sample"
    ball = np.pi * (r / 2) ** 2 if dim == 2 else 4.0 / 3.0 * np.pi * (r / 2) ** 3
    phi = n * ball / V
    lo_phi, hi_phi = (0.45, 0.6) if dim == 2 else (0.3, 0.42)
    assert lo_phi < phi < hi_phi, f"packing fraction {phi:.3f}"


def test_seeded(progs):
    a = sample(progs, 3, 0.03, (0.2, 0.2, 0.2), 8, seed=7)[4]
    b = sample(progs, 3, 0.03, (0.2, 0.2, 0.2), 8, seed=7)[4]
    c = sample(progs, 3, 0.03, (0.2, 0.2, 0.2), 8, seed=8)[4]
    assert np.array_equal(a, b)
    assert a.shape != c.shape or not np.array_equal(a, c)
