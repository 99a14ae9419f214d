"""Seeded scene builders: jittered particle lattices and the benchmark/test workloads.

Generator: SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number generators", OOPSLA 2014)
over the counter (seed << 32) + k, mapped to a 53-bit uniform in [0, 1). Base seed 42, as the reference's
samplers use (src/mpm.cpp:107, src/sampling/sampling_particles.hpp:13). Particle k of a lattice with
spacing h sits at its lattice point plus U(-0.25h, +0.25h) per axis, so it stays inside its lattice cell.
Particle volume is h^dim (= dx^dim / ppc as in sampling_particles.hpp:29,91), mass rho * volume,
velocity zero.
"""
import math
import numpy as np

import mpm_abi as A

MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(counter):
    z = (counter + np.uint64(0x9E3779B97F4A7C15)) & MASK64
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & MASK64
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & MASK64
    return z ^ (z >> np.uint64(31))


def uniform(seed, n, stream=0):
    """n uniforms in [0,1) from stream `stream` of seed `seed`."""
    with np.errstate(over="ignore"):
        base = (np.uint64(seed) << np.uint64(40)) + (np.uint64(stream) << np.uint64(36))
        c = base + np.arange(n, dtype=np.uint64)
        return (splitmix64(c) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def jittered_lattice(counts, h, origin, seed=42, jitter=0.25):
    """Lattice of prod(counts) points, spacing h, lower corner `origin`, jitter +-jitter*h per axis.
    Ordering: last axis fastest (x-major)."""
    dim = len(counts)
    idx = np.indices(counts).reshape(dim, -1).T.astype(np.float64)
    n = idx.shape[0]
    x = np.empty((n, dim))
    for d in range(dim):
        u = uniform(seed, n, stream=d)
        x[:, d] = origin[d] + (idx[:, d] + 0.5) * h + (2.0 * u - 1.0) * jitter * h
    return x


def column_scene(dim, width, height, depth, dx, ppc_per_axis, rho, seed=42, x0=0.0, y0=0.0, z0=0.0):
    """Block of material [x0, x0+width] x [y0, y0+height] (x [z0, z0+depth]) sampled as a jittered lattice with
    ppc_per_axis**dim particles per grid cell. Returns positions, particle volume and particle mass."""
    h = dx / ppc_per_axis
    counts = [int(round(width / h)), int(round(height / h))] + ([int(round(depth / h))] if dim == 3 else [])
    origin = [x0, y0] + ([z0] if dim == 3 else [])
    x = jittered_lattice(counts, h, origin, seed=seed)
    vol = h ** dim
    return x, vol, rho * vol


def dp_collapse(dim=3, n_per_axis=None, dx=0.005, ppc_axis=2, musl=True, friction=0.4, bc="SlipFree",
                end_frame=1, fps=1.0, seed=42, side=0.4, flip_ratio=-2.0, precision="fp64", plastic="DP"):
    """Granular column collapse with Drucker-Prager (30 deg), Hencky E=1e6, nu=0.3, rho=1550, gravity 9.81,
    quadratic B-spline, APIC (flip_ratio=-2). 3D default: 0.4 m cube, dx=0.005, 8 ppc (4,096,000 particles),
    MUSL, SlipFree floor with Coulomb mu=0.4. n_per_axis overrides the lattice size keeping the spacing."""
    h = dx / ppc_axis
    if n_per_axis is not None:
        side = n_per_axis * h
    x, vol, mass = column_scene(dim, side, side, side, dx, ppc_axis, 1550.0, seed=seed)
    g = [0.0, -9.81, 0.0]
    cfg = A.default_config(dim=dim, spline_degree=2, precision=precision, plastic_model=plastic, E=1e6, nu=0.3,
                           rho=1550.0, M=math.tan(math.radians(30.0)), q_cohesion=0.0, use_mises_q=0,
                           use_musl=1 if musl else 0, flip_ratio=flip_ratio, cfl=0.5, gravity=g, dx=dx,
                           particle_volume=vol, particle_mass=mass, end_frame=end_frame, fps=fps)
    plates = [A.make_plate(0.0, "bottom", bc, friction)]
    return cfg, plates, [], x


def column_2d(n_side=100, dx=0.01, ppc_axis=2, seed=42, end_frame=1, fps=1.0):
    """2D plane-strain granular column 0.5 x 0.5 m, dx=0.01, 4 ppc (10,000 particles), quadratic, USL, APIC,
    Drucker-Prager 30 deg, E=1e6, nu=0.3, rho=1550, NoSlip floor, gravity 9.81, CFL 0.5."""
    h = dx / ppc_axis
    side = n_side * h
    x, vol, mass = column_scene(2, side, side, 0, dx, ppc_axis, 1550.0, seed=seed)
    cfg = A.default_config(dim=2, spline_degree=2, plastic_model="DP", E=1e6, nu=0.3, rho=1550.0,
                           M=math.tan(math.radians(30.0)), q_cohesion=0.0, use_musl=0, flip_ratio=-2.0, cfl=0.5,
                           gravity=[0.0, -9.81], dx=dx, particle_volume=vol, particle_mass=mass,
                           end_frame=end_frame, fps=fps)
    plates = [A.make_plate(0.0, "bottom", "NoSlip", 0.0)]
    return cfg, plates, [], x
