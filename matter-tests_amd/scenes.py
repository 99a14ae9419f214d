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


def batch_collapses(count=128, seed=42, n_min=5_000, n_max=500_000, precision="split", dx=0.005, musl=False):
    """Independent granular column collapses of the batched workload: per scene, dimension 2 or 3 (even
    odds), particle count log-uniform in [n_min, n_max], aspect ratio height/width uniform in [0.5, 2],
    friction angle uniform in [20, 40] deg, and material Drucker-Prager or Perzyna-von Mises (even odds;
    yield stress q_y uniform in [1e3, 1e4] Pa, viscosity eta uniform in [1, 100] Pa s, given to the reference's
    Perzyna law as perzyna_visc = eta / q_y in seconds, with q_max = q_min = q_y and exponent 1). Quadratic
    B-spline, APIC, USL by default, 4 (2D) or 8 (3D) particles per cell, NoSlip floor in 2D and SlipFree
    floor with mu = tan(friction angle) in 3D, E = 1e6 Pa, nu = 0.3, rho = 1550. Draws come from stream
    1000 + k of the seed. Returns a list of (name, cfg, plates, objects, x)."""
    out = []
    h = dx / 2
    for k in range(count):
        u = uniform(seed, 6, stream=1000 + k)
        dim = 2 if u[0] < 0.5 else 3
        n_target = math.exp(math.log(n_min) + u[1] * (math.log(n_max) - math.log(n_min)))
        aspect = 0.5 + 1.5 * u[2]
        phi = 20.0 + 20.0 * u[3]
        model = "DP" if u[4] < 0.5 else "VMVisc"
        if dim == 3:
            nx = max(4, int(round((n_target / aspect) ** (1.0 / 3.0))))
            counts = [nx, max(4, int(round(aspect * nx))), nx]
        else:
            nx = max(4, int(round(math.sqrt(n_target / aspect))))
            counts = [nx, max(4, int(round(aspect * nx)))]
        x = jittered_lattice(counts, h, [0.0] * dim, seed=seed + 1 + k)
        vol = h ** dim
        params = dict(dim=dim, spline_degree=2, precision=precision, E=1e6, nu=0.3, rho=1550.0,
                      M=math.tan(math.radians(phi)), q_cohesion=0.0, use_musl=1 if musl else 0, flip_ratio=-2.0,
                      cfl=0.5, gravity=[0.0, -9.81, 0.0][:dim], dx=dx, particle_volume=vol,
                      particle_mass=1550.0 * vol, end_frame=1000, fps=1.0)
        if model == "VMVisc":
            qy = 1e3 + 9e3 * u[5]
            eta = 1.0 + 99.0 * uniform(seed, 7, stream=1000 + k)[6]
            params.update(plastic_model="VMVisc", q_max=qy, q_min=qy, perzyna_exp=1.0, perzyna_visc=eta / qy, xi=0.0)
        else:
            params.update(plastic_model="DP")
        cfg = A.default_config(**params)
        if dim == 3:
            plates = [A.make_plate(0.0, "bottom", "SlipFree", math.tan(math.radians(phi)))]
        else:
            plates = [A.make_plate(0.0, "bottom", "NoSlip", 0.0)]
        out.append((f"scene{k:03d}_{dim}d_{model}", cfg, plates, [], x))
    return out


def incline_flow(n_target=None, dx=0.004, ppc_axis=2, seed=42, precision="split", slope_deg=35.0, lattice=None):
    """Flow down an inclined plane with a rectangular bump (the analytic-plane configuration): cubic
    B-spline, AFLIP (flip_ratio = -1), mu(I)-Modified Cam-Clay (mu_s = 0.38, mu_2 = 0.64, I_0 = 0.279, grain
    diameter 1 mm, rho_s = 2500 kg/m^3, solid fraction 0.6 so rho = 1500 kg/m^3, MCC defaults p0 = 1000 Pa,
    beta = 0, xi = 0, ExpoImpl hardening), Hencky K = 1e7 Pa with nu = 0.3 (E = 3K(1 - 2 nu) = 1.2e7 Pa), MIBF
    friction, SlipStick floor and a SlipStick box bump, gravity tilted by the slope angle. The slab has the
    proportions 1.0 x 0.1 x 0.3 m (along the slope x, height y, width z); `lattice` gives its lattice counts
    directly, otherwise they are chosen so that the slab holds about n_target particles at ppc_axis^3
    particles per cell. Returns (cfg, plates, objects, x)."""
    h = dx / ppc_axis
    if lattice is None:
        nx = max(8, int(round((n_target / 0.03) ** (1.0 / 3.0))))
        lattice = [nx, max(4, int(round(0.1 * nx))), max(4, int(round(0.3 * nx)))]
    L = [c * h for c in lattice]
    x = jittered_lattice(lattice, h, [0.0, 0.0, 0.0], seed=seed)
    th = math.radians(slope_deg)
    g = [9.81 * math.sin(th), -9.81 * math.cos(th), 0.0]
    K, nu = 1e7, 0.3
    E = 3.0 * K * (1.0 - 2.0 * nu)
    vol = h ** 3
    rho = 0.6 * 2500.0
    cfg = A.default_config(dim=3, spline_degree=3, precision=precision, plastic_model="MCCMui", hardening_law="ExpoImpl",
                           mu_1=0.38, mu_2=0.64, I_ref=0.279, grain_diameter=1e-3, rho_s=2500.0, E=E, nu=nu, rho=rho,
                           use_mibf=1, flip_ratio=-1.0, gravity=g, dx=dx, particle_volume=vol, particle_mass=rho * vol,
                           p0=1000.0, beta=0.0, xi=0.0, cfl=0.5, end_frame=1000, fps=1.0)
    plates = [A.make_plate(0.0, "bottom", "SlipStick", 0.38)]
    bump = A.make_object("BoxRotated", "SlipStick", 0.38, center=(L[0] + 0.1 * L[0], 0.0, 0.5 * L[2]),
                         half_extents=(0.02 * L[0] + 2 * dx, 0.3 * L[1] + dx, L[2]), rotation=np.eye(3))
    return cfg, plates, [bump], x


def served_requests(count=48, rate=50.0, seed=42, n_min=10_000, n_max=2_000_000, steps_min=50, steps_max=500,
                    precision="split"):
    """Open-loop Poisson stream of scene requests: inter-arrival times exponential with mean 1/rate, each
    request a 3D collapse (Drucker-Prager, quadratic, APIC, MUSL) or an inclined flow (incline_flow) with even
    odds, particle count log-uniform in [n_min, n_max], steps uniform in [steps_min, steps_max]. Draws come
    from stream 5000 + k of the seed. Returns a list of (name, cfg, plates, objects, x, steps, arrival_s)."""
    out = []
    t = 0.0
    for k in range(count):
        u = uniform(seed, 4, stream=5000 + k)
        t += -math.log(1.0 - u[0]) / rate
        n_target = math.exp(math.log(n_min) + u[1] * (math.log(n_max) - math.log(n_min)))
        steps = steps_min + int(u[2] * (steps_max - steps_min + 1))
        if u[3] < 0.5:
            side = max(8, int(round(n_target ** (1.0 / 3.0))))
            cfg, plates, objs, x = dp_collapse(n_per_axis=side, musl=True, precision=precision, seed=seed + 100 + k,
                                               end_frame=1000)
            name = f"req{k:03d}_collapse"
        else:
            cfg, plates, objs, x = incline_flow(n_target, seed=seed + 100 + k, precision=precision)
            name = f"req{k:03d}_incline"
        out.append((name, cfg, plates, objs, x, steps, t))
    return out
