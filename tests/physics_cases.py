"""The reference's own physics tests (tests/tests.cpp), restated as scene builders + checks that run on
either library (the CPU oracle or the HIP engine) through the same C ABI.

The reference samples particles with tph_poisson (not available here, src/sampling/sampling_particles.hpp:23);
these builders use the seeded jittered lattice of scenes.py with the reference's particles-per-cell, so
each test keeps its physical tolerance but not its exact particle count.
"""
import math

import numpy as np

import mpm_abi as A
import scenes as S

G = 9.81


def lattice_block(dim, L, spacing, origin=(0.0, 0.0, 0.0), seed=42):
    counts = [max(1, int(round(L[d] / spacing))) for d in range(dim)]
    return S.jittered_lattice(counts, spacing, origin[:dim], seed=seed)


def run_to_end(e, chunk=20000):
    """simulate(): step until end_frame/fps (simulation.cpp:168-201)."""
    total = 0
    while True:
        rc, info = e.step(chunk)
        total += info.steps_done
        if rc:
            raise A.MPMError(rc, "step")
        if info.finished or info.steps_done == 0:
            return info, total


def analytic_slide(lib, bc, dim=3):
    """BoundaryTest.AnalyticSlipFree / AnalyticSlipStick (tests/tests.cpp:13-98): one particle released
    at (-1, 1) on the frictionless curve y = x^2 reaches v_x = sqrt(2 g) at the bottom, 1e-3 relative."""
    dx = 0.001
    vol = dx ** dim
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=1000.0 * vol, rho=1000.0,
                           end_frame=100, fps=100 / 0.5949238427, cfl=0.5, flip_ratio=-0.95, E=1e6, nu=0.3,
                           gravity=[0.0, -G, 0.0])
    e = A.Engine(lib, cfg, [], [A.make_object("Curve", bc, 0.0)])
    x = np.zeros((1, dim))
    x[0, 0], x[0, 1] = -1.0, 1.0
    e.set_particles(x)
    run_to_end(e)
    v = e.get_particles()["v"][0, 0]
    return abs(v - math.sqrt(2 * G)) / math.sqrt(2 * G)


def coulomb_incline(lib, bc, general, dim=3, spacing=None, Lz=None):
    """CoulombFrictionTest.{Plate,General}{SlipFree,SlipStick} (tests/tests.cpp:100-306): a 0.1 x 0.05 block on a
    floor with Coulomb mu = tan 15 deg under gravity tilted by 24 deg slides mean x = g t^2 (sin - mu cos)/2
    after t = 5 s, 1e-3 relative. The reference samples k_rad = 0.005 with Lz = 1 (3D default)."""
    theta = math.radians(24.0)
    mu = math.tan(math.radians(15.0))
    ppc_axis = 2
    spacing = spacing or 0.005
    dx = ppc_axis * spacing
    vol = spacing ** dim
    Lx, Ly = 0.1, 0.05
    Lz = Lz if Lz is not None else 1.0
    x = lattice_block(dim, [Lx, Ly, Lz], spacing)
    x[:, 0] -= 0.5 * Lx
    x[:, 1] += 0.5 * dx
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=1000.0 * vol, rho=1000.0, end_frame=5,
                           fps=1.0, cfl=0.5, flip_ratio=-0.95, E=1e6, nu=0.3,
                           gravity=[G * math.sin(theta), -G * math.cos(theta), 0.0], grid_reference_point=[0, 0, 0])
    plates, objs = [], []
    if general:
        objs = [A.make_object("Ground", bc, mu, param=0.0)]
    else:
        plates = [A.make_plate(0.0, "bottom", bc, mu)]
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(x)
    run_to_end(e)
    mean_x = e.get_particles()["x"][:, 0].mean()
    t = 5.0
    true_x = 0.5 * G * t * t * (math.sin(theta) - mu * math.cos(theta))
    return abs(mean_x - true_x) / true_x


def rotation_energy(lib, dim=3, spacing=0.01, end_frame=20, Lz=0.05):
    """EnergyTest.Rotation (tests/tests.cpp:543-603): rigid rotation with pure AFLIP (flip_ratio = -1), no
    gravity; kinetic energy may only decrease, by at most 1e-3 relative."""
    dx = 2 * spacing
    vol = spacing ** dim
    x = lattice_block(dim, [1.0, 1.0, Lz], spacing)
    x[:, 0] -= 0.5
    x[:, 1] -= 0.5
    v = np.zeros_like(x)
    v[:, 0] = -1.0 * x[:, 1] + 0.5
    v[:, 1] = 1.0 * x[:, 0] + 0.5
    e0 = 0.5 * np.sum(v[:, 0] ** 2 + v[:, 1] ** 2)
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=1550.0 * vol, rho=1550.0,
                           end_frame=end_frame, fps=1.0, cfl=0.5, flip_ratio=-1.0, E=1e6, nu=0.3)
    e = A.Engine(lib, cfg)
    e.set_particles(x, v=v)
    run_to_end(e)
    vv = e.get_particles()["v"]
    e1 = 0.5 * np.sum(vv ** 2)
    return (e0 - e1) / e0


def dp_collapse_runout(lib, model, dim=3, spacing=0.0015):
    """CollapseTest.DruckerPragerOne (DPSoft) / DruckerPragerTwo (DPVisc) (tests/tests.cpp:605-736): a
    0.20 x 0.15 x 0.10 column under gravity tilted by 10 deg, walls back/front/left SlipFree, NoSlip floor,
    70 frames at 50 fps; maximum x = 0.56 +- 0.011 m."""
    theta = math.radians(10.0)
    ppc_axis = 2
    dx = ppc_axis * spacing
    vol = spacing ** dim
    Lx, Ly, Lz = 0.20, 0.15, 0.10
    x = lattice_block(dim, [Lx, Ly, Lz], spacing)
    x[:, 1] += 0.5 * dx
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=1000.0 * vol, rho=1000.0, end_frame=70,
                           fps=50.0, cfl=0.5, flip_ratio=-0.95, E=1e6, nu=0.3, plastic_model=model, use_mises_q=0,
                           use_pradhana=1, xi=0.0, q_cohesion=0.0, M=math.tan(math.radians(30.0)), perzyna_visc=0.0,
                           perzyna_exp=1.0, gravity=[G * math.sin(theta), -G * math.cos(theta), 0.0],
                           grid_reference_point=[0, 0, 0])
    if dim == 3:
        plates = [A.make_plate(0.0, "back", "SlipFree"), A.make_plate(Lz, "front", "SlipFree"),
                  A.make_plate(0.0, "left", "SlipFree"), A.make_plate(0.0, "bottom", "NoSlip")]
    else:
        plates = [A.make_plate(0.0, "left", "SlipFree"), A.make_plate(0.0, "bottom", "NoSlip")]
    e = A.Engine(lib, cfg, plates)
    e.set_particles(x)
    run_to_end(e)
    return e.get_particles()["x"][:, 0].max()


def bulk_modulus(lib, dim=3, spacing=0.02, end_frame=100):
    """ElasticityTest.BulkModulus (tests/tests.cpp:476-541): plates compress a slab at 1e-3 m/s (10x slower in
    the first frame); K from the volume-averaged pressure against log J matches lambda + 2 mu / dim within 3 %."""
    ppc_axis = 2
    dx = ppc_axis * spacing
    vol = spacing ** dim
    Lx, Ly, Lz = 1.0, 1.0, 0.2
    x = lattice_block(dim, [Lx, Ly, Lz], spacing)
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=10000.0 * vol, rho=10000.0,
                           end_frame=end_frame, fps=1.0, cfl=0.6, flip_ratio=-0.95, E=1e6, nu=0.3)
    vel = 0.001
    plates = [A.make_plate(0 - 0.5 * dx, "bottom", "NoSlip", 0.0, v_object=(0.0, vel, 0.0), vmin_factor=10.0, load_factor=1.0),
              A.make_plate(Ly + 0.5 * dx, "top", "NoSlip", 0.0, v_object=(0.0, -vel, 0.0), vmin_factor=10.0, load_factor=1.0)]
    e = A.Engine(lib, cfg, plates)
    e.set_particles(x)
    run_to_end(e)
    p = e.get_particles()
    # computeAvgData (save_data.cpp:264-292) on the downloaded state
    E, nu = 1e6, 0.3
    lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    K = lam + 2 * mu / dim
    F = p["F"]
    kirch = np.zeros((dim, dim))
    Jsum = 0.0
    for i in range(F.shape[0]):
        w, Q = np.linalg.eigh(F[i] @ F[i].T)
        lnV = Q @ np.diag(0.5 * np.log(w)) @ Q.T
        tau = 2 * mu * lnV + lam * np.trace(lnV) * np.eye(dim)
        J = np.linalg.det(F[i]) * math.exp(p["eps_pl_vol"][i])
        Jsum += J
        kirch += tau * J
    kirch /= Jsum
    Javg = Jsum / F.shape[0]
    p_avg = -np.trace(kirch) / dim
    measured = p_avg / (-math.log(Javg))
    return abs(measured - K) / K


# ---------------------------------------------------------------------------------------------------------
# The two cases whose outcome depends on the particle sample, driven by an explicit sample: the host sampler's
# output (tests/cpp/sampler_check.cpp prints "N dx particle_volume particle_mass" and then the coordinates),
# so that the CPU oracle and the HIP engine run the same particles as tests/cpp/reference_tests.cpp.
def read_sample(text):
    lines = text.strip().splitlines()
    n, dx, vol, mass = lines[0].split()
    x = np.array([[float(v) for v in ln.split()] for ln in lines[1:]], dtype=np.float64)
    assert x.shape[0] == int(n)
    return x, float(dx), float(vol), float(mass)


def host_sample(exe, r, L, ppc, rho, seed=42):
    """Runs the sampler_check program `exe` (3D build) and returns (x, dx, particle_volume, particle_mass)."""
    import subprocess
    out = subprocess.run([exe, repr(r), repr(L[0]), repr(L[1]), repr(L[2]), repr(ppc), repr(rho), str(seed)],
                         capture_output=True, text=True, check=True).stdout
    return read_sample(out)


def mibf_slope(lib, sample, mibf, mises, n_threads=1):
    """One run of BoundaryTest.MIBF (tests/tests.cpp:308-474) on a given sample of the 0.1 x 0.05 x 0.05 block:
    DPVisc with zero viscosity, 32 deg slope, SlipFree floor with Coulomb mu = tan 30 deg, 0.5 s. The three
    runs of the test: plate friction (mibf = mises = False), MIBF grid friction (mibf), von Mises q with
    M = sqrt(3) tan 30 deg (mises). Returns the maximum x."""
    x0, dx, vol, mass = sample
    x = x0.copy()
    x[:, 0] -= 0.5 * 0.1
    x[:, 1] += 0.5 * dx
    friction = math.tan(math.radians(30.0))
    theta = math.radians(32.0)
    cfg = A.default_config(dim=3, dx=dx, particle_volume=vol, particle_mass=mass, rho=1000.0, end_frame=1, fps=2.0,
                           cfl=0.5, flip_ratio=-0.95, E=1e5, nu=0.3, n_threads=n_threads,
                           gravity=[G * math.sin(theta), -G * math.cos(theta), 0.0], grid_reference_point=[0, 0, 0],
                           plastic_model="DPVisc", use_pradhana=0, use_mises_q=int(mises), use_mibf=int(mibf),
                           M=math.sqrt(3.0) * friction if mises else friction, q_cohesion=0.0, perzyna_exp=1.0,
                           perzyna_visc=0.0)
    e = A.Engine(lib, cfg, [A.make_plate(0.0, "bottom", "SlipFree", friction)])
    e.set_particles(x)
    run_to_end(e)
    return float(e.get_particles()["x"][:, 0].max())


def bulk_modulus_on_sample(lib, sample, n_threads=1):
    """ElasticityTest.BulkModulus (tests/tests.cpp:476-541) on a given sample of the 1 x 1 x 0.2 slab: returns
    (measured K, lambda + 2 mu / 3, J average), the measured K from computeAvgData's Kirchhoff average
    (save_data.cpp:264-292) as the reference test computes it."""
    x, dx, vol, mass = sample
    cfg = A.default_config(dim=3, dx=dx, particle_volume=vol, particle_mass=mass, rho=10000.0, end_frame=100, fps=1.0,
                           cfl=0.6, flip_ratio=-0.95, E=1e6, nu=0.3, n_threads=n_threads)
    vel = 0.001
    plates = [A.make_plate(0 - 0.5 * dx, "bottom", "NoSlip", 0.0, v_object=(0.0, vel, 0.0), vmin_factor=10.0, load_factor=1.0),
              A.make_plate(1.0 + 0.5 * dx, "top", "NoSlip", 0.0, v_object=(0.0, -vel, 0.0), vmin_factor=10.0, load_factor=1.0)]
    e = A.Engine(lib, cfg, plates)
    e.set_particles(x)
    run_to_end(e)
    p = e.get_particles()
    E, nu, dim = 1e6, 0.3, 3
    lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    K = lam + 2 * mu / dim
    F = p["F"]
    w, Q = np.linalg.eigh(F @ F.transpose(0, 2, 1))
    lnV = np.einsum("nij,nj,nkj->nik", Q, 0.5 * np.log(w), Q)
    tau = 2 * mu * lnV + lam * np.trace(lnV, axis1=1, axis2=2)[:, None, None] * np.eye(dim)
    J = np.linalg.det(F) * np.exp(p["eps_pl_vol"])
    kirch = (tau * J[:, None, None]).sum(axis=0) / J.sum()
    Javg = J.sum() / F.shape[0]
    return -np.trace(kirch) / dim / (-math.log(Javg)), K, Javg


# ---------------------------------------------------------------------------------------------------------
# Level-set terrain (ObjectVdb, src/objects/object_vdb.hpp) on a lattice: the signed distance to the parabola
# y = x^2 (the curve of the reference's tests/curve.vdb), negative below it.
def parabola_sdf(xs, ys):
    """Signed distance of the points (xs[i], ys[j]) to y = x^2: the nearest curve point t solves
    2 t^3 + (1 - 2 y) t - x = 0 (a depressed cubic t^3 + p t + q = 0), taken over its real roots by Cardano's
    formula (one real root) or the trigonometric form (three)."""
    X, Y = np.meshgrid(np.asarray(xs, float), np.asarray(ys, float), indexing="ij")
    p = (1.0 - 2.0 * Y) / 2.0
    q = -X / 2.0
    disc = (q / 2.0) ** 2 + (p / 3.0) ** 3
    roots = []
    sq = np.sqrt(np.maximum(disc, 0.0))
    roots.append(np.where(disc > 0, np.cbrt(-q / 2.0 + sq) + np.cbrt(-q / 2.0 - sq), np.nan))
    with np.errstate(invalid="ignore", divide="ignore"):
        m = 2.0 * np.sqrt(np.maximum(-p / 3.0, 0.0))
        arg = np.clip(np.where(p < 0, (3.0 * q / (2.0 * p)) * np.sqrt(np.maximum(-3.0 / p, 0.0)), 0.0), -1.0, 1.0)
        th = np.arccos(arg) / 3.0
        for k in range(3):
            roots.append(np.where(disc <= 0, m * np.cos(th - 2.0 * np.pi * k / 3.0), np.nan))
    d = np.full(X.shape, np.inf)
    for t in roots:
        dd = np.sqrt((t - X) ** 2 + (t * t - Y) ** 2)
        d = np.where(np.isfinite(dd) & (dd < d), dd, d)
    return np.where(Y <= X * X, -d, d)


def curve_lattice(dim, h=0.005, x=(-1.3, 0.4), y=(-0.3, 1.3), z=(-0.02, 0.02)):
    """(phi, origin, spacing) of parabola_sdf on a lattice covering x, y (and z in 3D, where phi does not vary)."""
    xs = np.arange(x[0], x[1] + 0.5 * h, h)
    ys = np.arange(y[0], y[1] + 0.5 * h, h)
    phi = parabola_sdf(xs, ys)
    origin = [xs[0], ys[0]]
    if dim == 3:
        zs = np.arange(z[0], z[1] + 0.5 * h, h)
        phi = np.repeat(phi[:, :, None], zs.size, axis=2)
        origin.append(zs[0])
    return phi.astype(np.float32), origin, h


def sdf_slide(lib, bc="SlipFree", dim=3):
    """BoundaryTest.VDB (tests/tests_vdb.cpp:40-82): one particle released at (-1, 1) on the level set of the curve
    y = x^2 (frictionless) reaches v_x = sqrt(2 g) at the bottom, 1e-3 relative."""
    dx = 0.001
    vol = dx ** dim
    cfg = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=1000.0 * vol, rho=1000.0,
                           end_frame=100, fps=100 / 0.5949238427, cfl=0.5, flip_ratio=-0.95, E=1e6, nu=0.3,
                           gravity=[0.0, -G, 0.0])
    e = A.Engine(lib, cfg)
    phi, origin, h = curve_lattice(dim)
    e.add_sdf(phi, origin, h, bc, 0.0)
    x = np.zeros((1, dim))
    x[0, 0], x[0, 1] = -1.0, 1.0
    e.set_particles(x)
    run_to_end(e)
    v = e.get_particles()["v"][0, 0]
    return abs(v - math.sqrt(2 * G)) / math.sqrt(2 * G)
