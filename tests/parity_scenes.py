"""Seeded scenes shared by the parity tests: each returns (config, plates, objects, particle fields).

Every scene is small enough for the CPU oracle to finish 10 steps in about a second. Together they
exercise every plastic model and hardening law, every flip_ratio regime (PIC/FLIP, AFLIP, APIC), USL and
MUSL, quadratic and cubic splines, 2D and 3D, MIBF, moving plates, every plate type and BC, and every
analytic object kind.
"""
import math

import numpy as np

import mpm_abi as A
import scenes as S


def _block(dim, n_axis, h, origin, seed=42):
    counts = [n_axis[0], n_axis[1]] + ([n_axis[2]] if dim == 3 else [])
    return S.jittered_lattice(counts, h, origin, seed=seed)


def base_cfg(dim, dx, ppc_axis, rho=1550.0, **kw):
    h = dx / ppc_axis
    vol = h ** dim
    c = A.default_config(dim=dim, dx=dx, particle_volume=vol, particle_mass=rho * vol, rho=rho, **kw)
    return c


def dp_collapse(dim=3, musl=False, n=16, spline=2, bc="SlipFree", friction=0.4, flip=-2.0, model="DP", **kw):
    dx = 0.01
    h = dx / 2
    cfg = base_cfg(dim, dx, 2, spline_degree=spline, plastic_model=model, M=math.tan(math.radians(30)),
                   use_musl=int(musl), flip_ratio=flip, gravity=[0.0, -9.81, 0.0], end_frame=1, fps=1.0, **kw)
    x = _block(dim, [n, n, n], h, [0.0, 0.0, 0.0])
    return cfg, [A.make_plate(0.0, "bottom", bc, friction)], [], {"x": x}


def column_2d(n=40, musl=False):
    return dp_collapse(dim=2, musl=musl, n=n, bc="NoSlip", friction=0.0)


def incline_mccmui(n=14, hardening="ExpoImpl"):
    """Cubic B-spline, AFLIP, mu(I)-MCC, MIBF, SlipStick floor and a rotated box bump, 35 deg tilted gravity."""
    dx = 0.01
    th = math.radians(35.0)
    g = [9.81 * math.sin(th), -9.81 * math.cos(th), 0.0]
    cfg = base_cfg(3, dx, 2, rho=1500.0, spline_degree=3, plastic_model="MCCMui", hardening_law=hardening,
                   mu_1=0.38, mu_2=0.64, I_ref=0.279, grain_diameter=1e-3, rho_s=2500.0, E=1.2e7, nu=0.3,
                   use_mibf=1, flip_ratio=-1.0, gravity=g, end_frame=1, fps=1.0, p0=1000.0, xi=0.0, beta=0.0)
    x = _block(3, [2 * n, n // 2, n], dx / 2, [0.0, 0.012, 0.0])
    plates = [A.make_plate(0.0, "bottom", "SlipStick", 0.38)]
    objs = [A.make_object("BoxRotated", "SlipStick", 0.38, center=(0.12, 0.0, 0.035), half_extents=(0.01, 0.02, 0.05),
                          rotation=np.eye(3))]
    return cfg, plates, objs, {"x": x}


def model_block(model, dim=3, n=12, hardening="ExpoImpl", flip=-0.95, musl=False, **kw):
    """Block dropped onto a NoSlip floor with side walls, one plastic model."""
    dx = 0.01
    h = dx / 2
    g = [2.0, -9.81, 0.0]
    params = dict(q_max=300.0, q_min=100.0, M=math.tan(math.radians(30)), q_cohesion=50.0, p0=2000.0, beta=0.2,
                  xi=0.5, perzyna_visc=0.0, perzyna_exp=1.0, mu_1=0.38, mu_2=0.64)
    params.update(kw)
    cfg = base_cfg(dim, dx, 2, plastic_model=model, hardening_law=hardening, flip_ratio=flip, use_musl=int(musl),
                   gravity=g, end_frame=1, fps=1.0, E=2e5, nu=0.3, **params)
    x = _block(dim, [n, n, n], h, [0.0, 0.02, 0.0])
    plates = [A.make_plate(0.0, "bottom", "NoSlip"), A.make_plate(-0.005, "left", "SlipFree", 0.2),
              A.make_plate(n * h + 0.01, "right", "SlipStick", 0.3)]
    if dim == 3:
        plates += [A.make_plate(-0.004, "back", "SlipFree", 0.1), A.make_plate(n * h + 0.004, "front", "SlipStick", 0.2)]
    # initial state: stretched/sheared F so every branch of the return maps is reached in the first steps
    rng = np.random.default_rng(7)
    npart = x.shape[0]
    F = np.tile(np.eye(dim), (npart, 1, 1))
    F += 0.02 * rng.standard_normal((npart, dim, dim))
    F[: npart // 4] *= 1.01   # some particles in tension (DP tip branch)
    v = 0.3 * rng.standard_normal((npart, dim))
    fields = {"x": x, "F": F, "v": v}
    if model in ("DPSoft", "DPVisc", "DPMui") and not kw.get("use_mibf"):
        fields["eps_pl_vol_pradhana"] = 1e-4 * rng.standard_normal(npart)
    return cfg, plates, [], fields


def rotation_elastic(dim=2, n=24, flip=-1.0):
    """Rigid rotation, no gravity, Hencky elasticity (EnergyTest.Rotation, tests/tests.cpp:543-603)."""
    dx = 0.02
    h = dx / 2
    cfg = base_cfg(dim, dx, 2, plastic_model="NoPlasticity", flip_ratio=flip, gravity=[0, 0, 0], end_frame=1, fps=1.0)
    x = _block(dim, [n, n, max(2, n // 6)], h, [-n * h / 2, -n * h / 2, 0.0])
    v = np.zeros_like(x)
    v[:, 0] = -1.0 * x[:, 1] + 0.5
    v[:, 1] = 1.0 * x[:, 0] + 0.5
    return cfg, [], [], {"x": x, "v": v}


def neohookean_drop(n=10):
    cfg, plates, objs, f = dp_collapse(dim=3, n=n, model="NoPlasticity", bc="NoSlip")
    A.set_field(cfg, "elastic_model", "NeoHookean")
    return cfg, plates, objs, f


def object_scene(kind, bc, dim=3, friction=0.3, param=0.0, n=10):
    """Block falling onto one analytic object (src/objects/object_*.hpp)."""
    dx = 0.01
    h = dx / 2
    placements = {  # lower corner of the block so that the object surface is met within a few steps
        "Curve": [0.02, 0.0006, 0.0], "Ground": [0.0, param + 0.0006, 0.0], "Ramp": [0.02, -0.099, 0.0],
        "Gate": [-0.03, 0.0, 0.0], "Bump": [0.40, 0.045, 0.0], "Silo": [-0.03, -0.5, -0.03],
        "BoxRotated": [0.0, 0.011, 0.0],
    }
    if kind == "Silo":  # against the wall r = tanh(y) + 1 at y = -0.5 (r = 0.538), moving outwards
        placements["Silo"] = [0.52, -0.5, -0.025]
    cfg = base_cfg(dim, dx, 2, plastic_model="DP", M=0.6, flip_ratio=-2.0, gravity=[0.5, -9.81, 0.0], end_frame=1,
                   fps=1.0)
    x = _block(dim, [n, n, n], h, placements[kind])
    v = np.zeros_like(x)
    v[:, 1] = -2.0
    if kind == "Gate":
        v[:, 1] = 2.0
    if kind == "Silo":
        v[:, 0] = 1.5
    rot = None
    center, half = (0, 0, 0), (0, 0, 0)
    if kind == "BoxRotated":
        t = math.radians(20.0)
        rot = np.array([[math.cos(t), -math.sin(t), 0], [math.sin(t), math.cos(t), 0], [0, 0, 1.0]])[:dim, :dim]
        center, half = (0.025, 0.0, 0.025), (0.05, 0.01, 0.05)
    o = A.make_object(kind, bc, friction, param=param, center=center, half_extents=half, rotation=rot)
    return cfg, [], [o], {"x": x, "v": v}


def sdf_scene(bc, dim=3, friction=0.3, n=10):
    """Block falling onto the level set of the curve y = x^2 sampled on a lattice (ObjectVdb,
    src/objects/object_vdb.hpp), placed as object_scene places it on the analytic curve."""
    import physics_cases as PC
    cfg, plates, _, f = object_scene("Curve", bc, dim=dim, friction=friction, n=n)
    phi, origin, h = PC.curve_lattice(dim, h=0.004, x=(-0.1, 0.2), y=(-0.1, 0.15), z=(-0.05, 0.1))
    f["sdf"] = [(phi, origin, h, bc, friction)]
    return cfg, plates, [], f


def moving_plates(dim=3, n=10):
    """Plates moving with vmin_factor/load_factor (ElasticityTest.BulkModulus, tests/tests.cpp:476-541)."""
    dx = 0.01
    h = dx / 2
    cfg = base_cfg(dim, dx, 2, plastic_model="NoPlasticity", flip_ratio=-0.95, gravity=[0, 0, 0], end_frame=1,
                   fps=100.0, rho=10000.0)
    x = _block(dim, [n, n, n], h, [0.0, 0.0, 0.0])
    vel = 0.5
    plates = [A.make_plate(-0.5 * dx, "bottom", "NoSlip", 0.0, v_object=(0.0, vel, 0.0), vmin_factor=10.0, load_factor=1.0),
              A.make_plate(n * h + 0.5 * dx, "top", "NoSlip", 0.0, v_object=(0.0, -vel, 0.0), vmin_factor=10.0, load_factor=1.0)]
    return cfg, plates, [], {"x": x}


def gravity_ramp(dim=2, n=20):
    cfg, plates, objs, f = dp_collapse(dim=dim, n=n)
    A.set_field(cfg, "gravity_special", 1)
    A.set_field(cfg, "gravity_time", 0.002)
    return cfg, plates, objs, f


def periodic_layer(dim=2, spline=2, musl=False, flip=-2.0, cells_x=16, copies=1, model="DP", mibf=0, seed=11,
                   cells_y=6, cells_z=4, ppc_axis=2, precision="fp64"):
    """Layer of material filling the periodic interval 0 <= x < Lx (Lx = cells_x dx) on a SlipFree floor,
    moving in +x at about 5 m/s under gravity, with seeded velocity and deformation perturbations, run with
    periodic boundaries in x (pbc = 1; pbc.cpp, remesh.cpp:14-36, position_update.cpp:26-33).
    copies > 1: the same layer repeated `copies` times along x without periodic boundaries (the unrolled
    periodic problem), on a grid anchored at the origin so that its nodes sit at multiples of dx, as the
    periodic grid's do. Returns (cfg, plates, objects, fields, Lx)."""
    dx = 0.01
    h = dx / ppc_axis
    counts = [ppc_axis * cells_x, ppc_axis * cells_y] + ([ppc_axis * cells_z] if dim == 3 else [])
    base = _block(dim, counts, h, [0.0] * dim, seed=seed)
    rng = np.random.default_rng(seed)
    npart = base.shape[0]
    v = 0.2 * rng.standard_normal((npart, dim))
    v[:, 0] += 5.0
    F = np.tile(np.eye(dim), (npart, 1, 1)) + 0.01 * rng.standard_normal((npart, dim, dim))
    Lx = cells_x * dx
    cfg = base_cfg(dim, dx, ppc_axis, spline_degree=spline, plastic_model=model, M=math.tan(math.radians(30)),
                   use_musl=int(musl), flip_ratio=flip, gravity=[0.0, -9.81, 0.0], end_frame=1, fps=1.0,
                   grid_reference_point=[0.0, 0.0, 0.0], use_mibf=mibf, precision=precision)
    if copies == 1:
        cfg.pbc = 1
        cfg.Lx, cfg.Ly, cfg.Lz = Lx, counts[1] * h + 2 * dx, (counts[2] * h if dim == 3 else 1.0)
        x = base
    else:
        shift = np.zeros(dim)
        shift[0] = Lx
        x = np.concatenate([base + j * shift for j in range(copies)])
        v, F = np.tile(v, (copies, 1)), np.tile(F, (copies, 1, 1))
    return cfg, [A.make_plate(0.0, "bottom", "SlipFree", 0.3)], [], {"x": x, "v": v, "F": F}, Lx


def periodic(dim=2, spline=2, musl=False, flip=-2.0, **kw):
    cfg, plates, objs, f, _ = periodic_layer(dim, spline, musl, flip, **kw)
    return cfg, plates, objs, f


SCENES = {
    "dp3_usl_apic": lambda: dp_collapse(3, musl=False),
    "dp3_musl_apic": lambda: dp_collapse(3, musl=True),
    "dp3_musl_cubic": lambda: dp_collapse(3, musl=True, spline=3, n=12),
    "col2d_usl": lambda: column_2d(),
    "col2d_musl_cubic": lambda: dp_collapse(2, musl=True, n=30, spline=3, bc="SlipFree", friction=0.3),
    "incline_mccmui_mibf": lambda: incline_mccmui(),
    "vm": lambda: model_block("VM"),
    "dpsoft": lambda: model_block("DPSoft", xi=2.0),
    "vmvisc_closed": lambda: model_block("VMVisc", xi=0.0, perzyna_visc=0.01),
    "vmvisc_newton": lambda: model_block("VMVisc", xi=0.5, perzyna_visc=0.01, perzyna_exp=2.0),
    "dpvisc": lambda: model_block("DPVisc", perzyna_visc=0.01),
    "dpvisc_newton_mibf": lambda: model_block("DPVisc", perzyna_visc=0.01, perzyna_exp=2.0, use_mibf=1),
    "dp_apic_mibf": lambda: model_block("DP", use_mibf=1, flip=-2.0),  # APIC + MIBF: 5 values per node in P2G
    "dpmui": lambda: model_block("DPMui", q_cohesion=20.0),
    "mcc_nohard": lambda: model_block("MCC", hardening="NoHard"),
    "mcc_expoexpl": lambda: model_block("MCC", hardening="ExpoExpl"),
    "mcc_sinhexpl": lambda: model_block("MCC", hardening="SinhExpl"),
    "mcc_expoimpl": lambda: model_block("MCC", hardening="ExpoImpl"),
    "mcc_sinhimpl": lambda: model_block("MCC", hardening="SinhImpl"),
    "mccvisc_mibf": lambda: model_block("MCCVisc", hardening="ExpoImpl", perzyna_visc=0.01, use_mibf=1),
    "mccvisc_newton": lambda: model_block("MCCVisc", hardening="SinhImpl", perzyna_visc=0.01, perzyna_exp=2.0),
    "mccmui_sinhimpl_2d": lambda: model_block("MCCMui", dim=2, n=20, hardening="SinhImpl"),
    "dp2d_picflip": lambda: model_block("DP", dim=2, n=20, flip=0.95),
    "dp3_aflip_musl": lambda: model_block("DP", flip=-1.0, musl=True),
    "rotation2d_aflip": lambda: rotation_elastic(2),
    "rotation3d_aflip": lambda: rotation_elastic(3, n=16),
    "neohookean": lambda: neohookean_drop(),
    "obj_curve_slipfree": lambda: object_scene("Curve", "SlipFree"),
    "obj_curve_slipstick_2d": lambda: object_scene("Curve", "SlipStick", dim=2),
    "obj_ground_noslip": lambda: object_scene("Ground", "NoSlip", param=0.0),
    "obj_ramp_slipfree": lambda: object_scene("Ramp", "SlipFree"),
    "obj_gate_slipstick": lambda: object_scene("Gate", "SlipStick", param=0.05),
    "obj_bump_slipfree": lambda: object_scene("Bump", "SlipFree"),
    "obj_silo_slipstick": lambda: object_scene("Silo", "SlipStick", param=-1.0),
    "obj_box_rotated": lambda: object_scene("BoxRotated", "SlipFree"),
    "obj_sdf_curve_slipfree": lambda: sdf_scene("SlipFree"),
    "obj_sdf_curve_slipstick_2d": lambda: sdf_scene("SlipStick", dim=2),
    "moving_plates": lambda: moving_plates(),
    "gravity_ramp_2d": lambda: gravity_ramp(),
    "periodic2d_usl_apic": lambda: periodic(2),
    "periodic2d_musl_picflip": lambda: periodic(2, musl=True, flip=0.95),
    "periodic3d_usl_cubic_aflip_mibf": lambda: periodic(3, spline=3, flip=-1.0, mibf=1, cells_x=12),
    "periodic3d_musl_apic_vm": lambda: periodic(3, musl=True, model="VM", cells_x=12),
}


def build(lib, name, n_threads=1):
    cfg, plates, objs, fields = SCENES[name]()
    cfg.n_threads = n_threads
    e = A.Engine(lib, cfg, plates, objs)
    for phi, origin, h, bc, friction in fields.pop("sdf", []):
        e.add_sdf(phi, origin, h, bc, friction)
    x = fields.pop("x")
    e.set_particles(x, **fields)
    return e
