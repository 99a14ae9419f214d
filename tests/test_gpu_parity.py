"""Parity of the HIP engine (through the C ABI of include/mpm.h) with the CPU oracle on seeded scenes.

Bar (plain FP64 form): grid geometry bit-exact (node counts, origin and every node coordinate, i.e. the
cell-index frame of the step); particle x, v, F, B and plastic scalars, and grid mass / velocity / flip,
within 1e-10 norm-wise relative error after 1 and after 10 steps, each field measured against
max(its largest magnitude, its physical scale) (field_floors); NaNs, which the reference itself produces
in some states, at the same places; dt, time and plastic count equal.
The oracle runs single-threaded so its sums are in the reference's sequential order; the engine's
scatter uses atomics, so its sums differ from the oracle's in rounding only.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
PARTICLE_FIELDS = ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "eps_pl_vol_pradhana", "delta_gamma",
                   "viscosity", "muI")


def field_floors(op, dt, dx):
    """Physical scale of each field. B is measured against |v| dx/4 (B = C dx^2/4 with C ~ |v|/dx).
    The plastic history scalars are not part of the state bar (x, v, F, B, grid) and carry the
    velocity noise of reordered FP64 sums: a velocity difference of ~1e-13 m/s gives grad v differences
    ~1e-13/dx, hence strain increments that differ by ~1e-15 per step (dt grad v). Strains are therefore
    measured against 1e-3 (1e-10 of it is 1e-13), delta_gamma against 1e-3/dt, friction coefficients
    against 1 and mu(I) viscosities against 1e3 Pa s."""
    vmax = float(np.nanmax(np.abs(op["v"]))) if op["v"].size else 0.0
    return {"x": 0.0, "v": 1e-6, "F": 1.0, "Bmat": vmax * dx / 4, "eps_pl_dev": 1e-3, "eps_pl_vol": 1e-3,
            "eps_pl_vol_pradhana": 1e-3, "delta_gamma": 1e-3 / dt, "viscosity": 1e3, "muI": 1.0}


def compare_state(ge, oe, label, dt):
    gp, op = ge.get_particles(), oe.get_particles()
    fl = field_floors(op, dt, ge.cfg.dx)
    errs = {k: rel_err(gp[k], op[k], fl[k]) for k in PARTICLE_FIELDS}
    bad = {k: e for k, e in errs.items() if not e <= TOL}
    assert not bad, f"{label}: particle fields beyond {TOL}: {bad}"
    gg, og = ge.get_grid(), oe.get_grid()
    gi, oi = gg["info"], og["info"]
    assert (gi.Nx, gi.Ny, gi.Nz, gi.grid_nodes) == (oi.Nx, oi.Ny, oi.Nz, oi.grid_nodes), label
    assert (gi.xc, gi.yc, gi.zc) == (oi.xc, oi.yc, oi.zc), label
    for ax in ("x", "y", "z"):
        assert np.array_equal(gg[ax], og[ax]), f"{label}: node coordinates along {ax} differ"
    active = og["mass"] > 0
    assert np.array_equal(gg["mass"] > 0, active), f"{label}: set of active nodes differs"
    # grid bar: mass and momentum (m v), momentum increment (m flip), and m * friction. Velocities of
    # edge nodes whose mass comes from weights barely above 1e-25 are ratios of tiny sums; their momentum
    # is what the particles see.
    m_g, m_o = gg["mass"][active], og["mass"][active]
    e = rel_err(m_g, m_o)
    assert e <= TOL, f"{label}: grid mass rel err {e:.3e}"
    mom_o = m_o[:, None] * og["v"][active]
    mfloor = 1e-6 * float(m_o.max())
    # grid flip is formed only where something reads it (flip_ratio >= -1); APIC never does
    for k in ("v", "flip") if ge.cfg.flip_ratio >= -1 else ("v",):
        e = rel_err(m_g[:, None] * gg[k][active], m_o[:, None] * og[k][active], mfloor)
        assert e <= TOL, f"{label}: grid mass*{k} rel err {e:.3e}"
    if ge.cfg.use_mibf:
        e = rel_err(m_g * gg["friction"][active], m_o * og["friction"][active], float(m_o.max()))
        assert e <= TOL, f"{label}: grid mass*friction rel err {e:.3e}"
    del mom_o
    gpl, opl = ge.get_plates(), oe.get_plates()
    for a, b in zip(gpl, opl):
        assert (a.pos_object, a.pos_lower, a.pos_upper) == (b.pos_object, b.pos_lower, b.pos_upper), label
    return errs


@pytest.mark.parametrize("name", list(PS.SCENES))
def test_parity_1_and_10_steps(engine, oracle, name):
    print(f"[parity] scene={name}", flush=True)
    ge = PS.build(engine, name)
    oe = PS.build(oracle, name)
    for nsteps, label in ((1, "after 1 step"), (9, "after 10 steps")):
        rg, ig = ge.step(nsteps)
        ro, io = oe.step(nsteps)
        assert rg == ro == 0, f"{name}: status engine={A.status_names(rg)} oracle={A.status_names(ro)}"
        assert ig.steps_done == io.steps_done == nsteps
        assert ig.dt == io.dt and ig.time == io.time and ig.step == io.step and ig.frame == io.frame, name
        assert ig.plastic_count == io.plastic_count, f"{name}: plastic count {ig.plastic_count} vs {io.plastic_count}"
        compare_state(ge, oe, f"{name} {label}", io.dt)
    ge.close()
    oe.close()


def test_batched_steps_match_single_steps(engine):
    """mpm_step(ctx, 10) equals ten mpm_step(ctx, 1) calls (same kernels, device-side loop state)."""
    a = PS.build(engine, "dp3_musl_apic")
    b = PS.build(engine, "dp3_musl_apic")
    a.step(10)
    for _ in range(10):
        b.step(1)
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("x", "v", "F"):
        assert rel_err(pa[k], pb[k]) <= TOL


def test_frame_end_and_final_time(engine, oracle):
    """STOP_AT_FRAME_END stops after the step that completes a frame; the run ends at end_frame/fps."""
    cfg, plates, objs, f = PS.dp_collapse(dim=2, n=10)
    cfg.fps, cfg.end_frame = 2000.0, 2
    out = []
    for lib in (engine, oracle):
        e = A.Engine(lib, cfg, plates, objs)
        e.set_particles(f["x"])
        rc, i1 = e.step(1000, A.STEP_STOP_AT_FRAME_END)
        rc2, i2 = e.step(1000, A.STEP_STOP_AT_FRAME_END)
        rc3, i3 = e.step(1000, 0)
        out.append((i1.steps_done, i1.frame, i1.frame_ended, i2.steps_done, i2.frame, i2.finished, i3.steps_done, i2.time))
    assert out[0] == out[1]
    assert out[0][1] == 1 and out[0][4] == 2 and out[0][5] == 1 and out[0][6] == 0
    assert abs(out[0][7] - 2 / 2000.0) < 1e-15


def test_conservation_flag(engine):
    """checkMassConservation / checkMomentumConservation (simulation.cpp:292-333) pass at 1e-10."""
    e = PS.build(engine, "dp3_usl_apic")
    rc, info = e.step(5, A.STEP_CHECK_CONSERVATION)
    assert rc == 0 and not (info.status & A.ERR["CONSERVATION"])


@pytest.mark.parametrize("name", ["dp3_musl_apic", "col2d_usl", "incline_mccmui_mibf"])
def test_binning_bit_exact(engine, oracle, name):
    """Cell indices and particle-to-block binning are bit-exact: the device's floor((x - origin)/dx) - 1
    equals the same expression evaluated in binary64 on the host for the engine's own x and origin, the
    origin equals the oracle's bit for bit, and the block key follows from the clamped index."""
    ge = PS.build(engine, name)
    oe = PS.build(oracle, name)
    ge.step(3)
    oe.step(3)
    base, key = ge.get_binning()
    g, og = ge.get_grid(), oe.get_grid()
    info = g["info"]
    assert (info.xc, info.yc, info.zc) == (og["info"].xc, og["info"].yc, og["info"].zc)
    x = ge.get_particles()["x"]
    dim = x.shape[1]
    origin = np.array([info.xc, info.yc, info.zc][:dim])
    h1 = 1.0 / ge.cfg.dx
    ref = (np.floor((x - origin) * h1) - 1).astype(np.int64)
    np.testing.assert_array_equal(base, ref)
    n_axes = [info.Nx, info.Ny, info.Nz][:dim]
    edge = 4 if dim == 3 else 8
    nb = [(n + edge - 1) // edge for n in n_axes]
    # the block follows the first node of the non-zero window: P2G's clamped base, moved up by one node
    # for the quadratic spline when the base node's weight is exactly zero (|x - x_base| / dx >= 1.5)
    wb = np.maximum(ref, 0)
    if ge.cfg.spline_degree == 2:
        coords = [g["x"], g["y"], g["z"]][:dim]
        for d in range(dim):
            t0 = (x[:, d] - coords[d][wb[:, d]]) * h1
            wb[:, d] += (np.abs(t0) >= 1.5).astype(np.int64)
    b = np.minimum(wb // edge, np.array(nb) - 1)
    kref = (b[:, 0] * nb[1] + b[:, 1]) * nb[2] + b[:, 2] if dim == 3 else b[:, 0] * nb[1] + b[:, 1]
    np.testing.assert_array_equal(key, kref.astype(np.uint32))
    # the same cell index from the oracle's own state agrees wherever the positions agree to the bit
    xo = oe.get_particles()["x"]
    same = np.all(x == xo, axis=1)
    refo = (np.floor((xo - origin) * h1) - 1).astype(np.int64)
    np.testing.assert_array_equal(ref[same], refo[same])


@pytest.mark.parametrize("dim,precision", [(3, "fp64"), (2, "fp64"), (3, "split")])
def test_sparse_blocks(engine, oracle, dim, precision):
    """About one particle per block: the 256 particles of a k_defgrad workgroup lie in about 256 blocks, more
    than it copies to LDS, so most of them gather grad v from the grid arrays instead of a tile. Held to the
    oracle at the FP64 parity bar (plain form) or the split form's position bar."""
    print(f"[parity] sparse blocks dim={dim} {precision}", flush=True)
    dx = 0.01
    edge = 4 if dim == 3 else 8                      # cells per block edge
    n = 12 if dim == 3 else 20
    sp = (edge + 0.5) * dx
    idx = np.indices([n] * dim).reshape(dim, -1).T.astype(np.float64)
    x = 0.02 + idx * sp + 0.1 * dx * np.sin(np.arange(idx.size).reshape(idx.shape))
    cfg = PS.base_cfg(dim, dx, 2, plastic_model="DP", M=np.tan(np.radians(30)), flip_ratio=-2.0,
                      gravity=[0.0, -9.81, 0.0], end_frame=1, fps=1.0, precision=precision)
    v = 0.5 * np.cos(np.arange(x.size).reshape(x.shape))
    plates = [A.make_plate(0.0, "bottom", "NoSlip")]
    ge, oe = A.Engine(engine, cfg, plates), A.Engine(oracle, cfg, plates)
    ge.set_particles(x, v=v)
    oe.set_particles(x, v=v)
    rg, ig = ge.step(5)
    ro, io = oe.step(5)
    assert rg == ro == 0, (A.status_names(rg), A.status_names(ro))
    if precision == "fp64":
        compare_state(ge, oe, f"sparse blocks {dim}D", io.dt)
    else:
        gp, op = ge.get_particles(), oe.get_particles()
        err = float(np.linalg.norm(gp["x"] - op["x"]) / np.linalg.norm(op["x"]))
        assert err < 1e-5, err
        assert rel_err(gp["F"], op["F"], 1.0) < 1e-5
