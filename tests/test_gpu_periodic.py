"""Periodic boundaries on the GPU (pbc = 1: pbc.cpp:37-101, remesh.cpp:14-36, position_update.cpp:26-33).

The small periodic scenes of parity_scenes.SCENES are held to the oracle at the FP64 parity bar by
test_gpu_parity.py. Here:
* at a size the oracle would take minutes for, the size-independent property that pins the oracle: with
  USL a periodic run equals the middle copy of the layer repeated along x without periodic boundaries;
* blocks with more particles and periodic copies than the ordered scatter holds (the LDS-atomic path),
  against the oracle;
* the conservation checks count the periodic copies, as the reference's Np does while they exist;
* the split form against the oracle, with positions compared along the periodic axis modulo Lx;
* in a batch, and as requests of the server, a periodic scene's state equals the scene advanced alone.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
from conftest import rel_err
from test_gpu_parity import compare_state

pytestmark = pytest.mark.gpu


def make(lib, cfg, plates, objs, f):
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    return e


def test_unrolled_equality_at_scale(engine):
    """3D, quadratic, APIC, USL; 64 x 16 x 32 cells at 8 particles per cell (262,144 particles), unrolled
    over 5 copies (1,310,720 particles), 5 steps, plain FP64 form on the GPU."""
    kw = dict(dim=3, spline=2, flip=-2.0, cells_x=64, cells_y=16, cells_z=32)
    print("periodic vs unrolled, 262,144 particles x 5 copies, 5 steps", flush=True)
    cfg, plates, objs, f, Lx = PS.periodic_layer(**kw)
    per = make(engine, cfg, plates, objs, f)
    rc, ip = per.step(5)
    assert rc == 0, A.status_names(rc)
    cfg2, plates2, objs2, f2, _ = PS.periodic_layer(copies=5, **kw)
    unr = make(engine, cfg2, plates2, objs2, f2)
    rc, iu = unr.step(5)
    assert rc == 0, A.status_names(rc)
    assert ip.dt == iu.dt and ip.time == iu.time
    a, b = per.get_particles(), unr.get_particles()
    n = a["x"].shape[0]
    mid = {k: v[2 * n:3 * n] for k, v in b.items()}
    xm = mid["x"].copy()
    xm[:, 0] -= 2 * Lx
    crossed = int(np.count_nonzero((xm[:, 0] > Lx) | (xm[:, 0] < 0)))
    xm[xm[:, 0] > Lx, 0] -= Lx
    xm[xm[:, 0] < 0, 0] += Lx
    mid["x"] = xm
    assert crossed > 1000, crossed
    assert np.all((a["x"][:, 0] >= 0) & (a["x"][:, 0] <= Lx))
    for k, fl in (("x", Lx), ("v", 1.0), ("F", 1.0), ("Bmat", 1.0), ("eps_pl_dev", 1e-3)):
        err = rel_err(a[k], mid[k], fl)
        assert err < 1e-10, f"{k}: {err:.3e}"


@pytest.mark.parametrize("musl", [False, True], ids=["usl", "musl"])
def test_dense_blocks_with_copies(engine, oracle, musl):
    """64 particles per cell: a full 4x4x4-cell block holds 4096 particles and up to as many periodic copies,
    beyond the 2048 that the ordered scatter keeps in LDS, so P2G adds them with LDS atomics."""
    print(f"periodic dense blocks musl={musl}", flush=True)
    cfg, plates, objs, f, _ = PS.periodic_layer(dim=3, musl=musl, cells_x=8, cells_y=4, cells_z=4, ppc_axis=4)
    ge, oe = make(engine, cfg, plates, objs, f), make(oracle, cfg, plates, objs, f)
    for steps in (1, 2):
        rg, ig = ge.step(steps)
        ro, io = oe.step(steps)
        assert rg == ro == 0, (A.status_names(rg), A.status_names(ro))
        assert ig.dt == io.dt and ig.plastic_count == io.plastic_count
        compare_state(ge, oe, f"dense periodic, {steps}", io.dt)


def test_conservation_counts_the_copies(engine, oracle):
    cfg, plates, objs, f = PS.SCENES["periodic2d_usl_apic"]()
    for lib in (engine, oracle):
        e = make(lib, cfg, plates, objs, f)
        rc, info = e.step(4, A.STEP_CHECK_CONSERVATION)
        assert rc == 0 and not (info.status & A.ERR["CONSERVATION"]), A.status_names(info.status)


def test_split_form_against_oracle(engine, oracle):
    """Split form, 3D periodic layer, 60 steps: positions within 1e-5 of the layer length, compared along x
    as periodic distances (a particle within rounding of x = Lx may wrap in one run only)."""
    print("periodic split vs oracle", flush=True)
    cfg, plates, objs, f, Lx = PS.periodic_layer(dim=3, cells_x=12, precision="split")
    ge = make(engine, cfg, plates, objs, f)
    cfg64, *_ = PS.periodic_layer(dim=3, cells_x=12)
    oe = make(oracle, cfg64, plates, objs, f)
    for e in (ge, oe):
        rc, _ = e.step(60)
        assert rc == 0
    a, b = ge.get_particles()["x"], oe.get_particles()["x"]
    d = np.abs(a - b)
    d[:, 0] = np.minimum(d[:, 0], Lx - d[:, 0])
    err = float(np.max(d)) / Lx
    print(f"split periodic position error {err:.3e}", flush=True)
    assert err < 1e-5


def periodic_mix():
    """Periodic scenes of several launch classes (2D/3D, USL/MUSL, quadratic/cubic, MIBF, split form) and one
    scene without periodic boundaries that shares the other switches of a periodic one."""
    out = [(n, *PS.SCENES[n]()) for n in ("periodic2d_usl_apic", "periodic2d_musl_picflip", "periodic3d_musl_apic_vm",
                                          "periodic3d_usl_cubic_aflip_mibf")]
    cfg, plates, objs, f, _ = PS.periodic_layer(dim=3, cells_x=12, precision="split")
    out.append(("periodic3d_split", cfg, plates, objs, f))
    cfg, plates, objs, f, _ = PS.periodic_layer(dim=2, copies=2)  # the same layer unrolled, no periodic boundaries
    out.append(("unrolled2d", cfg, plates, objs, f))
    return out


FIELDS = ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "delta_gamma", "muI", "viscosity")


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "direct"])
def test_batched_periodic_equals_alone(engine, use_graph):
    """A batch of periodic scenes (and one without): each member's state equals the scene advanced alone,
    bit for bit, as for every other batch member."""
    print(f"periodic batch vs alone, graph={use_graph}", flush=True)
    scenes = periodic_mix()
    steps = 12
    alone = []
    for name, cfg, plates, objs, f in scenes:
        e = make(engine, cfg, plates, objs, f)
        rc, info = e.step(steps)
        assert rc == 0, (name, A.status_names(rc))
        alone.append((e, info))
    batch = A.Batch(engine, chunk_steps=5, use_graph=use_graph)
    members = []
    for name, cfg, plates, objs, f in scenes:
        e = make(engine, cfg, plates, objs, f)
        batch.add(e, steps)
        members.append(e)
    rc, binfo = batch.step(1000)
    assert rc == 0, A.status_names(rc)
    assert binfo.n_active == 0
    for (name, *_), m, (a, ia) in zip(scenes, members, alone):
        im, left = A.Batch.scene_info(m)
        assert left == 0 and (im.time, im.step, im.plastic_count) == (ia.time, ia.step, ia.plastic_count), name
        pm, pa = m.get_particles(), a.get_particles()
        for k in FIELDS:
            assert np.array_equal(pm[k], pa[k], equal_nan=True), f"{name}: {k} differs between batched and alone"
        gm, ga = m.get_grid(), a.get_grid()
        assert np.array_equal(gm["mass"], ga["mass"]) and np.array_equal(gm["v"], ga["v"]), name
    batch.close()


def test_served_periodic_equals_alone(engine):
    """Periodic requests arriving among others: each result equals the scene advanced alone, bit for bit."""
    print("periodic served vs alone", flush=True)
    scenes = periodic_mix()
    srv = A.Server(engine, chunk_steps=4, keep_results=True)
    reqs = [(cfg, p, o, f, 10 + 3 * i, 0.002 * i) for i, (_, cfg, p, o, f) in enumerate(scenes)]
    ids = [srv.submit(cfg, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"})
           for cfg, p, o, f, steps, t in reqs]
    rc, st = srv.run(60.0)
    assert rc == 0 and st.completed == len(reqs) and st.failed == 0, st.as_dict()
    for rid, (cfg, p, o, f, steps, _) in zip(ids, reqs):
        got, info, _ = srv.result(rid)
        e = make(engine, cfg, p, o, f)
        rc, rinfo = e.step(steps)
        assert rc == 0 and info.time == rinfo.time
        ref = e.get_particles()
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k], equal_nan=True), f"request {rid}: {k} differs from the scene alone"
    srv.close()
