"""Periodic boundaries on the GPU (pbc = 1: pbc.cpp:37-101, remesh.cpp:14-36, position_update.cpp:26-33).

The small periodic scenes of parity_scenes.SCENES are held to the oracle at the FP64 parity bar by
test_gpu_parity.py. Here:
* at a size the oracle would take minutes for, the size-independent property that pins the oracle: with
  USL a periodic run equals the middle copy of the layer repeated along x without periodic boundaries;
* blocks with more particles and periodic copies than the ordered scatter holds (the LDS-atomic path),
  against the oracle;
* the conservation checks count the periodic copies, as the reference's Np does while they exist;
* the split form against the oracle, with positions compared along the periodic axis modulo Lx;
* batches and the server refuse periodic scenes (they step alone).
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


def test_batches_refuse_periodic_scenes(engine):
    cfg, plates, objs, f = PS.SCENES["periodic2d_usl_apic"]()
    e = make(engine, cfg, plates, objs, f)
    b = A.Batch(engine)
    with pytest.raises(A.MPMError) as ex:
        b.add(e, 5)
    assert ex.value.code == A.ERR["UNSUPPORTED"]
    b.close()
    srv = A.Server(engine)
    with pytest.raises(A.MPMError) as ex:
        srv.submit(cfg, plates, objs, f["x"], 5, 0.0)
    assert ex.value.code == A.ERR["UNSUPPORTED"]
    srv.close()
    rc, _ = e.step(5)  # the scene itself steps alone
    assert rc == 0
