"""The FP32/FP64 split form in the configurations the workloads use it in, each held to the FP64 CPU oracle
(and to the plain GPU form where the size allows): the inclined flow (cubic B-spline, AFLIP, mu(I)-MCC, MIBF,
rotated-box bump), a Perzyna-Von Mises member inside a batch, a served request, and the 4,096,000-particle
Drucker-Prager collapse that the split form is specified on.

Bar (the split form's specification): particle positions within 1e-5 relative (norm-wise) after 100 steps.
The other fields' errors are printed, not asserted: binary32 transfers carry about 1e-7 relative error per
step, and velocity, F and the plastic scalars are compared at that level only through the position.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
import scenes as S

pytestmark = pytest.mark.gpu

POS_TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def make(lib, cfg, plates, objs, f, precision=None):
    c = type(cfg).from_buffer_copy(cfg)
    if precision is not None:
        c.precision = A.PRECISION[precision]
    e = A.Engine(lib, c, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    return e


def report(label, got, ref):
    errs = {k: rel(got[k], ref[k]) for k in ("x", "v", "F") if np.isfinite(ref[k]).all()}
    print(label, " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    return errs


def test_incline_split_vs_oracle_and_fp64(engine, oracle):
    """Inclined-flow configuration (cubic, AFLIP, mu(I)-MCC ExpoImpl, MIBF, SlipStick floor and bump) in the
    split form, 2,744 particles, 100 steps: against the oracle and against the plain GPU form."""
    steps = 100
    print(f"incline split vs oracle/fp64, {steps} steps", flush=True)
    cfg, plates, objs, f = PS.incline_mccmui()
    gs = make(engine, cfg, plates, objs, f, "split")
    gd = make(engine, cfg, plates, objs, f, "fp64")
    oe = make(oracle, cfg, plates, objs, f)
    for e in (gs, gd, oe):
        rc, info = e.step(steps)
        assert rc == 0 and info.steps_done == steps, A.status_names(rc)
    ps, pd, po = gs.get_particles(), gd.get_particles(), oe.get_particles()
    e_o = report("split vs oracle:", ps, po)
    e_d = report("split vs fp64 gpu:", ps, pd)
    report("fp64 gpu vs oracle:", pd, po)
    assert e_o["x"] < POS_TOL and e_d["x"] < POS_TOL


def test_perzyna_vm_split_member_of_a_batch_vs_oracle(engine, oracle):
    """A batch of split members (Perzyna-Von Mises, linear and Newton forms, and Drucker-Prager), each against
    the oracle after 100 steps of the same scene."""
    steps = 100
    scenes = [
        ("vmvisc_linear", *PS.model_block("VMVisc", n=10, q_max=2e3, q_min=2e3, perzyna_visc=0.01, xi=0.0)),
        ("vmvisc_newton", *PS.model_block("VMVisc", n=10, q_max=3e3, q_min=1e3, perzyna_visc=0.01, xi=0.5,
                                          perzyna_exp=2.0)),
        ("dp_musl", *PS.dp_collapse(dim=3, musl=True, n=12)),
    ]
    batch = A.Batch(engine, chunk_steps=8)
    members = []
    for name, cfg, plates, objs, f in scenes:
        e = make(engine, cfg, plates, objs, f, "split")
        batch.add(e, steps)
        members.append(e)
    print(f"split batch of {len(members)} vs oracle, {steps} steps", flush=True)
    rc, info = batch.step(steps)
    assert rc == 0 and info.n_active == 0, A.status_names(rc)
    for (name, cfg, plates, objs, f), m in zip(scenes, members):
        im, _ = A.Batch.scene_info(m)
        assert im.steps_done == steps and im.status & A.EXIT_MASK == 0, (name, A.status_names(im.status))
        oe = make(oracle, cfg, plates, objs, f)
        assert oe.step(steps)[0] == 0
        errs = report(f"{name} split member vs oracle:", m.get_particles(), oe.get_particles())
        assert errs["x"] < POS_TOL, name
    batch.close()


def test_served_split_requests_vs_oracle(engine, oracle):
    """Split-form requests through the request server, each result against the oracle run of its scene."""
    reqs = [
        (*PS.dp_collapse(dim=3, musl=True, n=12), 100, 0.0),
        (*PS.model_block("VMVisc", n=10, q_max=2e3, q_min=2e3, perzyna_visc=0.01, xi=0.0), 60, 0.002),
        (*PS.incline_mccmui(n=10), 50, 0.004),
    ]
    srv = A.Server(engine, chunk_steps=8, keep_results=True)
    ids = []
    for cfg, p, o, f, steps, t in reqs:
        c = type(cfg).from_buffer_copy(cfg)
        c.precision = A.PRECISION["split"]
        ids.append(srv.submit(c, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"}))
    print(f"served split requests vs oracle: {len(reqs)}", flush=True)
    rc, st = srv.run(60.0)
    assert rc == 0 and st.completed == len(reqs), A.status_names(rc)
    for rid, (cfg, p, o, f, steps, _) in zip(ids, reqs):
        got, info, _ = srv.result(rid)
        assert info.steps_done == steps
        oe = make(oracle, cfg, p, o, f)
        assert oe.step(steps)[0] == 0
        errs = report(f"request {rid} vs oracle:", got, oe.get_particles())
        assert errs["x"] < POS_TOL, rid
    srv.close()


def test_split_gate_at_4m_particles(engine):
    """The split form's specification: the 3D Drucker-Prager collapse at 4,096,000 particles (0.4 m cube,
    dx 0.005 m, 8 particles per cell, MUSL, APIC, SlipFree floor mu 0.4), 100 steps, position within 1e-5
    relative of the plain FP64 form on the GPU."""
    steps = 100
    print("split vs fp64 at 4,096,000 particles, 100 steps", flush=True)
    cfg, plates, objs, x = S.dp_collapse(precision="fp64")
    assert x.shape[0] == 4_096_000
    out = {}
    for form in ("fp64", "split"):
        c = type(cfg).from_buffer_copy(cfg)
        c.precision = A.PRECISION[form]
        e = A.Engine(engine, c, plates, objs)
        e.set_particles(x)
        rc, info = e.step(steps)
        assert rc == 0 and info.steps_done == steps, A.status_names(rc)
        p = e.get_particles()
        out[form] = (p["x"], p["v"], info)
        e.close()
        print(f"{form}: t={info.time:.6e} plastic={info.plastic_count}", flush=True)
    ex = rel(out["split"][0], out["fp64"][0])
    ev = rel(out["split"][1], out["fp64"][1])
    print(f"4M gate: position rel err {ex:.3e}, velocity rel err {ev:.3e}", flush=True)
    assert ex < POS_TOL
