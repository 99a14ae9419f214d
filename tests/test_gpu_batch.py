"""The batched form (mpm_batch_*): many independent scenes advanced by one launch per kernel and class.

Bar: every member's state after a batched run equals, bit for bit, the state of the same scene advanced
alone by mpm_step (the step has no floating-point atomics and no sum mixes scenes), with the same time,
dt, step and frame; budgets are honoured per member; graph replay and direct launches agree; members may
join and leave between calls. A multi-scene run is also held to the CPU oracle at the FP64 parity bar.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
import scenes as S

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "delta_gamma", "muI", "viscosity")


def mixed_scenes():
    """Scenes of several launch classes: 3D/2D, quadratic/cubic, MUSL/USL, APIC/AFLIP/FLIP, several materials,
    both precision forms, MIBF, objects."""
    out = []
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=14)
    out.append(("dp3_musl", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=False, n=10)
    out.append(("dp3_usl", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=12)
    cfg.precision = A.PRECISION["split"]
    out.append(("dp3_musl_split", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.column_2d(n=30)
    out.append(("col2d", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.incline_mccmui(n=10)
    out.append(("incline_mccmui", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.model_block("VMVisc", n=10, q_max=2e3, q_min=2e3, perzyna_visc=0.01)
    out.append(("vmvisc", cfg, plates, objs, f))
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=False, n=9)  # second member of the dp3_usl class
    cfg.M = 0.5
    out.append(("dp3_usl_b", cfg, plates, objs, f))
    return out


def make(lib, cfg, plates, objs, f):
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    return e


def assert_same(a, b, label):
    pa, pb = a.get_particles(), b.get_particles()
    for k in FIELDS:
        assert np.array_equal(pa[k], pb[k], equal_nan=True), f"{label}: {k} differs between batched and alone"


@pytest.mark.parametrize("use_graph", [True, False], ids=["graph", "direct"])
def test_batch_equals_alone(engine, use_graph):
    print(f"batch vs alone, graph={use_graph}", flush=True)
    scenes = mixed_scenes()
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
    assert binfo.n_scenes == len(scenes) and binfo.n_active == 0
    assert binfo.particle_steps == sum(e.n for e in members) * steps
    for (name, *_), m, (a, ia) in zip(scenes, members, alone):
        im, left = A.Batch.scene_info(m)
        assert left == 0 and im.steps_done == steps, name
        assert (im.time, im.dt, im.step, im.frame) == (ia.time, ia.dt, ia.step, ia.frame), name
        assert im.plastic_count == ia.plastic_count, name
        assert_same(m, a, name)
        gm, ga = m.get_grid(), a.get_grid()
        assert np.array_equal(gm["mass"], ga["mass"]) and np.array_equal(gm["v"], ga["v"]), name
    batch.close()


def test_batch_budgets_and_membership(engine):
    """Members with different budgets; a member that joins later and one that leaves early; results equal
    those of the same scenes advanced alone by the same number of steps."""
    print("batch budgets and membership", flush=True)
    (_, c1, p1, o1, f1), (_, c2, p2, o2, f2), (_, c3, p3, o3, f3) = mixed_scenes()[:3]
    batch = A.Batch(engine, chunk_steps=4)
    a = make(engine, c1, p1, o1, f1)
    b = make(engine, c2, p2, o2, f2)
    batch.add(a, 7)
    batch.add(b, 3)
    rc, _ = batch.step(5)           # a: 5 of 7, b: 3 of 3
    assert rc == 0
    assert A.Batch.scene_info(a)[1] == 2 and A.Batch.scene_info(b)[1] == 0
    c = make(engine, c3, p3, o3, f3)
    batch.add(c, 6)
    batch.remove(b)
    rc, info = batch.step(100)      # a: 2 more, c: 6
    assert rc == 0 and info.n_scenes == 2
    for e, cfg, plates, objs, f, total in ((a, c1, p1, o1, f1, 7), (b, c2, p2, o2, f2, 3), (c, c3, p3, o3, f3, 6)):
        ref = make(engine, cfg, plates, objs, f)
        rc, _ = ref.step(total)
        assert rc == 0
        assert_same(e, ref, f"budget {total}")
    # a removed member is an ordinary context again
    rc, _ = b.step(2)
    assert rc == 0
    batch.close()


def test_member_refuses_mpm_step(engine):
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=8)
    batch = A.Batch(engine)
    e = make(engine, cfg, plates, objs, f)
    batch.add(e, 2)
    rc, _ = e.step(1)
    assert rc == A.ERR["INVALID_ARGUMENT"]
    batch.close()
    rc, _ = e.step(1)
    assert rc == 0


def test_batch_vs_oracle(engine, oracle):
    """Two 3D collapses of different sizes in one batch, 10 steps, against the oracle at the FP64 bar."""
    print("batch vs oracle", flush=True)
    from test_gpu_parity import compare_state
    batch = A.Batch(engine, chunk_steps=10)
    pairs = []
    for n in (12, 16):
        cfg, plates, objs, x = S.dp_collapse(n_per_axis=n, musl=True)
        g = make(engine, cfg, plates, objs, {"x": x})
        o = A.Engine(oracle, cfg, plates, objs)
        o.set_particles(x)
        batch.add(g, 10)
        pairs.append((g, o))
    rc, _ = batch.step(10)
    assert rc == 0
    for g, o in pairs:
        rc, io = o.step(10)
        assert rc == 0
        ig, _ = A.Batch.scene_info(g)
        compare_state(g, o, f"batched {g.n} particles", io.dt)
        assert ig.time == io.time
    batch.close()


def test_batch_grid_growth(engine):
    """A scene thrown sideways grows its grid past the headroom of a chunk: the batch reallocates that
    member and resumes; the result equals the scene advanced alone."""
    print("batch grid growth", flush=True)
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=8)
    v = np.zeros_like(f["x"])
    v[:, 0] = 30.0
    f = dict(f, v=v)
    other = PS.dp_collapse(dim=3, musl=False, n=10)
    batch = A.Batch(engine, chunk_steps=32)
    e = make(engine, cfg, plates, objs, f)
    o = make(engine, *other[:3], other[3])
    batch.add(e, 60)
    batch.add(o, 60)
    rc, _ = batch.step(60)
    assert rc == 0, A.status_names(rc)
    ref = make(engine, cfg, plates, objs, f)
    rc, _ = ref.step(60)
    assert rc == 0
    assert_same(e, ref, "thrown scene")
    batch.close()


def test_autotune_keeps_results(engine):
    """The chunk autotuner advances the members by ordinary steps: afterwards every member equals the scene
    advanced alone by the same number of steps, and the chosen length is one of the candidates."""
    print("batch autotune", flush=True)
    scenes = mixed_scenes()[:4]
    batch = A.Batch(engine, chunk_steps=16)
    members = []
    for name, cfg, plates, objs, f in scenes:
        e = make(engine, cfg, plates, objs, f)
        batch.add(e, 200)
        members.append(e)
    chosen, rate = batch.autotune(180)
    assert chosen in (4, 8, 16, 32) and rate > 0
    rc, _ = batch.step(200)  # finish the budgets
    assert rc == 0
    for (name, cfg, plates, objs, f), m in zip(scenes, members):
        info, left = A.Batch.scene_info(m)
        assert left == 0
        ref = make(engine, cfg, plates, objs, f)
        rc, _ = ref.step(200)
        assert rc == 0
        assert_same(m, ref, f"autotuned {name}")
    batch.close()
