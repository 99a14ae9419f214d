"""The served form (mpm_server_*): requests arriving over time, admitted by device footprint, advanced
together by continuous batching. Bar: every request's final state equals, bit for bit, the state of the
same scene advanced alone by mpm_step for the same number of steps; requests overlap on the device;
admission honours the footprint budget; latencies are reported."""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "delta_gamma", "muI", "viscosity")


def requests():
    """(cfg, plates, objects, fields, steps, arrival_s) of several classes and sizes."""
    out = []
    cfg, p, o, f = PS.dp_collapse(dim=3, musl=True, n=14)
    out.append((cfg, p, o, f, 30, 0.0))
    cfg, p, o, f = PS.column_2d(n=30)
    out.append((cfg, p, o, f, 45, 0.0))
    cfg, p, o, f = PS.incline_mccmui(n=10)
    cfg.precision = A.PRECISION["split"]
    out.append((cfg, p, o, f, 20, 0.005))
    cfg, p, o, f = PS.dp_collapse(dim=3, musl=False, n=10)
    out.append((cfg, p, o, f, 25, 0.01))
    cfg, p, o, f = PS.model_block("VMVisc", n=10, q_max=2e3, q_min=2e3, perzyna_visc=0.01)
    out.append((cfg, p, o, f, 15, 0.02))
    cfg, p, o, f = PS.dp_collapse(dim=3, musl=True, n=12)
    out.append((cfg, p, o, f, 40, 0.03))
    return out


def alone(engine, cfg, plates, objs, f, steps):
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    rc, info = e.step(steps)
    assert rc == 0
    return e.get_particles(), info


def test_served_results_equal_alone(engine):
    print("served vs alone", flush=True)
    reqs = requests()
    srv = A.Server(engine, chunk_steps=4, keep_results=True)
    ids = [srv.submit(cfg, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"})
           for cfg, p, o, f, steps, t in reqs]
    rc, st = srv.run(60.0)
    print(st.as_dict(), flush=True)
    assert rc == 0
    assert st.completed == len(reqs) and st.failed == 0
    assert st.peak_resident >= 2, "requests never overlapped on the device"
    assert 0 < st.step_ms_p50 <= st.step_ms_p99
    assert 0 < st.request_ms_p50 <= st.request_ms_p99
    assert st.particle_steps == sum(f["x"].shape[0] * steps for _, _, _, f, steps, _ in reqs)
    for rid, (cfg, p, o, f, steps, _) in zip(ids, reqs):
        got, info, lat = srv.result(rid)
        ref, rinfo = alone(engine, cfg, p, o, f, steps)
        assert info.steps_done == steps and info.time == rinfo.time and lat > 0
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k], equal_nan=True), f"request {rid}: {k} differs from the scene alone"
    srv.close()


def test_admission_rejects_what_cannot_fit(engine):
    """A budget below every request's footprint: each request fails with an out-of-memory status instead of
    waiting forever; a budget of one scene at a time serves them one after another."""
    print("admission budget", flush=True)
    reqs = requests()[:3]
    srv = A.Server(engine, hbm_budget_bytes=1, chunk_steps=4)
    for cfg, p, o, f, steps, t in reqs:
        srv.submit(cfg, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"})
    rc, st = srv.run(30.0)
    assert rc == 0 and st.failed == len(reqs) and st.completed == 0
    srv.close()
    srv = A.Server(engine, max_scenes=1, chunk_steps=4)
    for cfg, p, o, f, steps, t in reqs:
        srv.submit(cfg, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"})
    rc, st = srv.run(60.0)
    assert rc == 0 and st.completed == len(reqs) and st.peak_resident == 1
    srv.close()


def test_time_limit_requeues_then_a_second_run_serves_all(engine):
    """A run stopped by its time limit returns the requests it had admitted (waiting for preparation,
    prepared or resident) to the queue with their inputs; a second run serves them from the start, and every
    result equals the scene advanced alone."""
    print("served: time limit, then resume", flush=True)
    reqs = requests()
    cfg, p, o, f = PS.dp_collapse(dim=3, musl=True, n=40)  # 64,000 particles: still running at the limit
    reqs.append((cfg, p, o, f, 60, 0.0))
    srv = A.Server(engine, chunk_steps=4, keep_results=True)
    ids = [srv.submit(cfg, p, o, f["x"], steps, t, **{k: v for k, v in f.items() if k != "x"})
           for cfg, p, o, f, steps, t in reqs]
    rc1, st1 = srv.run(0.001)
    assert rc1 == 0 and st1.failed == 0 and st1.completed < len(reqs)
    rc2, st2 = srv.run(60.0)
    print(st1.completed, st2.completed, flush=True)
    assert rc2 == 0 and st2.failed == 0 and st1.completed + st2.completed == len(reqs)
    for rid, (cfg, p, o, f, steps, _) in zip(ids, reqs):
        got, info, lat = srv.result(rid)
        ref, rinfo = alone(engine, cfg, p, o, f, steps)
        assert info.steps_done == steps and info.time == rinfo.time
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k], equal_nan=True), f"request {rid}: {k} differs from the scene alone"
    srv.close()
