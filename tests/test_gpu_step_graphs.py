"""HIP-graph replay of single-context step batches (mpm_set_step_graphs).

Bar: a small scene stepped with graph replay equals, bit for bit, the same scene stepped with ordinary launches,
after every call of a sequence that replays each captured graph several times with host synchronisations and
downloads in between (batches of 32 and 16 steps, both layout parities, a flag change, a grid that grows), in
both precision forms, and a graph-replayed run still matches the CPU oracle at the parity bar.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
from conftest import rel_err

pytestmark = pytest.mark.gpu

FIELDS = ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "delta_gamma")


def load(e, f):
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})


def pair(cfg, plates, objs, f):
    out = []
    for graphs in (True, False):
        e = A.Engine(A.product_library(), cfg, plates, objs)
        e.set_step_graphs(graphs)
        load(e, f)
        out.append(e)
    return out


@pytest.mark.parametrize("precision", ["fp64", "split"])
@pytest.mark.parametrize("scene", ["column_2d", "dp3_musl"])
def test_graph_replay_equals_ordinary_launches(scene, precision):
    if scene == "column_2d":
        cfg, plates, objs, f = PS.column_2d(n=30)
    else:
        cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=12)
    cfg.precision = A.PRECISION[precision]
    g, e = pair(cfg, plates, objs, f)
    # 48 = 32 + 16 (two graphs), then 33 (32 + 1: the odd tail takes ordinary launches and flips the parity),
    # then 64 with another flag set (the conservation check in the plain form; the split form's binary32 grid
    # sums fail its 1e-10 bar by design, so there the stop at frame end, not reached here), then 48 again
    flag = A.STEP_CHECK_CONSERVATION if precision == "fp64" else A.STEP_STOP_AT_FRAME_END
    plan = [(48, 0), (48, 0), (33, 0), (64, flag), (48, 0), (48, 0)]
    for k, (steps, flags) in enumerate(plan):
        rg, ig = g.step(steps, flags)
        re_, ie = e.step(steps, flags)
        assert rg == re_ and ig.steps_done == ie.steps_done == steps, (k, rg, re_)
        assert (ig.time, ig.dt, ig.step, ig.plastic_count) == (ie.time, ie.dt, ie.step, ie.plastic_count), k
        pg, pe = g.get_particles(), e.get_particles()
        for f in FIELDS:
            np.testing.assert_array_equal(pg[f], pe[f], err_msg=f"{scene} {precision} call {k} field {f}")
        gg, ge = g.get_grid(), e.get_grid()
        for f in ("mass", "v"):
            np.testing.assert_array_equal(gg[f], ge[f], err_msg=f"{scene} {precision} call {k} grid {f}")


def test_graph_replayed_run_matches_the_oracle():
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=12)
    g = A.Engine(A.product_library(), cfg, plates, objs)
    load(g, f)
    o = A.Engine(A.oracle_library(), cfg, plates, objs)
    load(o, f)
    for _ in range(2):  # the second call replays the graph captured by the first
        rg, ig = g.step(32)
        ro, io = o.step(32)
        assert rg == ro == 0 and ig.time == io.time
    pg, po = g.get_particles(), o.get_particles()
    for f, floor in (("x", 1.0), ("v", 1e-6), ("F", 1.0)):
        assert rel_err(pg[f], po[f], floor) < 1e-10, f
