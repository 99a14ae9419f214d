"""mpm_step_noalloc: the step without device allocation (grid, block arrays and staging tiles in the
device's static scratch) against mpm_step on the same seeded input.

The engine sums every node and particle value in a fixed order (no floating-point atomics), so the two
entry points give equal bits. Two calls in a row share the device's scratch: the second scene is smaller
than the first, so stale contents of the first call would show up in the second's results.
"""
import numpy as np
import pytest

import mpm_abi as A
import scenes as S


def _pair(lib, precision, n_axis, musl):
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=n_axis, precision=precision, musl=musl)
    a = A.Engine(lib, cfg, plates, objs)
    b = A.Engine(lib, cfg, plates, objs)
    a.set_particles(x)
    b.set_particles(x)
    return a, b


def _same(a, b, what):
    pa, pb = a.get_particles(), b.get_particles()
    for k in ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol", "delta_gamma"):
        assert np.array_equal(pa[k], pb[k]), f"{what}: {k} differs"
    ga, gb = a.get_grid(), b.get_grid()
    assert ga["info"].grid_nodes == gb["info"].grid_nodes, what
    for k in ("mass", "v"):
        assert np.array_equal(ga[k], gb[k]), f"{what}: grid {k} differs"


@pytest.mark.gpu
@pytest.mark.parametrize("precision,musl", [("fp64", True), ("split", False)])
def test_noalloc_step_equals_step(engine, precision, musl):
    for n_axis, steps in ((20, 30), (12, 17)):  # the second call is the smaller scene
        print(f"noalloc vs step: {precision} musl={musl} {n_axis}^3 particles, {steps} steps", flush=True)
        a, b = _pair(engine, precision, n_axis, musl)
        ra, ia = a.step(steps)
        rb, ib = b.step_noalloc(steps)
        assert ra == 0 and rb == 0, (A.status_names(ra), A.status_names(rb))
        assert ia.steps_done == ib.steps_done == steps
        assert ia.time == ib.time and ia.dt == ib.dt and ia.plastic_count == ib.plastic_count
        _same(a, b, f"{precision} {n_axis}^3")
        a.close()
        b.close()


@pytest.mark.gpu
def test_noalloc_rejects_a_grid_beyond_the_scratch_then_steps_normally(engine):
    """A scene whose grid cannot fit the scratch is refused before any step; the context stays usable."""
    print("noalloc scratch limit", flush=True)
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=4, precision="fp64", musl=True)
    far = x.copy()
    far[:, :] += 3.0  # a second cluster 3 m away: a dense grid of several hundred nodes per axis
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(np.vstack([x, far]))
    rc, info = e.step_noalloc(5)
    assert rc & A.ERR["SCRATCH_LIMIT"], A.status_names(rc)
    assert info.steps_done == 0 and info.time == 0.0
    e.close()
    a, b = _pair(engine, "split", 10, True)
    assert a.step(5)[0] == 0 and b.step_noalloc(5)[0] == 0
    _same(a, b, "after a refused call")
