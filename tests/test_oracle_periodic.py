"""Periodic boundaries of the oracle (pbc.cpp:37-101, remesh.cpp:14-36, position_update.cpp:26-33,
simulation.cpp:219-273), pinned by what the mathematics fixes rather than by retyping the formulas:

* the grid of a periodic run is the fixed `remeshFixed(4)` grid: origin (-5 dx, -dx, -5 dx), node counts of
  the accumulating arange, unchanged from step to step;
* unrolling: with USL, a periodic run equals the middle copy of the same layer repeated along x without
  periodic boundaries, inside the domain of dependence (each step reaches about 3 dx for the quadratic
  spline, 4 dx for the cubic). The ghost band of 4 dx covers every node a real particle reads, so the
  equality holds to rounding; a ghost at the wrong place, a missing field on the ghost, a wrong wrap or a
  ghost band that is too narrow each break it;
* MUSL: the reference deletes the ghosts before G2P, so its MUSL re-scatter has no ghosts. Away from the
  x boundaries the periodic run still equals the unrolled one, but near them the re-scattered velocity
  lacks the ghosts' momentum, and the deformation gradient updated from its gradient differs by far more
  than rounding.
"""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
from conftest import rel_err


def run(lib, cfg, plates, objs, f, steps):
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    rc, info = e.step(steps)
    assert rc == 0, A.status_names(rc)
    return e, info


def unrolled_pair(oracle, steps, copies=5, **kw):
    cfg, plates, objs, f, Lx = PS.periodic_layer(**kw)
    per, ip = run(oracle, cfg, plates, objs, f, steps)
    cfg2, plates2, objs2, f2, _ = PS.periodic_layer(copies=copies, **kw)
    unr, iu = run(oracle, cfg2, plates2, objs2, f2, steps)
    assert ip.dt == iu.dt and ip.time == iu.time  # same dt sequence (both at dt_max)
    a, b = per.get_particles(), unr.get_particles()
    n = a["x"].shape[0]
    j = copies // 2
    mid = {k: v[j * n:(j + 1) * n] for k, v in b.items()}
    xm = mid["x"].copy()
    xm[:, 0] -= j * Lx
    wrapped = (xm[:, 0] > Lx) | (xm[:, 0] < 0)
    xm[xm[:, 0] > Lx, 0] -= Lx
    xm[xm[:, 0] < 0, 0] += Lx
    mid["x"] = xm
    return a, mid, Lx, int(wrapped.sum()), per


def test_fixed_grid(oracle):
    cfg, plates, objs, f, Lx = PS.periodic_layer(dim=3, cells_x=12)
    e = A.Engine(oracle, cfg, plates, objs)
    e.set_particles(f["x"], v=f["v"], F=f["F"])
    dx = cfg.dx
    # accumulating arange, src/tools.hpp:291-296
    def arange(a, b, h):
        out, v = [], a
        while v < b:
            out.append(v)
            v += h
        return np.array(out)
    expect = {"x": arange(-dx * 5, Lx + 6 * dx, dx), "y": arange(-dx, cfg.Ly + 6 * dx, dx),
              "z": arange(-dx * 5, cfg.Lz + 6 * dx, dx)}
    for steps in (1, 3):
        rc, _ = e.step(1 if steps == 1 else 2)
        assert rc == 0
        g = e.get_grid()
        for k in ("x", "y", "z"):
            assert np.array_equal(g[k], expect[k]), (k, steps)
        assert (g["info"].xc, g["info"].yc, g["info"].zc) == (-5 * dx, -dx, -5 * dx)


@pytest.mark.parametrize("kw", [
    dict(dim=2, spline=2, flip=-2.0),
    dict(dim=2, spline=2, flip=0.95),
    dict(dim=2, spline=3, flip=-1.0, mibf=1),
    dict(dim=3, spline=2, flip=-2.0, cells_x=12),
    dict(dim=3, spline=3, flip=-1.0, mibf=1, cells_x=12),
], ids=["2d_apic", "2d_picflip", "2d_cubic_aflip_mibf", "3d_apic", "3d_cubic_aflip_mibf"])
def test_usl_equals_unrolled(oracle, kw):
    steps = 5
    a, b, Lx, wrapped, _ = unrolled_pair(oracle, steps, **kw)
    assert wrapped > 0, "no particle crossed the periodic boundary"
    assert np.all((a["x"][:, 0] >= 0) & (a["x"][:, 0] <= Lx))
    floors = {"x": Lx, "v": 1.0, "F": 1.0, "Bmat": 1.0, "eps_pl_dev": 1e-3, "eps_pl_vol": 1e-3, "delta_gamma": 1e-6,
              "muI": 1.0}
    for k, fl in floors.items():
        err = rel_err(a[k], b[k], fl)
        assert err < 1e-10, f"{k}: {err:.3e}"
    # the layer really deforms plastically and spreads stress across the boundary
    assert np.max(a["eps_pl_dev"]) > 1e-4


def test_musl_ghosts_stop_before_the_rescatter(oracle):
    cfg, *_ = PS.periodic_layer(dim=2, musl=True)
    a, b, Lx, _, _ = unrolled_pair(oracle, 1, dim=2, musl=True)
    dx = cfg.dx
    x0 = a["x"][:, 0]
    inner = (x0 > 5 * dx) & (x0 < Lx - 5 * dx)
    band = (x0 < 1.0 * dx) | (x0 > Lx - 1.0 * dx)
    assert inner.sum() > 0 and band.sum() > 0
    for k, fl in (("x", Lx), ("v", 1.0), ("F", 1.0)):
        assert rel_err(a[k][inner], b[k][inner], fl) < 1e-10, k
    # positions and particle velocities come from G2P (ghosts present; musl.cpp:13-25 sets v before the
    # re-scatter) and agree everywhere
    assert rel_err(a["x"], b["x"], Lx) < 1e-10
    assert rel_err(a["v"], b["v"], 1.0) < 1e-10
    # the re-scattered grid velocity near the boundary lacks the ghosts' momentum, so the deformation
    # gradient updated from its gradient differs there
    dF = np.max(np.abs(a["F"][band] - b["F"][band]))
    assert dF > 1e-3, dF
