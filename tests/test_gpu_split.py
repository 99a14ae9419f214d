"""The FP32/FP64 split form (MPM_PRECISION_SPLIT): transfers, grid values and the APIC affine matrix in
binary32; position, deformation gradient, return mapping and the step reductions in binary64.

Its error is stated against the FP64 results: the plain GPU form (itself held to the oracle at 1e-10) on
the same seeded input, and the CPU oracle directly at a small size. The bar is the one the split form is
specified with: particle positions within 1e-5 relative (norm-wise) after 100 steps of a 3D
Drucker-Prager collapse with APIC and MUSL.
"""
import numpy as np
import pytest

import mpm_abi as A
import scenes as S

POS_TOL = 1e-5


def run(lib, precision, steps, n_axis, **kw):
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=n_axis, precision=precision, **kw)
    e = A.Engine(lib, cfg, plates, objs)
    e.set_particles(x)
    rc, info = e.step(steps)
    assert rc == 0, A.status_names(rc)
    assert info.steps_done == steps
    return e, e.get_particles(), info


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.gpu
def test_split_vs_fp64_gpu_100_steps(engine):
    """3D DP collapse, MUSL + APIC, 64,000 particles, 100 steps: split vs plain FP64 on the GPU."""
    print("split vs fp64, 40^3 particles, 100 steps", flush=True)
    _, p64, i64 = run(engine, "fp64", 100, 40)
    _, p32, i32 = run(engine, "split", 100, 40)
    ex = rel(p32["x"], p64["x"])
    ev = float(np.linalg.norm(p32["v"] - p64["v"]) / max(np.linalg.norm(p64["v"]), 1e-30))
    print(f"position rel err {ex:.3e}, velocity rel err {ev:.3e}, time {i32.time} vs {i64.time}")
    assert ex < POS_TOL, f"position rel err {ex:.3e}"
    assert abs(i32.time - i64.time) <= 1e-6 * i64.time


@pytest.mark.gpu
def test_split_vs_oracle_100_steps(engine, oracle):
    """Split form against the CPU oracle (FP64) directly, 8,000 particles, 100 steps."""
    print("split vs oracle, 20^3 particles, 100 steps", flush=True)
    _, p32, _ = run(engine, "split", 100, 20)
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=20)
    oe = A.Engine(oracle, cfg, plates, objs)
    oe.set_particles(x)
    rc, _ = oe.step(100)
    assert rc == 0
    po = oe.get_particles()
    ex = rel(p32["x"], po["x"])
    print(f"position rel err vs oracle {ex:.3e}")
    assert ex < POS_TOL


@pytest.mark.gpu
@pytest.mark.parametrize("musl", [True, False], ids=["musl", "usl"])
def test_split_deterministic(engine, musl):
    """Two split-form runs of the same input give the same bits (no floating-point atomics in the step)."""
    print(f"split determinism musl={musl}", flush=True)
    _, a, _ = run(engine, "split", 20, 24, musl=musl)
    _, b, _ = run(engine, "split", 20, 24, musl=musl)
    for k in ("x", "v", "F", "Bmat"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.gpu
def test_fp64_deterministic(engine):
    """Two plain-form runs of the same input give the same bits."""
    print("fp64 determinism", flush=True)
    _, a, _ = run(engine, "fp64", 20, 24)
    _, b, _ = run(engine, "fp64", 20, 24)
    for k in ("x", "v", "F", "Bmat"):
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.gpu
def test_split_2d_usl(engine, oracle):
    """2D plane-strain column (USL, APIC, NoSlip floor): split vs oracle after 100 steps."""
    print("split 2d usl vs oracle", flush=True)
    cfg, plates, objs, x = S.column_2d(n_side=60)
    cfg32 = type(cfg).from_buffer_copy(cfg)
    cfg32.precision = A.PRECISION["split"]
    ge = A.Engine(engine, cfg32, plates, objs)
    oe = A.Engine(oracle, cfg, plates, objs)
    ge.set_particles(x)
    oe.set_particles(x)
    assert ge.step(100)[0] == 0 and oe.step(100)[0] == 0
    ex = rel(ge.get_particles()["x"], oe.get_particles()["x"])
    print(f"2d position rel err {ex:.3e}")
    assert ex < POS_TOL
