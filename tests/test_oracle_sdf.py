"""The level-set object (ObjectVdb, src/objects/object_vdb.hpp:17-94, here on a dense lattice) on the oracle:
  * a sampled plane phi = y - y0 is ObjectGround(y0): same inside() away from the surface, normal exactly +y;
  * the signed distance of the curve y = x^2 gives the curve's normal at (-0.5, 0.25) within 1e-4 and the same
    inside() (BoundaryTest.CalcNormal, tests/tests_vdb.cpp:18-38);
  * a particle sliding on that level set reaches sqrt(2 g) within 1e-3 (BoundaryTest.VDB, tests_vdb.cpp:40-82);
  * outside the lattice the object is absent."""
import ctypes as C
import math

import numpy as np
import pytest

import mpm_abi as A
import physics_cases as PC
from returnmap import dptr

rng = np.random.default_rng(11)


def evaluate(oracle, e, k, X):
    inside, n = C.c_int(), np.zeros(e.dim)
    x = np.ascontiguousarray(X, dtype=np.float64)
    assert oracle.lib.oracle_object_eval(e.h, k, dptr(x), C.byref(inside), dptr(n)) == 0
    return bool(inside.value), n


def engine_with(oracle, dim, objects=()):
    return A.Engine(oracle, A.default_config(dim=dim, dx=0.01, particle_volume=1e-6, particle_mass=1e-3), [], objects)


@pytest.mark.parametrize("dim", [2, 3])
def test_sampled_plane_is_ground(oracle, dim):
    y0, h = 0.0137, 0.01
    shape = (31, 21) + ((11,) if dim == 3 else ())
    origin = [-0.15, -0.1, -0.05][:dim]
    ys = origin[1] + h * np.arange(shape[1])
    phi = np.broadcast_to((ys - y0)[None, :, None] if dim == 3 else (ys - y0)[None, :], shape).astype(np.float32)
    e = engine_with(oracle, dim, [A.make_object("Ground", "SlipFree", 0.0, param=y0)])
    e.add_sdf(phi, origin, h)
    lo, hi = np.array(origin), np.array(origin) + h * (np.array(shape) - 1)
    for _ in range(300):
        X = lo + (hi - lo) * rng.random(dim)
        ins_g, n_g = evaluate(oracle, e, 0, X)
        ins_s, n_s = evaluate(oracle, e, 1, X)
        if abs(X[1] - y0) > 1e-6:
            assert ins_s == ins_g
        np.testing.assert_array_equal(n_s, n_g)


@pytest.mark.parametrize("dim", [2, 3])
def test_curve_level_set_normal_and_inside(oracle, dim):
    phi, origin, h = PC.curve_lattice(dim)
    e = engine_with(oracle, dim, [A.make_object("Curve", "SlipFree", 0.0)])
    e.add_sdf(phi, origin, h)
    X = np.array([-0.5, 0.25, 0.0][:dim])
    ins_c, n_c = evaluate(oracle, e, 0, X)
    ins_s, n_s = evaluate(oracle, e, 1, X)
    assert ins_s == ins_c
    assert abs(n_s[0] - n_c[0]) + abs(n_s[1] - n_c[1]) < 1e-4
    for _ in range(300):
        # on the surface, where a distance field's gradient is the curve's normal
        x = rng.uniform(-1.1, 0.3)
        X = np.array([x, x * x, rng.uniform(-0.01, 0.01)][:dim])
        n_c = evaluate(oracle, e, 0, X)[1]
        n_s = evaluate(oracle, e, 1, X)[1]
        assert np.linalg.norm(n_s - n_c) < 1e-3, X
        # off the surface: the same side (inside below y = x^2)
        X[1] = x * x + rng.choice([-1.0, 1.0]) * rng.uniform(1e-4, 0.02)
        assert evaluate(oracle, e, 1, X)[0] == evaluate(oracle, e, 0, X)[0], X


@pytest.mark.parametrize("dim", [2, 3])
def test_outside_the_lattice_the_object_is_absent(oracle, dim):
    phi = -np.ones((4, 4) + ((4,) if dim == 3 else ()), dtype=np.float32)   # solid everywhere it is defined
    e = engine_with(oracle, dim)
    e.add_sdf(phi, [0.0] * dim, 0.1)
    assert evaluate(oracle, e, 0, np.full(dim, 0.15))[0]
    ins, n = evaluate(oracle, e, 0, np.array([0.31, 0.1, 0.1][:dim]))
    assert not ins and not n.any()


@pytest.mark.parametrize("dim", [2, 3])
def test_slide_on_the_curve_level_set(oracle, dim):
    assert PC.sdf_slide(oracle, "SlipFree", dim) < 1e-3
