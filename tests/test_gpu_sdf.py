"""The level-set object (mpm_add_sdf_object, the dense-lattice form of ObjectVdb, src/objects/object_vdb.hpp) on
the engine: BoundaryTest.VDB (tests/tests_vdb.cpp:40-82), a particle sliding on the level set of the curve
y = x^2 to sqrt(2 g) within 1e-3, in 2D and 3D; a batch member with a level set equals the scene alone. The
1e-10 parity with the oracle is in test_gpu_parity.py (obj_sdf_* scenes)."""
import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
import physics_cases as PC

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim", [2, 3])
def test_slide_on_the_curve_level_set(engine, dim):
    print(f"sdf slide {dim}d", flush=True)
    assert PC.sdf_slide(engine, "SlipFree", dim) < 1e-3


def test_level_set_member_of_a_batch_equals_alone(engine):
    print("sdf batch", flush=True)
    alone = PS.build(engine, "obj_sdf_curve_slipfree")
    assert alone.step(12)[0] == 0
    m = PS.build(engine, "obj_sdf_curve_slipfree")
    other = PS.build(engine, "dp3_musl_apic")
    b = A.Batch(engine, chunk_steps=5)
    b.add(m, 12)
    b.add(other, 12)
    rc, _ = b.step(12)
    assert rc == 0
    pa, pm = alone.get_particles(), m.get_particles()
    for k in ("x", "v", "F"):
        assert np.array_equal(pa[k], pm[k]), k
    b.close()
