"""The reference's run-stopping errors on the oracle: each failure scene raises its bit on the first step, the
failing step completes but time and the step counter do not advance (simulate() returns after advanceStep when
exit is set, simulation.cpp:175-177), and an invalid boundary condition stops the run at the first collision
(boundary_collision.cpp:68,248,378)."""
import numpy as np
import pytest

import failure_scenes as FS
import mpm_abi as A


@pytest.mark.parametrize("name", list(FS.FAILURES))
def test_error_stops_after_the_failing_step(oracle, name):
    e = FS.build(oracle, name)
    rc, info = e.step(5)
    want = A.ERR[FS.FAILURES[name][1]]
    assert rc & want, A.status_names(rc)
    assert info.steps_done == 1 and info.time == 0.0 and info.step == 0


def test_invalid_bc_stops_at_the_first_collision(oracle):
    cfg = A.default_config(dim=3, dx=0.01, particle_volume=1e-6, particle_mass=1e-3, gravity=[0, -9.81, 0])
    obj = A.make_object("Ground", "NoSlip", 0.0, param=0.0)
    obj.bc = 7
    e = A.Engine(oracle, cfg, [], [obj])
    e.set_particles(np.array([[0.0, 0.001, 0.0], [0.004, 0.002, 0.003]]))
    rc, info = e.step(5)
    assert rc & A.ERR["INVALID_ARGUMENT"] and info.steps_done == 1 and info.time == 0.0
