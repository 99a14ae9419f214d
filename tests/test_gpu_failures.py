"""The reference's run-stopping errors on the engine, against the oracle: the same bits on the same step, the
failing step's particle state (x, v at the parity bar; F where finite), time and step counter not advanced
(simulation.cpp:175-177). Errors that the reference does not have: a NaN in the particle state stops the run
before the step that would use it (MPM_ERR_NONFINITE), and invalid enumerations are refused when a plate or
object is added (the reference stops at the first collision, boundary_collision.cpp:68). A failing member of a
batch stops alone."""
import numpy as np
import pytest

import failure_scenes as FS
import mpm_abi as A
import parity_scenes as PS
from conftest import rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", list(FS.FAILURES))
def test_error_bits_and_state_match_the_oracle(engine, oracle, name):
    print(f"failure scene {name}", flush=True)
    g, o = FS.build(engine, name), FS.build(oracle, name)
    rg, ig = g.step(5)
    ro, io = o.step(5)
    want = A.ERR[FS.FAILURES[name][1]]
    assert rg & want and rg == ro, (A.status_names(rg), A.status_names(ro))
    assert ig.steps_done == io.steps_done == 1 and ig.time == io.time == 0.0 and ig.step == io.step == 0
    pg, po = g.get_particles(), o.get_particles()
    assert rel_err(pg["x"], po["x"]) < 1e-10
    assert rel_err(pg["v"], po["v"], floor=1e-6) < 1e-10
    fin = np.isfinite(po["F"]).all(axis=(1, 2)) & np.isfinite(pg["F"]).all(axis=(1, 2))
    assert rel_err(pg["F"][fin], po["F"][fin], floor=1.0) < 1e-8


def test_nan_velocity_stops_before_the_first_step(engine):
    print("NaN velocity", flush=True)
    cfg, plates, objs, f = PS.dp_collapse(dim=3, n=8)
    v = np.zeros_like(f["x"])
    v[17, 1] = np.nan
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(f["x"], v=v)
    rc, info = e.step(5)
    assert rc & A.ERR["NONFINITE"] and info.steps_done == 0 and info.time == 0.0 and info.step == 0


def test_nan_deformation_gradient_stops_the_run(engine):
    """A NaN in F reaches the stress and from there the grid and the velocities; the run stops at the start of the
    first step whose velocities or positions are not finite. An off-diagonal NaN enters the stress only once the
    deformation update has spread it over F (one step later), so the run stops after one or two completed
    steps, with time and step counter those of the completed steps."""
    print("NaN F", flush=True)
    cfg, plates, objs, f = PS.dp_collapse(dim=3, n=8)
    F = np.tile(np.eye(3), (f["x"].shape[0], 1, 1))
    F[40, 0, 1] = np.nan
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(f["x"], F=F)
    rc, info = e.step(5)
    assert rc & A.ERR["NONFINITE"] and 1 <= info.steps_done <= 2 and info.step == info.steps_done and info.time > 0.0


def test_supersonic_particle_warns_and_steps_on(engine, oracle):
    """updateDt (update_dt.cpp:14-17): a particle at or above the elastic wave speed sqrt(E / rho) raises the
    warning bit and the step is taken, with the oracle's dt and the same bits."""
    print("supersonic warning", flush=True)
    c_wave = (1e6 / 1000.0) ** 0.5
    cfg = A.default_config(dim=3, dx=0.01, particle_volume=1e-6, particle_mass=1e-3, rho=1000.0, E=1e6,
                           flip_ratio=-2.0, fps=1.0)
    x = np.array([[0.0, 0.0, 0.0], [0.003, 0.002, 0.001], [0.006, 0.001, 0.004]])
    v = np.tile([1.2 * c_wave, 0.0, 0.0], (3, 1))
    out = []
    for lib in (engine, oracle):
        e = A.Engine(lib, cfg, [])
        e.set_particles(x, v=v)
        out.append(e.step(1))
    (rg, ig), (ro, io) = out
    assert rg == ro == 0 and ig.status == io.status and ig.status & A.ERR["WARN_SUPERSONIC"], A.status_names(ig.status)
    assert ig.steps_done == io.steps_done == 1 and ig.dt == io.dt and ig.time == io.time


def test_invalid_enumerations_are_refused_when_added(engine):
    cfg = A.default_config(dim=3, dx=0.01, particle_volume=1e-6, particle_mass=1e-3)
    for make in (lambda: A.make_object("Ground", "NoSlip", 0.0), lambda: A.make_plate(0.0, "bottom", "NoSlip")):
        bad = make()
        bad.bc = 7
        with pytest.raises(A.MPMError) as ex:
            A.Engine(engine, cfg, [bad] if hasattr(bad, "plate_type") else [], [] if hasattr(bad, "plate_type") else [bad])
        assert ex.value.code == A.ERR["INVALID_ARGUMENT"]
    obj = A.make_object("Ground", "NoSlip", 0.0)
    obj.kind = 42
    with pytest.raises(A.MPMError) as ex:
        A.Engine(engine, cfg, [], [obj])
    assert ex.value.code == A.ERR["INVALID_ARGUMENT"]


def test_a_failing_batch_member_stops_alone(engine):
    print("failing batch member", flush=True)
    bad = FS.build(engine, "perzyna_negative")
    cfg, plates, objs, f = PS.dp_collapse(dim=3, musl=True, n=10)
    good = A.Engine(engine, cfg, plates, objs)
    good.set_particles(f["x"])
    alone = A.Engine(engine, cfg, plates, objs)
    alone.set_particles(f["x"])
    assert alone.step(12)[0] == 0
    b = A.Batch(engine, chunk_steps=4)
    b.add(bad, 12)
    b.add(good, 12)
    rc, _ = b.step(12)
    ib, left_b = A.Batch.scene_info(bad)
    ig, left_g = A.Batch.scene_info(good)
    assert ib.status & A.ERR["PERZYNA_NEGATIVE"] and ib.steps_done == 1 and ib.time == 0.0 and left_b == 0
    assert ig.steps_done == 12 and ig.status & A.EXIT_MASK == 0
    assert np.array_equal(good.get_particles()["x"], alone.get_particles()["x"])
    b.close()
