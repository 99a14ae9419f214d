"""The CPU oracle against the reference's own physics tests (tests/tests.cpp) at sizes the CPU finishes in
seconds, the APIC affine invariant, conservation, and the golden fixtures it wrote."""
import os

import numpy as np
import pytest

import mpm_abi as A
import invariants as I
import parity_scenes as PS
import physics_cases as P

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("bc", ["SlipFree", "SlipStick"])
def test_analytic_slide_on_curve(oracle, bc, dim):
    """BoundaryTest.AnalyticSlipFree/SlipStick, tests/tests.cpp:13-98 (1 particle, same sizes)."""
    assert P.analytic_slide(oracle, bc, dim=dim) < 1e-3


@pytest.mark.parametrize("general", [False, True])
@pytest.mark.parametrize("bc", ["SlipFree", "SlipStick"])
def test_coulomb_incline_2d(oracle, bc, general):
    """CoulombFrictionTest.*, tests/tests.cpp:100-306, in plane strain (the reference's 2D branches)."""
    assert P.coulomb_incline(oracle, bc, general, dim=2) < 1e-3


def test_rotation_energy_2d(oracle):
    """EnergyTest.Rotation, tests/tests.cpp:543-603, 2D and 2 s instead of 20 s."""
    loss = P.rotation_energy(oracle, dim=2, spacing=0.02, end_frame=2)
    assert 0.0 <= loss <= 1e-3


@pytest.mark.parametrize("dim,spline,musl", [(3, 2, False), (3, 2, True), (2, 2, False), (3, 3, False), (2, 3, True)])
def test_apic_affine_invariant(oracle, dim, spline, musl):
    assert I.check_affine_step(oracle, dim, spline, musl)


def test_conservation_checks_pass(oracle):
    """checkMassConservation / checkMomentumConservation (simulation.cpp:292-333) at 1e-10."""
    e = PS.build(oracle, "dp3_usl_apic")
    rc, info = e.step(5, A.STEP_CHECK_CONSERVATION)
    assert rc == 0


@pytest.mark.parametrize("name", ["dp3_musl_apic", "col2d_usl", "incline_mccmui_mibf", "mcc_sinhimpl"])
def test_oracle_reproduces_golden(oracle, name):
    """Regression pin: the committed fixtures (tests/golden/make_golden.py, single-threaded oracle) are
    reproduced bit for bit."""
    path = os.path.join(GOLDEN, f"{name}.npz")
    gold = np.load(path)
    e = PS.build(oracle, name)
    rc, info = e.step(int(gold["steps"]))
    assert rc == 0
    p = e.get_particles()
    for k in ("x", "v", "F", "Bmat", "eps_pl_dev", "eps_pl_vol"):
        np.testing.assert_array_equal(p[k], gold[k], err_msg=k)
    assert info.dt == float(gold["dt"]) and info.plastic_count == int(gold["plastic_count"])
