"""Independent pins of the oracle functions that the reference's own tests do not fix: each check follows from
the mathematics of the function (flow rule, yield surface, hardening law, geometry, a closed form, a brute-force
search or a special case that reduces to an already pinned routine) rather than re-typing the oracle's code,
and is chosen so that a plausible restatement mistake (a wrong sign, a dropped term, the old instead of the
new hardening variable, a swapped prefactor) fails it.

Return maps (src/simulation/plasticity.cpp, src/plasticity_helpers/*.hpp) are driven one particle at a time
through oracle_plasticity_single; p, q and the Hencky strains are measured from the returned F.
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS
import returnmap as RM
from returnmap import consts, dptr, pq, project

rng = np.random.default_rng(77)
I_EPS_DEV, I_EPS_VOL, I_PRAD, I_DGAMMA, I_VISC, I_MUI = range(6)


def trial(dim, vol, dev):
    return RM.trial(dim, vol, dev, rng)


def dev_norm(F):
    h = np.log(np.linalg.svd(F, compute_uv=False))
    return np.linalg.norm(h - h.mean())


# ------------------------------------------------------------------------------------------------ MCC
def hardened_p0(law, cfg, K, eps_old, eps_new):
    """Consolidation pressure of each hardening law (the explicit laws use the plastic volumetric strain of
    the step's start, the implicit ones the returned strain): p0 exp(-xi eps) (Expo*), K sinh(-xi eps +
    asinh(p0 / K)) (Sinh*), floored as the reference floors them."""
    p0, xi, tol = cfg.p0, cfg.xi, cfg.stress_tolerance
    if law == "NoHard":
        return p0
    if law == "ExpoExpl":
        return max(tol, p0 * math.exp(-xi * eps_old))
    if law == "SinhExpl":
        return max(tol, K * math.sinh(-xi * eps_old + math.asinh(p0 / K)))
    if law == "ExpoImpl":
        return p0 * math.exp(-xi * eps_new)
    return max(0.01, K * math.sinh(-xi * eps_new + math.asinh(p0 / K)))


MCC_TRIALS = [(-2.0e-3, 6e-3), (-3.5e-3, 4e-3), (-1.0e-3, 5e-3), (0.4e-3, 3e-3)]


@pytest.mark.parametrize("law", ["NoHard", "ExpoExpl", "SinhExpl", "ExpoImpl", "SinhImpl"])
@pytest.mark.parametrize("dim", [2, 3])
def test_mcc_return_is_normal_to_the_hardened_ellipse(oracle, law, dim):
    """MCC (plasticity.cpp:487-621 with mcc_rma_explicit.hpp / mcc_rma_implicit_*_onevar.hpp), xi > 0 and a
    nonzero plastic volumetric strain at the start:
      * the returned (p, q) lies on the ellipse M^2 (p - p_c)(p + beta p_c) + q^2 = 0 of the hardening law's
        p_c (explicit laws: p_c of the old strain; implicit laws: of the returned strain);
      * associative flow: the plastic multiplier read from the pressure, (p_t - p) / (K dy/dp), equals the one
        read from the deviator, (q_t - q) / (2 mu' q) with mu' = rma_prefac mu (explicit) or mu (implicit);
      * bookkeeping: the volumetric plastic strain grows by (p - p_t) / K and the deviatoric one by the
        deviatoric Hencky strain removed, |dev h_trial| - |dev h_new|."""
    # xi per law so that the step's plastic volumetric strain (about 1e-3) moves p_c by a few per cent
    M, p0, beta, eps0 = 1.2, 2000.0, 0.3, -5e-4
    xi = {"NoHard": 0.0, "ExpoExpl": 20.0, "ExpoImpl": 20.0, "SinhExpl": 0.1, "SinhImpl": 0.1}[law]
    cfg = A.default_config(dim=dim, dx=0.01, plastic_model="MCC", hardening_law=law, M=M, p0=p0, beta=beta, xi=xi)
    c = consts(cfg)
    K, mu = c["K"], c["mu"]
    explicit = law in ("NoHard", "ExpoExpl", "SinhExpl")
    checked = 0
    for vol, dv in MCC_TRIALS:
        F, h = trial(dim, vol, dv)
        p_t, q_t, _, _ = pq(F, cfg, K, c["e_mu"])
        Fn, sc, plastic, err = project(oracle, cfg, F, scalars=[0.0, eps0, 0, 0, 0, 0])
        assert err == 0
        if not plastic:
            continue
        p, q, _, _ = pq(Fn, cfg, K, c["e_mu"])
        eps_new = sc[I_EPS_VOL]
        pc = hardened_p0(law, cfg, K, eps0, eps_new)
        assert -beta * pc < p < pc, "returned pressure outside the ellipse's range"
        y = M * M * (p - pc) * (p + beta * pc) + q * q
        assert abs(y) < 1e-7 * pc * pc, (vol, dv, y)
        assert abs((eps_new - eps0) - (p - p_t) / K) < 1e-12 * max(1.0, abs(p - p_t) / K)
        assert abs(sc[I_EPS_DEV] - (dev_norm(F) - dev_norm(Fn))) < 1e-12
        dydp = M * M * (2 * p + (beta - 1) * pc)
        g_p = (p_t - p) / (K * dydp)
        g_q = (q_t - q) / ((c["rma"] if explicit else 1.0) * mu * 2 * q)
        tol = 2e-3 if explicit else 2e-2
        assert g_p > 0 and abs(g_p - g_q) <= tol * abs(g_p), (vol, dv, g_p, g_q)
        checked += 1
    assert checked >= 3


def test_mcc_hardening_variable_enters_with_its_sign(oracle):
    """A larger accumulated compaction (eps_pl_vol < 0 in the reference's sign) raises p_c = p0 exp(-xi eps): the
    same trial returns to a larger ellipse, hence a larger q."""
    cfg = A.default_config(dim=3, dx=0.01, plastic_model="MCC", hardening_law="ExpoImpl", M=1.2, p0=2000.0, beta=0.3, xi=20.0)
    c = consts(cfg)
    F, _ = trial(3, -2e-3, 6e-3)
    q_soft = pq(project(oracle, cfg, F, scalars=[0, 0.004, 0, 0, 0, 0])[0], cfg, c["K"], c["e_mu"])[1]
    q_hard = pq(project(oracle, cfg, F, scalars=[0, -0.004, 0, 0, 0, 0])[0], cfg, c["K"], c["e_mu"])[1]
    assert q_hard > q_soft * 1.05


# ------------------------------------------------------------------------------------------------ MCCMui
@pytest.mark.parametrize("law", ["ExpoImpl", "SinhImpl", "ExpoExpl"])
def test_mccmui_rate_dependent_overlay(oracle, law):
    """mu(I)-MCC (plasticity.cpp:405-484): after the MCC return to (p, q_s) with M = mu_1, the deviatoric
    stress is raised by the mu(I) rheology of Jop, Forterre & Pouliquen: with gdot = delta_gamma / dt and the
    inertial number I = gdot d / sqrt(p_s / rho_s) (p_s = p + beta p_c + tol),
      muI = mu_1 + (mu_2 - mu_1) / (I_0 / I + 1),
      q - q_s = (muI - mu_1) sqrt((p_c - p) p_s) = viscosity * gdot,
      q = q_trial - f_mu dt gdot (the stress update of the step)."""
    mu1, mu2, I0, d, rho_s, p0 = 0.38, 0.64, 0.279, 1e-3, 2500.0, 1000.0
    # the sinh law with K = 1e7 keeps p_c near p0 only for xi eps << asinh(p0 / K) = 1e-4
    xi, eps0 = (0.2, 1e-4) if law == "SinhImpl" else (2.0, 1e-3)
    dt = 2e-5
    cfg = A.default_config(dim=3, dx=0.004, plastic_model="MCCMui", hardening_law=law, mu_1=mu1, mu_2=mu2, I_ref=I0,
                           grain_diameter=d, rho_s=rho_s, p0=p0, beta=0.0, xi=xi, E=1.2e7, nu=0.3)
    c = consts(cfg)
    K, tol = c["K"], cfg.stress_tolerance
    checked = 0
    # trials about 1.5-4x above the ellipse, away from its apex (p_t 400-600 Pa, p_c about 1 kPa)
    for vol, dv in ((-5e-5, 5e-5), (-6e-5, 8e-5), (-4e-5, 6e-5), (-5e-5, 1.2e-4)):
        F, _ = trial(3, vol, dv)
        p_t, q_t, _, _ = pq(F, cfg, K, c["e_mu"])
        Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt, scalars=[0.0, eps0, 0, 0, 0, 0])
        assert err == 0 and plastic == 1
        p, q, _, _ = pq(Fn, cfg, K, c["e_mu"])
        eps_new = sc[I_EPS_VOL]
        if law == "ExpoImpl":
            pc = max(p0 * math.exp(-xi * eps_new), p + tol)
        elif law == "SinhImpl":
            pc = max(K * math.sinh(-xi * eps_new + math.asinh(p0 / K)), p + tol)
        else:
            pc = max(tol, p0 * math.exp(-xi * eps0))
        qs = mu1 * math.sqrt((pc - p) * p)        # the MCC ellipse at p (beta = 0), M = mu_1
        gdot = sc[I_DGAMMA]
        if q <= qs * (1 + 1e-12):
            continue                                 # rate-independent return (q_trial within tolerance)
        ps = p + tol
        I = gdot * d / math.sqrt(ps / rho_s)
        # gdot is read back as (q_trial - q) / (f_mu dt), which loses digits to the cancellation; a structural
        # mistake (the inertial number without sqrt(rho_s), mu_1 and mu_2 swapped, p for p_s) is off by >1e-3
        assert abs(sc[I_MUI] - (mu1 + (mu2 - mu1) / (I0 / I + 1.0))) < 1e-7
        S = math.sqrt((pc - p) * ps)
        assert abs((q - qs) - (sc[I_MUI] - mu1) * S) < 1e-8 * q
        assert abs(sc[I_VISC] * gdot - (q - qs)) < 1e-8 * q
        assert abs(q - (q_t - c["f_mu"] * dt * gdot)) < 1e-9 * q_t
        checked += 1
    assert checked >= 2


def test_mccmui_elastic_state_keeps_mu1(oracle):
    cfg = A.default_config(dim=3, dx=0.004, plastic_model="MCCMui", hardening_law="ExpoImpl", mu_1=0.38, mu_2=0.64,
                           p0=1000.0, E=1.2e7, nu=0.3)
    F, _ = trial(3, -1e-5, 1e-6)
    Fn, sc, plastic, err = project(oracle, cfg, F, dt=2e-5, scalars=[0, 0, 0, 0, 0.7, 0.9])
    assert plastic == 0 and err == 0 and sc[I_MUI] == 0.38 and sc[I_VISC] == 0.0
    np.testing.assert_array_equal(Fn, F)


# ------------------------------------------------------------------------------------------------ DPSoft
def _dp_cfg(model, dim=3, **kw):
    base = dict(dim=dim, dx=0.01, plastic_model=model, M=math.tan(math.radians(30)), q_cohesion=200.0, use_pradhana=0, xi=0.0)
    base.update(kw)
    return A.default_config(**base)


@pytest.mark.parametrize("dim", [2, 3])
def test_dpsoft_without_softening_is_drucker_prager(oracle, dim):
    """DPSoft (plasticity.cpp:87-162) with xi = 0 and no Pradhana shift has the fixed tip -q_cohesion / M and the
    cone q = M p + q_cohesion of DP (plasticity.cpp:51-85, pinned by the cone/tip tests): same F and strains."""
    for vol, dv in ((-2e-3, 8e-3), (5e-3, 1e-3), (-1e-3, 1e-5), (-3e-3, 2e-3), (1e-4, 4e-3)):
        F, _ = trial(dim, vol, dv)
        a = project(oracle, _dp_cfg("DP", dim), F, scalars=[0.01, 0.002, 0, 0, 0, 0])
        b = project(oracle, _dp_cfg("DPSoft", dim), F, scalars=[0.01, 0.002, 0, 0, 0, 0])
        np.testing.assert_array_equal(a[0], b[0])
        np.testing.assert_array_equal(a[1][[I_EPS_DEV, I_EPS_VOL, I_DGAMMA]], b[1][[I_EPS_DEV, I_EPS_VOL, I_DGAMMA]])
        assert a[2] == b[2]


def test_dpsoft_softened_cone_tip_and_pradhana_shift(oracle):
    """DPSoft with softening (xi > 0) and the Pradhana shift p_shift = -K eps_pradhana:
      * tip (p_t + p_shift at or below the softened tip p_tip(eps) = -q_c / M exp(-xi eps_dev)): an isotropic
        state at p = p_tip(eps_dev of the step's start); eps_vol and eps_pradhana grow by (p - p_t) / K; all
        deviatoric Hencky strain becomes plastic;
      * cone: no volume change (p = p_t); the state lies on the cone softened by the plastic strain that the
        unsoftened return would give, q = M (p + p_shift - p_tip(eps~)) with eps~ = eps_dev + |dev h| - q_y / e_mu
        (q_y = M (p_t + p_shift - p_tip(eps_dev))); eps_dev grows by the deviatoric strain removed; the Pradhana
        strain resets;
      * elastic: the Pradhana strain resets and delta_gamma is zero."""
    M, qc, xi = math.tan(math.radians(30)), 200.0, 30.0
    cfg = _dp_cfg("DPSoft", 3, q_cohesion=qc, xi=xi, use_pradhana=1)
    c = consts(cfg)
    K, e_mu = c["K"], c["e_mu"]
    ptip0 = -qc / M
    eps_dev0, prad0 = 0.01, -2e-5
    shift = -K * prad0
    seen = set()
    for vol, dv in ((-2e-3, 6e-3), (-8e-4, 3e-3), (2e-4, 1e-3), (-3e-3, 5e-4), (1e-3, 2e-3), (-1e-3, 1e-5)):
        F, h = trial(3, vol, dv)
        p_t, q_t, _, _ = pq(F, cfg, K, e_mu)
        Fn, sc, plastic, err = project(oracle, cfg, F, scalars=[eps_dev0, 0.0, prad0, 0.5, 0, 0])
        p, q, _, _ = pq(Fn, cfg, K, e_mu)
        ptip = ptip0 * math.exp(-xi * eps_dev0)
        qy = M * (p_t + shift - ptip)
        if p_t + shift <= ptip:
            seen.add("tip")
            assert abs(p - ptip) < 1e-9 * abs(ptip) and q < 1e-9 * abs(ptip)
            assert abs(sc[I_EPS_VOL] - (p - p_t) / K) < 1e-12 and abs(sc[I_PRAD] - prad0 - (p - p_t) / K) < 1e-12
            assert abs(sc[I_EPS_DEV] - eps_dev0 - dev_norm(F)) < 1e-12
        elif q_t <= qy:
            seen.add("elastic")
            assert plastic == 0 and sc[I_PRAD] == 0.0 and sc[I_DGAMMA] == 0.0
        else:
            eps_t = eps_dev0 + dev_norm(F) - qy / e_mu
            ptip_t = ptip0 * math.exp(-xi * eps_t)
            if p_t + shift < ptip_t:
                seen.add("softened tip")
                assert abs(p - ptip_t) < 1e-9 * abs(ptip_t)
                continue
            seen.add("cone")
            assert abs(p - p_t) < 1e-9 * abs(p_t)
            assert abs(q - M * (p + shift - ptip_t)) < 1e-9 * q
            assert abs((sc[I_EPS_DEV] - eps_dev0) - (q_t - q) / e_mu) < 1e-12
            assert sc[I_PRAD] == 0.0
    assert {"tip", "elastic", "cone"} <= seen, seen


# ------------------------------------------------------------------------------------------------ DPVisc / MCCVisc
@pytest.mark.parametrize("mibf", [0, 1])
def test_dpvisc_linear_rate_law_and_mibf_friction(oracle, mibf):
    """DPVisc with exponent 1 (plasticity.cpp:284-290): the Perzyna overstress law q = q_y (1 + eta gdot) with
    q_y = M (p_t + p_shift) + q_c and gdot = delta_gamma; the pressure is unchanged. With MIBF the particle's
    friction coefficient is that of the returned state, q = muI (p + p_shift) + q_c (:327-332). With eta = 0 the
    return is the DP cone."""
    M, qc, eta, dt = math.tan(math.radians(30)), 50.0, 0.02, 1e-4
    for visc in (eta, 0.0):
        cfg = _dp_cfg("DPVisc", 3, q_cohesion=qc, perzyna_visc=visc, use_mibf=mibf, use_pradhana=1)
        c = consts(cfg)
        K, e_mu = c["K"], c["e_mu"]
        for vol, dv in ((-2e-3, 6e-3), (-1e-3, 3e-3), (-4e-3, 8e-3)):
            prad0 = 1e-5
            F, _ = trial(3, vol, dv)
            p_t, q_t, _, _ = pq(F, cfg, K, e_mu)
            Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt, scalars=[0, 0, prad0, 0, 0, 0.3])
            assert err == 0 and plastic == 1
            p, q, _, _ = pq(Fn, cfg, K, e_mu)
            shift = -K * prad0
            qy = M * (p_t + shift) + qc
            assert abs(p - p_t) < 1e-9 * abs(p_t)
            assert abs(q - qy * (1 + visc * sc[I_DGAMMA])) < 1e-9 * q
            assert abs(sc[I_DGAMMA] * dt - (q_t - q) / c["f_mu"]) < 1e-9 * sc[I_DGAMMA] * dt
            if mibf:
                assert abs(q - (sc[I_MUI] * (p + shift) + qc)) < 1e-9 * q
            else:
                assert sc[I_MUI] == 0.3


@pytest.mark.parametrize("mibf", [0, 1])
def test_mccvisc_overstress_on_the_hardened_ellipse(oracle, mibf):
    """MCCVisc with exponent 1 (plasticity.cpp:540-610): after the MCC return to (p, q_s), q = q_s (1 + eta gdot)
    (gdot = delta_gamma); q_s lies on the ellipse of the returned hardening strain. With MIBF the friction
    coefficient is q / sqrt((p + beta p_c)(p_c - p)). With eta = 0 the result is MCC's."""
    M, p0, beta, xi, eta, dt = 1.1, 2000.0, 0.2, 10.0, 0.05, 1e-4
    kw = dict(dim=3, dx=0.01, M=M, p0=p0, beta=beta, xi=xi, hardening_law="ExpoImpl", use_mibf=mibf)
    for vol, dv in ((-2e-3, 6e-3), (-3e-3, 4e-3), (-1e-3, 5e-3)):
        F, _ = trial(3, vol, dv)
        cfg = A.default_config(plastic_model="MCCVisc", perzyna_visc=eta, **kw)
        c = consts(cfg)
        K = c["K"]
        p_t, q_t, _, _ = pq(F, cfg, K, c["e_mu"])
        Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt, scalars=[0, 0.001, 0, 0, 0, 0])
        assert err == 0 and plastic == 1
        p, q, _, _ = pq(Fn, cfg, K, c["e_mu"])
        pc = p0 * math.exp(-xi * sc[I_EPS_VOL])
        qs = M * math.sqrt((pc - p) * (beta * pc + p))
        assert q > qs and abs(q - qs * (1 + eta * sc[I_DGAMMA])) < 1e-8 * q
        if mibf:
            assert abs(sc[I_MUI] - q / math.sqrt((p + beta * pc) * (pc - p))) < 1e-10
        ref = project(oracle, A.default_config(plastic_model="MCC", **kw), F, dt=dt, scalars=[0, 0.001, 0, 0, 0, 0])
        zero = project(oracle, A.default_config(plastic_model="MCCVisc", perzyna_visc=0.0, **kw), F, dt=dt,
                       scalars=[0, 0.001, 0, 0, 0, 0])
        np.testing.assert_allclose(zero[0], ref[0], rtol=0, atol=1e-15)


# ------------------------------------------------------------------------------------------------ objects
def object_eval(oracle, e, k, X):
    inside, n = C.c_int(), np.zeros(e.dim)
    x = np.ascontiguousarray(X, dtype=np.float64)
    assert oracle.lib.oracle_object_eval(e.h, k, dptr(x), C.byref(inside), dptr(n)) == 0
    return bool(inside.value), n


# Level-set functions of the shipped objects, negative inside, from each object's geometry
# (src/objects/object_*.hpp): the region below a bump / ramp profile, above the gate's parabola, outside the
# silo's surface of revolution r = tanh(y) + 1, below the curve y = x^2.
LEVEL_SETS = {
    "Bump": lambda X, prm: X[1] - 0.0475 / math.cosh(25 * (X[0] - 0.43)),
    "Ramp": lambda X, prm: X[1] - 0.1 * math.tanh(-100 * X[0]),
    "Gate": lambda X, prm: (prm + 100 * X[0] ** 2) - X[1],
    "Curve": lambda X, prm: X[1] - X[0] ** 2,
    "Ground": lambda X, prm: X[1] - prm,
    "Silo": lambda X, prm: (math.tanh(X[1]) + 1) - math.hypot(X[0], X[2]),
}
SAMPLE_BOX = {"Bump": ((0.25, 0.6), (-0.02, 0.08)), "Ramp": ((-0.05, 0.05), (-0.15, 0.15)), "Gate": ((-0.05, 0.05), (0.0, 0.3)),
              "Curve": ((-1.0, 1.0), (-0.2, 1.0)), "Ground": ((-1.0, 1.0), (-0.5, 0.5)), "Silo": ((-2.0, 2.0), (-0.9, 1.0))}


@pytest.mark.parametrize("kind", ["Bump", "Ramp", "Gate", "Curve", "Ground", "Silo"])
@pytest.mark.parametrize("dim", [2, 3])
def test_object_normal_is_the_level_set_gradient(oracle, kind, dim):
    """inside() agrees with the sign of the object's level-set function phi and normal() is grad phi / |grad phi|,
    the gradient taken by central differences (the outward normal of the boundary phi = 0)."""
    if kind == "Silo" and dim == 2:
        pytest.skip("ObjectSilo is 3D only")
    param = {"Gate": 0.016, "Ground": 0.05, "Silo": -0.8}.get(kind, 0.0)
    e = A.Engine(oracle, A.default_config(dim=dim, dx=0.01, particle_volume=1e-6, particle_mass=1e-3), [],
                 [A.make_object(kind, "SlipFree", 0.0, param=param)])
    phi = LEVEL_SETS[kind]
    (x0, x1), (y0, y1) = SAMPLE_BOX[kind]
    h = 1e-7
    for _ in range(400):
        X = np.array([rng.uniform(x0, x1), rng.uniform(y0, y1), rng.uniform(-1.0, 1.0)][:dim])
        Xp = X if dim == 3 else np.append(X, 0.0)
        inside, n = object_eval(oracle, e, 0, X)
        f = phi(Xp, param)
        if abs(f) > 1e-9:
            want = f <= 0 if kind in ("Silo", "Curve", "Ground") else f < 0
            if kind == "Silo" and Xp[1] < param:
                want = False
            assert inside == want, (X, f)
        g = np.array([(phi(Xp + h * np.eye(3)[d], param) - phi(Xp - h * np.eye(3)[d], param)) / (2 * h) for d in range(dim)])
        if kind == "Silo" and math.hypot(Xp[0], Xp[2]) < 1e-3:
            continue
        np.testing.assert_allclose(n, g / np.linalg.norm(g), rtol=0, atol=2e-6)


@pytest.mark.parametrize("dim", [2, 3])
def test_box_rotated_geometry(oracle, dim):
    """ObjectBoxRotated (object_box.hpp:6-44): inside() is the box |R^T (X - c)|_i <= h_i; normal() is the
    rotated unit axis of the local coordinate whose distance to its face, |h_i - |l_i||, is largest (the
    reference's choice), signed by that coordinate. Brute force over the axes; rotation covariance."""
    c = np.array([0.1, -0.2, 0.05][:dim])
    hx = np.array([0.2, 0.05, 0.1][:dim])
    R = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
    e = A.Engine(oracle, A.default_config(dim=dim, dx=0.01, particle_volume=1e-6, particle_mass=1e-3), [],
                 [A.make_object("BoxRotated", "NoSlip", 0.0, center=c, half_extents=hx, rotation=R)])
    for _ in range(500):
        local = rng.uniform(-1.3, 1.3, dim) * hx
        X = c + R @ local
        inside, n = object_eval(oracle, e, 0, X)
        assert inside == bool(np.all(np.abs(local) <= hx))
        dist = np.abs(hx - np.abs(local))
        k = int(np.argmax(dist))
        if np.sort(dist)[-1] - np.sort(dist)[-2] < 1e-12:
            continue
        want = np.zeros(dim)
        want[k] = 1.0 if local[k] > 0 else -1.0
        np.testing.assert_allclose(n, R @ want, rtol=0, atol=1e-14)


# ------------------------------------------------------------------------------------------------ MIBF grid friction
@pytest.mark.parametrize("model,fill", [("DPVisc", "M"), ("MCCMui", "mu_1")])
def test_mibf_grid_friction_of_a_uniform_field_is_that_value(oracle, model, fill):
    """P2G's MIBF friction (p2g.cpp:45-46,92-93) is the weight average sum w mu_p / sum w: with every particle at
    the same mu (the value simulate() fills in, simulation.cpp:138-145) it is that mu at every node with mass."""
    cfg, plates, objs, f = PS.model_block(model, n=8, use_mibf=1)
    mu = getattr(cfg, fill)
    e = A.Engine(oracle, cfg, plates, objs)
    e.set_particles(f["x"], F=f["F"], v=f["v"])
    assert e.step(1)[0] == 0
    g = e.get_grid()
    m = g["mass"] > 0
    assert m.sum() > 100
    assert np.max(np.abs(g["friction"][m] - mu)) <= 8 * np.finfo(float).eps * mu


# ------------------------------------------------------------------------------------------------ remesh
def snapped(extreme, init, base, dx, lower):
    """remeshFixedCont's snap (remesh.cpp:227-280) by brute force in exact arithmetic: the grid edge moves from
    its initial position `base` by whole cells, as many as fit in the distance the particle extreme moved from its
    initial value less 1e-8 dx^2 (the reference's guard), away from the material when it expanded and towards it
    when it contracted."""
    dist = Fraction(extreme) - Fraction(init)
    grow = dist > 0 if not lower else dist < 0
    dist = abs(dist) - Fraction(1e-8) * Fraction(dx) * Fraction(dx)
    k = 0
    while (k + 1) * Fraction(dx) <= dist:
        k += 1
    sign = (1 if grow else -1) * (-1 if lower else 1)
    return base + sign * k * dx


@pytest.mark.parametrize("dim", [2, 3])
def test_remesh_snap_by_brute_force(oracle, dim):
    """Step by step, the grid of each step (origin, node count) is the initial box snapped by whole cells to the
    particle bounds of the step's start: edges at low_init + k dx and high_init + k dx (found by integer search),
    and node counts those of the accumulating arange of [low, high + dx)."""
    cfg, plates, objs, f = PS.dp_collapse(dim=dim, n=10, bc="NoSlip", friction=0.0)
    A.set_field(cfg, "gravity", [3.0, -9.81, 1.0])
    v = np.zeros_like(f["x"])
    v[:, 0] = np.linspace(-2.0, 2.0, v.shape[0])
    v[:, 1] = 1.0
    e = A.Engine(oracle, cfg, plates, objs)
    e.set_particles(f["x"], v=v)
    dx = cfg.dx
    x = f["x"]
    mn0, mx0 = x.min(axis=0), x.max(axis=0)
    lo0, hi0 = mn0 - 2 * dx, mx0 + 2 * dx
    out = np.zeros(4096)
    for step in range(12):
        x = e.get_particles()["x"]
        mn, mx = x.min(axis=0), x.max(axis=0)
        assert e.step(1)[0] == 0
        g = e.get_grid()
        axes = [g["x"], g["y"]] + ([g["z"]] if dim == 3 else [])
        for d in range(dim):
            low = snapped(mn[d], mn0[d], lo0[d], dx, True) if step else lo0[d]
            high = snapped(mx[d], mx0[d], hi0[d], dx, False) if step else hi0[d]
            assert axes[d][0] == low, (step, d)
            n = oracle.lib.oracle_arange(low, high + dx, dx, dptr(out), out.size)
            assert len(axes[d]) == n, (step, d)
            assert axes[d][0] <= mn[d] - dx and axes[d][-1] >= mx[d] + dx   # every window fits


# ------------------------------------------------------------------------------------------------ plates
def test_plate_schedule_closed_form(oracle):
    """moveObjects / ObjectPlate::move (object_plate.hpp:136-162): a plate moves at v / vmin_factor while the
    step's start time is below load_factor * frame_dt, then at v; a bottom plate's position follows v_y and its
    extent v_x. With constant dt the position after N steps is pos0 + dt (n_slow / vmin + N - n_slow) v with
    n_slow = ceil(t_load / dt)."""
    vx, vy, vmin, load, fps = 0.2, 0.5, 4.0, 0.3, 10.0
    cfg = A.default_config(dim=3, dx=0.01, particle_volume=1e-6, particle_mass=1e-3, rho=1000.0, fps=fps, end_frame=1)
    plate = A.make_plate(-1.0, "bottom", "NoSlip", 0.0, pos_lower=-2.0, pos_upper=2.0, v_object=(vx, vy, 0.0),
                         vmin_factor=vmin, load_factor=load)
    e = A.Engine(oracle, cfg, [plate])
    e.set_particles(np.array([[0.0, 0.0, 0.0], [0.005, 0.004, 0.003]]))
    dt = 0.5 * 0.01 / math.sqrt(1e6 / 1000.0)                # dt_max: the particles are at rest
    N = int((1.0 / fps) / dt) - 2                             # inside the first frame: every dt equal
    rc, info = e.step(N)
    assert rc == 0 and info.steps_done == N and info.dt == dt
    n_slow = math.ceil(load / fps / dt)
    travel = dt * (n_slow / vmin + (N - n_slow))
    p = e.get_plates()[0]
    assert abs(p.pos_object - (-1.0 + vy * travel)) < 1e-13
    assert abs(p.pos_lower - (-2.0 + vx * travel)) < 1e-13 and abs(p.pos_upper - (2.0 + vx * travel)) < 1e-13
    assert p.v_object[1] == vy                                # after the load phase the full velocity


# ------------------------------------------------------------------------------------------------ updateDt
def test_update_dt_closed_forms(oracle):
    """updateDt (update_dt.cpp:5-66): dt = min(cfl dx / |v|max, dt_max = cfl_elastic dx / sqrt(E / rho)), clamped to
    the frame end and the final time; a particle at or above the elastic wave speed raises a warning only."""
    dx, E, rho = 0.01, 1e6, 1000.0
    c_wave = math.sqrt(E / rho)
    base = dict(dim=3, dx=dx, particle_volume=1e-6, particle_mass=1e-3, rho=rho, E=E, flip_ratio=-2.0)
    x = np.array([[0.0, 0.0, 0.0], [0.003, 0.002, 0.001], [0.006, 0.001, 0.004]])
    v0 = np.array([3.0, -1.0, 2.0])
    # free flight: the CFL bound of the uniform speed (cfl 0.05, below dt_max for |v| = 3.7 m/s)
    e = A.Engine(oracle, A.default_config(fps=1.0, cfl=0.05, **base))
    e.set_particles(x, v=np.tile(v0, (3, 1)))
    rc, info = e.step(1)
    assert rc == 0 and info.dt == 0.05 * dx / math.sqrt(14.0)
    # at rest: dt_max
    e = A.Engine(oracle, A.default_config(fps=1.0, cfl_elastic=0.4, **base))
    e.set_particles(x)
    rc, info = e.step(1)
    assert info.dt == 0.4 * dx / c_wave
    # frame end: fps such that a frame is 2.5 steps of dt_max; the third step ends the frame exactly
    dtm = 0.5 * dx / c_wave
    e = A.Engine(oracle, A.default_config(fps=1.0 / (2.5 * dtm), end_frame=3, **base))
    e.set_particles(x)
    dts = []
    for _ in range(3):
        rc, info = e.step(1)
        dts.append(info.dt)
    assert dts[0] == dts[1] == dtm and abs(dts[2] - 0.5 * dtm) < 1e-12 * dtm and info.frame == 1 and info.frame_ended
    # supersonic: a warning, not an exit
    e = A.Engine(oracle, A.default_config(fps=1.0, **base))
    e.set_particles(x, v=np.tile([1.2 * c_wave, 0.0, 0.0], (3, 1)))
    rc, info = e.step(1)
    assert rc == 0 and info.status & A.ERR["WARN_SUPERSONIC"] and info.steps_done == 1
