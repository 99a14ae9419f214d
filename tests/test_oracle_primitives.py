"""Pins of the CPU oracle's primitives against closed forms, textbook routines and defining equations.

Each check is chosen so that a plausible restatement mistake (wrong branch bound, dropped term, transposed
factor, wrong prefactor) fails it. None of them re-types the oracle's own formula:
  * B-splines: exact values at knots, partition of unity, linear reproduction, the APIC second moment
    that fixes apicDinverse, and dN = d/du N by finite differences;
  * arange: the accumulation known answer;
  * SVD: LAPACK (numpy.linalg.svd) singular values, reconstruction, orthogonality;
  * Kirchhoff stress: eigen-decomposition route tau = 2 mu ln V + lambda tr(ln V) I, rotation covariance;
  * return maps: the yield surface / flow rule each projection must satisfy.
"""
import ctypes as C
import math

import numpy as np
import pytest

import mpm_abi as A
import returnmap as RM
from returnmap import consts, dptr, pq, project

rng = np.random.default_rng(20240611)


# ---------------------------------------------------------------- B-splines (src/tools.hpp:108-188)
@pytest.mark.parametrize("u,val", [(0.0, 0.75), (0.5, 0.5), (-0.5, 0.5), (1.0, 0.125), (-1.0, 0.125), (1.5, 0.0),
                                   (-1.5, 0.0), (2.0, 0.0)])
def test_quadratic_knots(oracle, u, val):
    assert oracle.lib.oracle_spline_N(2, u) == val


@pytest.mark.parametrize("u,val", [(0.0, 2.0 / 3.0), (1.0, 1.0 / 6.0), (-1.0, 1.0 / 6.0), (2.0, 0.0), (-2.0, 0.0)])
def test_cubic_knots(oracle, u, val):
    assert abs(oracle.lib.oracle_spline_N(3, u) - val) < 1e-16


@pytest.mark.parametrize("deg,u,val", [(2, 0.0, 0.0), (2, 0.5, -1.0), (2, -0.5, 1.0), (2, 1.0, -0.5), (2, 1.5, 0.0),
                                       (3, 0.0, 0.0), (3, 1.0, -0.5), (3, -1.0, 0.5), (3, 2.0, 0.0)])
def test_spline_derivative_knots(oracle, deg, u, val):
    assert abs(oracle.lib.oracle_spline_dN(deg, u) - val) < 1e-15


@pytest.mark.parametrize("deg", [2, 3])
def test_spline_moments(oracle, deg):
    """For the 4-node window used by every phase (base = floor(u)-1, p2g.cpp:21): sum N = 1,
    sum (x_i-x_p) N = 0, sum dN = 0, -sum u dN = 1 (sum x_i grad w^T = I), and the second moment
    sum u^2 N = 1/4 (quadratic) or 1/3 (cubic), whose inverse is apicDinverse*dx^2 (simulation.cpp:95,98)."""
    N, dN = oracle.lib.oracle_spline_N, oracle.lib.oracle_spline_dN
    second = {2: 0.25, 3: 1.0 / 3.0}[deg]
    for f in np.concatenate([rng.random(200), [0.0, 0.5, 0.25, 0.999999]]):
        us = [1.0 + f - a for a in range(4)]
        w = [N(deg, u) for u in us]
        d = [dN(deg, u) for u in us]
        assert abs(sum(w) - 1.0) < 1e-15
        assert abs(sum(u * wi for u, wi in zip(us, w))) < 1e-15
        assert abs(sum(d)) < 1e-14
        assert abs(-sum(u * di for u, di in zip(us, d)) - 1.0) < 1e-14
        assert abs(sum(u * u * wi for u, wi in zip(us, w)) - second) < 1e-15


@pytest.mark.parametrize("deg", [2, 3])
def test_spline_derivative_is_derivative(oracle, deg):
    N, dN = oracle.lib.oracle_spline_N, oracle.lib.oracle_spline_dN
    h = 1e-6
    for u in rng.uniform(-2.2, 2.2, 300):
        if min(abs(abs(u) - k) for k in (0.5, 1.0, 1.5, 2.0)) < 1e-4:
            continue
        fd = (N(deg, u + h) - N(deg, u - h)) / (2 * h)
        assert abs(fd - dN(deg, u)) < 1e-8


def test_arange_accumulates(oracle):
    """arange (tools.hpp:291-296) accumulates value += step: for [0,1) step 0.1 the tenth sum is
    0.9999999999999999 < 1, so it yields 11 values where numpy.arange (start + i*step) yields 10."""
    out = np.zeros(64)
    n = oracle.lib.oracle_arange(0.0, 1.0, 0.1, dptr(out), 64)
    assert n == 11 and len(np.arange(0.0, 1.0, 0.1)) == 10
    assert out[10] == 0.9999999999999999
    assert out[3] == 0.30000000000000004
    np.testing.assert_array_equal(out[:n], np.cumsum([0.0] + [0.1] * 10))


# ---------------------------------------------------------------- SVD (Eigen JacobiSVD restatement)
def svd_cases(dim):
    cases = [np.eye(dim), np.diag(np.arange(1, dim + 1, dtype=float)[::-1]), -np.eye(dim),
             np.diag([2.0, -3.0, 0.5][:dim]), np.eye(dim) + 1e-9 * rng.standard_normal((dim, dim)),
             np.zeros((dim, dim))]
    r = rng.standard_normal((dim, dim))
    cases.append(r)
    cases.append(r * 1e-200)
    cases.append(r * 1e200)
    u, _, vt = np.linalg.svd(rng.standard_normal((dim, dim)))
    cases.append(u @ np.diag([3.0, 3.0, 1.0][:dim]) @ vt)          # repeated singular value
    cases.append(u @ np.diag([2.0, 1e-12, 0.0][:dim]) @ vt)        # rank deficient
    cases.append(u @ np.diag([1.0 + 1e-4, 1.0 - 2e-4, 1.0][:dim]) @ vt)  # typical elastic F
    for _ in range(40):
        cases.append(rng.standard_normal((dim, dim)) * 10.0 ** rng.uniform(-3, 3))
    return cases


@pytest.mark.parametrize("dim", [2, 3])
def test_svd_against_lapack(oracle, dim):
    for Am in svd_cases(dim):
        a = np.asfortranarray(Am).flatten(order="F").copy()
        U, S, V = np.zeros(dim * dim), np.zeros(dim), np.zeros(dim * dim)
        oracle.lib.oracle_svd(dim, dptr(a), dptr(U), dptr(S), dptr(V))
        U = U.reshape(dim, dim, order="F")
        V = V.reshape(dim, dim, order="F")
        s_ref = np.linalg.svd(Am, compute_uv=False)
        scale = max(np.abs(Am).max(), 1e-300)
        assert np.all(S >= 0) and np.all(np.diff(S) <= 0), "not sorted descending / negative"
        assert np.max(np.abs(S - s_ref)) <= 1e-14 * max(s_ref[0], 1e-300) * 4
        assert np.max(np.abs(U @ np.diag(S) @ V.T - Am)) <= 1e-14 * scale * 8
        assert np.max(np.abs(U.T @ U - np.eye(dim))) < 1e-14
        assert np.max(np.abs(V.T @ V - np.eye(dim))) < 1e-14


# ---------------------------------------------------------------- Kirchhoff stress (simulation.hpp:227-238)
def tau_from_eigh(F, mu, lam):
    """tau = 2 mu ln V + lambda tr(ln V) I with V = sqrt(F F^T), via the symmetric eigensolver."""
    w, Q = np.linalg.eigh(F @ F.T)
    lnV = Q @ np.diag(0.5 * np.log(w)) @ Q.T
    return 2 * mu * lnV + lam * np.trace(lnV) * np.eye(F.shape[0])


def kirchhoff(oracle, cfg, F):
    dim = F.shape[0]
    f = F.flatten(order="F").copy()
    t = np.zeros(dim * dim)
    assert oracle.lib.oracle_kirchhoff(C.byref(cfg), dptr(f), dptr(t)) == 0
    return t.reshape(dim, dim, order="F")


@pytest.mark.parametrize("dim", [2, 3])
def test_hencky_kirchhoff(oracle, dim):
    cfg = A.default_config(dim=dim, dx=0.01, E=1e6, nu=0.3)
    E, nu = 1e6, 0.3
    lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    s = np.array([1.01, 0.98, 1.003][:dim])
    t = kirchhoff(oracle, cfg, np.diag(s))
    np.testing.assert_allclose(np.diag(t), 2 * mu * np.log(s) + lam * np.log(s).sum(), rtol=1e-13)
    for _ in range(30):
        F = np.eye(dim) + 0.05 * rng.standard_normal((dim, dim))
        t = kirchhoff(oracle, cfg, F)
        ref = tau_from_eigh(F, mu, lam)
        assert np.max(np.abs(t - ref)) <= 1e-9 * np.max(np.abs(ref))
        # objectivity: tau(R F Q) = R tau(F) R^T
        R = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
        Q = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
        t2 = kirchhoff(oracle, cfg, R @ F @ Q)
        assert np.max(np.abs(t2 - R @ t @ R.T)) <= 1e-9 * np.max(np.abs(t))


def test_neohookean_kirchhoff(oracle):
    cfg = A.default_config(dim=3, dx=0.01, E=1e6, nu=0.3, elastic_model="NeoHookean")
    lam, mu = 0.3e6 / (1.3 * 0.4), 1e6 / 2.6
    for _ in range(20):
        F = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
        ref = mu * (F @ F.T - np.eye(3)) + lam * np.log(np.linalg.det(F)) * np.eye(3)
        assert np.max(np.abs(kirchhoff(oracle, cfg, F) - ref)) <= 1e-9 * np.max(np.abs(ref))


# ---------------------------------------------------------------- return maps (plasticity.cpp)
def trial(dim, vol, dev_mag):
    return RM.trial(dim, vol, dev_mag, rng)


@pytest.mark.parametrize("dim", [2, 3])
def test_von_mises_projects_onto_q_max(oracle, dim):
    cfg = A.default_config(dim=dim, dx=0.01, plastic_model="VM", q_max=500.0)
    c = consts(cfg)
    F, h = trial(dim, -1e-4, 5e-3)
    Fn, sc, plastic, err = project(oracle, cfg, F)
    p, q, hn, dev = pq(Fn, cfg, c["K"], c["e_mu"])
    assert err == 0 and plastic == 1
    assert abs(q - 500.0) < 1e-9 * 500
    assert abs(hn.sum() - h.sum()) < 1e-13                      # volumetric strain untouched
    dg = np.linalg.norm(h - h.mean()) - 500.0 / c["e_mu"]
    assert abs(sc[0] - dg) < 1e-15 and abs(sc[3] - c["d"] * dg / 1e-4) < 1e-9
    Fe, sce, pe, _ = project(oracle, cfg, trial(dim, 0.0, 1e-4)[0])
    assert pe == 0 and sce[3] == 0                              # elastic: nothing changes


@pytest.mark.parametrize("dim", [2, 3])
def test_drucker_prager_cone_and_tip(oracle, dim):
    M, coh = math.tan(math.radians(30)), 100.0
    cfg = A.default_config(dim=dim, dx=0.01, plastic_model="DP", M=M, q_cohesion=coh)
    c = consts(cfg)
    F, h = trial(dim, -2e-3, 8e-3)                              # compression, large shear
    Fn, sc, plastic, err = project(oracle, cfg, F)
    p, q, hn, _ = pq(Fn, cfg, c["K"], c["e_mu"])
    assert plastic == 1 and abs(q - (M * p + coh)) < 1e-9 * q
    assert abs(p - (-c["K"] * h.sum())) < 1e-9 * abs(p)         # no volumetric change on the cone
    F, h = trial(dim, 5e-3, 1e-3)                               # tension: tip return
    Fn, sc, plastic, err = project(oracle, cfg, F)
    s = np.linalg.svd(Fn, compute_uv=False)
    p_tip = -coh / M
    np.testing.assert_allclose(np.log(s), -p_tip / (c["K"] * dim), rtol=0, atol=1e-15)
    assert abs(sc[1] - (p_tip - (-c["K"] * h.sum())) / c["K"]) < 1e-12 * abs(sc[1])


def test_perzyna_linear_relation(oracle):
    """VMVisc, exponent 1: q_new = q_y (1 + eta dgamma/dt) (closed form of plasticity.cpp:188-194)."""
    dt, eta, qy = 1e-4, 0.02, 400.0
    cfg = A.default_config(dim=3, dx=0.01, plastic_model="VMVisc", q_max=qy, q_min=qy, perzyna_visc=eta)
    c = consts(cfg)
    F, h = trial(3, 0.0, 6e-3)
    Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt)
    p, q, hn, _ = pq(Fn, cfg, c["K"], c["e_mu"])
    dg = sc[3] * dt
    assert abs(q - qy * (1 + eta * dg / dt)) < 1e-8 * q


@pytest.mark.parametrize("model", ["VMVisc", "DPVisc"])
def test_perzyna_newton_residual(oracle, model):
    """Exponent 2: the Newton solution satisfies (q_t - f dg) (dt/(eta dg + dt))^2 = q_y within the
    reference's stopping tolerance 1e-2 (plasticity.cpp:211-213)."""
    dt, eta = 1e-4, 0.02
    cfg = A.default_config(dim=3, dx=0.01, plastic_model=model, q_max=400.0, q_min=400.0, perzyna_visc=eta,
                           perzyna_exp=2.0, M=0.5, q_cohesion=0.0, use_pradhana=0)
    c = consts(cfg)
    F, h = trial(3, -1e-3, 6e-3)
    Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt)
    assert err == 0
    p_t, q_t, _, _ = pq(F, cfg, c["K"], c["e_mu"])
    dg = sc[3] * dt
    qy = 400.0 if model == "VMVisc" else 0.5 * p_t
    res = (q_t - c["f_mu"] * dg) * (dt / (eta * dg + dt)) ** 2 - qy
    assert abs(res) < 1e-2


def test_dp_mu_of_I_flow_rule(oracle):
    """DPMui (plasticity.cpp:376-401): q_new = q_y + (mu_2-mu_1) p gdot / (gdot + I_ref/(d sqrt(rho_s)) sqrt(p_s)),
    muI = mu_1 + (mu_2-mu_1)/(Q sqrt(p_s)/gdot + 1)."""
    dt = 1e-4
    cfg = A.default_config(dim=3, dx=0.01, plastic_model="DPMui", mu_1=0.38, mu_2=0.64, q_cohesion=0.0, use_pradhana=0)
    c = consts(cfg)
    Q = 0.279 / (1e-3 * math.sqrt(2500.0))
    F, h = trial(3, -1e-3, 8e-3)
    Fn, sc, plastic, err = project(oracle, cfg, F, dt=dt)
    p_t, q_t, _, _ = pq(F, cfg, c["K"], c["e_mu"])
    p, q, _, _ = pq(Fn, cfg, c["K"], c["e_mu"])
    gdot = sc[3]
    ps = p_t + 1e-5
    qy = 0.38 * p_t
    assert abs(q - (q_t - c["f_mu"] * dt * gdot)) < 1e-7 * q
    assert abs(q - (qy + 0.26 * p_t * gdot / (gdot + Q * math.sqrt(ps)))) < 1e-7 * q
    assert abs(sc[5] - (0.38 + 0.26 / (Q * math.sqrt(ps) / gdot + 1.0))) < 1e-12


@pytest.mark.parametrize("hard", ["NoHard", "ExpoExpl", "SinhExpl", "ExpoImpl", "SinhImpl"])
@pytest.mark.parametrize("dim", [2, 3])
def test_mcc_return_lands_on_ellipse(oracle, hard, dim):
    """Every MCC return map ends with q = M sqrt((p0-p)(beta p0 + p)) (mcc_rma_*.hpp), so the returned
    state satisfies the yield function y = M^2 (p-p0)(p+beta p0) + q^2 = 0, and keeps the deviatoric direction."""
    M, p0, beta = 1.2, 2000.0, 0.3
    cfg = A.default_config(dim=dim, dx=0.01, plastic_model="MCC", hardening_law=hard, M=M, p0=p0, beta=beta, xi=0.0)
    c = consts(cfg)
    for vol, dev in ((-1e-3, 6e-3), (-4e-3, 2e-3), (1e-3, 3e-3)):
        F, h = trial(dim, vol, dev)
        Fn, sc, plastic, err = project(oracle, cfg, F)
        assert err == 0
        p, q, hn, dvn = pq(Fn, cfg, c["K"], c["e_mu"])
        y = M * M * (p - p0) * (p + beta * p0) + q * q
        assert plastic == 1 and abs(y) < 1e-9 * p0 * p0, (vol, dev, y)
        dv = h - h.mean()
        if np.linalg.norm(dvn) > 1e-12:
            assert np.allclose(dvn / np.linalg.norm(dvn), dv / np.linalg.norm(dv), atol=1e-9)


# ---------------------------------------------------------------- boundary conditions (boundary_collision.cpp)
def bc_engine(oracle, dim, plates=(), objects=(), **kw):
    cfg = A.default_config(dim=dim, dx=0.01, particle_volume=1e-6, particle_mass=1e-3, **kw)
    return A.Engine(oracle, cfg, plates, objects)


def collide(oracle, e, X, v, dt=1e-3, fr=0.0):
    Xa, va = np.array(X, dtype=float), np.array(v, dtype=float)
    oracle.lib.oracle_boundary_single(e.h, dt, dptr(Xa), dptr(va), fr)
    return va


@pytest.mark.parametrize("dim", [2, 3])
def test_plate_noslip_takes_plate_velocity(oracle, dim):
    e = bc_engine(oracle, dim, plates=[A.make_plate(0.0, "bottom", "NoSlip", v_object=(0.3, 0.1, -0.2))])
    v = collide(oracle, e, [0.0, -0.001, 0.0][:dim], [1.0, -2.0, 0.5][:dim])
    np.testing.assert_array_equal(v, [0.3, 0.1, -0.2][:dim])


@pytest.mark.parametrize("dim", [2, 3])
def test_plate_coulomb_slip_and_stick(oracle, dim):
    mu = 0.4
    e = bc_engine(oracle, dim, plates=[A.make_plate(0.0, "bottom", "SlipFree", mu)])
    v = collide(oracle, e, [0.0, 0.0001, 0.0][:dim], [1.0, -0.5, 0.0][:dim])   # predicted inside, sliding
    assert v[1] == 0.0 and abs(v[0] - (1.0 - mu * 0.5)) < 1e-15
    v = collide(oracle, e, [0.0, 0.0001, 0.0][:dim], [0.1, -0.5, 0.0][:dim])   # sticks: |vt| < mu |vn|
    assert np.all(v == 0.0)
    v = collide(oracle, e, [0.0, -0.0001, 0.0][:dim], [1.0, 0.5, 0.0][:dim])   # separating: SlipFree leaves it
    np.testing.assert_array_equal(v, [1.0, 0.5, 0.0][:dim])
    v = collide(oracle, e, [0.0, 0.01, 0.0][:dim], [1.0, -0.5, 0.0][:dim])     # not colliding
    np.testing.assert_array_equal(v, [1.0, -0.5, 0.0][:dim])
    es = bc_engine(oracle, dim, plates=[A.make_plate(0.0, "bottom", "SlipStick", mu)])
    v = collide(oracle, es, [0.0, -0.01, 0.0][:dim], [1.0, 0.5, 0.0][:dim])    # SlipStick: normal always removed
    assert v[1] == 0.0 and abs(v[0] - (1.0 - mu * 0.5)) < 1e-15


def test_plate_extent_limits(oracle):
    e = bc_engine(oracle, 3, plates=[A.make_plate(0.0, "bottom", "NoSlip", pos_lower=-0.1, pos_upper=0.1)])
    np.testing.assert_array_equal(collide(oracle, e, [0.2, -0.001, 0.0], [1.0, -1.0, 0.0]), [1.0, -1.0, 0.0])
    np.testing.assert_array_equal(collide(oracle, e, [0.05, -0.001, 0.0], [1.0, -1.0, 0.0]), [0.0, 0.0, 0.0])


def test_plates_chain_at_a_corner(oracle):
    """Plates apply in order and each sees the previous result (boundary_collision.cpp:257)."""
    e = bc_engine(oracle, 3, plates=[A.make_plate(0.0, "bottom", "SlipFree", 0.0), A.make_plate(0.0, "left", "SlipFree", 0.0)])
    v = collide(oracle, e, [-0.001, -0.001, 0.0], [-1.0, -1.0, 0.7])
    np.testing.assert_array_equal(v, [0.0, 0.0, 0.7])


@pytest.mark.parametrize("dim", [2, 3])
def test_curve_slipfree_removes_normal_component(oracle, dim):
    e = bc_engine(oracle, dim, objects=[A.make_object("Curve", "SlipFree", 0.0)])
    X = np.array([0.3, 0.3 * 0.3 - 1e-4, 0.0][:dim])
    v = np.array([1.0, -2.0, 0.3][:dim])
    dt = 1e-6
    vn = collide(oracle, e, X, v, dt=dt)
    xs = X[0] + dt * v[0]                                       # normal at the predicted position
    n = np.array([-2 * xs, 1.0, 0.0][:dim])
    n /= np.linalg.norm(n)
    assert abs(vn @ n) < 1e-15
    np.testing.assert_allclose(vn, v - (v @ n) * n, rtol=0, atol=1e-15)


def test_object_noslip_and_ground_normal(oracle):
    e = bc_engine(oracle, 3, objects=[A.make_object("Ground", "SlipStick", 0.0, param=0.0)])
    vn = collide(oracle, e, [0.0, -0.01, 0.0], [1.0, 2.0, 3.0])
    np.testing.assert_array_equal(vn, [1.0, 0.0, 3.0])         # n = (0,1,0); z component kept
    e = bc_engine(oracle, 3, objects=[A.make_object("Ground", "NoSlip", 0.0, param=0.0)])
    np.testing.assert_array_equal(collide(oracle, e, [0.0, -0.01, 0.0], [1.0, 2.0, 3.0]), [0.0, 0.0, 0.0])
