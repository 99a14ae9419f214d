"""Size-independent invariants of one MPM step, usable on the oracle and on the engine.

APIC affine preservation (Jiang et al., "The affine particle-in-cell method", SIGGRAPH 2015): with particle
velocities v_p = A x_p + b and B_p = A / apicDinverse, stress-free F = I, no gravity and no boundaries,
P2G gives every active node v_i = A x_i + b; G2P returns pic = A x_p + b and B = A D = A / apicDinverse
(D = dx^2/4 quadratic, dx^2/3 cubic: the spline's second moment); grad v = sum v_i (x) grad w_i = A, so
F becomes I + dt A. These follow from the B-spline moment identities alone, not from the code under test.
"""
import numpy as np

import mpm_abi as A
import scenes as S


def affine_scene(dim, spline, musl, n=12, seed=5):
    dx = 0.01
    h = dx / 2
    vol = h ** dim
    cfg = A.default_config(dim=dim, spline_degree=spline, dx=dx, particle_volume=vol, particle_mass=1000 * vol,
                           rho=1000.0, plastic_model="NoPlasticity", flip_ratio=-2.0, use_musl=int(musl),
                           gravity=[0, 0, 0], end_frame=1, fps=1.0, E=1e6, nu=0.3)
    counts = [n] * dim
    x = S.jittered_lattice(counts, h, [0.013] * dim, seed=seed)
    rng = np.random.default_rng(seed)
    Am = 0.5 * rng.standard_normal((dim, dim))
    b = rng.standard_normal(dim)
    aD = (4.0 if spline == 2 else 3.0) / (dx * dx)
    v = x @ Am.T + b
    B = np.tile(Am / aD, (x.shape[0], 1, 1))
    return cfg, x, v, B, Am, b, aD


def check_affine_step(lib, dim, spline, musl, tol=1e-11):
    cfg, x, v, B, Am, b, aD = affine_scene(dim, spline, musl)
    e = A.Engine(lib, cfg)
    e.set_particles(x, v=v, Bmat=B)
    rc, info = e.step(1)
    assert rc == 0
    dt = info.dt
    g = e.get_grid()
    act = g["mass"] > 0
    axes = [g["x"], g["y"]] + ([g["z"]] if dim == 3 else [])
    idx = np.array(np.unravel_index(np.arange(g["mass"].size), [len(a) for a in axes])).T
    Xi = np.stack([axes[d][idx[:, d]] for d in range(dim)], axis=1)
    vref = Xi @ Am.T + b
    if not musl:  # with MUSL the grid holds the re-scattered velocity, which is the same affine field
        pass
    scale = np.abs(vref[act]).max()
    assert np.abs(g["v"][act] - vref[act]).max() <= tol * scale, "grid velocity is not the affine field"
    p = e.get_particles()
    vp = x @ Am.T + b
    assert np.abs(p["v"] - vp).max() <= tol * np.abs(vp).max(), "pic != A x_p + b"
    assert np.abs(p["Bmat"] - Am / aD).max() <= tol * np.abs(Am / aD).max(), "B != A D"
    Fref = np.eye(dim) + dt * Am
    assert np.abs(p["F"] - Fref).max() <= tol, "F != I + dt A"
    assert np.abs(p["x"] - (x + dt * vp)).max() <= tol * np.abs(x).max(), "x != x + dt pic"
    return True
