"""Helpers of the return-map tests: one call of the oracle's plasticity() on a single particle, and the
Hencky invariants of a deformation gradient (pressure p = -K tr ln V, deviatoric q = e_mu |dev ln V|)."""
import ctypes as C
import math

import numpy as np


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


SCALAR_INDEX = dict(eps_pl_dev=0, eps_pl_vol=1, eps_pl_vol_pradhana=2, delta_gamma=3, viscosity=4, muI=5)


def project(oracle, cfg, F, dt=1e-4, scalars=None):
    """plasticity() on one particle with trial deformation F; returns (F_out, scalars_out, plastic, err).
    scalars: eps_pl_dev, eps_pl_vol, eps_pl_vol_pradhana, delta_gamma, viscosity, muI (in and out)."""
    dim = cfg.dim
    f = np.asarray(F, dtype=float).flatten(order="F").copy()
    out = np.zeros(dim * dim)
    sc = np.zeros(6) if scalars is None else np.array(scalars, dtype=float)
    plastic = C.c_uint32()
    err = oracle.lib.oracle_plasticity_single(C.byref(cfg), dt, dptr(f), dptr(sc), dptr(out), C.byref(plastic))
    return out.reshape(dim, dim, order="F"), sc, plastic.value, err


def pq(F, cfg, K, e_mu):
    """(p, q, Hencky strain, its deviator) of F from its singular values."""
    s = np.linalg.svd(F, compute_uv=False)
    h = np.log(s)
    dev = h - h.mean()
    return -K * h.sum(), e_mu * np.linalg.norm(dev), h, dev


def consts(cfg):
    """Lame constants and the q / flow prefactors of simulation.cpp:94-145 for cfg."""
    E, nu, dim = cfg.E, cfg.nu, cfg.dim
    lam, mu = nu * E / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    K = lam + 2 * mu / dim
    qp = math.sqrt(1.5) if cfg.use_mises_q else 1 / math.sqrt(2)
    dp = math.sqrt(2 / 3) if cfg.use_mises_q else math.sqrt(2)
    return dict(lam=lam, mu=mu, K=K, e_mu=2 * qp * mu, f_mu=2 * qp / dp * mu, d=dp, rma=2 * qp * qp)


def trial(dim, vol, dev_mag, rng):
    """Rotated trial deformation with Hencky trace `vol` and deviatoric magnitude `dev_mag`."""
    d = np.array([1.0, -0.4, -0.6][:dim])
    d = d - d.mean()
    d = d / np.linalg.norm(d) * dev_mag
    h = d + vol / dim
    R = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
    Q = np.linalg.qr(rng.standard_normal((dim, dim)))[0]
    return R @ np.diag(np.exp(h)) @ Q, h
