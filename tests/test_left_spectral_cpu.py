"""The deformation update's left spectral decomposition (mpm_math.hpp: left_spectral, recompose_left), built for
the host from the same header, against LAPACK's dgesvd (numpy). The update takes it instead of the two-sided
SVD when it is accepted: b = F F^T = U diag(lam) U^T gives the Hencky strain ln sqrt(lam) and the stress on U,
and the plastic update F' = U diag(exp h) W^T is formed as U diag(exp(h) / sqrt lam) U^T F.

Checked: lam = s^2 (s from dgesvd) to 1e-14 of lam_max, U orthogonal and U diag(lam) U^T = F F^T, F' equal to
U_svd diag(s exp(shift)) W_svd^T and the Hencky stress of h = ln s + shift equal to the one on U_svd, with an
isotropic shift (-0.3 times the deviatoric part of ln s, the shape of a Drucker-Prager projection), so that
equal singular values get equal shifts and the result does not depend on the basis of a repeated one, nor on the
order in which the decomposition leaves the eigenpairs (it does not sort them). And the
acceptance rule: singular values within a factor 100 of each other and none below 1e-4 (the two-sided SVD
takes the rest); every deformation gradient near a rotation (granular flow) is accepted.
"""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "left_spectral_check.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"  # same compiler family and fused rounding as the device build


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("lsp") / "left_spectral_check")
    assert os.path.exists(CLANG), f"{CLANG} is missing"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-mfma", "-ffp-contract=off", SRC, "-o", exe], check=True)
    return exe


def matrices(dim, rng):
    near, other = [], []
    for k in range(300):
        q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        amp = (0.001, 0.05, 0.2)[k % 3]
        near.append(q @ (np.eye(dim) + amp * rng.standard_normal((dim, dim))))
    refl = np.eye(dim)
    refl[0, 0] = -1.0
    near += [refl @ m for m in near[:50]]
    near.append(np.eye(dim))
    near.append(np.diag([1.2] + [0.9] * (dim - 1)))
    near.append(np.diag([0.95] * dim))
    other += [rng.standard_normal((dim, dim)) for _ in range(200)]
    v = rng.standard_normal(dim)
    other.append(np.outer(v, v) + 1e-9 * rng.standard_normal((dim, dim)))
    other.append(np.diag([1.0] + [2e-3] * (dim - 1)))
    other.append(1e-5 * np.eye(dim))
    return near, other


def shift_of(s):
    h = np.log(s)
    return -0.3 * (h - h.mean())


def run(prog, mats):
    lines = []
    for m in mats:
        lines.append(f"{m.shape[0]} " + " ".join(format(float(x), ".17g") for x in m.flatten(order="F")) + "\n")
    res = subprocess.run([prog], input="".join(lines), capture_output=True, text=True, check=True).stdout.split("\n")
    out = []
    for m, line in zip(mats, res):
        d = m.shape[0]
        v = [float(x) for x in line.split()]
        ok = v[0] == 1.0
        lam = np.array(v[1:1 + d])
        U = np.array(v[1 + d:1 + d + d * d]).reshape(d, d, order="F")
        Fn = np.array(v[1 + d + d * d:1 + d + 2 * d * d]).reshape(d, d, order="F")
        tau = np.array(v[1 + d + 2 * d * d:])
        out.append((ok, lam, U, Fn, tau))
    return out


def packed(T, dim):
    return np.array([T[0, 0], T[1, 1], T[2, 2], T[0, 1], T[0, 2], T[1, 2]]) if dim == 3 else \
        np.array([T[0, 0], T[1, 1], T[0, 1]])


@pytest.mark.parametrize("dim", [3, 2])
def test_left_spectral_against_lapack(prog, dim):
    rng = np.random.default_rng(21 + dim)
    near, other = matrices(dim, rng)
    mats = near + other
    accepted = 0
    for idx, (m, (ok, lam, U, Fn, tau)) in enumerate(zip(mats, run(prog, mats))):
        Ur, s, Wt = np.linalg.svd(m)
        ratio = s[-1] / s[0]
        if idx < len(near):
            assert ok, f"near-rotation matrix {idx} rejected (s = {s})"
        elif ratio < 0.99e-2 or s[-1] < 0.99e-4:
            assert not ok, f"matrix {idx} accepted with s = {s}"
        if not ok:
            continue
        accepted += 1
        assert np.max(np.abs(np.sort(lam)[::-1] - s ** 2)) <= 1e-14 * s[0] ** 2, (idx, lam, s ** 2)
        assert np.max(np.abs(U.T @ U - np.eye(dim))) <= 1e-13
        assert np.max(np.abs(U @ np.diag(lam) @ U.T - m @ m.T)) <= 1e-14 * s[0] ** 2
        sh = shift_of(s)
        F_ref = Ur @ np.diag(s * np.exp(sh)) @ Wt
        assert np.max(np.abs(Fn - F_ref)) <= 1e-12 * np.max(np.abs(F_ref)), (idx, np.max(np.abs(Fn - F_ref)))
        h = np.log(s) + sh
        d = 2.0 * h + 1.5 * h.sum()
        tau_ref = packed(Ur @ np.diag(d) @ Ur.T, dim)
        assert np.max(np.abs(tau - tau_ref)) <= 1e-12 * max(1.0, np.max(np.abs(tau_ref))), (idx, tau, tau_ref)
    assert accepted >= len(near)
