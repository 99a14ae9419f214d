"""The engine's SVD (mpm_math.hpp: svd, two-sided Jacobi with Eigen's JacobiSVD conventions), built for the
host from the same header, at the plain form's tolerance (2 eps) and the split form's (1e-12), against
LAPACK's dgesvd (numpy): singular values, reconstruction A = U diag(s) W^T, orthogonality, descending order.
Matrices: rotations times stretches near the identity (the deformation gradients of a granular flow),
general random matrices, reflections (det < 0), a diagonal with a repeated value and a nearly rank-one
matrix."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "device_svd_check.cpp")


# The engine compiles with -ffp-contract=off and lets products and sums fuse only inside the functions marked
# `#pragma clang fp contract(on)` (the Jacobi rotations, 3x3 products, Hencky stress). The host build uses
# the same compiler family and flags, with FMA instructions enabled, so that those functions round here as
# they do on the device (g++ ignores the clang pragma and would check an unfused variant).
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.fixture(scope="module")
def prog(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("svd") / "device_svd_check")
    assert os.path.exists(CLANG), f"{CLANG} is missing"
    subprocess.run([CLANG, "-O2", "-std=c++17", "-mfma", "-ffp-contract=off", SRC, "-o", exe], check=True)
    return exe


def matrices(dim, rng):
    out = []
    for k in range(300):
        q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        amp = (0.001, 0.05, 0.5)[k % 3]
        out.append(q @ (np.eye(dim) + amp * rng.standard_normal((dim, dim))))
    out += [rng.standard_normal((dim, dim)) for _ in range(100)]
    refl = np.eye(dim)
    refl[0, 0] = -1.0
    out += [refl @ m for m in out[:50]]
    out.append(np.diag([1.2] + [0.9] * (dim - 1)))
    v = rng.standard_normal(dim)
    out.append(np.outer(v, v) + 1e-9 * rng.standard_normal((dim, dim)))
    return out


def run(prog, mats):
    text = "".join(f"{m.shape[0]} " + " ".join(format(float(x), ".17g") for x in m.flatten(order="F")) + "\n" for m in mats)
    res = subprocess.run([prog], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    out = []
    for m, line in zip(mats, res):
        d = m.shape[0]
        v = np.array([float(x) for x in line.split()])
        per = d + 2 * d * d
        routines = []
        for r in range(2):
            c = v[r * per:(r + 1) * per]
            s = c[:d]
            U = c[d:d + d * d].reshape(d, d, order="F")
            W = c[d + d * d:].reshape(d, d, order="F")
            routines.append((s, U, W))
        out.append(routines)
    return out


@pytest.mark.parametrize("dim", [3, 2])
def test_engine_svd_against_lapack(prog, dim):
    rng = np.random.default_rng(11 + dim)
    mats = matrices(dim, rng)
    for m, routines in zip(mats, run(prog, mats)):
        s_ref = np.linalg.svd(m, compute_uv=False)
        scale = s_ref[0]
        for name, (s, U, W) in zip(("tol 2eps", "tol 1e-12"), routines):
            assert np.all(np.diff(s) <= 0) and s[-1] >= 0, name
            assert np.max(np.abs(s - s_ref)) <= 1e-13 * scale, (name, s, s_ref)
            rec = U @ np.diag(s) @ W.T
            assert np.max(np.abs(rec - m)) <= 1e-10 * scale, (name, np.max(np.abs(rec - m)))
            assert np.max(np.abs(U.T @ U - np.eye(dim))) <= 1e-10, name
            assert np.max(np.abs(W.T @ W - np.eye(dim))) <= 1e-10, name
