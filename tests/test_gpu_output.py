"""Frame output of the reference's saveParticleData / saveGridData (save_data.cpp:7-259): the engine's
binary PLY files carry the reference's properties in its order, the particle state of the context, and
the stress measures p, q, Je computed from the oracle's Kirchhoff stress (HenckyPiola(F) F^T) of the same
deformation gradients. The files are read back with a PLY reader written from the format's description
(header lines, then rows of little-endian float64 values)."""
import ctypes as C

import numpy as np
import pytest

import mpm_abi as A
import parity_scenes as PS

pytestmark = pytest.mark.gpu


def read_ply(path):
    raw = open(path, "rb").read()
    end = raw.index(b"end_header\n") + len(b"end_header\n")
    lines = raw[:end].decode().splitlines()
    assert lines[0] == "ply" and lines[1] == "format binary_little_endian 1.0"
    rows, names = None, []
    for ln in lines[2:]:
        f = ln.split()
        if f[0] == "element":
            assert f[1] == "vertex"
            rows = int(f[2])
        elif f[0] == "property":
            assert f[1] == "double"
            names.append(f[2])
    data = np.frombuffer(raw[end:], dtype="<f8")
    assert data.size == rows * len(names)
    return names, data.reshape(rows, len(names))


def stress_measures_from_oracle(oracle, cfg, F):
    """p, q, Je per particle from the oracle's tau = HenckyPiola(F) F^T (save_data.cpp:16-38)."""
    D = cfg.dim
    out = np.zeros((F.shape[0], 3))
    tau = np.zeros(D * D)
    for k in range(F.shape[0]):
        Fc = np.ascontiguousarray(F[k].T).ravel()  # column-major
        rc = oracle.lib.oracle_kirchhoff(C.byref(cfg), Fc.ctypes.data_as(C.POINTER(C.c_double)),
                                         tau.ctypes.data_as(C.POINTER(C.c_double)))
        assert rc == 0
        t = tau.reshape(D, D).T
        p = -np.trace(t) / D
        dev = t + p * np.eye(D)
        c = 1.5 if cfg.use_mises_q else 0.5
        out[k] = (p, np.sqrt(c * np.sum(dev * dev)), np.linalg.det(F[k]))
    return out


@pytest.mark.parametrize("model,mibf", [("DPMui", 0), ("DPVisc", 1), ("NoPlasticity", 0)])
def test_particle_file(engine, oracle, tmp_path, model, mibf):
    print(f"particle PLY {model} mibf={mibf}", flush=True)
    cfg, plates, objs, f = PS.model_block(model, n=8, use_mibf=mibf, perzyna_visc=0.01)
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(f["x"], **{k: v for k, v in f.items() if k != "x"})
    rc, _ = e.step(5)
    assert rc == 0
    path = str(tmp_path / "particles_f0.ply")
    assert engine.lib.mpm_write_particles_ply(e.h, path.encode()) == 0
    names, rows = read_ply(path)
    expect = ["x", "y", "z", "vx", "vy", "vz", "p", "q", "Je"]
    if model != "NoPlasticity":
        expect += ["eps_pl_vol", "eps_pl_dev", "delta_gamma"]
        if model in ("DPMui", "MCCMui") or (model in ("DPVisc", "MCCVisc") and mibf):
            expect += ["muI"]
        if model in ("DPMui", "MCCMui"):
            expect += ["viscosity"]
    assert names == expect
    st = e.get_particles()
    col = {nm: rows[:, i] for i, nm in enumerate(names)}
    assert np.array_equal(np.stack([col["x"], col["y"], col["z"]], 1), st["x"])
    assert np.array_equal(np.stack([col["vx"], col["vy"], col["vz"]], 1), st["v"])
    for k in ("eps_pl_vol", "eps_pl_dev", "delta_gamma", "muI", "viscosity"):
        if k in col:
            assert np.array_equal(col[k], st[k]), k
    ref = stress_measures_from_oracle(oracle, cfg, st["F"])
    for i, k in enumerate(("p", "q", "Je")):
        scale = max(float(np.max(np.abs(ref[:, i]))), 1e-12)
        err = float(np.max(np.abs(col[k] - ref[:, i]))) / scale
        assert err < 1e-12, f"{k}: {err:.3e}"


def test_grid_file(engine, tmp_path):
    print("grid PLY", flush=True)
    cfg, plates, objs, f = PS.incline_mccmui(n=8)
    e = A.Engine(engine, cfg, plates, objs)
    e.set_particles(f["x"])
    rc, _ = e.step(3)
    assert rc == 0
    path = str(tmp_path / "grid_f0.ply")
    assert engine.lib.mpm_write_grid_ply(e.h, path.encode()) == 0
    names, rows = read_ply(path)
    assert names == ["x", "y", "z", "vx", "vy", "vz", "mass", "friction"]
    g = e.get_grid()
    X = np.stack(np.meshgrid(g["x"], g["y"], g["z"], indexing="ij"), -1).reshape(-1, 3)  # ind(i,j,k) order
    assert np.array_equal(rows[:, 0:3], X)
    assert np.array_equal(rows[:, 3:6], g["v"])
    assert np.array_equal(rows[:, 6], g["mass"])
    act = g["mass"] > 0
    assert np.array_equal(rows[act, 7], g["friction"][act])
