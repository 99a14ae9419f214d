#!/usr/bin/env python3
"""Benchmark of the MPM timestep engine on one MI355X.

Workload (default): 3D granular column collapse, 0.4 m cube, dx = 0.005 m, 8 particles per cell on a
seeded jittered lattice (4,096,000 particles), quadratic B-spline, MUSL, APIC, Hencky + Drucker-Prager
(30 deg, E = 1e6 Pa, nu = 0.3, rho = 1550 kg/m^3), SlipFree floor with Coulomb mu = 0.4, gravity 9.81,
plain FP64 form. A step is one advanceStep() of all particles (+ the time bookkeeping of simulate()).

Prints one JSON line: throughput in particle-updates/s (particles x steps / device-synchronised wall time of
the timed steps), the roofline of the dominant kernel from per-kernel HIP-event durations over the timed
region, the parity of the same configuration against the CPU oracle at a small size, and the oracle timed
on the host cores on a bounded sample of the same workload (a reported baseline, not a target).

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "matter-tests_amd"))

import numpy as np  # noqa: E402

import mpm_abi as A  # noqa: E402
import scenes as S  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published peak
STEP_BYTES_PER_PARTICLE_UPDATE = 438  # 2*(24+24+72+72) + 40 + (2*(8+24) + 2*24)/8
METRIC = "particle-updates/s 3D DP-APIC FP32 vs FP64 @stated err; served p50/p99 step ms"


def algorithmic_bytes(kernel, cfg, n_particles, active_nodes, grid_nodes, blocks=None, tile_nodes=None):
    """Compulsory HBM bytes of one launch: particle state the kernel must read and write once, plus node
    data it must read/write once (the staged block tiles count as node data: each tile is written by the
    scatter and read once by the node pass). FP64 (8 B) throughout in the plain form."""
    D = cfg.dim
    f8 = 8
    apic = cfg.flip_ratio < 0
    flip = cfg.flip_ratio >= -1
    tau = 6 if D == 3 else 3
    mibf = bool(cfg.use_mibf)
    nv = 1 + D + (D if flip else 0) + (1 if mibf else 0)
    staged = (blocks or 0) * (tile_nodes or 0)
    if kernel == "k_p2g":
        part = D + D + (D * D if apic else 0) + tau + (1 if mibf else 0)
        return f8 * (part * n_particles + nv * staged) + 4 * 3 * n_particles  # + perm read/write, ids
    if kernel == "k_grid_update":
        out = 1 + D + (D if flip else 0) + (1 if mibf else 0)
        return f8 * (nv * staged + out * grid_nodes)
    if kernel == "k_g2p_musl":
        part_r = D + (D if flip else 0)
        part_w = D + (D * D if apic else 0) + D
        return f8 * ((part_r + part_w) * n_particles + (D + (D if flip else 0)) * active_nodes)
    if kernel == "k_musl_scatter":
        return f8 * ((D + D + (D * D if apic else 0)) * n_particles + D * staged)
    if kernel == "k_musl_update":
        return f8 * (D * staged + (1 + D) * grid_nodes)
    if kernel in ("k_defgrad", "k_g2p_usl"):
        part_r = D + D * D + 2 + (D if kernel == "k_g2p_usl" else 0)
        part_w = D * D + tau + 3 + (D + D + (D * D if apic else 0) if kernel == "k_g2p_usl" else 0)
        node = D + (D if (kernel == "k_g2p_usl" and flip) else 0)
        return f8 * ((part_r + part_w) * n_particles + node * active_nodes)
    if kernel == "k_bin_count":
        return f8 * D * n_particles + 4 * n_particles
    if kernel == "k_bin_scatter":
        return 4 * 2 * n_particles
    return 0


def parity_check(steps=10, n_axis=20):
    """Same configuration at n_axis^3 particles: engine vs single-threaded oracle after `steps` steps."""
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=n_axis, musl=True)
    out = {}
    ge = A.Engine(A.product_library(), cfg, plates, objs)
    oe = A.Engine(A.oracle_library(), cfg, plates, objs)
    ge.set_particles(x)
    oe.set_particles(x)
    ge.step(steps)
    oe.step(steps)
    gp, op = ge.get_particles(), oe.get_particles()
    for k in ("x", "v", "F"):
        out[k] = float(np.max(np.abs(gp[k] - op[k])) / max(np.max(np.abs(op[k])), 1e-300))
    gg, og = ge.get_grid(), oe.get_grid()
    act = og["mass"] > 0
    out["grid_mass"] = float(np.max(np.abs(gg["mass"][act] - og["mass"][act])) / np.max(og["mass"][act]))
    out["grid_v"] = float(np.max(np.abs(gg["v"][act] - og["v"][act])) / max(np.max(np.abs(og["v"][act])), 1e-300))
    out.update(particles=int(x.shape[0]), steps=steps, reference="CPU oracle, 1 thread")
    return out


def cpu_baseline(cfg, plates, objs, x, steps=3):
    """The oracle (FP64 OpenMP restatement of the reference loops) on all host cores (at most 16), one
    warm-up step then `steps` timed steps of the same 4,096,000-particle workload."""
    lib = A.oracle_library()
    threads = max(1, min(16, os.cpu_count() or 1))
    ccfg = type(cfg).from_buffer_copy(cfg)
    ccfg.n_threads = threads
    e = A.Engine(lib, ccfg, plates, objs)
    e.set_particles(x)
    info = A.StepInfo()
    lib.lib.oracle_time_steps(e.h, 1, info)
    secs = lib.lib.oracle_time_steps(e.h, steps, info)
    n = x.shape[0]
    return {"value": n * steps / secs, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps (after 1 warm-up step) of the same {n}-particle scene, {secs:.2f} s"}


COPY_BYTES = 256 * 1024 * 1024
TRAFFIC_FILE = os.path.join(REPO, "profiles", "c2_fp64_traffic.txt")


def copy_calibration():
    """Three device copies of a 256 MiB buffer, so that a counter run (rocprofv3 --pmc FETCH_SIZE or
    WRITE_SIZE) records what the counter reads for a copy of known size next to the engine's kernels."""
    import torch
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def summarize_pmc(fetch_csv, write_csv):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 counter-collection CSVs (FETCH_SIZE, WRITE_SIZE,
    both KiB per dispatch), each scaled by the factor that makes the 256 MiB copy read/write 256 MiB."""
    import csv
    import collections

    def per_kernel(path, counter):
        agg = collections.defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                agg[r["Kernel_Name"]].append((int(r["Grid_Size"]), float(r["Counter_Value"]) * 1024.0))
        return agg

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")

    def copy_factor(agg, which):
        big = [v for k, lst in agg.items() if "copyBuffer" in k for g, v in lst if v > COPY_BYTES / 4]
        if not big:
            raise SystemExit(f"no 256 MiB copy dispatch found in the {which} run (run it with --calibrate-copy)")
        return COPY_BYTES / (sum(big) / len(big))

    ff, fw = copy_factor(fetch, "FETCH_SIZE"), copy_factor(write, "WRITE_SIZE")
    lines = [f"# copy calibration: FETCH_SIZE x {ff:.3f}, WRITE_SIZE x {fw:.3f} (256 MiB device copy)",
             "# kernel  dispatches  hbm_read_bytes_per_dispatch  hbm_write_bytes_per_dispatch"]
    for k in sorted(fetch):
        if not k.startswith("void mpm::") and not k.startswith("mpm::"):
            continue
        name = k.split("mpm::", 1)[1].split("<")[0].split("(")[0]
        fv = [v for _, v in fetch[k]]
        wv = [v for _, v in write.get(k, [])]
        if not wv:
            continue
        lines.append(f"{name} {len(fv)} {ff * sum(fv) / len(fv):.0f} {fw * sum(wv) / len(wv):.0f}")
    return "\n".join(lines)


def read_traffic(kernel):
    """HBM bytes per launch (read + write) of `kernel` from the committed counter summary, or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        f = line.split()
        if len(f) == 4 and f[0] == kernel:
            return float(f[2]) + float(f[3])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n-axis", type=int, default=None, help="lattice points per axis (default 160: 4,096,000)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="run a 256 MiB device copy first (counter runs)")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print the per-kernel HBM traffic summary of two rocprofv3 --pmc runs and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported (multi-GPU sharding "
              "is a design in DESIGN.md, not a measured path)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        copy_calibration()
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=args.n_axis, musl=True)
    n = x.shape[0]
    eng = A.Engine(A.product_library(), cfg, plates, objs)
    eng.set_particles(x)
    if args.warmup:
        rc, _ = eng.step(args.warmup)
        if rc:
            raise A.MPMError(rc, "warm-up steps")
    eng.set_profiling(True)
    t0 = time.perf_counter()
    rc, info = eng.step(args.steps)  # returns after the device finished (stream synchronised)
    t1 = time.perf_counter()
    if rc or info.steps_done != args.steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], f"timed steps ({info.steps_done}/{args.steps} done)")
    kt = eng.kernel_times()
    eng.set_profiling(False)
    wall = t1 - t0
    value = n * args.steps / wall
    g = eng.get_grid()
    active = int(np.count_nonzero(g["mass"] > 0))
    grid_nodes = int(g["info"].grid_nodes)
    _, keys = eng.get_binning()
    blocks = int(np.unique(keys).size)  # non-empty blocks of the last step
    tile_nodes = (4 + (3 if cfg.spline_degree == 2 else 4) - 1) ** 3 if cfg.dim == 3 else (8 + (3 if cfg.spline_degree == 2 else 4) - 1) ** 2
    kernels = {}
    for name, (ms, launches) in kt.items():
        avg_ms = ms / max(launches, 1)
        b = algorithmic_bytes(name, cfg, n, active, grid_nodes, blocks, tile_nodes)
        kernels[name] = {"avg_ms": avg_ms, "launches": launches, "total_ms": ms,
                         "algorithmic_bytes": b, "GBps": (b / (avg_ms * 1e-3) / 1e9) if avg_ms > 0 and b else None}
    dom = max(kernels, key=lambda k: kernels[k]["total_ms"])
    d = kernels[dom]
    traffic = read_traffic(dom)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (d["GBps"] / HBM_PEAK_GBS) if d["GBps"] else None, "traffic": traffic,
                "traffic_source": os.path.relpath(TRAFFIC_FILE, REPO) if traffic is not None else None,
                "avg_ms": d["avg_ms"], "algorithmic_bytes_per_launch": d["algorithmic_bytes"]}
    # whole step against the compulsory traffic of one particle-update (x, v, F, B read and written, plastic
    # scalars, grid share): 438 B in 3D FP64 APIC Drucker-Prager at 8 particles per cell
    step_bytes = STEP_BYTES_PER_PARTICLE_UPDATE * n
    step_gbs = step_bytes / (wall / args.steps) / 1e9
    step_roofline = {"bytes_per_particle_update": STEP_BYTES_PER_PARTICLE_UPDATE, "achieved": step_gbs,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": step_gbs / HBM_PEAK_GBS}
    result = {
        "metric": METRIC, "value": value, "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded jittered lattice, SplitMix64 seed 42)",
        "config": {"workload": "3D granular column collapse, Drucker-Prager 30 deg, APIC, MUSL, plain FP64",
                   "particles": n, "dx": cfg.dx, "ppc": 8, "spline": "quadratic", "boundary": "SlipFree floor mu=0.4",
                   "grid_nodes_last_step": grid_nodes, "active_nodes_last_step": active, "blocks_last_step": blocks,
                   "precision": "fp64"},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "kernels": kernels,
    }
    if not args.no_parity:
        result["error_vs_fp64_oracle"] = parity_check()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(cfg, plates, objs, x)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
