#!/usr/bin/env python3
"""Benchmark of the MPM timestep engine on one MI355X.

Headline workload: 3D granular column collapse, 0.4 m cube, dx = 0.005 m, 8 particles per cell on a seeded
jittered lattice (4,096,000 particles), quadratic B-spline, MUSL, APIC, Hencky + Drucker-Prager (30 deg,
E = 1e6 Pa, nu = 0.3, rho = 1550 kg/m^3), SlipFree floor with Coulomb mu = 0.4, gravity 9.81, in the
FP32/FP64 split form (`value`), with the plain FP64 form beside it (`fp64`). A step is one advanceStep() of
all particles (+ the time bookkeeping of simulate()).

Prints one JSON line: throughput in particle-updates/s (particles x steps / device-synchronised wall time of
the timed steps), the roofline of the dominant kernel from per-kernel HIP-event durations over the timed
region, the split form's position error against the plain form, the parity of the same configuration
against the CPU oracle at a small size, the oracle timed on the host cores on a bounded sample of the same
workload (a reported baseline, not a target), and the extra workloads: the 2D column collapse run whole on the
GPU and on the oracle (`c1`), the inclined flow (`c3`), 128 batched scenes (`batched`) and a served Poisson request
stream (`served`, with p50/p99 step and request latency). Each of c3, batched and served carries the error of
its own configuration against the FP64 oracle after 100 steps at a few thousand particles per scene
(`error_vs_fp64_oracle`).

    python bench.py [--gpus 1] [--steps 100] [--warmup 5] [--no-cpu-baseline] [--extra c1,c3,batched,served]
"""
import argparse
import ctypes as C
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
METRIC = "particle-updates/s 3D DP-APIC FP32 vs FP64 @stated err; served p50/p99 step ms"


def algorithmic_bytes(kernel, cfg, n_particles, active_nodes, grid_nodes, blocks=None, tile_nodes=None):
    """Compulsory HBM bytes of one launch: particle state the kernel must read and write once, plus node
    data it must read/write once (the staged block tiles count as node data: each tile is written by the
    scatter and read once by the node pass). Position, deformation gradient and plastic scalars are 8 B;
    velocity, affine matrix, Kirchhoff stress and node values are 8 B in the plain form, 4 B in the split form."""
    D = cfg.dim
    f8 = 8
    ft = 4 if cfg.precision == A.PRECISION["split"] else 8
    apic = cfg.flip_ratio < 0
    flip = cfg.flip_ratio >= -1
    tau = 6 if D == 3 else 3
    mibf = bool(cfg.use_mibf)
    nv = 1 + D + (D if flip else 0) + (1 if mibf else 0)
    staged = (blocks or 0) * (tile_nodes or 0)
    B = D * D if apic else 0
    if kernel == "k_p2g":
        part = f8 * D + ft * (D + B + tau) + (f8 if mibf else 0)
        return part * n_particles + ft * nv * staged + 4 * 3 * n_particles  # + permutation read/write, ids
    if kernel == "k_grid_update":
        out = 1 + D + (D if flip else 0) + (1 if mibf else 0)
        return ft * (nv * staged + out * grid_nodes)
    if kernel == "k_g2p":
        part_r = f8 * D + (ft * D if flip else 0)
        part_w = f8 * D + ft * (B + D)
        return (part_r + part_w) * n_particles + ft * (D + (D if flip else 0)) * active_nodes
    if kernel == "k_musl_scatter":
        return (f8 * D + ft * (D + B)) * n_particles + ft * D * staged
    if kernel == "k_g2p_musl":  # k_g2p, and the re-scatter's tiles without its second pass over the particles
        part_r = f8 * D + (ft * D if flip else 0)
        part_w = f8 * D + ft * (B + D)
        return (part_r + part_w) * n_particles + ft * (D + (D if flip else 0)) * active_nodes + ft * D * staged
    if kernel == "k_musl_update":
        return ft * (D * staged + (1 + D) * grid_nodes)
    if kernel == "k_defgrad":
        part_r = f8 * (D + D * D + 2)
        part_w = f8 * (D * D + 3) + ft * tau
        return (part_r + part_w) * n_particles + ft * D * active_nodes
    if kernel == "k_bin_count":
        return f8 * D * n_particles + 4 * n_particles
    if kernel == "k_bin_scatter":
        return 4 * 2 * n_particles
    return 0


def step_bytes_per_particle_update(cfg):
    """Compulsory bytes of one particle-update of the 3D APIC Drucker-Prager step at 8 particles per cell:
    x, v, F, B read and written once, the plastic scalars (read eps_pl_dev, eps_pl_vol; write those and
    delta_gamma), and the grid's share (mass and momentum written and read, velocity written and read, per
    8 particles). Plain form 438 B; split form (float v, B, node values) 335 B."""
    ft = 4 if cfg.precision == A.PRECISION["split"] else 8
    D = cfg.dim
    return 2 * (8 * D + ft * D + 8 * D * D + ft * D * D) + 40 + (2 * (ft + ft * D) + 2 * ft * D) / 8


def parity_check(prec, n_axis=20):
    """Same configuration at n_axis^3 particles: engine vs single-threaded oracle after 10 steps (plain
    form) or 100 steps (split form)."""
    steps = 10 if prec == "fp64" else 100
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=n_axis, musl=True)
    out = {}
    gcfg = type(cfg).from_buffer_copy(cfg)
    gcfg.precision = A.PRECISION[prec]
    ge = A.Engine(A.product_library(), gcfg, plates, objs)
    oe = A.Engine(A.oracle_library(), cfg, plates, objs)
    ge.set_particles(x)
    oe.set_particles(x)
    ge.step(steps)
    oe.step(steps)
    gp, op = ge.get_particles(), oe.get_particles()
    for k in ("x", "v", "F"):
        out[k] = float(np.max(np.abs(gp[k] - op[k])) / max(np.max(np.abs(op[k])), 1e-300))
    out["x_norm_rel"] = float(np.linalg.norm(gp["x"] - op["x"]) / np.linalg.norm(op["x"]))
    gg, og = ge.get_grid(), oe.get_grid()
    if gg["mass"].shape == og["mass"].shape:
        act = og["mass"] > 0
        out["grid_mass"] = float(np.max(np.abs(gg["mass"][act] - og["mass"][act])) / np.max(og["mass"][act]))
        mom_g, mom_o = gg["mass"][act, None] * gg["v"][act], og["mass"][act, None] * og["v"][act]
        out["grid_momentum"] = float(np.max(np.abs(mom_g - mom_o)) / max(np.max(np.abs(mom_o)), 1e-300))
    out.update(particles=int(x.shape[0]), steps=steps, reference="CPU oracle, 1 thread")
    return out


def field_errors(got, ref):
    """Norm-wise relative errors of position, velocity and F, and the largest position difference in m."""
    out = {f"{k}_norm_rel": float(np.linalg.norm(got[k] - ref[k]) / max(np.linalg.norm(ref[k]), 1e-300))
           for k in ("x", "v", "F")}
    out["x_max_abs_m"] = float(np.max(np.abs(got["x"] - ref["x"])))
    return out


def oracle_run(cfg, plates, objs, x, steps):
    """The same scene on the FP64 oracle (min(16, host cores) threads) for `steps` steps: its particle state."""
    ocfg = type(cfg).from_buffer_copy(cfg)
    ocfg.precision = A.PRECISION["fp64"]
    ocfg.n_threads = max(1, min(16, os.cpu_count() or 1))
    oe = A.Engine(A.oracle_library(), ocfg, plates, objs)
    oe.set_particles(x)
    rc, info = oe.step(steps)
    if rc or info.steps_done != steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "oracle steps")
    return oe.get_particles()


def worst_of(per_scene):
    """Summary over scenes: the largest of each error, and the scenes' own errors."""
    keys = per_scene[0][1].keys()
    return {"max_over_scenes": {k: max(e[k] for _, e in per_scene) for k in keys},
            "scenes": {name: {k: e[k] for k in ("x_norm_rel", "x_max_abs_m")} for name, e in per_scene}}


PARITY_STEPS = 100  # the split form's specification: positions within 1e-5 relative after 100 steps


def incline_parity():
    """The inclined-flow configuration (cubic, AFLIP, mu(I)-MCC, MIBF, bump) on a 40 x 6 x 12 lattice
    (2,880 particles), split form on the GPU against the FP64 oracle after 100 steps."""
    cfg, plates, objs, x = S.incline_flow(lattice=[40, 6, 12], ppc_axis=0.004 / 0.0015536, precision="split")
    e = A.Engine(A.product_library(), cfg, plates, objs)
    e.set_particles(x)
    rc, info = e.step(PARITY_STEPS)
    if rc or info.steps_done != PARITY_STEPS:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "c3 parity steps")
    out = field_errors(e.get_particles(), oracle_run(cfg, plates, objs, x, PARITY_STEPS))
    out.update(particles=int(x.shape[0]), steps=PARITY_STEPS, reference="CPU oracle, FP64",
               scene="same configuration, 40 x 6 x 12 lattice")
    return out


def batched_parity(count=8):
    """The batched workload's scene generator at 2k-6k particles (8 scenes, both materials and dimensions),
    one split-form batch on the GPU, each member against its FP64 oracle run after 100 steps."""
    scs = S.batch_collapses(count=count, n_min=2_000, n_max=6_000)
    L = A.product_library()
    b = A.Batch(L, chunk_steps=16, use_graph=True)
    engs = []
    for name, cfg, plates, objs, x in scs:
        e = A.Engine(L, cfg, plates, objs)
        e.set_particles(x)
        b.add(e, PARITY_STEPS)
        engs.append(e)
    rc, info = b.step(PARITY_STEPS)
    if rc or info.n_active != 0:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "batched parity steps")
    per = [(name, field_errors(e.get_particles(), oracle_run(cfg, plates, objs, x, PARITY_STEPS)))
           for (name, cfg, plates, objs, x), e in zip(scs, engs)]
    b.close()
    out = worst_of(per)
    out.update(particles=int(sum(s[4].shape[0] for s in scs)), steps=PARITY_STEPS, reference="CPU oracle, FP64",
               scene=f"batch of {count} scenes of the same generator at 2k-6k particles")
    return out


def served_parity(count=4):
    """The served workload's request generator at 2k-6k particles and 100 steps (4 requests, collapses and
    inclined flows), through the server, each result against its FP64 oracle run."""
    reqs = S.served_requests(count=count, rate=200.0, n_min=2_000, n_max=6_000, steps_min=PARITY_STEPS,
                             steps_max=PARITY_STEPS)
    srv = A.Server(A.product_library(), chunk_steps=8, use_graph=True, keep_results=True)
    ids = [srv.submit(cfg, plates, objs, x, steps, t) for _, cfg, plates, objs, x, steps, t in reqs]
    rc, st = srv.run(120.0)
    if rc or st.completed != count:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "served parity run")
    per = []
    for rid, (name, cfg, plates, objs, x, steps, _) in zip(ids, reqs):
        got, info, _ = srv.result(rid)
        if info.steps_done != steps:
            raise A.MPMError(A.ERR["NOT_READY"], f"served parity request {name}")
        per.append((name, field_errors(got, oracle_run(cfg, plates, objs, x, steps))))
    srv.close()
    out = worst_of(per)
    out.update(particles=int(sum(q[4].shape[0] for q in reqs)), steps=PARITY_STEPS, reference="CPU oracle, FP64",
               scene=f"{count} requests of the same generator at 2k-6k particles")
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
    t0, t1 = (C.c_double * 4)(), (C.c_double * 4)()
    lib.lib.oracle_phase_times(e.h, t0)
    secs = lib.lib.oracle_time_steps(e.h, steps, info)
    lib.lib.oracle_phase_times(e.h, t1)
    n = x.shape[0]
    phases = {k: (t1[i] - t0[i]) * 1e3 / steps for i, k in enumerate(("p2g", "euler", "g2p", "defgrad"))}
    return {"value": n * steps / secs, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps (after 1 warm-up step) of the same {n}-particle scene, {secs:.2f} s",
            "phase_ms_per_step": phases, "host_cpu": host_cpu(), "host_cores": os.cpu_count()}


def host_cpu():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


COPY_BYTES = 256 * 1024 * 1024
# HBM (TCC-to-fabric) requests by size: read bytes 32 RDREQ_32B + 64 RDREQ_64B + 128 RDREQ_128B, write bytes
# 32 (WRREQ - WRREQ_64B) + 64 WRREQ_64B. (FETCH_SIZE counts 128-byte reads through TCC_BUBBLE, which stays zero
# on gfx950, so it books them as 64 bytes.) No calibration: a 256 MiB device copy in the same run checks them.
READ_COUNTERS = {"TCC_EA0_RDREQ_32B_sum": 32, "TCC_EA0_RDREQ_64B_sum": 64, "TCC_EA0_RDREQ_128B_sum": 128}
WRITE_COUNTERS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum")


def traffic_file(workload, prec):
    return os.path.join(REPO, "profiles", f"{workload}_{prec}_traffic.txt")


def copy_calibration():
    """Three device copies of a 256 MiB buffer, so that a counter run records the request counts of a copy of
    known size next to the engine's kernels (the check of the byte formulas)."""
    import torch
    a = torch.ones(COPY_BYTES // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b


def summarize_pmc(read_csv, write_csv, form):
    """Per-kernel HBM bytes per dispatch from two rocprofv3 counter-collection CSVs (the read-request counters
    READ_COUNTERS, then the write-request counters WRITE_COUNTERS), for the kernels of one precision form
    ("split": float transfer type, "fp64": double; kernels without one count for both). Each kernel's value is
    the median over its dispatches, which is a steady-state launch: the run's first step scatters from the
    unsorted initial layout and reads several times as much."""
    import csv
    import collections
    import statistics

    def per_dispatch(path):
        out = collections.defaultdict(lambda: collections.defaultdict(float))
        names = {}
        for r in csv.DictReader(open(path)):
            d = int(r["Dispatch_Id"])
            names[d] = r["Kernel_Name"]
            out[d][r["Counter_Name"]] += float(r["Counter_Value"])
        return names, out

    rn, rv = per_dispatch(read_csv)
    wn, wv = per_dispatch(write_csv)
    rbytes = {d: sum(w * c.get(k, 0.0) for k, w in READ_COUNTERS.items()) for d, c in rv.items()}
    wbytes = {d: 32.0 * (c.get(WRITE_COUNTERS[0], 0.0) - c.get(WRITE_COUNTERS[1], 0.0)) + 64.0 * c.get(WRITE_COUNTERS[1], 0.0)
              for d, c in wv.items()}
    rc = [rbytes[d] for d, k in rn.items() if "copyBuffer" in k and rbytes[d] > COPY_BYTES / 4]
    wc = [wbytes[d] for d, k in wn.items() if "copyBuffer" in k and wbytes[d] > COPY_BYTES / 4]
    if not rc or not wc:
        raise SystemExit("no 256 MiB copy dispatch in the counter runs (run them with --calibrate-copy)")
    tf = {"split": "float", "fp64": "double"}[form]

    def grouped(names, values):
        g = collections.defaultdict(list)
        for d, k in names.items():
            if not (k.startswith("void mpm::") or k.startswith("mpm::")):
                continue
            args = k.split("<", 1)[1] if "<" in k else ""
            if ("float" in args or "double" in args) and tf not in args:
                continue
            g[k.split("mpm::", 1)[1].split("<")[0].split("(")[0]].append(values[d])
        return g

    gr, gw = grouped(rn, rbytes), grouped(wn, wbytes)
    lines = [f"# bytes per dispatch from request counts (32/64/128-byte reads, 32/64-byte writes); check: the 256 MiB "
             f"copy reads {statistics.median(rc) / COPY_BYTES:.4f} x and writes {statistics.median(wc) / COPY_BYTES:.4f} x "
             "its size",
             "# per kernel: median over its dispatches (the first step scatters from the unsorted initial layout)",
             "# kernel  dispatches  hbm_read_bytes_per_dispatch  hbm_write_bytes_per_dispatch"]
    for k in sorted(gr):
        if k in gw:
            lines.append(f"{k} {len(gr[k])} {statistics.median(gr[k]):.0f} {statistics.median(gw[k]):.0f}")
    return "\n".join(lines)


def read_traffic(kernel, workload, prec):
    """HBM bytes per launch (read + write) of `kernel` from the committed counter summary, or None."""
    path = traffic_file(workload, prec)
    if not os.path.exists(path):
        return None
    for line in open(path):
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
    ap.add_argument("--precision", choices=["both", "split", "fp64"], default="both",
                    help="forms to run; the first one run (split when both) gives the headline value")
    ap.add_argument("--only", choices=["c3", "batched"], default=None,
                    help="run only this extra workload and print its object (counter and trace runs)")
    ap.add_argument("--extra", default="c1,c3,batched,served",
                    help="comma-separated extra workloads reported beside the headline: c1 (the 2D column collapse, "
                         "GPU and oracle, 200 steps), c3 (inclined flow), batched "
                         "(128 scenes), served (Poisson request stream); empty for none")
    ap.add_argument("--summarize-pmc", nargs=3, metavar=("READ_CSV", "WRITE_CSV", "FORM"),
                    help="print the per-kernel HBM traffic summary of two rocprofv3 --pmc runs (request counters) "
                         "for one form (split or fp64) and exit")
    args = ap.parse_args()
    if args.summarize_pmc:
        print(summarize_pmc(*args.summarize_pmc))
        return
    if args.gpus != 1:
        print(f"bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported (sharding the particles "
              "over several GPUs is designed but not built or measured)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        copy_calibration()
    if args.only:
        print(json.dumps({"c3": incline_run, "batched": batched_run}[args.only](args)), flush=True)
        return
    forms = ["split", "fp64"] if args.precision == "both" else [args.precision]
    runs = {}
    for prec in forms:
        runs[prec] = timed_run(prec, args)
    head = runs[forms[0]]
    cfg, n = head["cfg"], head["n"]
    result = {
        "metric": METRIC, "value": head["value"], "unit": "particle-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32/f64 split" if forms[0] == "split" else "f64",
        "data": "synthetic (seeded jittered lattice, SplitMix64 seed 42)",
        "config": {"workload": "3D granular column collapse, Drucker-Prager 30 deg, APIC, MUSL",
                   "particles": n, "dx": cfg.dx, "ppc": 8, "spline": "quadratic", "boundary": "SlipFree floor mu=0.4",
                   "grid_nodes_last_step": head["grid_nodes"], "active_nodes_last_step": head["active"],
                   "blocks_last_step": head["blocks"], "precision": forms[0]},
        "roofline": head["roofline"],
        "step_roofline": head["step_roofline"],
        "pcie_inclusive": head["pcie_inclusive"],
        "kernels": head["kernels"],
    }
    if "split" in runs and "fp64" in runs:
        xs, xd = runs["split"]["x"], runs["fp64"]["x"]
        result["fp64"] = {k: runs["fp64"][k] for k in ("value", "ms_per_step", "roofline", "step_roofline", "pcie_inclusive",
                                                       "kernels")}
        result["split_speedup_vs_fp64"] = runs["split"]["value"] / runs["fp64"]["value"]
        result["split_error_vs_fp64_gpu"] = {
            "position_rel": float(np.linalg.norm(xs - xd) / np.linalg.norm(xd)),
            "position_max_abs_m": float(np.max(np.abs(xs - xd))),
            "after_steps": args.warmup + args.steps, "particles": n,
            "note": "same seeded input, both forms on the GPU; the plain form is held to the CPU oracle at 1e-10"}
    extra = [w for w in args.extra.split(",") if w]
    if "c1" in extra:
        result["c1"] = column2d_run(args)
    if "c3" in extra:
        result["c3"] = incline_run(args)
        if not args.no_parity:
            result["c3"]["error_vs_fp64_oracle"] = incline_parity()
    if "batched" in extra:
        result["batched"] = batched_run(args)
        if not args.no_parity:
            result["batched"]["error_vs_fp64_oracle"] = batched_parity()
    if "served" in extra:
        result["served"] = served_run(args)
        if not args.no_parity:
            result["served"]["error_vs_fp64_oracle"] = served_parity()
    if not args.no_parity:
        result["error_vs_fp64_oracle"] = {prec: parity_check(prec) for prec in forms}
    if not args.no_cpu_baseline:
        plates, objs, x = head["scene"]
        cfg64 = type(cfg).from_buffer_copy(cfg)
        cfg64.precision = A.PRECISION["fp64"]
        result["cpu_baseline"] = cpu_baseline(cfg64, plates, objs, x)
    print(json.dumps(result), flush=True)


def column2d_run(args, steps=200):
    """The 2D plane-strain column collapse that the reference runs on the CPU: 0.5 x 0.5 m, dx = 0.01 m, 4 particles
    per cell (10,000), quadratic, USL, APIC, Drucker-Prager 30 deg, NoSlip floor, 200 steps from rest, plain FP64.
    The whole run on the GPU (one mpm_step call, after a 32-step run of the same scene on another context has
    loaded the 2D kernels' code objects) and on the oracle (min(16, host cores) threads), and the difference of
    the final states."""
    cfg, plates, objs, x = S.column_2d()
    n = x.shape[0]
    out = {"workload": "2D plane-strain column collapse, DP 30 deg, APIC, USL, NoSlip floor, 200 steps",
           "particles": n, "precision": "fp64", "steps": steps, "unit": "particle-updates/s"}
    # the process's first launches of the 2D kernels load their code objects: done once, on a context of its own
    warm = A.Engine(A.product_library(), cfg, plates, objs)
    warm.set_particles(x)
    warm.step(32)
    del warm
    eng = A.Engine(A.product_library(), cfg, plates, objs)
    eng.set_particles(x)
    t0 = time.perf_counter()
    rc, info = eng.step(steps)
    t1 = time.perf_counter()
    if rc or info.steps_done != steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "2D column steps")
    out.update(value=n * steps / (t1 - t0), ms_per_step=1e3 * (t1 - t0) / steps)
    g = eng.get_particles()
    lib = A.oracle_library()
    threads = max(1, min(16, os.cpu_count() or 1))
    ocfg = type(cfg).from_buffer_copy(cfg)
    ocfg.n_threads = threads
    oe = A.Engine(lib, ocfg, plates, objs)
    oe.set_particles(x)
    oinfo = A.StepInfo()
    secs = lib.lib.oracle_time_steps(oe.h, steps, oinfo)
    o = oe.get_particles()
    out["oracle"] = {"value": n * steps / secs, "cores": threads, "seconds": secs}
    out["final_state_vs_oracle"] = {
        "x_rel": float(np.linalg.norm(g["x"] - o["x"]) / np.linalg.norm(o["x"])),
        "x_max_abs_m": float(np.max(np.abs(g["x"] - o["x"]))),
        "time_equal": bool(info.time == oinfo.time), "plastic_count": [int(info.plastic_count), int(oinfo.plastic_count)]}
    return out


def incline_run(args):
    """Flow down a 35 deg incline with a bump, 7,954,688 particles (644 x 64 x 193 lattice at 1.5536 mm, about 17 per
    cell at dx = 0.004 m), cubic B-spline, AFLIP, mu(I)-MCC, MIBF, SlipStick floor and bump, split form
    (float transfers, binary64 return map): W warm-up steps, then K timed steps."""
    cfg, plates, objs, x = S.incline_flow(lattice=[644, 64, 193], ppc_axis=0.004 / 0.0015536, precision="split")
    n = x.shape[0]
    eng = A.Engine(A.product_library(), cfg, plates, objs)
    eng.set_particles(x)
    del x
    # per-kernel times from the warm-up steps after the first (which sizes the buffers); the timed steps run
    # without events
    first = 1 if args.warmup > 1 else 0
    for k, steps in enumerate((first, args.warmup - first)):
        if k == 1:
            eng.set_profiling(True)  # clears the sums
        if steps:
            rc, _ = eng.step(steps)
            if rc:
                raise A.MPMError(rc, "c3 warm-up")
    warm_kt = eng.kernel_times() or {}
    eng.set_profiling(False)
    t0 = time.perf_counter()
    rc, info = eng.step(args.steps)
    t1 = time.perf_counter()
    if rc or info.steps_done != args.steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "c3 timed steps")
    wall = t1 - t0
    kernels, roofline, (grid_nodes, active, blocks) = kernel_table(eng, cfg, n, warm_kt, {}, "c3", "split")
    return {"workload": "3D flow down a 35 deg incline with a rectangular bump, mu(I)-MCC, cubic, AFLIP, MIBF",
            "particles": n, "precision": "split", "value": n * args.steps / wall, "unit": "particle-updates/s",
            "ms_per_step": 1e3 * wall / args.steps, "steps": args.steps, "plastic_count_last_step": info.plastic_count,
            "grid_nodes_last_step": grid_nodes, "active_nodes_last_step": active, "blocks_last_step": blocks,
            "roofline": roofline, "kernels": kernels}


def batched_run(args, count=128):
    """128 independent 2D/3D collapses (scenes.batch_collapses: 5k-500k particles log-uniform, aspect 0.5-2,
    friction angle 20-40 deg, Drucker-Prager or Perzyna-von Mises, quadratic APIC USL, split form) in one batch:
    W warm-up steps, then K steps of every scene timed around one synchronised mpm_batch_step; the same scenes
    advanced one after another with mpm_step over the same K steps give the unbatched rate."""
    L = A.product_library()
    scs = S.batch_collapses(count=count)
    engs = []
    for name, cfg, plates, objs, x in scs:
        e = A.Engine(L, cfg, plates, objs)
        e.set_particles(x)
        engs.append(e)
    del scs
    total = sum(e.n for e in engs)
    b = A.Batch(L, chunk_steps=16, use_graph=True)
    tune_steps = 180
    for e in engs:
        b.add(e, tune_steps + args.warmup + args.steps)
    chunk, tuned_rate = b.autotune(tune_steps)  # ordinary steps of the scenes, the same number for each
    if args.warmup:
        rc, _ = b.step(args.warmup)
        if rc:
            raise A.MPMError(rc, "batched warm-up")
    t0 = time.perf_counter()
    rc, info = b.step(args.steps)
    t1 = time.perf_counter()
    if rc or info.particle_steps != total * args.steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], "batched timed steps")
    wall = t1 - t0
    b.close()
    t2 = time.perf_counter()
    for e in engs:
        rc, _ = e.step(args.steps)
        if rc:
            raise A.MPMError(rc, "unbatched steps")
    t3 = time.perf_counter()
    dims = [e.dim for e in engs]
    return {"workload": f"{count} independent 2D/3D collapse scenes, DP / Perzyna-VM, quadratic APIC USL",
            "scenes": count, "scenes_3d": int(sum(d == 3 for d in dims)), "particles": total, "precision": "split",
            "value": total * args.steps / wall, "unit": "particle-updates/s", "ms_per_step": 1e3 * wall / args.steps,
            "steps": args.steps, "launch_classes": info.n_classes, "chunk_steps": chunk, "hip_graph": True,
            "autotune": {"candidates": [4, 8, 16, 32], "chosen": chunk, "rate_at_chosen": tuned_rate},
            "unbatched_value": total * args.steps / (t3 - t2), "batching_speedup": (t3 - t2) / wall}


def served_run(args, count=48):
    """Open-loop Poisson stream (50 requests/s) of 48 requests, each a 3D collapse or an inclined flow with
    10k-2M particles (log-uniform) asking for 50-500 steps, split form, served by mpm_server (footprint
    admission, continuous batching in chunks of 8 steps, HIP graphs)."""
    reqs = S.served_requests(count=count)
    srv = A.Server(A.product_library(), chunk_steps=8, use_graph=True)
    for name, cfg, plates, objs, x, steps, t in reqs:
        srv.submit(cfg, plates, objs, x, steps, t)
    offered = sum(q[4].shape[0] * q[5] for q in reqs)
    last_arrival = reqs[-1][6]
    del reqs
    rc, st = srv.run(600.0)
    srv.close()
    if rc or st.completed != count:
        raise A.MPMError(rc or A.ERR["NOT_READY"], f"served run ({st.completed}/{count} completed)")
    d = st.as_dict()
    d.update(workload=f"Poisson stream of {count} collapse / inclined-flow requests at 50 req/s, 10k-2M particles, "
                      "50-500 steps each", precision="split", unit="particle-updates/s", value=st.throughput,
             offered_particle_updates=offered, arrival_window_s=last_arrival)
    return d


def kernel_table(eng, cfg, n, warm_kt, live_kt, workload, prec):
    """Per-kernel average durations (the timed steps' where measured there, else the warm-up's), algorithmic
    bytes and achieved GB/s on the engine's last-step grid, and the roofline of the kernel that took the most
    time in the timed steps (with its counter-measured traffic from profiles/<workload>_<prec>_traffic.txt)."""
    kt = dict(warm_kt)
    kt.update(live_kt)
    g = eng.get_grid()
    active = int(np.count_nonzero(g["mass"] > 0))
    grid_nodes = int(g["info"].grid_nodes)
    _, keys = eng.get_binning()
    blocks = int(np.unique(keys).size)  # non-empty blocks of the last step
    del g, keys
    w = 3 if cfg.spline_degree == 2 else 4
    tile_nodes = (4 + w - 1) ** 3 if cfg.dim == 3 else (8 + w - 1) ** 2
    kernels = {}
    for name, (ms, launches) in kt.items():
        avg_ms = ms / max(launches, 1)
        b = algorithmic_bytes(name, cfg, n, active, grid_nodes, blocks, tile_nodes)
        kernels[name] = {"avg_ms": avg_ms, "launches": launches, "total_ms": ms,
                         "algorithmic_bytes": b, "GBps": (b / (avg_ms * 1e-3) / 1e9) if avg_ms > 0 and b else None,
                         "timed_over": "timed steps" if name in live_kt else "warm-up steps"}
    src = live_kt or warm_kt
    dom = max(src, key=lambda k: src[k][0])
    d = kernels[dom]
    traffic = read_traffic(dom, workload, prec)
    roofline = {"bound": "hbm", "kernel": dom, "achieved": d["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (d["GBps"] / HBM_PEAK_GBS) if d["GBps"] else None, "traffic": traffic,
                "traffic_source": os.path.relpath(traffic_file(workload, prec), REPO) if traffic is not None else None,
                "avg_ms": d["avg_ms"], "algorithmic_bytes_per_launch": d["algorithmic_bytes"]}
    return kernels, roofline, (grid_nodes, active, blocks)


def timed_run(prec, args):
    """W warm-up steps (the last W-1 with every kernel timed by HIP events), then K steps timed by the host clock
    around one synchronised mpm_step call, with HIP events around the dominant kernel over the K steps."""
    cfg, plates, objs, x = S.dp_collapse(n_per_axis=args.n_axis, musl=True, precision=prec)
    n = x.shape[0]
    eng = A.Engine(A.product_library(), cfg, plates, objs)
    tu = time.perf_counter()
    eng.set_particles(x)  # host arrays -> device layout (synchronous)
    upload_s = time.perf_counter() - tu
    # Warm-up steps with every kernel timed: the per-kernel table and the choice of the dominant kernel. The timed
    # steps then record HIP events around the dominant kernel only (two per step instead of two per launch, which
    # would add about 5 % to the step), and its average duration over them gives the roofline.
    warm_kt = {}
    if args.warmup:
        rc, _ = eng.step(1)  # the first step (grid set-up, first launches) is not representative
        if rc:
            raise A.MPMError(rc, "warm-up steps")
        eng.set_profiling(True)
        if args.warmup > 1:
            rc, _ = eng.step(args.warmup - 1)
            if rc:
                raise A.MPMError(rc, "warm-up steps")
        warm_kt = eng.kernel_times() or {}
    if warm_kt:
        eng.set_profiling_kernel(max(warm_kt, key=lambda k: warm_kt[k][0]))
    else:
        eng.set_profiling(True)  # no profiled warm-up: every kernel is timed over the timed steps
    t0 = time.perf_counter()
    rc, info = eng.step(args.steps)  # returns after the device finished (stream synchronised)
    t1 = time.perf_counter()
    if rc or info.steps_done != args.steps:
        raise A.MPMError(rc or A.ERR["NOT_READY"], f"timed steps ({info.steps_done}/{args.steps} done)")
    live_kt = eng.kernel_times()
    eng.set_profiling(False)
    wall = t1 - t0
    kernels, roofline, (grid_nodes, active, blocks) = kernel_table(eng, cfg, n, warm_kt, live_kt, "c2", prec)
    bpu = step_bytes_per_particle_update(cfg)
    step_gbs = bpu * n / (wall / args.steps) / 1e9
    step_roofline = {"bytes_per_particle_update": bpu, "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_gbs / HBM_PEAK_GBS}
    td = time.perf_counter()
    state = eng.get_particles()  # device layout -> host arrays in the caller's order (synchronous)
    download_s = time.perf_counter() - td
    # The same K steps with the host hand-over of the boundary counted: particle state in before them and out
    # after them. Reported beside the value, never as it.
    pcie = {"value": n * args.steps / (wall + upload_s + download_s), "upload_s": upload_s, "download_s": download_s,
            "note": "particles x K / (K steps + mpm_set_particles + mpm_get_particles of the full state)"}
    out = dict(cfg=cfg, n=n, value=n * args.steps / wall, ms_per_step=wall * 1e3 / args.steps, kernels=kernels,
               roofline=roofline, step_roofline=step_roofline, grid_nodes=grid_nodes, active=active, blocks=blocks,
               x=state["x"], scene=(plates, objs, x), pcie_inclusive=pcie)
    del eng
    return out


if __name__ == "__main__":
    main()
