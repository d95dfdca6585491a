"""Throughput benchmark of the MI355X FDTD step (one JSON line on stdout).

    python bench.py [--gpus 1] [--steps 200] [--warmup 20] [--workload enclosure_512]

Workload (default): the 512^3 bandwidth-bound solve — UniformGrid 512^3 at
1 mm, air, PML 16, rigid SDF box enclosure (300x200x200 cells, 20-cell walls)
with a cylindrical port (radius 12, length 60) and a radius-40 ADE absorber
sphere (two seeded Debye poles), 500 Hz pulse, 8 seeded probes. A "step" is one
full time step of that solve (velocity, ADE, rigid, pressure, PML, source,
probes). Inputs are resident in HBM before the timed region.

metric: Gcell-updates/s = cells * steps / time / 1e9, plus the HBM roofline of
the dominant kernel (the fused step) at 33 algorithmic bytes per cell-update
(p, vx, vy, vz read + written once in float32, 1 flag byte), measured with HIP
events on the solver's stream over the timed region. cpu_baseline: the
reference's own native kernels (compiled from source into oracle/_ref) or the
oracle restatement, timed on this host's cores on a bounded sample of the same
workload.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s + %HBM roofline (512^3); sweep solves/s with p50/p99 latency"
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 33


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str):
    from strata_fdtd_amd import workloads as W
    if name == "enclosure_512":
        return W.enclosure_512()
    if name == "plain_100":
        return W.plain_100()
    if name == "resonators_nu":
        return W.resonators_nu()
    if name.startswith("enclosure_"):
        return W.enclosure_512(n=int(name.split("_")[1]))
    raise SystemExit(f"unknown workload {name}")


def traffic_from_profiles(workload: str):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"fused_step_{workload}_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return float(d["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(solver, seconds_budget: float = 20.0):
    """Time the reference native kernels (or the oracle port) on this host."""
    import numpy as np
    from oracle import cpu
    from oracle.from_solver import problem_from_solver

    kind = "reference" if cpu.have_ref() else "port"
    impl = "ref" if kind == "reference" else "oracle"
    if not os.path.exists(cpu.ORACLE_SO):
        cpu.build(ref=False)
    n_short, n_long = 1, 4
    pb = problem_from_solver(solver, n_long + 1)
    cells = int(np.prod(solver.shape))
    t0 = time.perf_counter()
    cpu.run(pb, n_short, impl=impl)
    t1 = time.perf_counter()
    per_call_short = t1 - t0
    # scale the long run to the budget from the short one
    est_step = max(per_call_short / max(n_short, 1), 1e-3)
    n_long = int(max(3, min(40, seconds_budget / est_step)))
    pb = problem_from_solver(solver, n_long + 1)
    t2 = time.perf_counter()
    cpu.run(pb, n_long, impl=impl)
    t3 = time.perf_counter()
    per_step = ((t3 - t2) - per_call_short) / (n_long - n_short)
    if per_step <= 0:
        per_step = (t3 - t2) / n_long
    threads = cpu.num_threads(impl if impl == "oracle" else "ref")
    return {
        "value": round(cells / per_step / 1e9, 4),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{n_long} steps of the same {solver.shape[0]}x{solver.shape[1]}x{solver.shape[2]} "
                  f"workload (setup cost removed by differencing a {n_short}-step run), "
                  f"OMP_NUM_THREADS={threads}",
        "seconds_per_step": round(per_step, 4),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="enclosure_512")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--graph-chunk", type=int, default=50)
    ap.add_argument("--tiling", default=None, help="rows,planes override")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for rocprof byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); multi-GPU slab "
              "decomposition is a design note, not part of this build", file=sys.stderr)
        return 2

    try:
        os.sched_setaffinity  # noqa: B018
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count() or 1
    os.environ.setdefault("OMP_NUM_THREADS", str(ncpu))

    import torch
    import strata_fdtd_amd  # noqa: F401

    wl = build_workload(args.workload)
    log(f"workload {wl.name}: {wl.description}")
    t_setup = time.perf_counter()
    kw = {}
    if args.tiling:
        r, p = (int(v) for v in args.tiling.split(","))
        kw["tiling"] = (r, p)
    solver = wl.build(**kw)
    solver._prepare(args.warmup + args.steps)
    torch.cuda.synchronize()
    log(f"setup {time.perf_counter() - t_setup:.1f}s; warmup {args.warmup} steps")
    solver.step_n(args.warmup, graph_chunk=args.graph_chunk)
    torch.cuda.synchronize()
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    stream = solver._stream
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    ev0.record(stream)
    solver.step_n(args.steps, graph_chunk=args.graph_chunk)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    dev_ms = ev0.elapsed_time(ev1)
    health = solver.check_health()
    log(f"timed {args.steps} steps: device {dev_ms:.2f} ms, wall {wall * 1e3:.2f} ms, health {health}")

    cells = wl.cells
    ms_step = dev_ms / args.steps
    value = cells * args.steps / (dev_ms / 1e3) / 1e9
    achieved = BYTES_PER_CELL * cells / (ms_step / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(wl.name),
        "bytes_per_cell_update": BYTES_PER_CELL,
        "kernel": "k_fused_step (+ per-step ADE density and counter kernels inside the timed window)",
    }
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene, no external data)",
        "config": {"workload": wl.name, "shape": list(wl.shape), "cells": cells,
                   "description": wl.description, "wall_ms": round(wall * 1e3, 2),
                   "gcells_20B_convention_GBs": round(20 * cells / (ms_step / 1e3) / 1e9, 1)},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            line["cpu_baseline"] = cpu_baseline(solver)
        except Exception as exc:  # the GPU number stands; report why the baseline is missing
            line["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
