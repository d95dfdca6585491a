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

import numpy as np

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


# solves up to this many cells (33 B each) fit the 256 MiB Infinity Cache and
# are swept with ordinary stores; larger ones with non-temporal stores
CACHE_RESIDENT_CELLS = (256 << 20) // BYTES_PER_CELL


def kernel_name(kname: str, rows: int, stack: int, uniform: bool, cells: int) -> str:
    """Template instance of the sweep that runs for this solve, as rocprofv3
    names it."""
    nu = "false" if uniform else "true"
    if stack != 1:
        return f"{kname}<{rows},{nu},{stack}>"
    ix = "uint32_t" if cells < (1 << 30) else "long"
    nts = "false" if cells <= CACHE_RESIDENT_CELLS else "true"
    return f"{kname}<{rows},{nu},{ix},{nts}>"


def traffic_from_profiles(workload: str, kernel: str):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, copy-kernel calibrated), and a note.
    The bytes are returned only when the summary was taken on this workload
    and on the kernel instance this run launches; otherwise None and why."""
    path = os.path.join(ROOT, "profiles", f"fused_step_{workload}_traffic.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        nbytes = float(d["bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None, f"no PMC summary for {workload} under profiles/"
    if d.get("workload") != workload or d.get("kernel") != kernel:
        return None, (f"PMC summary is of {d.get('kernel')} on {d.get('workload')}, "
                      f"this run launches {kernel} on {workload}")
    return nbytes, f"rocprofv3 FETCH_SIZE + WRITE_SIZE, {os.path.relpath(path, ROOT)}"


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
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), model)
    except OSError:
        pass
    return {
        "value": round(cells / per_step / 1e9, 4),
        "unit": "Gcell-updates/s",
        "cores": threads,
        "kind": kind,
        "sample": f"{n_long} steps of the same {solver.shape[0]}x{solver.shape[1]}x{solver.shape[2]} "
                  f"workload (setup cost removed by differencing a {n_short}-step run), "
                  f"OMP_NUM_THREADS={threads}, {model}, g++ -O3 -march=x86-64-v2 -ffp-contract=off -fopenmp",
        "seconds_per_step": round(per_step, 4),
    }


GOLDEN = os.path.join(ROOT, "tests", "golden")
FULL_LENGTH = ("plain_100", "resonators_nu", "enclosure_512")


def _rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def full_length_parity(name: str) -> dict:
    """Run a workload over its stated step count on the GPU and compare with
    its committed full-length fixture (tests/golden/full_<name>.npz, data
    written once by the CPU oracle in float32 and float64): probe traces bit
    for bit, SHA-256 of the final fields' bytes, and the error against the
    float64 run (probe traces, final pressure on a 16-cell sub-lattice)."""
    import hashlib
    from strata_fdtd_amd import workloads as W
    g = np.load(os.path.join(GOLDEN, f"full_{name}.npz"), allow_pickle=False)
    wl = {"plain_100": W.plain_100, "enclosure_512": W.enclosure_512, "resonators_nu": W.resonators_nu}[name]()
    steps = int(g["steps"])
    if steps != wl.steps:
        raise ValueError(f"{name}: fixture has {steps} steps, the workload {wl.steps}")
    s = wl.build()
    s.step_n(steps)
    tr = np.stack(list(s.get_probe_data().values()), axis=1)
    out = {"steps": steps, "trace_bitexact": bool(np.array_equal(tr, g["trace_f32"])),
           "trace_rel_l2_vs_oracle_f32": _rel_l2(tr, g["trace_f32"]),
           "trace_rel_l2_vs_f64": _rel_l2(tr, g["trace_f64"])}
    same = True
    for k in ("p", "vx", "vy", "vz"):
        a = np.ascontiguousarray(getattr(s, k))
        same &= hashlib.sha256(a.tobytes()).hexdigest() == str(g[f"sha256_{k}"])
        if k == "p":
            sub = int(g["sub"])
            out["p_sub_rel_l2_vs_f64"] = _rel_l2(a[::sub, ::sub, ::sub], g["p_sub_f64"])
    out["fields_bitexact"] = bool(same)
    return out


def parity_bench():
    """Full-length parity of the three solve workloads: each runs its stated
    step count and is compared with committed oracle fixtures
    (tests/golden/full_*.npz: probe traces and field hashes of the float32
    oracle, probe traces and a pressure sub-lattice of the float64 oracle).
    Reports bit-exactness and the error against float64."""
    out = {}
    for name in FULL_LENGTH:
        if not os.path.exists(os.path.join(GOLDEN, f"full_{name}.npz")):
            out[name] = {"error": "no fixture"}
            continue
        t = time.perf_counter()
        out[name] = full_length_parity(name)
        out[name]["seconds"] = round(time.perf_counter() - t, 1)
        log(f"parity {name}: {out[name]}")
    return out


def serving_bench(n_requests: int = 64, rate: float = 40.0, chunk: int = 100, steps=(2000, 8000),
                  batched_only: bool = False):
    """Continuous batching vs one-at-a-time on the seeded Poisson design sweep
    (step counts drawn from `steps`; fewer only for short profiling runs).
    batched_only: run the default batched server alone (profiling its kernel)."""
    import torch
    from strata_fdtd_amd.serving import SweepServer, serve_one_at_a_time, sweep_requests
    reqs = sweep_requests(n=n_requests, rate=rate, steps=tuple(steps))
    # warm both paths (code objects, allocator) on a tiny stream first
    warm = sweep_requests(n=3, rate=1e6, seed=99, steps=(20, 40), dims=(64,))
    SweepServer(chunk=10).serve(warm, realtime=False)
    serve_one_at_a_time(warm, realtime=False)
    torch.cuda.synchronize()
    log(f"serving: {len(reqs)} requests, Poisson rate {rate}/s, "
        f"{sum(r.cells * r.steps for r in reqs) / 1e9:.1f} Gcell-updates in total")
    # default server: cache-resident active set (6M cells), least work first
    # with an age credit
    srv = SweepServer(chunk=chunk)
    batched = srv.serve(reqs, realtime=True)
    srv.close()
    log(f"serving batched: {batched.summary()}")
    torch.cuda.empty_cache()
    if batched_only:
        return {"requests": len(reqs), "arrival_rate_per_s": rate, "chunk_steps": chunk,
                "request_steps": list(steps),
                "gcell_updates_total": round(sum(r.cells * r.steps for r in reqs) / 1e9, 2),
                "batched": batched.summary(), "note": "default batched server only (profiling run)"}
    # least work first without the age credit
    srv = SweepServer(chunk=chunk, policy="sjf")
    sjf = srv.serve(reqs, realtime=True)
    srv.close()
    log(f"serving batched, least work first: {sjf.summary()}")
    torch.cuda.empty_cache()
    # every arrival admitted at once, in arrival order
    srv = SweepServer(chunk=chunk, max_cells=None, policy="fifo")
    fifo = srv.serve(reqs, realtime=True)
    srv.close()
    log(f"serving batched, all admitted (fifo): {fifo.summary()}")
    torch.cuda.empty_cache()
    single = serve_one_at_a_time(reqs, realtime=True)
    log(f"serving one-at-a-time: {single.summary()}")
    # spot-check equality of results in the mix and alone
    same = all(
        all(np.array_equal(a.traces[k], b.traces[k]) for k in a.traces)
        for a, b in zip(sorted(batched.results, key=lambda r: r.rid), sorted(single.results, key=lambda r: r.rid)))
    return {
        "requests": len(reqs), "arrival_rate_per_s": rate, "chunk_steps": chunk, "request_steps": list(steps),
        "gcell_updates_total": round(sum(r.cells * r.steps for r in reqs) / 1e9, 2),
        "batched": batched.summary(),
        "batched_policy": "<= 6M active cells (cache-resident), least work first with an age credit",
        "batched_least_work_first": sjf.summary(),
        "batched_all_admitted_fifo": fifo.summary(),
        "one_at_a_time": single.summary(),
        "speedup_solves_per_s": round(batched.solves_per_s / single.solves_per_s, 3),
        "results_identical_alone_and_in_mix": bool(same),
    }


def adjoint_bench(n: int = 192, steps: int = 2000, fd_voxels: int = 0):
    """Gradient cost (forward + checkpointed adjoint) vs a forward solve."""
    import torch
    from strata_fdtd_amd import FDTDSolver, PML, UniformGrid
    from strata_fdtd_amd.adjoint import design_problem_192, design_ratio, forward_traces, probe_traces
    pb, rho, K, target = design_problem_192(n=n, steps=steps)

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return time.perf_counter() - t, out

    def median_of(k, fn):
        runs = [timed(fn) for _ in range(k)]
        runs.sort(key=lambda t: t[0])
        return runs[k // 2]

    forward_traces(pb, rho, K)  # warm-up
    t_fwd_cell, _ = timed(lambda: forward_traces(pb, rho, K, impl="cell"))
    plain = FDTDSolver(grid=UniformGrid(pb.shape, pb.resolution), c=pb.c, rho=pb.rho, backend="mi355x")
    plain.add_boundary(PML(depth=pb.pml_depth))
    plain.add_source(pb.source)
    for q, p in enumerate(pb.probes):
        plain.add_probe(f"p{q}", p)
    plain._prepare(steps)
    t_plain, _ = timed(lambda: plain.step_n(steps))

    def grad(history):
        r = rho.clone().requires_grad_(True)
        k = K.clone().requires_grad_(True)
        tr = probe_traces(r, k, pb, history=history)
        loss = ((tr - target) ** 2).sum()
        loss.backward()
        return loss.item(), r.grad, k.grad

    grad("all")  # warm-up (history allocation, code objects)
    torch.cuda.empty_cache()
    # forward and gradient timed in alternation, so that both see the same
    # allocator state and buffer placement; medians of 5 each
    fwd_runs, grad_runs = [], []
    for _ in range(5):
        fwd_runs.append(timed(lambda: forward_traces(pb, rho, K)))
        grad_runs.append(timed(lambda: grad("all")))
    t_fwd = sorted(t for t, _ in fwd_runs)[2]
    grad_runs.sort(key=lambda t: t[0])
    t_grad, (loss, g_r, g_k) = grad_runs[2]
    torch.cuda.empty_cache()
    t_grad_ck, (loss_ck, g_r_ck, g_k_ck) = median_of(3, lambda: grad("checkpoint"))
    same = bool(torch.equal(g_r, g_r_ck) and torch.equal(g_k, g_k_ck))
    from strata_fdtd_amd.adjoint import history_bytes
    out = {
        "shape": list(pb.shape), "steps": steps, "design_box": list(pb.box_shape), "checkpoint": pb.checkpoint,
        "timing": "medians of 5 alternating forward / gradient runs; median of 3 checkpointed gradients",
        "forward_s": round(t_fwd, 4), "forward_cell_kernels_s": round(t_fwd_cell, 4),
        "forward_plain_fused_s": round(t_plain, 4),
        "gradient_s": round(t_grad, 4), "gradient_history": "all steps in HBM",
        "history_gb": round(history_bytes(pb) / 1e9, 2),
        "gradient_over_forward": round(t_grad / t_fwd, 3),
        "gradient_checkpoint_recompute_s": round(t_grad_ck, 4),
        "gradient_checkpoint_over_forward": round(t_grad_ck / t_fwd, 3),
        "checkpoint_and_history_gradients_identical": same,
        "gradient_over_plain_forward": round(t_grad / t_plain, 3),
        "loss": loss, "grad_norm_rho": float(g_r.norm()), "grad_norm_K": float(g_k.norm()),
    }
    if fd_voxels:
        rng = np.random.default_rng(17)
        rows = []

        def loss64(r_, k_):
            tr, _ = forward_traces(pb, r_, k_, dtype=torch.float64)
            return float(((tr - target.double()) ** 2).sum())

        for v in range(fd_voxels):
            which = "rho" if v % 2 == 0 else "K"
            base, grad_t = (rho, g_r) if which == "rho" else (K, g_k)
            idx = tuple(int(rng.integers(0, s)) for s in pb.box_shape)
            ref = pb.rho if which == "rho" else pb.K0
            b64 = base.double()
            plus, minus = b64.clone(), b64.clone()
            plus[idx] *= 1 + 1e-4
            minus[idx] *= 1 - 1e-4
            # the step that reaches the kernels: difference of the float32 ratios
            dr = float(design_ratio(plus, ref)[idx].double() - design_ratio(minus, ref)[idx].double())
            lp = loss64(plus if which == "rho" else rho.double(), plus if which == "K" else K.double())
            lm = loss64(minus if which == "rho" else rho.double(), minus if which == "K" else K.double())
            fd = (lp - lm) / (dr * ref)
            g = float(grad_t[idx])
            # per voxel, and against the whole gradient's L2 norm (the float32
            # adjoint's rounding floor is a fraction of the norm, so voxels with a
            # tiny gradient show a large per-voxel relative error)
            rows.append({"param": which, "voxel": list(idx), "adjoint": g, "fd": fd,
                         "rel_err": abs(g - fd) / max(abs(fd), 1e-300),
                         "err_over_grad_norm": abs(g - fd) / float(grad_t.norm())})
        errs = np.array([r["rel_err"] for r in rows])
        nerrs = np.array([r["err_over_grad_norm"] for r in rows])
        out["finite_differences"] = {"voxels": fd_voxels, "h_rel": 1e-4, "fd_precision": "f64",
                                     "adjoint_precision": "f32",
                                     "median_rel_err": float(np.median(errs)), "max_rel_err": float(errs.max()),
                                     "max_err_over_grad_norm": float(nerrs.max()), "rows": rows}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="enclosure_512")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--graph-chunk", type=int, default=0,
                    help="steps per captured hipGraph replay (0: kernels enqueued one by one, the faster)")
    ap.add_argument("--tiling", default=None, help="rows,planes[,stack] override")
    ap.add_argument("--no-serving", action="store_true", help="skip the batched-serving sweep")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint cost measurement")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the full-length parity runs against the oracle fixtures")
    ap.add_argument("--adjoint-fd", type=int, default=0, help="finite-difference check on N voxels")
    ap.add_argument("--serving-steps", default="2000,8000",
                    help="step-count range of the served requests (2000,8000; fewer for short profiling runs)")
    ap.add_argument("--serving-batched-only", action="store_true",
                    help="serve with the default batched server only (no comparison legs; profiling)")
    ap.add_argument("--adjoint-steps", type=int, default=2000,
                    help="steps of the adjoint workload (2000; fewer for short profiling runs)")
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
        kw["tiling"] = tuple(int(v) for v in args.tiling.split(","))
    solver = wl.build(**kw)
    solver._prepare(args.warmup + 2 * args.steps)
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
    # the dominant kernel alone: HIP events around each fused sweep launch, on
    # the solver's stream, over a second window of the same length
    sweep_ms, eager_ms = solver.time_sweeps(args.steps)
    # the practical ceiling of the same bytes: a stencil-free streaming pass
    # over the solve's own buffers, with the sweep's store kind
    stream_ms = solver.stream_ceiling(20)
    rows, planes, stack = solver.tiling
    # the same pass in the sweep's own strip order (lane = k, rows x planes)
    march_ms = solver.stream_ceiling(20, order="march") if wl.cells < (1 << 30) and stack == 1 else None
    torch.cuda.synchronize()
    kname = "k_fused_step" if stack == 1 else "k_fused_stack"
    log(f"sweep kernel {kname} rows={rows} planes={planes} stack={stack}: "
        f"{sweep_ms / args.steps * 1e3:.1f} us/launch; eager step {eager_ms / args.steps * 1e3:.1f} us")

    cells = wl.cells
    ms_step = dev_ms / args.steps
    value = cells * args.steps / (dev_ms / 1e3) / 1e9
    ms_sweep = sweep_ms / args.steps
    achieved = BYTES_PER_CELL * cells / (ms_sweep / 1e3) / 1e9
    kinst = kernel_name(kname, rows, stack, solver.grid.is_uniform, cells)
    traffic, traffic_note = traffic_from_profiles(wl.name, kinst)
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "traffic_source": traffic_note,
        "kernel": kinst,
        "algorithmic_bytes_per_launch": BYTES_PER_CELL * cells,
        "bytes_per_cell_update": BYTES_PER_CELL,
        "avg_launch_us": round(ms_sweep * 1e3, 1),
        "whole_step_us": round(ms_step * 1e3, 1),
        "whole_step_frac": round(BYTES_PER_CELL * cells / (ms_step / 1e3) / 1e9 / HBM_PEAK_GBS, 4),
        "stream_ceiling": {
            "achieved": round(BYTES_PER_CELL * cells / (stream_ms / 1e3) / 1e9, 1),
            "avg_pass_us": round(stream_ms * 1e3, 1),
            "sweep_frac_of_ceiling": round(stream_ms / ms_sweep, 4),
            "what": "stencil-free pass reading p, vx, vy, vz and the flag byte and writing the four fields, "
                    "on the same buffers with the same store kind, 16-byte accesses in memory order",
        },
    }
    if march_ms is not None:
        roofline["stream_ceiling_march"] = {
            "achieved": round(BYTES_PER_CELL * cells / (march_ms / 1e3) / 1e9, 1),
            "avg_pass_us": round(march_ms * 1e3, 1),
            "sweep_frac_of_ceiling": round(march_ms / ms_sweep, 4),
            "what": "the same stencil-free pass in the sweep's own order (one lane per k cell, "
                    f"{rows}-row x {planes}-plane strips, same block remap): what the sweep adds to it is "
                    "its stencil and halo traffic",
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
    del solver
    torch.cuda.empty_cache()
    if not args.no_parity:
        try:
            line["parity"] = parity_bench()
        except Exception as exc:
            line["parity"] = {"error": f"{type(exc).__name__}: {exc}"}
        torch.cuda.empty_cache()
    if not args.no_serving:
        try:
            line["serving"] = serving_bench(steps=tuple(int(v) for v in args.serving_steps.split(",")),
                                            batched_only=args.serving_batched_only)
        except Exception as exc:
            line["serving"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_adjoint:
        try:
            line["adjoint"] = adjoint_bench(steps=args.adjoint_steps, fd_voxels=args.adjoint_fd)
            log(f"adjoint: {json.dumps({k: v for k, v in line['adjoint'].items() if k != 'finite_differences'})}")
        except Exception as exc:
            line["adjoint"] = {"error": f"{type(exc).__name__}: {exc}"}
        torch.cuda.empty_cache()
    if not args.no_cpu:
        log("cpu baseline ...")
        try:
            line["cpu_baseline"] = cpu_baseline(wl.build(**kw))
        except Exception as exc:  # the GPU number stands; report why the baseline is missing
            line["cpu_baseline"] = {"error": f"{type(exc).__name__}: {exc}"}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
