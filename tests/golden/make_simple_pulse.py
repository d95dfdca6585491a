"""Fixture from the reference's sample output file (strata-web/public/samples/
simple-pulse.h5, written by the reference's HDF5 writer, io/hdf5.py, from a
run of its "python" backend, solver version 0.1.0).

    python tests/golden/make_simple_pulse.py

The run: a 40^3 uniform grid at 2 mm, air (c = 343, rho = 1.2, all cells air,
rigid domain walls, no PML), a 20 kHz Gaussian pulse (bandwidth 40 kHz) at
(20, 20, 20), probes at (25, 20, 20) and (35, 20, 20), 63 steps of 3.198 us,
and 13 pressure snapshots. The snapshots are the states after 1, 6, 11, ...,
61 steps (the first equals the CPU oracle's state after one step bit for bit,
each later one is closest to that step by five orders of magnitude), which is
the cadence of a snapshot every 5 steps (FDTDSolver.enable_snapshots(5)).

The file is read with the minimal parser beside this script (h5py is not
installed here). Stored in ``simple_pulse.npz``: the run's parameters, the
recorded probe traces, the 13 snapshots on the sub-lattice of every second
cell per axis, a SHA-256 of the oracle's float32 probe traces, the file's
layout (objects, shapes, attribute names) with the attributes that describe
the run (grid, simulation, source), and how each dataset is stored (chunk
shapes, maximum extents, fill-value and filter messages), for comparison with
a file the product writes.
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)

import hdf5_read as H  # noqa: E402
from helpers import Scenario, oracle_problem  # noqa: E402
from oracle import cpu, tables as T  # noqa: E402

SAMPLE = "/root/reference/strata-web/public/samples/simple-pulse.h5"
SUB = 2
EVERY = 5


def main() -> None:
    f = H.File(SAMPLE)
    r = f.root
    grid, sim, src = r["grid"].attrs, r["simulation"].attrs, r["sources/source_0"].attrs
    meta = r["metadata"].attrs
    shape = tuple(int(v) for v in grid["shape"])
    if not bool(grid["is_uniform"]) or str(src["type"]) != "point":
        raise SystemExit("unexpected sample layout")
    names = sorted(r["probes"].keys())
    probes = [tuple(int(v) for v in r[f"probes/{n}"].attrs["position"]) for n in names]
    traces = np.stack([r[f"probes/{n}"].read() for n in names], axis=1).astype(np.float32)
    snaps = r["fields/pressure"].read().astype(np.float32)
    geo = r["materials/geometry"].read().astype(bool)
    steps = int(sim["num_steps"])

    cpu.build(ref=False)
    sc = Scenario(shape=shape, dx=float(grid["resolution"]), c=float(sim["c"]), rho=float(sim["rho"]),
                  pml_depth=0, air=geo, sources=[("point", tuple(int(v) for v in src["position"]),
                                                   float(src["frequency"]))], probes=probes)
    pb = oracle_problem(sc, steps)
    if pb.dt != float(sim["timestep"]):
        raise SystemExit(f"time step {pb.dt} differs from the recorded {sim['timestep']}")
    w = T.waveform_table(T.accumulated_times(pb.dt, steps), float(src["frequency"]), float(src["bandwidth"]), 1.0)
    pb.sources = [(0, pb.sources[0][1], 0, w)]
    _, tr = cpu.run(pb, steps)
    snap_steps = np.arange(snaps.shape[0]) * EVERY + 1
    sub = (slice(None, None, SUB),) * 3
    out = dict(
        shape=np.array(shape, np.int64), dx=np.array(float(grid["resolution"])), c=np.array(float(sim["c"])),
        rho=np.array(float(sim["rho"])), dt=np.array(float(sim["timestep"])), steps=np.array(steps, np.int64),
        source_position=np.array(src["position"], np.int64), source_frequency=np.array(float(src["frequency"])),
        source_bandwidth=np.array(float(src["bandwidth"])), probe_names=np.array(names),
        probe_positions=np.array(probes, np.int64), geometry_bits=np.packbits(geo.ravel()),
        recorded_traces=traces, snapshot_every=np.array(EVERY, np.int64), snapshot_steps=snap_steps,
        sub=np.array(SUB, np.int64), snapshots_sub=np.ascontiguousarray(snaps[(slice(None),) + sub]),
        snapshot0=snaps[0], backend=np.array(str(meta["backend"])),
        oracle_traces_sha256=np.array(hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest()))
    # the file's layout (every object's kind, shape and attribute names) and
    # the run-describing attributes, for a product-written file to match
    layout = {}
    for opath, node in H.walk(r):
        kind = "dataset" if isinstance(node, H.Dataset) else "group"
        layout[opath] = {"kind": kind, "attrs": sorted(node.attrs),
                         "shape": list(node.shape) if kind == "dataset" else None}
    values = {}
    for g in ("grid", "simulation", "sources/source_0"):
        values[g] = {k: (v.tolist() if isinstance(v, np.ndarray) else (v.item() if hasattr(v, "item") else v))
                     for k, v in r[g].attrs.items()}
    out["layout_json"] = np.array(json.dumps(layout, sort_keys=True))
    out["storage_json"] = np.array(json.dumps(H.storage(r), sort_keys=True))
    out["attrs_json"] = np.array(json.dumps(values, sort_keys=True))
    path = os.path.join(HERE, "simple_pulse.npz")
    np.savez_compressed(path, **out)
    print(f"{shape}, {steps} steps, probes {names}, {snaps.shape[0]} snapshots; wrote {path} "
          f"({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
