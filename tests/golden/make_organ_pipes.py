"""Organ-pipe fixtures from the reference's own recorded demo runs.

    python tests/golden/make_organ_pipes.py

The reference ships three finished organ-pipe runs as data
(shared/demos/organ-pipes/{closed,half-open,open}-pipe/): per run the grid,
time step, source and probe positions (*_metadata.json), the air mask as
np.packbits of the C-order bool array (*_geometry.bin, io/legacy.py:158-185),
the recorded probe traces (*_probes.json; none for the half-open pipe) and the
resonant modes found in them (*_mode_analysis.json). The boundary is not
recorded: the closed pipe is reproduced with rigid domain walls only and the
open and half-open pipes with a PML of depth 10; no other depth (4-20 tried)
reproduces the open pipe's traces (rel-L2 0.05-0.3 instead of 2e-6).

This script reads those files (the reference tree is present only in the
build container), runs the CPU oracle in float32 over each run's full step
count, and writes ``organ_pipes.npz`` beside it: the inputs, the recorded
traces (float32, exactly as recorded), a SHA-256 of the oracle's float32
traces and the recorded mode analysis. Takes about a minute on 8 cores.
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

from helpers import Scenario, oracle_problem  # noqa: E402
from oracle import cpu, tables as T  # noqa: E402

DEMOS = "/root/reference/shared/demos/organ-pipes"
PIPES = {"closed": ("closed-pipe", 0), "half_open": ("half-open-pipe", 10), "open": ("open-pipe", 10)}


def _load(name: str):
    d = os.path.join(DEMOS, name)
    with open(os.path.join(d, f"{name}_metadata.json")) as fh:
        meta = json.load(fh)
    with open(os.path.join(d, f"{name}_mode_analysis.json")) as fh:
        modes = json.load(fh)
    probes = None
    path = os.path.join(d, f"{name}_probes.json")
    if os.path.exists(path):
        with open(path) as fh:
            probes = json.load(fh)
    bits = np.fromfile(os.path.join(d, f"{name}_geometry.bin"), dtype=np.uint8)
    return meta, modes, probes, bits


def scenario(meta, bits, pml: int):
    """(Scenario, source bandwidth, probe names, steps) of a recorded run."""
    shape = tuple(int(v) for v in meta["grid"]["shape"])
    air = np.unpackbits(bits)[:int(np.prod(shape))].reshape(shape).astype(bool)
    src = meta["sources"][0]
    names = list(meta["probes"].keys())
    sc = Scenario(shape=shape, dx=float(meta["grid"]["resolution"]), c=float(meta["physics"]["c"]),
                  rho=float(meta["physics"]["rho"]), pml_depth=pml, air=air,
                  sources=[("point", tuple(int(v) for v in src["position"]), float(src["frequency"]))],
                  probes=[tuple(int(v) for v in meta["probes"][k]["position"]) for k in names])
    return sc, float(src["bandwidth"]), names, int(meta["simulation"]["step_count"])


def oracle_traces(sc, bandwidth: float, steps: int) -> np.ndarray:
    pb = oracle_problem(sc, steps)
    times = T.accumulated_times(pb.dt, steps)
    f = sc.sources[0][2]
    pb.sources = [(0, pb.sources[0][1], 0, T.waveform_table(times, f, bandwidth, 1.0))]
    _, tr = cpu.run(pb, steps)
    return tr, pb.dt


def main() -> None:
    cpu.build(ref=False)
    out = {}
    for key, (name, pml) in PIPES.items():
        meta, modes, probes, bits = _load(name)
        sc, bw, names, steps = scenario(meta, bits, pml)
        tr, dt = oracle_traces(sc, bw, steps)
        if dt != float(meta["simulation"]["dt"]):
            raise SystemExit(f"{name}: time step {dt} differs from the recorded {meta['simulation']['dt']}")
        src = meta["sources"][0]
        p = f"{key}_"
        out[p + "shape"] = np.array(sc.shape, np.int64)
        out[p + "dx"] = np.array(sc.dx)
        out[p + "c"] = np.array(sc.c)
        out[p + "rho"] = np.array(sc.rho)
        out[p + "dt"] = np.array(float(meta["simulation"]["dt"]))
        out[p + "steps"] = np.array(steps, np.int64)
        out[p + "pml"] = np.array(pml, np.int64)
        out[p + "source_position"] = np.array(src["position"], np.int64)
        out[p + "source_frequency"] = np.array(float(src["frequency"]))
        out[p + "source_bandwidth"] = np.array(bw)
        out[p + "probe_names"] = np.array(names)
        out[p + "probe_positions"] = np.array(sc.probes, np.int64)
        out[p + "geometry_bits"] = bits
        # the oracle's float32 traces, by hash (the GPU run must equal them bit for bit)
        out[p + "oracle_traces_sha256"] = np.array(hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest())
        if probes is not None:
            rec = np.array([probes["probes"][k]["data"] for k in names], dtype=np.float64).T
            rec32 = rec.astype(np.float32)
            if not np.array_equal(rec32.astype(np.float64), rec):
                raise SystemExit(f"{name}: recorded traces are not float32 values")
            out[p + "recorded_traces"] = rec32
            out[p + "sample_rate"] = np.array(float(probes["sample_rate"]))
        ap = modes["analysis_parameters"]
        out[p + "analysis_probe"] = np.array(names.index(modes["analysis_probe"]), np.int64)
        out[p + "n_fft"] = np.array(int(ap["n_fft"]), np.int64)
        out[p + "transient_skip"] = np.array(int(ap["transient_skip_samples"]), np.int64)
        out[p + "expected_modes_hz"] = np.array(modes["expected_modes_hz"], np.float64)
        out[p + "measured_modes_hz"] = np.array(
            [v if isinstance(v, (int, float)) else np.nan for v in modes["measured_modes_hz"]], np.float64)
        print(f"{name}: {sc.shape}, {steps} steps, PML {pml}, {len(names)} probes", flush=True)
    path = os.path.join(HERE, "organ_pipes.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes)")


if __name__ == "__main__":
    main()
