"""Full-length oracle runs of the three solve workloads (fixtures).

    python tests/golden/make_full_length.py [plain_100] [resonators_nu] [enclosure_512]

Runs the CPU oracle (oracle/fdtd_oracle.cpp, pinned bit for bit to the
reference's own native kernels) in float32 and in float64 over the stated
step counts of
  plain_100      100^3, PML 10, 1 kHz pulse at the centre, one probe, 1000 steps;
  enclosure_512  512^3, PML 16, rigid SDF enclosure + port + two-pole ADE
                 sphere, 500 Hz pulse, 8 probes, 4000 steps;
  resonators_nu  320x320x480 nonuniform plate with 6x6 Helmholtz resonators,
                 PML 16, 2 kHz plane source, 16 probes, 3000 steps,
and stores ``full_<name>.npz`` beside this script: the float32 and float64
probe traces, a SHA-256 of each final float32 field's bytes, the final p of
both runs on a coarse sub-lattice (every 16th cell per axis) and their norms.
The scene comes from the same builders the workloads use
(strata_fdtd_amd.workloads), voxelised by the host voxeliser (pinned to the
reference's sdf.py masks) and described to the oracle through
tests/helpers.py, so no GPU is needed. The GPU test then runs the product
solver over the same steps and compares bit for bit. Takes about an hour on
8 cores (the float64 run is the longer).
"""

from __future__ import annotations

import hashlib
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import Scenario, oracle_problem  # noqa: E402
from oracle import cpu  # noqa: E402

SUB = 16          # sub-lattice stride of the stored final pressure
CHUNK = 250       # steps per oracle call (progress lines) of scenes without ADE poles


def _poles(material) -> list:
    out = []
    for p in material.poles:
        if p.is_debye:
            out.append(dict(kind="debye", target=p.target, tau=p.tau, dchi=p.delta_chi))
        else:
            out.append(dict(kind="lorentz", target=p.target, w0=p.omega_0, gamma=p.gamma, dchi=p.delta_chi))
    return out


def scenario(name: str):
    """(Scenario, steps) of a workload, built from its product-side scene."""
    from strata_fdtd_amd import UniformGrid
    from strata_fdtd_amd import workloads as W
    if name == "enclosure_512":
        geo, src, probes = W.enclosure_512_scene(512)
        air, ids = geo.voxelize_with_materials(UniformGrid((512, 512, 512), 1e-3))
        (region, material, mid), = geo._material_regions
        sc = Scenario(shape=(512, 512, 512), pml_depth=16, air=air, material=ids,
                      materials={mid: _poles(material)}, sources=[("point", tuple(src), 500.0)],
                      probes=[tuple(p) for p in probes])
        return sc, 4000
    if name == "plain_100":
        sc = Scenario(shape=(100, 100, 100), pml_depth=10, sources=[("point", (50, 50, 50), 1000.0)],
                      probes=[(75, 50, 50)])
        return sc, 1000
    if name == "resonators_nu":
        grid, air_geo, probes, src_index = W.resonators_nu_scene()
        air = air_geo.voxelize(grid)
        sc = Scenario(shape=grid.shape, pml_depth=16, air=air, sources=[("plane", 2, src_index, 2000.0)],
                      probes=[tuple(p) for p in probes], coords=(grid.x_coords, grid.y_coords, grid.z_coords))
        return sc, 3000
    raise SystemExit(f"unknown workload {name}")


def run(pb, steps, impl, label):
    dtype = np.float64 if impl == "oracle64" else np.float32
    st = cpu.new_state(pb.shape, dtype)
    traces = []
    t0 = time.time()
    # the oracle keeps ADE pole state within one call only: scenes with
    # poles run all their steps in a single call
    chunk = steps if pb.poles else CHUNK
    for s0 in range(0, steps, chunk):
        n = min(chunk, steps - s0)
        st, tr = cpu.run(pb, n, state=st, impl=impl, step0=s0)
        traces.append(tr.copy())
        print(f"{label} {impl}: {s0 + n}/{steps} steps, {time.time() - t0:.0f} s", flush=True)
    return st, np.concatenate(traces)


def main(names) -> None:
    cpu.build(ref=False)
    for name in names:
        sc, steps = scenario(name)
        pb = oracle_problem(sc, steps)
        print(f"{name}: {sc.shape}, {steps} steps, {len(sc.probes)} probes, "
              f"{sum(len(v) for v in sc.materials.values())} poles", flush=True)
        st32, tr32 = run(pb, steps, "oracle", name)
        sub = (slice(None, None, SUB),) * 3
        out = dict(steps=steps, shape=np.array(sc.shape), sub=SUB, trace_f32=tr32,
                   p_sub_f32=np.ascontiguousarray(st32["p"][sub]))
        for k in ("p", "vx", "vy", "vz"):
            out[f"sha256_{k}"] = np.array(hashlib.sha256(np.ascontiguousarray(st32[k]).tobytes()).hexdigest())
            out[f"norm_{k}_f32"] = np.array(float(np.linalg.norm(st32[k].astype(np.float64))))
        del st32
        st64, tr64 = run(pb, steps, "oracle64", name)
        out["trace_f64"] = tr64
        out["p_sub_f64"] = np.ascontiguousarray(st64["p"][sub])
        for k in ("p", "vx", "vy", "vz"):
            out[f"norm_{k}_f64"] = np.array(float(np.linalg.norm(st64[k])))
        del st64
        path = os.path.join(HERE, f"full_{name}.npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path} ({os.path.getsize(path)} bytes)", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:] or ["plain_100", "resonators_nu", "enclosure_512"])
