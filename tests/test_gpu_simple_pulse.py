"""The product against the reference's sample output file
(strata-web/public/samples/simple-pulse.h5; tests/golden/simple_pulse.npz),
through the drop-in API with snapshots (FDTDSolver.enable_snapshots(5),
run(duration)): the probe traces equal the CPU oracle's bit for bit (by
SHA-256) and are within 1e-6 relative L2 of the recorded traces; the first
snapshot equals the recorded one bit for bit, and every snapshot is within
1e-6 of the recorded one on its sub-lattice, at the recorded times.
"""

import hashlib

import numpy as np
import pytest

from test_simple_pulse import air_mask, load, rel_l2

pytestmark = pytest.mark.gpu


def test_product_reproduces_recorded_simple_pulse():
    import strata_fdtd_amd as sfa
    d = load()
    shape = tuple(int(v) for v in d["shape"])
    steps = int(d["steps"])
    print(f"[simple pulse] {shape}, {steps} steps", flush=True)
    s = sfa.FDTDSolver(grid=sfa.UniformGrid(shape, float(d["dx"])), c=float(d["c"]), rho=float(d["rho"]),
                       backend="mi355x")
    assert s.dt == float(d["dt"])
    s.set_geometry(air_mask(d))
    s.add_source(sfa.GaussianPulse(position=tuple(int(v) for v in d["source_position"]),
                                   frequency=float(d["source_frequency"]),
                                   bandwidth=float(d["source_bandwidth"])))
    names = [str(n) for n in d["probe_names"]]
    for name, pos in zip(names, d["probe_positions"]):
        s.add_probe(name, tuple(int(v) for v in pos))
    s.enable_snapshots(int(d["snapshot_every"]))
    s.run(duration=steps * s.dt * (1 - 1e-9))
    assert s.step_count == steps
    got = s.get_probe_data()
    tr = np.stack([got[n] for n in names], axis=1).astype(np.float32)
    assert hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest() == str(d["oracle_traces_sha256"])
    rec = d["recorded_traces"]
    for q, n in enumerate(names):
        assert rel_l2(tr[:, q], rec[:, q]) < 1e-6, n
    snaps = s.get_snapshots()
    assert len(snaps) == d["snapshots_sub"].shape[0]
    sub = (slice(None, None, int(d["sub"])),) * 3
    # a snapshot carries the time at the start of the step it follows
    # (t_0 = 0, t_{m+1} = t_m + dt accumulated in float64)
    times = np.concatenate(([0.0], np.cumsum(np.full(steps - 1, s.dt))))
    for t, ((when, p), m) in enumerate(zip(snaps, d["snapshot_steps"])):
        assert when == times[int(m) - 1], (t, when, times[int(m) - 1])
        if t == 0:
            assert np.array_equal(p, d["snapshot0"])
        assert rel_l2(p[sub], d["snapshots_sub"][t]) < 1e-6, t


def test_run_writes_reference_result_file(tmp_path):
    """run(duration, output_file=..., snapshot_interval=5) on the same run
    writes an HDF5 file in the reference's result layout: read back with the
    parser that reads the reference's own sample, it has the sample's
    objects, shapes and attribute names, its dataset storage (chunk shapes,
    maximum extents, gzip level 4), the sample's grid, simulation and
    source attributes, the probe traces the solver returns, the pressure
    snapshots (the first equal to the recorded one bit for bit, all within
    1e-6 of the recorded ones) and the air mask."""
    import json
    import os
    import sys
    import strata_fdtd_amd as sfa
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    import hdf5_read as H
    d = load()
    shape = tuple(int(v) for v in d["shape"])
    steps = int(d["steps"])
    s = sfa.FDTDSolver(grid=sfa.UniformGrid(shape, float(d["dx"])), c=float(d["c"]), rho=float(d["rho"]),
                       backend="mi355x")
    s.set_geometry(air_mask(d))
    s.add_source(sfa.GaussianPulse(position=tuple(int(v) for v in d["source_position"]),
                                   frequency=float(d["source_frequency"]),
                                   bandwidth=float(d["source_bandwidth"])))
    names = [str(n) for n in d["probe_names"]]
    for name, pos in zip(names, d["probe_positions"]):
        s.add_probe(name, tuple(int(v) for v in pos))
    path = tmp_path / "result.h5"
    s.run(duration=steps * s.dt * (1 - 1e-9), output_file=str(path), snapshot_interval=int(d["snapshot_every"]))
    f = H.File(str(path))
    layout = {}
    for opath, node in H.walk(f.root):
        kind = "dataset" if isinstance(node, H.Dataset) else "group"
        layout[opath] = {"kind": kind, "attrs": sorted(node.attrs),
                         "shape": list(node.shape) if kind == "dataset" else None}
    assert layout == json.loads(str(d["layout_json"]))
    assert H.storage(f.root) == json.loads(str(d["storage_json"]))
    want = json.loads(str(d["attrs_json"]))
    for g, attrs in want.items():
        got = f.root[g].attrs
        for k, v in attrs.items():
            gv = got[k]
            gv = gv.tolist() if isinstance(gv, np.ndarray) else (gv.item() if hasattr(gv, "item") else gv)
            assert gv == v, (g, k, gv, v)
    probes = s.get_probe_data()
    for n in names:
        assert np.array_equal(f.root[f"probes/{n}"].read(), probes[n].astype(np.float32))
    p = f.root["fields/pressure"].read()
    assert np.array_equal(p[0], d["snapshot0"])
    sub = (slice(None),) + (slice(None, None, int(d["sub"])),) * 3
    for t in range(p.shape[0]):
        assert rel_l2(p[sub][t], d["snapshots_sub"][t]) < 1e-6, t
    assert np.array_equal(f.root["materials/geometry"].read(), air_mask(d).astype(np.uint8))
    assert f.root["metadata"].attrs["backend"] == "mi355x"
    # read back through the reference's reader interface
    with sfa.HDF5ResultReader(path) as rd:
        assert rd.get_num_snapshots() == p.shape[0] and sorted(rd.get_probe_names()) == sorted(names)
        assert np.array_equal(rd.load_timestep(0), d["snapshot0"])
        for n in names:
            assert np.array_equal(rd.load_probe(n), probes[n].astype(np.float32))
        assert np.array_equal(rd.load_geometry(), air_mask(d))
        assert rd.get_metadata()["simulation"]["num_steps"] == steps
