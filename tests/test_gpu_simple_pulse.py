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
