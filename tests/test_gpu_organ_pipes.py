"""The product against the reference's own recorded organ-pipe runs
(tests/golden/organ_pipes.npz, from shared/demos/organ-pipes): each pipe is
built through the drop-in API (FDTDSolver, UniformGrid, set_geometry with
the recorded air mask, PML, GaussianPulse, add_probe) and run for its
recorded step count on the GPU. Its probe traces equal the CPU oracle's bit
for bit (by SHA-256), are within 1e-5 relative L2 of the recorded traces at
every probe where the reference recorded them, and give the recorded
resonant modes.
"""

import hashlib

import numpy as np
import pytest

import organ_pipes as OP

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("key", OP.PIPES)
def test_product_reproduces_recorded_organ_pipe(key):
    import strata_fdtd_amd as sfa
    d = OP.load(key)
    shape = tuple(int(v) for v in d["shape"])
    steps = int(d["steps"])
    print(f"[organ pipe] {key}: {shape}, {steps} steps, PML {int(d['pml'])}", flush=True)
    s = sfa.FDTDSolver(grid=sfa.UniformGrid(shape, float(d["dx"])), c=float(d["c"]), rho=float(d["rho"]),
                       backend="mi355x")
    assert s.dt == float(d["dt"])
    s.set_geometry(OP.air_mask(d))
    if int(d["pml"]):
        s.add_boundary(sfa.PML(depth=int(d["pml"])))
    s.add_source(sfa.GaussianPulse(position=tuple(int(v) for v in d["source_position"]),
                                   frequency=float(d["source_frequency"]),
                                   bandwidth=float(d["source_bandwidth"])))
    names = [str(n) for n in d["probe_names"]]
    for name, pos in zip(names, d["probe_positions"]):
        s.add_probe(name, tuple(int(v) for v in pos))
    s.step_n(steps)
    got = s.get_probe_data()
    tr = np.stack([got[n] for n in names], axis=1).astype(np.float32)
    assert tr.shape == (steps, len(names))
    assert hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest() == str(d["oracle_traces_sha256"])
    if "recorded_traces" in d:
        rec = d["recorded_traces"]
        errs = [OP.rel_l2(tr[:, q], rec[:, q]) for q in range(len(names))]
        print(f"{key}: rel-L2 vs recorded per probe {['%.2e' % e for e in errs]}", flush=True)
        assert max(errs) < 1e-5, errs
    m = OP.modes(tr[:, int(d["analysis_probe"])], d)
    meas = d["measured_modes_hz"]
    assert np.allclose(m[~np.isnan(meas)], meas[~np.isnan(meas)], rtol=0, atol=1e-6), (m, meas)
