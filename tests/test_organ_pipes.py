"""The oracle against the reference's own recorded runs: the three organ
pipes of shared/demos/organ-pipes (tests/golden/organ_pipes.npz).

* closed pipe (130x44x44 at 2 mm, rigid walls, 9117 steps) and open pipe
  (180x76x76, PML 10, 11924 steps): the oracle's float32 probe traces are
  within 1e-5 relative L2 of the recorded ones at every probe (4-8e-6 and
  2-4e-6: the recording came from a backend whose float32 rounding differs
  from the native kernels' from the third step on), and the resonant modes
  found in them are the recorded "measured" modes, bin for bin;
* half-open pipe (165x76x76, PML 10, 22864 steps; no traces recorded): the
  modes found in the oracle's trace are the recorded measured modes.
"""

import hashlib

import numpy as np
import pytest

import organ_pipes as OP
from helpers import Scenario, oracle_problem
from oracle import cpu, tables as T


def _oracle(d):
    shape = tuple(int(v) for v in d["shape"])
    sc = Scenario(shape=shape, dx=float(d["dx"]), c=float(d["c"]), rho=float(d["rho"]),
                  pml_depth=int(d["pml"]), air=OP.air_mask(d),
                  sources=[("point", tuple(int(v) for v in d["source_position"]), float(d["source_frequency"]))],
                  probes=[tuple(int(v) for v in p) for p in d["probe_positions"]])
    steps = int(d["steps"])
    pb = oracle_problem(sc, steps)
    assert pb.dt == float(d["dt"])
    w = T.waveform_table(T.accumulated_times(pb.dt, steps), float(d["source_frequency"]),
                         float(d["source_bandwidth"]), 1.0)
    pb.sources = [(0, pb.sources[0][1], 0, w)]
    _, tr = cpu.run(pb, steps)
    return tr


@pytest.mark.parametrize("key", OP.PIPES)
def test_oracle_reproduces_recorded_organ_pipe(oracle_lib, key):
    d = OP.load(key)
    tr = _oracle(d)
    assert hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest() == str(d["oracle_traces_sha256"])
    if "recorded_traces" in d:
        rec = d["recorded_traces"]
        assert rec.shape == tr.shape
        for q, name in enumerate(d["probe_names"]):
            err = OP.rel_l2(tr[:, q], rec[:, q])
            assert err < 1e-5, (key, str(name), err)
        # the analysis reproduces the recorded measured modes from the recording itself
        got = OP.modes(rec[:, int(d["analysis_probe"])], d)
        meas = d["measured_modes_hz"]
        assert np.allclose(got[~np.isnan(meas)], meas[~np.isnan(meas)], rtol=0, atol=1e-6)
    got = OP.modes(tr[:, int(d["analysis_probe"])], d)
    meas = d["measured_modes_hz"]
    assert np.any(~np.isnan(meas))
    assert np.allclose(got[~np.isnan(meas)], meas[~np.isnan(meas)], rtol=0, atol=1e-6), (key, got, meas)
