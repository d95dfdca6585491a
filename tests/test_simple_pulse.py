"""The oracle against the reference's sample output file
(strata-web/public/samples/simple-pulse.h5, a run of the reference's "python"
backend; tests/golden/simple_pulse.npz): a 40^3 grid at 2 mm with rigid
walls, a 20 kHz pulse and 63 steps. The oracle's state after one step equals
the recorded first snapshot bit for bit; its later states (every 5 steps) are
within 1e-6 relative L2 of the recorded snapshots (on the recorded
sub-lattice) and its probe traces within 1e-6 of the recorded ones: the
waveform's float64 exp and sin were evaluated by another platform's math
library, which rounds a few samples differently in the last bit.
"""

import hashlib
import os

import numpy as np

from helpers import Scenario, oracle_problem
from oracle import cpu, tables as T

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simple_pulse.npz")


def load() -> dict:
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def air_mask(d) -> np.ndarray:
    shape = tuple(int(v) for v in d["shape"])
    return np.unpackbits(d["geometry_bits"])[:int(np.prod(shape))].reshape(shape).astype(bool)


def rel_l2(a, b) -> float:
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_oracle_reproduces_recorded_simple_pulse(oracle_lib):
    d = load()
    shape = tuple(int(v) for v in d["shape"])
    steps = int(d["steps"])
    sc = Scenario(shape=shape, dx=float(d["dx"]), c=float(d["c"]), rho=float(d["rho"]), air=air_mask(d),
                  sources=[("point", tuple(int(v) for v in d["source_position"]), float(d["source_frequency"]))],
                  probes=[tuple(int(v) for v in p) for p in d["probe_positions"]])
    pb = oracle_problem(sc, steps)
    assert pb.dt == float(d["dt"])
    w = T.waveform_table(T.accumulated_times(pb.dt, steps), float(d["source_frequency"]),
                         float(d["source_bandwidth"]), 1.0)
    pb.sources = [(0, pb.sources[0][1], 0, w)]
    sub = (slice(None, None, int(d["sub"])),) * 3
    st = cpu.new_state(shape)
    traces = []
    done = 0
    for t, m in enumerate(d["snapshot_steps"]):
        st, tr = cpu.run(pb, int(m) - done, state=st, step0=done)
        traces.append(tr)
        done = int(m)
        if t == 0:
            assert np.array_equal(st["p"], d["snapshot0"])
        err = rel_l2(st["p"][sub], d["snapshots_sub"][t])
        assert err < 1e-6, (t, int(m), err)
    st, tr = cpu.run(pb, steps - done, state=st, step0=done)
    traces.append(tr)
    tr = np.concatenate(traces)
    assert hashlib.sha256(np.ascontiguousarray(tr).tobytes()).hexdigest() == str(d["oracle_traces_sha256"])
    rec = d["recorded_traces"]
    for q in range(rec.shape[1]):
        assert rel_l2(tr[:, q], rec[:, q]) < 1e-6, (q, rel_l2(tr[:, q], rec[:, q]))
