"""The reference's recorded organ-pipe runs (tests/golden/organ_pipes.npz,
written by tests/golden/make_organ_pipes.py from the reference's
shared/demos/organ-pipes data): loading, the scene, and the mode analysis.

Mode analysis, as the recorded *_mode_analysis.json files report it (their
"measured" frequencies are bins of this spectrum, and it reproduces every one
of them from the recorded traces): drop the first `transient_skip` samples of
the analysis probe, apply a Hann window, take |rfft| with n_fft points at the
recorded sample rate (1/dt) and, for each expected mode, the frequency of the
largest bin within +-10% of it.
"""

from __future__ import annotations

import os

import numpy as np

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "organ_pipes.npz")
PIPES = ("closed", "half_open", "open")


def load(key: str) -> dict:
    with np.load(FIXTURE, allow_pickle=False) as z:
        return {k[len(key) + 1:]: z[k] for k in z.files if k.startswith(key + "_")}


def air_mask(d: dict) -> np.ndarray:
    shape = tuple(int(v) for v in d["shape"])
    return np.unpackbits(d["geometry_bits"])[:int(np.prod(shape))].reshape(shape).astype(bool)


def modes(trace: np.ndarray, d: dict) -> np.ndarray:
    x = np.asarray(trace, np.float64)[int(d["transient_skip"]):]
    x = x * np.hanning(len(x))
    n_fft = int(d["n_fft"])
    spec = np.abs(np.fft.rfft(x, n=n_fft))
    freq = np.fft.rfftfreq(n_fft, float(d["dt"]))
    out = []
    for fe in d["expected_modes_hz"]:
        win = (freq > 0.9 * fe) & (freq < 1.1 * fe)
        out.append(freq[np.argmax(np.where(win, spec, 0.0))])
    return np.array(out)


def rel_l2(a, b) -> float:
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))
