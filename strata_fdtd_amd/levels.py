"""Sound levels of recorded waveforms: IEC 61672-1 frequency weightings,
SPL, Leq and exponentially time-weighted levels (the measurement side of the
reference's microphones, core/solver.py:1175-1335 with
analysis/weighting.py).

Weightings. IEC 61672-1 defines the A and C weightings by analog transfer
functions with real poles at f1 = 20.598997 Hz, f2 = 107.65265 Hz,
f3 = 737.86223 Hz and f4 = 12194.217 Hz, normalised to 0 dB at 1 kHz:

    C(s) = k_C s^2 / ((s + w1)^2 (s + w4)^2)
    A(s) = k_A s^4 / ((s + w1)^2 (s + w2) (s + w3) (s + w4)^2)

They are mapped to the recording's sample rate with the bilinear transform
and run as second-order sections (scipy.signal). At FDTD sample rates
(hundreds of kHz) the frequency warping of the transform is negligible over
the audio band.

Levels. SPL = 20 log10(p_rms / p_ref) (-inf for silence); Leq is the SPL of
the most recent `duration` seconds (or of all of it); the time-weighted level
runs p^2 through a one-pole average y[n] = y[n-1] + a (p[n]^2 - y[n-1]),
a = 1 - exp(-1 / (fs tau)), with tau 125 ms (fast), 1 s (slow) or 35 ms
(impulse, its attack), and returns 10 log10(max(y, 1e-20) / p_ref^2).
"""

from __future__ import annotations

from functools import lru_cache

import numpy as np
from numpy.typing import NDArray

P_REF = 2e-5
IEC_POLES_HZ = (20.598997, 107.65265, 737.86223, 12194.217)
TIME_CONSTANTS = {"fast": 0.125, "slow": 1.0, "impulse": 0.035}


def _analog_zpk(weighting: str):
    w1, w2, w3, w4 = (2 * np.pi * f for f in IEC_POLES_HZ)
    if weighting == "A":
        zeros, poles = [0.0] * 4, [-w1, -w1, -w2, -w3, -w4, -w4]
    else:
        zeros, poles = [0.0] * 2, [-w1, -w1, -w4, -w4]
    # gain giving |H(j 2 pi 1000)| = 1
    s = 2j * np.pi * 1000.0
    h = np.prod([s - z for z in zeros]) / np.prod([s - p for p in poles])
    return np.array(zeros), np.array(poles), 1.0 / abs(h)


@lru_cache(maxsize=64)
def weighting_sos(weighting, fs: float):
    """Second-order sections of the A or C weighting at sample rate fs; None
    for Z (or no) weighting."""
    if weighting in ("Z", None):
        return None
    if weighting not in ("A", "C"):
        raise ValueError(f"Unknown weighting type: {weighting!r}. Valid options: 'A', 'C', 'Z', or None")
    from scipy import signal
    z, p, k = _analog_zpk(weighting)
    zd, pd, kd = signal.bilinear_zpk(z, p, k, fs)
    return signal.zpk2sos(zd, pd, kd)


def apply_weighting(waveform, fs: float, weighting) -> NDArray[np.float64]:
    """The waveform filtered by the weighting (float64, same length)."""
    sos = weighting_sos(weighting, float(fs))
    x = np.asarray(waveform, dtype=np.float64)
    if sos is None:
        return x
    from scipy import signal
    return signal.sosfilt(sos, x)


def weighting_response(weighting, fs: float, freqs) -> NDArray[np.float64]:
    """Magnitude response in dB of the digital weighting at the frequencies."""
    sos = weighting_sos(weighting, float(fs))
    f = np.asarray(freqs, dtype=np.float64)
    if sos is None:
        return np.zeros_like(f)
    from scipy import signal
    _, h = signal.sosfreqz(sos, worN=f, fs=fs)
    return 20 * np.log10(np.maximum(np.abs(h), 1e-20))


def spl(waveform, p_ref: float = P_REF) -> float:
    x = np.asarray(waveform, dtype=np.float64)
    rms = np.sqrt(np.mean(x**2))
    return -np.inf if rms == 0 else float(20 * np.log10(rms / p_ref))


def leq(waveform, fs: float, duration: float | None = None, p_ref: float = P_REF) -> float:
    x = np.asarray(waveform)
    if duration is not None:
        n = int(duration * fs)
        if n < len(x):
            x = x[len(x) - n:]
    return spl(x, p_ref)


def time_weighted_level(waveform, fs: float, time_constant: str = "fast", p_ref: float = P_REF) -> NDArray[np.float64]:
    from scipy import signal
    if time_constant not in TIME_CONSTANTS:
        raise ValueError(f"time_constant must be one of {tuple(TIME_CONSTANTS)}, got {time_constant!r}")
    a = 1 - np.exp(-1 / (fs * TIME_CONSTANTS[time_constant]))
    p2 = np.asarray(waveform, dtype=np.float64) ** 2
    y = signal.lfilter([a], [1.0, -(1 - a)], p2)
    return 10 * np.log10(np.maximum(y, 1e-20) / p_ref**2)
