"""Sources and probes (core/solver.py:152-207, 758-781).

The waveform is evaluated on the host in float64 at the accumulated simulation
time (``t += dt`` per step, core/solver.py:2072) and uploaded as a table; the
fused step kernel adds column ``s`` to the source cells with one float64 add
and one rounding, as ``p[i,j,k] += waveform_val`` does in numpy.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Literal

import numpy as np
from numpy.typing import NDArray


@dataclass
class GaussianPulse:
    """Gaussian-modulated sinusoid (core/solver.py:152-207).

    w(t) = A exp(-(t-t0)^2 / (2 s^2)) sin(2 pi f (t-t0)), s = 1/(pi bw),
    t0 = 4 s, bw = 2 f by default.
    """

    position: tuple[int, int, int] | dict
    frequency: float
    bandwidth: float | None = None
    amplitude: float = 1.0
    source_type: Literal["point", "plane"] = "point"

    def __post_init__(self):
        if self.bandwidth is None:
            self.bandwidth = 2.0 * self.frequency

    def waveform(self, t: NDArray[np.floating], dt: float) -> NDArray[np.floating]:
        sigma = 1.0 / (np.pi * self.bandwidth)
        t0 = 4.0 * sigma
        envelope = np.exp(-((t - t0) ** 2) / (2 * sigma**2))
        carrier = np.sin(2 * np.pi * self.frequency * (t - t0))
        return self.amplitude * envelope * carrier


@dataclass
class Probe:
    """Pressure probe at a grid index (core/solver.py:758-781)."""

    name: str
    position: tuple[int, int, int]
    data: list[float] = field(default_factory=list)

    def record(self, pressure: float) -> None:
        self.data.append(pressure)

    def get_data(self) -> NDArray[np.floating]:
        return np.array(self.data, dtype=np.float32)

    def clear(self) -> None:
        self.data.clear()


def accumulated_times(dt: float, n: int, t_start: float = 0.0) -> NDArray[np.float64]:
    """Times t_0 = t_start, t_{m+1} = t_m + dt, accumulated in float64 one step at
    a time exactly as ``FDTDSolver.step`` advances ``_time``."""
    out = np.empty(n, dtype=np.float64)
    t = float(t_start)
    for m in range(n):
        out[m] = t
        t += dt
    return out


def waveform_table(source, times: NDArray[np.float64], dt: float) -> NDArray[np.float64]:
    """Per-step waveform values, each from its own one-element call, the way
    ``_inject_sources`` samples the source (core/solver.py:2415-2416)."""
    out = np.empty(len(times), dtype=np.float64)
    for m, t in enumerate(times):
        out[m] = source.waveform(np.array([t]), dt)[0]
    return out
