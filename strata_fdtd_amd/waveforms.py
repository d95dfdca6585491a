"""Audio-file source waveforms (core/waveforms.py:33-231): a recording read
from a WAV file (other formats through soundfile, when it is installed),
converted to float32 in [-1, 1), one channel or the channel mean, trimmed to
[start_time, start_time + duration), resampled to the simulation's rate
1/dt with scipy's FFT resampler, and sampled per step at index int(t / dt)
(zero past the end, or wrapped when looping), times `amplitude`.

Any object with ``.waveform(t, dt)`` drives a source; the solver samples it
once per step into the device source table.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path
from typing import Literal

import numpy as np
from numpy.typing import NDArray


@dataclass
class AudioFileWaveform:
    filepath: str | Path
    amplitude: float = 1.0
    channel: int | Literal["mix"] = "mix"
    start_time: float = 0.0
    duration: float | None = None
    loop: bool = False
    _samples: NDArray[np.floating] | None = field(default=None, init=False, repr=False)
    _native_sr: int | None = field(default=None, init=False, repr=False)
    _resampled: dict = field(default_factory=dict, init=False, repr=False)

    def __post_init__(self) -> None:
        path = Path(self.filepath)
        if not path.exists():
            raise FileNotFoundError(f"Audio file not found: {path}")
        if path.suffix.lower() == ".wav":
            from scipy.io import wavfile
            sr, data = wavfile.read(path)
        else:
            try:
                import soundfile
            except ImportError as e:
                raise ImportError(f"soundfile package required for {path.suffix} files. "
                                  "Install with: pip install soundfile") from e
            data, sr = soundfile.read(path)
        self._native_sr = sr
        scale = {np.dtype(np.int16): (0.0, 32768.0), np.dtype(np.int32): (0.0, 2147483648.0),
                 np.dtype(np.uint8): (128.0, 128.0)}.get(data.dtype)
        if scale is not None:
            off, div = scale
            data = data.astype(np.float32)
            data = (data - off) / div if off else data / div
        else:
            data = data.astype(np.float32)
        if data.ndim > 1:
            if self.channel == "mix":
                data = np.mean(data, axis=1)
            else:
                if self.channel >= data.shape[1]:
                    raise ValueError(f"Channel {self.channel} requested but file only has {data.shape[1]} channels")
                data = data[:, self.channel]
        first = int(self.start_time * sr)
        if first >= len(data):
            raise ValueError(f"start_time {self.start_time}s is beyond end of file ({len(data) / sr:.2f}s)")
        data = data[first:first + int(self.duration * sr)] if self.duration is not None else data[first:]
        self._samples = data.astype(np.float32)

    def _at_rate(self, dt: float) -> NDArray[np.float32]:
        rate = int(round(1.0 / dt))
        if rate not in self._resampled:
            if rate == self._native_sr:
                self._resampled[rate] = self._samples
            else:
                from scipy import signal
                n = int(len(self._samples) * rate / self._native_sr)
                self._resampled[rate] = signal.resample(self._samples, n).astype(np.float32)
        return self._resampled[rate]

    def waveform(self, t, dt: float) -> NDArray[np.floating]:
        x = self._at_rate(dt)
        idx = (t / dt).astype(np.int64)
        if self.loop:
            return self.amplitude * x[idx % len(x)]
        ok = (idx >= 0) & (idx < len(x))
        out = np.zeros_like(t, dtype=np.float32)
        out[ok] = x[idx[ok]]
        return self.amplitude * out

    @property
    def duration_seconds(self) -> float:
        if self._samples is None or self._native_sr is None:
            return 0.0
        return len(self._samples) / self._native_sr

    @property
    def native_sample_rate(self) -> int:
        return self._native_sr or 0

    @property
    def num_samples(self) -> int:
        return 0 if self._samples is None else len(self._samples)
