"""Audio-file waveforms (core/waveforms.py:33-231) on WAV files written
here: integer and float sample formats, channel selection and mixing,
trimming, looping, resampling to 1/dt, and the errors."""

import numpy as np
import pytest
from scipy.io import wavfile

from strata_fdtd_amd import AudioFileWaveform


def test_formats_channels_and_sampling(tmp_path):
    rng = np.random.default_rng(1)
    st = (rng.uniform(-1, 1, (480, 2)) * 32767).astype(np.int16)
    path = tmp_path / "s.wav"
    wavfile.write(path, 48000, st)
    w = AudioFileWaveform(path)
    want = np.mean(st.astype(np.float32) / 32768.0, axis=1).astype(np.float32)
    assert w.native_sample_rate == 48000 and w.num_samples == 480 and w.duration_seconds == 0.01
    dt = 1 / 48000
    t = np.arange(500) * dt + dt * 1e-6                     # index int(t / dt) = n
    got = w.waveform(t, dt)
    assert got.dtype == np.float32
    assert np.array_equal(got[:480], want) and np.all(got[480:] == 0)
    right = AudioFileWaveform(path, channel=1, amplitude=2.0, start_time=0.001, duration=0.002)
    seg = st[48:48 + 96, 1].astype(np.float32) / 32768.0
    assert np.array_equal(right.waveform(t[:96], dt), (2.0 * seg).astype(np.float32))
    looped = AudioFileWaveform(path, channel=0, loop=True)
    assert np.array_equal(looped.waveform(t, dt)[480:], looped.waveform(t, dt)[:20])
    u8 = (rng.integers(0, 256, 100)).astype(np.uint8)
    wavfile.write(tmp_path / "u.wav", 48000, u8)
    assert np.array_equal(AudioFileWaveform(tmp_path / "u.wav")._samples, (u8.astype(np.float32) - 128) / 128.0)


def test_resampling_and_errors(tmp_path):
    x = np.sin(2 * np.pi * 1000 * np.arange(441) / 44100).astype(np.float32)
    wavfile.write(tmp_path / "f.wav", 44100, x)
    w = AudioFileWaveform(tmp_path / "f.wav")
    dt = 1 / 96000
    from scipy import signal
    res = signal.resample(x, int(441 * 96000 / 44100)).astype(np.float32)
    t = np.arange(len(res)) * dt + dt * 1e-6
    assert np.array_equal(w.waveform(t, dt), res)
    # the resampled tone keeps its frequency: zero crossings every 48 samples
    assert abs(res[48 * 3] - x[0]) < 0.05
    with pytest.raises(FileNotFoundError):
        AudioFileWaveform(tmp_path / "missing.wav")
    with pytest.raises(ValueError):
        AudioFileWaveform(tmp_path / "f.wav", start_time=1.0)
    wavfile.write(tmp_path / "st.wav", 44100, np.zeros((10, 2), np.int16))
    with pytest.raises(ValueError):
        AudioFileWaveform(tmp_path / "st.wav", channel=2)
