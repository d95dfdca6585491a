"""Microphones recorded on the device vs the oracle applied to the solver's
own fields after every step (bit for bit, float32 waveforms), plus the
reference's recording tests (tests/test_microphone.py of the reference:
sample counts, time axis, clear, reset)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(shape=(24, 26, 28), dx=2e-3, geometry=True):
    from strata_fdtd_amd import PML, FDTDSolver, GaussianPulse
    s = FDTDSolver(shape=shape, resolution=dx, c=343.0, rho=1.2, backend="mi355x")
    if geometry:
        air = np.ones(shape, bool)
        air[14:17, 4:20, 10:13] = False
        s.set_geometry(air)
    s.add_boundary(PML(depth=4))
    s.add_source(GaussianPulse(position=(8, 12, 14), frequency=12000.0))
    return s


PATTERNS = {
    "omni": dict(pattern="omni"),
    "cardioid": dict(pattern="cardioid", direction=(1.0, 0.5, 0.0)),
    "figure8": dict(pattern="figure8", direction=(0.0, 1.0, 0.0)),
    "hyper": dict(pattern="hypercardioid", direction=(-0.3, 0.9, 0.2)),
    "custom": dict(pattern=lambda th: 0.7 + 0.3 * np.cos(th), direction=(0.2, 0.2, 0.9)),
}


def _positions(shape, dx):
    rng = np.random.default_rng(4)
    pos = [tuple(rng.uniform(0.0, np.array(shape) - 1.0001) * dx) for _ in range(len(PATTERNS))]
    pos[0] = (0.0, 0.0, 0.0)              # velocity gathers clamp at the low edges
    pos[1] = (5 * dx, 6 * dx, 7 * dx)     # on a grid point
    # numpy float64 positions make the reference sum in float64; Python
    # floats make it sum float32 products of float32-rounded weights
    pos[3] = tuple(float(v) for v in pos[3])
    return pos


def test_device_microphones_equal_oracle_every_step():
    from oracle import microphones as OM
    s = _solver()
    pos = _positions(s.shape, s.dx)
    for (name, kw), p in zip(PATTERNS.items(), pos):
        s.add_microphone(position=p, name=name, **kw)
    want = {n: [] for n in PATTERNS}
    for step in range(40):
        if step < 20:
            s.step()
        else:
            s.step_n(1)
        for (name, kw), p in zip(PATTERNS.items(), pos):
            want[name].append(OM.record(s.p, s.vx, s.vy, s.vz, p, s.dx, pattern=kw["pattern"],
                                        direction=kw.get("direction", (1.0, 0.0, 0.0)), rho=s.rho, c=s.c))
    for name in PATTERNS:
        got = s.microphones[name].get_waveform()
        ref = np.array(want[name], dtype=np.float32)
        assert np.array_equal(got, ref), (name, np.abs(got - ref).max())
        assert np.abs(ref).max() > 0


def test_graph_chunks_equal_single_steps():
    a = _solver()
    b = _solver()
    pos = _positions(a.shape, a.dx)
    for (name, kw), p in zip(PATTERNS.items(), pos):
        a.add_microphone(position=p, name=name, **kw)
        b.add_microphone(position=p, name=name, **kw)
    for _ in range(60):
        a.step()
    b.step_n(60, graph_chunk=16)
    for name in PATTERNS:
        assert np.array_equal(a.microphones[name].get_waveform(), b.microphones[name].get_waveform())
        assert np.array_equal(a.microphones[name].get_time_axis(), b.microphones[name].get_time_axis())


def test_microphone_added_mid_run_and_reference_recording_tests():
    from strata_fdtd_amd import FDTDSolver, Microphone
    s = FDTDSolver(shape=(20, 20, 20), resolution=5e-3, c=343.0, rho=1.2, backend="mi355x")
    mic1 = s.add_microphone(position=(0.03, 0.05, 0.05), name="mic1")
    for _ in range(10):
        s.step()
    assert len(mic1) == 10
    np.testing.assert_allclose(mic1.get_time_axis(), np.arange(10) * s.dt, rtol=1e-10)
    mic2 = s.add_microphone(Microphone(position=(0.07, 0.05, 0.05), name="mic2", pattern="cardioid"))
    s.step_n(5)
    assert len(mic1) == 15 and len(mic2) == 5
    assert mic2.get_time_axis()[0] == mic1.get_time_axis()[10]
    assert mic1.get_waveform().dtype == np.float32
    assert mic1.get_sample_rate() == pytest.approx(1.0 / s.dt)
    auto = s.add_microphone(position=(0.05, 0.05, 0.05))
    assert auto.name == "mic_2"
    with pytest.raises(ValueError):
        s.add_microphone(position=(0.05, 0.05, 0.05), name="mic1")
    mic1.clear()
    assert len(mic1) == 0
    s.step()
    assert len(mic1) == 1 and len(mic2) == 6
    s.reset()
    assert len(mic1) == 0 and len(mic2) == 0 and len(auto) == 0
    s.step_n(3)
    assert len(mic1) == 3


def test_to_wav(tmp_path):
    import wave
    s = _solver(geometry=False)
    mic = s.add_microphone(position=(0.02, 0.025, 0.03), name="m")
    s.step_n(200)
    path = str(tmp_path / "m.wav")
    mic.to_wav(path, sample_rate=44100)
    with wave.open(path, "rb") as fh:
        assert fh.getnchannels() == 1 and fh.getsampwidth() == 2 and fh.getframerate() == 44100
        assert fh.getnframes() == int(200 * 44100 * s.dt)
