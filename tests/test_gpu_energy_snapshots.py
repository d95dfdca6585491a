"""Energy tracking, snapshots and waveform objects on the device solver.

compute_energy (core/solver.py:2689-2706): 1/2 sum_air p^2/(rho c^2) dV +
1/2 rho sum_air |v|^2 dV with dV = dx^3. Pinned by closed forms (uniform p,
uniform v) and by a float64 restatement over the oracle's fields; the
reference sums float32 arrays (numpy pairwise), which the device's float64
sum matches to float32 rounding. run(track_energy=True) records (step, time,
energy) after every sample interval, starting with the state before the run
(:2551-2555, :2073-2075); energy_report (:2723-2779) classifies the change:
a closed rigid box holds its energy ("stable"), PML absorbs it
("decaying"), and a Courant number above 1 makes the scheme unstable
("growing", with the drift warning of :2586-2596 when asked for).
enable_snapshots (:2056-2068) keeps (time at the start of step n, p after step
n) and cell-centred velocities every `interval` steps; they must equal the
oracle's fields bit for bit. Any object with .waveform(t, dt) and a point or
plane position is a source (:2413-2433): sampled per step at the accumulated
time, a windowed sine and a chirp give the oracle's fields bit for bit, in
the sweep and in the per-cell tables.
"""

import warnings

import numpy as np
import pytest

from helpers import Scenario, make_solver, oracle_problem, seeded_scenario
from oracle import cpu, tables as T

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")
RHO, C = 1.204, 343.0


def _energy64(state, air, dx, rho=RHO, c=C):
    """Float64 restatement of compute_energy over host fields."""
    p, vx, vy, vz = (np.asarray(state[k], np.float64) for k in FIELDS)
    dV = dx**3
    pe = 0.5 * np.sum(p[air] ** 2) / (rho * c**2) * dV
    ke = 0.5 * rho * np.sum((vx**2 + vy**2 + vz**2)[air]) * dV
    return pe + ke


def _energy_ref32(state, air, dx, rho=RHO, c=C):
    """The reference's own arithmetic: numpy sums of float32 arrays."""
    dV = dx**3
    pe = 0.5 * np.sum(state["p"][air] ** 2) / (rho * c**2) * dV
    v2 = state["vx"] ** 2 + state["vy"] ** 2 + state["vz"] ** 2
    ke = 0.5 * rho * np.sum(v2[air]) * dV
    return float(pe + ke)


def test_compute_energy_closed_forms():
    sc = seeded_scenario(41, shape=(24, 26, 30), pml=0, geometry=True)
    air = sc.air
    n_air = int(air.sum())
    s = make_solver(sc)
    s.p[...] = 1.0
    e = s.compute_energy()
    want = 0.5 * n_air / (RHO * C**2) * 1e-9
    assert abs(e - want) <= 1e-12 * want
    s.p[...] = 0.0
    s.vy[...] = 2.0
    e = s.compute_energy()
    want = 0.5 * RHO * 4.0 * n_air * 1e-9
    assert abs(e - want) <= 1e-12 * want


def test_compute_energy_after_steps_equals_oracle_fields(oracle_lib):
    sc = seeded_scenario(42, shape=(30, 28, 70), pml=4, geometry=True, ade=True)
    n = 60
    st, _ = cpu.run(oracle_problem(sc, n), n)
    s = make_solver(sc)
    s.step_n(n)
    e = s.compute_energy()
    e64 = _energy64(st, sc.air, sc.dx)
    assert e64 > 0
    assert abs(e - e64) <= 1e-12 * e64
    assert abs(e - _energy_ref32(st, sc.air, sc.dx)) <= 1e-5 * e64


def _blob_solver(shape=(28, 30, 32), pml=0, courant=0.95, warn=False):
    import strata_fdtd_amd as sfa
    s = sfa.FDTDSolver(grid=sfa.UniformGrid(shape, 1e-3), c=C, rho=RHO, courant=courant, backend="mi355x",
                       warn_energy_drift=warn)
    if pml:
        s.add_boundary(sfa.PML(depth=pml))
    i, j, k = (n // 2 for n in shape)
    g = np.indices(shape)
    r2 = (g[0] - i) ** 2 + (g[1] - j) ** 2 + (g[2] - k) ** 2
    s.p[...] = np.exp(-r2 / 50.0).astype(np.float32)
    return s


def test_energy_history_and_report(oracle_lib):
    s = _blob_solver()
    p0 = s.p.copy()
    n = 80
    s.run(duration=n * s.dt * (1 - 1e-9), track_energy=True)
    hist = s.get_energy_history()
    assert len(hist) == n + 1
    # oracle fields step by step: the same energies in float64
    sc = Scenario(shape=s.shape)
    pb = oracle_problem(sc, n)
    st = cpu.new_state(s.shape)
    st["p"][...] = p0
    air = np.ones(s.shape, bool)
    t = 0.0
    for m, (step, time, e) in enumerate(hist):
        assert step == m and time == t
        e64 = _energy64(st, air, 1e-3)
        assert abs(e - e64) <= 1e-12 * e64, m
        if m < n:
            st, _ = cpu.run(pb, 1, state=st, step0=m)
            t += s.dt
    rep = s.energy_report()
    assert rep["n_samples"] == n + 1 and rep["initial_energy"] == hist[0][2]
    e = np.array([h[2] for h in hist])
    assert rep["max_energy"] == e.max() and rep["min_energy"] == e.min()
    assert rep["energy_change_percent"] == (e[-1] - e[0]) / e[0] * 100
    # a smooth blob in a closed rigid box: within the 1% band after 80 steps
    # (the staggered-time energy swings by about 5% in between)
    assert rep["conservation_status"] == "stable", rep


def test_energy_report_decaying_with_pml():
    s = _blob_solver(pml=6)
    s.run(duration=400 * s.dt * (1 - 1e-9), track_energy=True, energy_sample_interval=20)
    hist = s.get_energy_history()
    assert [h[0] for h in hist] == list(range(0, 401, 20))
    rep = s.energy_report()
    assert rep["conservation_status"] == "decaying" and rep["energy_change_percent"] < -50, rep


def test_energy_report_growing_and_drift_warning_above_courant_one():
    s = _blob_solver(shape=(24, 24, 24), courant=1.3, warn=True)
    with warnings.catch_warnings(record=True) as caught:
        warnings.simplefilter("always")
        s.run(duration=30 * s.dt * (1 - 1e-9), track_energy=True, energy_sample_interval=10)
    rep = s.energy_report()
    assert np.isfinite(rep["final_energy"])
    assert rep["conservation_status"] == "growing" and rep["energy_change_percent"] > 1.0, rep
    msgs = [str(w.message) for w in caught if issubclass(w.category, UserWarning)]
    assert any(m.startswith("Energy drift detected") and "Status: growing" in m for m in msgs), msgs


def test_snapshots_equal_oracle_fields(oracle_lib):
    sc = seeded_scenario(43, shape=(26, 28, 68), pml=4, geometry=True, plane=True)
    n, every = 23, 5
    s = make_solver(sc)
    s.enable_snapshots(every, capture_velocity=True)
    s.run(duration=n * s.dt * (1 - 1e-9))
    snaps, vsnaps = s.get_snapshots(), s.get_velocity_snapshots()
    steps = list(range(0, n, every))
    assert len(snaps) == len(steps) == len(vsnaps)
    pb = oracle_problem(sc, n)
    st = cpu.new_state(sc.shape)
    times = T.accumulated_times(s.dt, n)
    done = 0
    for (t, p), (tv, cx, cy, cz), m in zip(snaps, vsnaps, steps):
        st, _ = cpu.run(pb, m + 1 - done, state=st, step0=done)
        done = m + 1
        assert t == times[m] and tv == times[m]
        assert np.array_equal(p, st["p"]), m
        for c, k, ax in ((cx, "vx", 0), (cy, "vy", 1), (cz, "vz", 2)):
            v = st[k]
            want = np.zeros_like(v)
            hi = [slice(None)] * 3
            lo = [slice(None)] * 3
            hi[ax], lo[ax] = slice(1, None), slice(None, -1)
            want[tuple(hi)] = 0.5 * (v[tuple(lo)] + v[tuple(hi)])
            first = [slice(None)] * 3
            first[ax] = 0
            want[tuple(first)] = 0.5 * v[tuple(first)]
            assert np.array_equal(c, want), (m, k)


class WindowedSine:
    """A waveform object of the reference's protocol: .waveform(t, dt)."""

    def __init__(self, position, frequency, cycles=4.0, source_type="point", amplitude=1.0):
        self.position = position
        self.frequency = frequency
        self.duration = cycles / frequency
        self.source_type = source_type
        self.amplitude = amplitude

    def waveform(self, t, dt):
        t = np.asarray(t, dtype=np.float64)
        win = np.where(t < self.duration, np.sin(np.pi * t / self.duration) ** 2, 0.0)
        return self.amplitude * win * np.sin(2 * np.pi * self.frequency * t)


class Chirp(WindowedSine):
    def waveform(self, t, dt):
        t = np.asarray(t, dtype=np.float64)
        f = self.frequency * (1.0 + t / self.duration)
        return np.where(t < self.duration, self.amplitude * np.sin(2 * np.pi * f * t), 0.0)


def test_waveform_objects_bitexact(oracle_lib):
    """A windowed sine at a point (sweep), a chirp on an x plane (sweep) and a
    windowed sine on a z plane (per-cell tables), against the oracle fed the
    same waveform objects sampled one step at a time."""
    sc = seeded_scenario(44, shape=(26, 30, 70), pml=4, geometry=True)
    sc.sources = []
    n = 50
    pt = (13, 15, 35)
    srcs = [WindowedSine(pt, 7000.0), Chirp({"axis": 0, "index": 9}, 5000.0, source_type="plane", amplitude=0.5),
            WindowedSine({"axis": 2, "index": 20}, 6000.0, cycles=3.0, source_type="plane", amplitude=0.25)]
    s = make_solver(sc)
    for src in srcs:
        s.add_source(src)
    s.step_n(n)
    pb = oracle_problem(sc, n)
    times = T.accumulated_times(s.dt, n)
    table = [np.array([src.waveform(np.array([t]), s.dt)[0] for t in times]) for src in srcs]
    pb.sources = [(0, int(np.ravel_multi_index(pt, sc.shape)), 0, table[0]),
                  (1, 0, 9, table[1]), (1, 2, 20, table[2])]
    st, tr = cpu.run(pb, n)
    got = s.get_probe_data()
    for q in range(len(sc.probes)):
        assert np.array_equal(got[f"probe{q}"], tr[:, q]), q
    for k in FIELDS:
        assert np.array_equal(getattr(s, k), st[k]), k
    assert np.abs(tr).max() > 0


@pytest.mark.parametrize("shape", [(17, 19, 23), (160, 170, 203)])
def test_energy_sums_tail_and_many_blocks(shape):
    """fdtd_ctx_energy over a cell count that is not a multiple of four (the
    scalar tail) and over more cells than one pass of its fixed grid covers:
    a call repeats to the bit and equals the float64 sums."""
    sc = seeded_scenario(45, shape=shape, pml=0, geometry=True)
    s = make_solver(sc)
    rng = np.random.default_rng(7)
    state = {k: rng.standard_normal(shape).astype(np.float32) for k in FIELDS}
    for k in FIELDS:
        setattr(s, k, state[k])
    e1 = s.compute_energy()
    e2 = s.compute_energy()
    assert e1 == e2
    e64 = _energy64(state, sc.air, sc.dx)
    assert abs(e1 - e64) <= 1e-12 * e64


def test_run_progress_and_callback_equal_plain_run():
    """run(progress=True) (device runs of at most 256 steps between bar
    updates) and run(callback=...) (one step at a time, the callback after
    each with the step index) give the probe traces and fields of a plain run
    bit for bit."""
    sc = seeded_scenario(46, shape=(24, 26, 40), pml=4, geometry=True)
    n = 600
    seen = []
    results = []
    for kw in ({}, {"progress": True}, {"callback": seen.append}):
        s = make_solver(sc)
        s.run(duration=n * s.dt * (1 - 1e-9), **kw)
        assert s.step_count == n
        results.append((s.get_probe_data(), {k: getattr(s, k).copy() for k in FIELDS}))
    assert seen == list(range(n))
    base_tr, base_f = results[0]
    for tr, f in results[1:]:
        for name in base_tr:
            assert np.array_equal(tr[name], base_tr[name]), name
        for k in FIELDS:
            assert np.array_equal(f[k], base_f[k]), k


def test_audio_file_waveform_source_bitexact(oracle_lib, tmp_path):
    """A point source and an x-plane source driven by AudioFileWaveform
    (core/waveforms.py; a WAV file resampled to 1/dt) give the oracle's
    fields and traces bit for bit when the oracle is fed the same samples
    at the accumulated times."""
    from scipy.io import wavfile
    import strata_fdtd_amd as sfa
    rng = np.random.default_rng(9)
    wavfile.write(tmp_path / "a.wav", 48000, (rng.uniform(-1, 1, 400) * 32767).astype(np.int16))
    sc = seeded_scenario(47, shape=(24, 26, 40), pml=4, geometry=True)
    sc.sources = []
    n = 60
    pt = (12, 13, 20)
    s = make_solver(sc)
    a = sfa.AudioFileWaveform(tmp_path / "a.wav", amplitude=0.5)
    b = sfa.AudioFileWaveform(tmp_path / "a.wav", amplitude=0.25, loop=True)

    class _Src:
        def __init__(self, w, position, source_type):
            self.w, self.position, self.source_type = w, position, source_type

        def waveform(self, t, dt):
            return self.w.waveform(t, dt)

    srcs = [_Src(a, pt, "point"), _Src(b, {"axis": 0, "index": 7}, "plane")]
    for src in srcs:
        s.add_source(src)
    s.step_n(n)
    pb = oracle_problem(sc, n)
    times = T.accumulated_times(s.dt, n)
    tabs = [np.array([src.waveform(np.array([t]), s.dt)[0] for t in times], dtype=np.float64) for src in srcs]
    assert np.abs(tabs[0]).max() > 0
    pb.sources = [(0, int(np.ravel_multi_index(pt, sc.shape)), 0, tabs[0]), (1, 0, 7, tabs[1])]
    st, tr = cpu.run(pb, n)
    got = s.get_probe_data()
    for q in range(len(sc.probes)):
        assert np.array_equal(got[f"probe{q}"], tr[:, q]), q
    for k in FIELDS:
        assert np.array_equal(getattr(s, k), st[k]), k


def test_library_materials_bitexact(oracle_lib):
    """Library materials in a solve: a porous absorber (fiberglass, Debye
    density and modulus poles from the JCA fit) and a damping solid (butyl,
    modulus poles), each with its own rho_inf and K_inf in the ADE
    corrections; fields and traces equal the oracle's bit for bit."""
    import strata_fdtd_amd as sfa
    sc = seeded_scenario(48, shape=(24, 20, 30), pml=3, geometry=True, ade=True)
    sc.materials = {}
    sc.material_objects = {1: sfa.get_material("fiberglass_48kg"), 2: sfa.get_material("butyl_rubber")}
    n = 24
    s = make_solver(sc)
    s.step_n(n)
    st, tr = cpu.run(oracle_problem(sc, n), n)
    got = s.get_probe_data()
    for q in range(len(sc.probes)):
        assert np.array_equal(got[f"probe{q}"], tr[:, q], equal_nan=True), q
    for k in FIELDS:
        assert np.array_equal(getattr(s, k), st[k], equal_nan=True), k
    assert np.isfinite(tr).any()
