"""Pin the CPU oracle (runs without a GPU).

1. Golden vectors written from the reference's own native kernels
   (tests/golden/*.npz, see make_golden.py): bit-exact.
2. Where the reference checkout is present: the reference kernels compiled
   from source (oracle/_ref) on fresh seeded scenarios and single phases:
   bit-exact.
3. Closed forms and physics the mathematics fixes: time step, PML sigma_max,
   Gaussian pulse values, Debye/Lorentz recurrences, the 1D plane-wave special
   case against a textbook leapfrog, wave speed, rigid-cavity mode frequency,
   p = 0 in solids, energy in a closed rigid box.
4. Planted mistakes (in the oracle's inputs, on the CPU) that those checks catch.
"""

import ast
import glob
import os

import numpy as np
import pytest

from helpers import Scenario, oracle_problem, rel_l2, seeded_scenario
from oracle import cpu, tables as T

HERE = os.path.dirname(os.path.abspath(__file__))
# step fixtures (the voxel-mask fixture is read by tests/test_voxel_golden.py,
# the full-length ones by the GPU workload tests, the reference's recorded
# runs by tests/test_organ_pipes.py, tests/test_simple_pulse.py and their GPU
# counterparts)
GOLDEN = sorted(p for p in glob.glob(os.path.join(HERE, "golden", "*.npz"))
                if not os.path.basename(p).startswith(("voxel_", "full_", "organ_", "simple_pulse", "boundary_", "material_")))
FIELDS = ("p", "vx", "vy", "vz")


# ---------------------------------------------------------------- 1. golden
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_equals_reference_golden_vectors(oracle_lib, path):
    g = np.load(path, allow_pickle=False)
    kw = dict(ast.literal_eval(str(g["kwargs"])))  # sorted (key, value) pairs written by make_golden.py
    sc = seeded_scenario(int(g["seed"]), shape=tuple(int(v) for v in g["shape"]), **kw)
    steps = int(g["steps"])
    state, trace = cpu.run(oracle_problem(sc, steps), steps)
    assert np.array_equal(trace, g["trace"])
    for k in FIELDS:
        assert np.array_equal(state[k], g[k]), k


def test_golden_vectors_present():
    assert len(GOLDEN) >= 4


# ------------------------------------------------- 2. reference build, fresh
needs_ref = pytest.mark.skipif(not cpu.have_ref(), reason="oracle/_ref not built (no reference checkout)")


@needs_ref
@pytest.mark.parametrize("seed,kw", [
    (101, dict(geometry=True, pml=4, plane=True)),
    (102, dict(geometry=True, pml=4, ade=True)),
    (103, dict(geometry=True, pml=4, nonuniform=True)),
    (104, dict(geometry=False, pml=0)),
    (105, dict(geometry=True, pml=4, ade=True, nonuniform=True)),
])
def test_oracle_equals_reference_build(oracle_lib, seed, kw):
    sc = seeded_scenario(seed, shape=(22, 19, 32), **kw)
    pb = oracle_problem(sc, 35)
    a, ta = cpu.run(pb, 35)
    b, tb = cpu.run(pb, 35, impl="ref")
    assert np.array_equal(ta, tb)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_single_phases_equal_reference(oracle_lib):
    rng = np.random.default_rng(5)
    nx, ny, nz = 9, 11, 13
    f = {k: rng.standard_normal((nx, ny, nz)).astype(np.float32) for k in FIELDS}
    air = (rng.random((nx, ny, nz)) > 0.2).astype(np.uint8)
    a = {k: v.copy() for k, v in f.items()}
    b = {k: v.copy() for k, v in f.items()}
    cpu.phase("velocity_f32", a["p"], a["vx"], a["vy"], a["vz"], nx, ny, nz, -0.01, None, None, None)
    cpu.phase("update_velocity", b["p"], b["vx"], b["vy"], b["vz"], nx, ny, nz, -0.01, impl="ref")
    cpu.phase("rigid_f32", a["vx"], a["vy"], a["vz"], air, nx, ny, nz)
    cpu.phase("apply_rigid", b["vx"], b["vy"], b["vz"], air, nx, ny, nz, impl="ref")
    cpu.phase("pressure_f32", a["p"], a["vx"], a["vy"], a["vz"], air, nx, ny, nz, -30.0, None, None, None)
    cpu.phase("update_pressure", b["p"], b["vx"], b["vy"], b["vz"], air, nx, ny, nz, -30.0, impl="ref")
    dec = [np.exp(-rng.random(n).astype(np.float32)) for n in (nx, ny, nz)]
    cpu.phase("pml_velocity_f32", a["vx"], a["vy"], a["vz"], *dec, nx, ny, nz)
    cpu.phase("apply_pml_velocity", b["vx"], b["vy"], b["vz"], *dec, nx, ny, nz, impl="ref")
    cpu.phase("pml_pressure_f32", a["p"], *dec, nx, ny, nz)
    cpu.phase("apply_pml_pressure", b["p"], *dec, nx, ny, nz, impl="ref")
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    da = np.zeros((nx, ny, nz), np.float32)
    db = np.zeros((nx, ny, nz), np.float32)
    cpu.phase("divergence_f32", da, a["vx"], a["vy"], a["vz"], nx, ny, nz, 1000.0, None, None, None)
    cpu.phase("compute_divergence", db, b["vx"], b["vy"], b["vz"], nx, ny, nz, 1000.0, impl="ref")
    assert np.array_equal(da, db)
    ic = [(950 + 100 * rng.random(n)).astype(np.float32) for n in (nx, ny, nz)]
    ifc = [(950 + 100 * rng.random(n - 1)).astype(np.float32) for n in (nx, ny, nz)]
    cpu.phase("velocity_f32", a["p"], a["vx"], a["vy"], a["vz"], nx, ny, nz, -0.01, *ifc)
    cpu.phase("update_velocity_nonuniform", b["p"], b["vx"], b["vy"], b["vz"], nx, ny, nz, -0.01, *ifc,
              impl="ref")
    cpu.phase("pressure_f32", a["p"], a["vx"], a["vy"], a["vz"], air, nx, ny, nz, -30.0, *ic)
    cpu.phase("update_pressure_nonuniform", b["p"], b["vx"], b["vy"], b["vz"], air, nx, ny, nz, -30.0, *ic,
              impl="ref")
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_pml_decay_numpy_vs_reference_expf(oracle_lib):
    """The oracle's decay table equals the native initialize_pml output bit for bit."""
    import ctypes as C
    dt = T.time_step(1e-3, 343.0)
    sig = T.pml_profile_uniform(64, 10, 3, T.pml_sigma_max(10, 3, 343.0, 1e-3))
    ours = T.pml_decay(sig, dt)
    out = np.zeros(64, np.float32)
    lib = cpu._lib("ref")
    fp = C.POINTER(C.c_float)
    lib.ref_initialize_pml(sig.ctypes.data_as(fp), None, None, 64, 2, 2, C.c_float(dt),
                           out.ctypes.data_as(fp), None, None)
    assert ours.dtype == np.float32
    assert np.array_equal(ours, out)
    assert np.all(ours[10:-10] == 1.0)


# ------------------------------------------------------------ 3. closed forms
def test_time_step_closed_form():
    assert T.time_step(1e-3, 343.0) == pytest.approx(1.599075e-6, rel=1e-6)


def test_pml_sigma_max_closed_form():
    assert T.pml_sigma_max(10, 3, 343.0, 1e-3) == pytest.approx(947_744, rel=1e-6)
    assert T.pml_sigma_max(16, 3, 343.0, 1e-3) == pytest.approx(592_340, rel=1e-6)


def test_gaussian_pulse_closed_form():
    f = 1000.0
    sigma = 1.0 / (np.pi * 2 * f)
    t0 = 4 * sigma
    assert T.gaussian_pulse(np.array([t0]), f)[0] == pytest.approx(0.0, abs=1e-15)
    tq = t0 + 0.25 / f
    want = np.exp(-((0.25 / f) ** 2) / (2 * sigma**2))
    assert T.gaussian_pulse(np.array([tq]), f)[0] == pytest.approx(want, rel=1e-12)


def test_debye_fixed_point_and_lorentz_frequency():
    dt = 1.6e-6
    al, be = T.debye_coefficients(2e-5, 0.4, dt)
    j = 0.0
    for _ in range(20000):
        j = al * j + be * 1.0
    assert j == pytest.approx(0.4, rel=1e-9)          # J -> dchi * source
    w0, g = 2 * np.pi * 3000, 200.0
    a, b, _d = T.lorentz_coefficients(w0, g, 0.5, dt)
    r = np.roots([1, -a, -b])                          # J_{n+1} = a J_n + b J_{n-1}
    wd = abs(np.angle(r[0])) / dt
    assert wd == pytest.approx(np.sqrt(w0**2 - (g / 2) ** 2), rel=1e-3)
    assert abs(r[0]) == pytest.approx(np.exp(-g * dt / 2), rel=1e-6)


def _leapfrog_1d(p0, n_steps, cv, cp):
    """Textbook 1D staggered leapfrog with rigid ends (float64)."""
    p = p0.astype(np.float64).copy()
    v = np.zeros_like(p)
    for _ in range(n_steps):
        v[:-1] += cv * (p[1:] - p[:-1])
        div = v.copy()
        div[1:] -= v[:-1]
        p += cp * div
    return p


def test_plane_wave_reduces_to_1d_leapfrog(oracle_lib):
    nx, ny, nz = 40, 5, 6
    sc = Scenario(shape=(nx, ny, nz))
    pb = oracle_problem(sc, 60)
    st = cpu.new_state(sc.shape, np.float64)
    x = np.arange(nx)
    p0 = np.exp(-0.5 * ((x - 15) / 3.0) ** 2)
    st["p"][:] = p0[:, None, None]
    st, _ = cpu.run(pb, 60, state=st, impl="oracle64")
    want = _leapfrog_1d(p0, 60, pb.cv, pb.cp)
    assert np.abs(st["p"] - want[:, None, None]).max() < 1e-12
    assert np.abs(st["vy"]).max() == 0 and np.abs(st["vz"]).max() == 0


def test_planted_sign_error_is_caught(oracle_lib):
    """A flipped velocity coefficient (a plausible mistake) breaks the 1D pin."""
    nx, ny, nz = 40, 5, 6
    sc = Scenario(shape=(nx, ny, nz))
    pb = oracle_problem(sc, 60)
    pb.cv = -pb.cv
    st = cpu.new_state(sc.shape, np.float64)
    x = np.arange(nx)
    p0 = np.exp(-0.5 * ((x - 15) / 3.0) ** 2)
    st["p"][:] = p0[:, None, None]
    st, _ = cpu.run(pb, 60, state=st, impl="oracle64")
    pb2 = oracle_problem(sc, 60)
    want = _leapfrog_1d(p0, 60, pb2.cv, pb2.cp)
    assert np.abs(st["p"] - want[:, None, None]).max() > 1e-3


def test_wave_speed(oracle_lib):
    """Arrival of a pulse at distance d matches d / c within 5% (reference test_fdtd.py:82-131)."""
    n = 80
    sc = Scenario(shape=(n, 24, 24), sources=[("point", (10, 12, 12), 8000.0)],
                  probes=[(10, 12, 12), (60, 12, 12)], pml_depth=6)
    steps = 500
    pb = oracle_problem(sc, steps)
    _, tr = cpu.run(pb, steps)
    dt = T.time_step(1e-3, 343.0)
    lag = np.argmax(np.abs(tr[:, 1])) - np.argmax(np.abs(tr[:, 0]))
    c_meas = 50e-3 / (lag * dt)
    assert c_meas == pytest.approx(343.0, rel=0.05)


def test_rigid_cavity_fundamental_mode(oracle_lib):
    """Closed rigid box: the lowest axial mode is c / (2 Lx) within 2% (reference test_fdtd.py:251-305)."""
    nx = 60
    sc = Scenario(shape=(nx, 4, 4), probes=[(3, 2, 2)])
    steps = 6000
    pb = oracle_problem(sc, steps)
    st = cpu.new_state(sc.shape)
    st["p"][5, :, :] = 1.0
    _, tr = cpu.run(pb, steps, state=st)
    dt = T.time_step(1e-3, 343.0)
    spec = np.abs(np.fft.rfft(tr[:, 0] - tr[:, 0].mean(), n=1 << 16))
    freqs = np.fft.rfftfreq(1 << 16, dt)
    f1 = 343.0 / (2 * nx * 1e-3)
    band = (freqs > 0.5 * f1) & (freqs < 1.5 * f1)
    f_meas = freqs[band][np.argmax(spec[band])]
    assert f_meas == pytest.approx(f1, rel=0.02)


def test_pressure_zero_in_solids_and_energy_in_closed_box(oracle_lib):
    sc = seeded_scenario(9, shape=(20, 22, 24), pml=0, geometry=True)
    sc.sources = []
    pb = oracle_problem(sc, 1)
    st = cpu.new_state(sc.shape, np.float64)
    st["p"][10, 11, 12] = 1.0
    st["p"][sc.air == 0] = 0.0
    dt = T.time_step(1e-3, 343.0)
    rho, c = 1.204, 343.0

    def energy(s):
        air = sc.air
        return (0.5 * np.sum(s["p"][air] ** 2) / (rho * c**2)
                + 0.5 * rho * np.sum((s["vx"] ** 2 + s["vy"] ** 2 + s["vz"] ** 2)[air]))

    es = []
    for _ in range(40):
        st, _ = cpu.run(pb, 10, state=st, impl="oracle64")
        assert np.all(st["p"][~sc.air] == 0.0)
        es.append(energy(st))
    es = np.array(es)
    assert (es.max() - es.min()) / es.mean() < 0.1
    _ = dt


def test_f32_oracle_close_to_f64(oracle_lib):
    sc = seeded_scenario(4, shape=(30, 28, 26), pml=4, geometry=True)
    pb = oracle_problem(sc, 200)
    _, t32 = cpu.run(pb, 200)
    _, t64 = cpu.run(pb, 200, impl="oracle64")
    assert rel_l2(t32, t64) < 1e-4


def test_ade_run_refuses_continuation(oracle_lib):
    """The ADE pole accumulators start at zero in every oracle call, so
    splitting an ADE run over calls changes its result: the oracle refuses a
    continuation (step0 > 0) of a problem with poles and accepts one without."""
    sc = seeded_scenario(12, shape=(20, 18, 24), pml=3, geometry=True, ade=True)
    pb = oracle_problem(sc, 40)
    assert pb.poles
    whole, tw = cpu.run(pb, 40)
    st, t1 = cpu.run(pb, 20)
    with pytest.raises(ValueError, match="ADE poles"):
        cpu.run(pb, 20, state=st, step0=20)
    # what a continuation would give: the poles restarted at step 20
    pb_plain = oracle_problem(seeded_scenario(12, shape=(20, 18, 24), pml=3, geometry=True), 40)
    a, _ = cpu.run(pb_plain, 20)
    a, _ = cpu.run(pb_plain, 20, state=a, step0=20)
    b, _ = cpu.run(pb_plain, 40)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert np.abs(whole["p"]).max() > 0 and np.array_equal(tw[:20], t1)


@needs_ref
def test_oracle_equals_reference_with_long_lists(oracle_lib):
    """12 point sources (two cells with two each), 100 probes, 10 poles in
    three materials: the oracle equals the reference kernels bit for bit (the
    scene the GPU limits test runs)."""
    from test_gpu_limits import _many
    sc = _many(31)
    pb = oracle_problem(sc, 25)
    a, ta = cpu.run(pb, 25)
    b, tb = cpu.run(pb, 25, impl="ref")
    assert np.array_equal(ta, tb)
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k


@needs_ref
def test_ade_phases_equal_reference(oracle_lib):
    """The oracle's per-phase ADE functions equal the reference's own
    update_ade_* / apply_ade_* kernels (ade.cpp:25-473) bit for bit, Debye and
    Lorentz, density and modulus, poles out of field order."""
    from helpers import ade_phase_case
    f, mat, poles, rho_inf, K_inf = ade_phase_case(9)
    nx, ny, nz = mat.shape
    a = {k: v.copy() for k, v in f.items()}
    b = {k: v.copy() for k, v in f.items()}
    n = len(poles)
    for impl, d in (("oracle", a), ("ref", b)):
        cpu.phase("update_ade_density_debye", d["Jd"], d["p"], mat, poles, n, nx, ny, nz, impl=impl)
        cpu.phase("update_ade_density_lorentz", d["Jl"], d["Jlp"], d["p"], mat, poles, n, nx, ny, nz, impl=impl)
        cpu.phase("apply_ade_velocity_correction", d["vx"], d["vy"], d["vz"], d["Jd"], d["Jl"], mat, poles, n,
                  rho_inf, 1.6e-6, 1000.0, nx, ny, nz, impl=impl)
        cpu.phase("update_ade_modulus_debye", d["Jd"], d["div"], mat, poles, n, nx, ny, nz, impl=impl)
        cpu.phase("update_ade_modulus_lorentz", d["Jl"], d["Jlp"], d["div"], mat, poles, n, nx, ny, nz, impl=impl)
        cpu.phase("apply_ade_pressure_correction", d["p"], d["Jd"], d["Jl"], mat, poles, n, K_inf, 1.6e-6, nx, ny, nz,
                  impl=impl)
    for k in f:
        assert np.array_equal(a[k], b[k]), k
        if k != "div":
            assert not np.array_equal(a[k], f[k]), k   # every output array was touched
