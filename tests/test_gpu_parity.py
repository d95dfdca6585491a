"""GPU parity of the MI355X kernels against the CPU oracle (through the C ABI).

Bar: bit-exact for the per-phase kernels and the fused step on the seeded
scenarios (the kernels keep the reference's operation order and are built
without FMA contraction); rel-L2 <= 1e-5 on the 100^3 workload after 1000
steps (the stated float32 tolerance), plus the float64 error reported.
"""

import numpy as np
import pytest

from helpers import Scenario, make_solver, oracle_problem, rel_l2, seeded_scenario
from oracle import cpu

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")


def _torch():
    import torch
    return torch


def _same(got, want, what):
    got = np.asarray(got)
    assert got.shape == want.shape, what
    if not np.array_equal(got, want):
        diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
        raise AssertionError(f"{what}: {int((diff > 0).sum())} cells differ, max |diff| {diff.max():.3e}, "
                             f"first at {np.unravel_index(int(np.argmax(diff > 0)), diff.shape)}")


def _dev(a):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_phase_kernels_bitexact(oracle_lib):
    """Each per-phase entry point equals the oracle phase on random data."""
    from strata_fdtd_amd import _native
    torch = _torch()
    rng = np.random.default_rng(7)
    nx, ny, nz = 17, 23, 70
    shp = (nx, ny, nz)
    f = {k: rng.standard_normal(shp).astype(np.float32) for k in FIELDS}
    air = (rng.random(shp) > 0.1).astype(np.uint8)
    st = torch.cuda.current_stream().cuda_stream
    # flags from the device builder
    flags = torch.empty(shp, dtype=torch.uint8, device="cuda")
    _native.call("fdtd_build_flags", _native.ptr(_dev(air)), None, _native.ptr(flags), nx, ny, nz, st)
    # velocity (uniform)
    d = {k: _dev(v) for k, v in f.items()}
    h = {k: v.copy() for k, v in f.items()}
    cv = np.float32(-0.00123)
    _native.call("fdtd_update_velocity", *[_native.ptr(d[k]) for k in FIELDS], nx, ny, nz, float(cv), st)
    cpu.phase("velocity_f32", h["p"], h["vx"], h["vy"], h["vz"], nx, ny, nz, float(cv), None, None, None)
    for k in FIELDS:
        _same(d[k].cpu().numpy(), h[k], k)
    # rigid
    _native.call("fdtd_apply_rigid", _native.ptr(d["vx"]), _native.ptr(d["vy"]), _native.ptr(d["vz"]),
                 _native.ptr(flags), nx, ny, nz, st)
    cpu.phase("rigid_f32", h["vx"], h["vy"], h["vz"], air, nx, ny, nz)
    for k in ("vx", "vy", "vz"):
        _same(d[k].cpu().numpy(), h[k], k)
    # pressure
    cp = np.float32(-23.7)
    _native.call("fdtd_update_pressure", *[_native.ptr(d[k]) for k in FIELDS], _native.ptr(flags), nx, ny,
                 nz, float(cp), st)
    cpu.phase("pressure_f32", h["p"], h["vx"], h["vy"], h["vz"], air, nx, ny, nz, float(cp), None, None, None)
    _same(d["p"].cpu().numpy(), h["p"], "p")
    # PML
    dec = [np.exp(-rng.random(n).astype(np.float32)) for n in shp]
    dd = [_dev(x) for x in dec]
    _native.call("fdtd_apply_pml_velocity", _native.ptr(d["vx"]), _native.ptr(d["vy"]), _native.ptr(d["vz"]),
                 *[_native.ptr(x) for x in dd], nx, ny, nz, st)
    _native.call("fdtd_apply_pml_pressure", _native.ptr(d["p"]), *[_native.ptr(x) for x in dd], nx, ny, nz, st)
    cpu.phase("pml_velocity_f32", h["vx"], h["vy"], h["vz"], *dec, nx, ny, nz)
    cpu.phase("pml_pressure_f32", h["p"], *dec, nx, ny, nz)
    for k in FIELDS:
        _same(d[k].cpu().numpy(), h[k], k)
    # divergence (uniform and nonuniform)
    div = torch.empty(shp, dtype=torch.float32, device="cuda")
    hdiv = np.empty(shp, np.float32)
    _native.call("fdtd_compute_divergence", _native.ptr(div), _native.ptr(d["vx"]), _native.ptr(d["vy"]),
                 _native.ptr(d["vz"]), nx, ny, nz, 1000.0, None, None, None, st)
    cpu.phase("divergence_f32", hdiv, h["vx"], h["vy"], h["vz"], nx, ny, nz, 1000.0, None, None, None)
    _same(div.cpu().numpy(), hdiv, "div")
    ic = [(900 + 200 * rng.random(n)).astype(np.float32) for n in shp]
    ic_d = [_dev(x) for x in ic]  # keep the device tables alive until the kernels ran
    _native.call("fdtd_compute_divergence", _native.ptr(div), _native.ptr(d["vx"]), _native.ptr(d["vy"]),
                 _native.ptr(d["vz"]), nx, ny, nz, 0.0, *[_native.ptr(x) for x in ic_d], st)
    cpu.phase("divergence_f32", hdiv, h["vx"], h["vy"], h["vz"], nx, ny, nz, 0.0, *ic)
    _same(div.cpu().numpy(), hdiv, "div")
    # nonuniform velocity / pressure
    ifc = [(900 + 200 * rng.random(n - 1)).astype(np.float32) for n in shp]
    ifc_d = [_dev(x) for x in ifc]
    _native.call("fdtd_update_velocity_nonuniform", *[_native.ptr(d[k]) for k in FIELDS], nx, ny, nz,
                 float(cv), *[_native.ptr(x) for x in ifc_d], st)
    cpu.phase("velocity_f32", h["p"], h["vx"], h["vy"], h["vz"], nx, ny, nz, float(cv), *ifc)
    _native.call("fdtd_update_pressure_nonuniform", *[_native.ptr(d[k]) for k in FIELDS], _native.ptr(flags),
                 nx, ny, nz, float(cp), *[_native.ptr(x) for x in ic_d], st)
    cpu.phase("pressure_f32", h["p"], h["vx"], h["vy"], h["vz"], air, nx, ny, nz, float(cp), *ic)
    for k in FIELDS:
        _same(d[k].cpu().numpy(), h[k], k)


VARIANTS = {
    "plain": dict(geometry=False, pml=0),
    "pml": dict(geometry=False, pml=4),
    "geometry_pml": dict(geometry=True, pml=4),
    "plane_source": dict(geometry=True, pml=3, plane=True),
    "ade": dict(geometry=True, pml=4, ade=True),
    "nonuniform": dict(geometry=True, pml=4, nonuniform=True),
    "nonuniform_ade": dict(geometry=True, pml=3, ade=True, nonuniform=True),
}


@pytest.mark.parametrize("name", sorted(VARIANTS))
# (rows, planes[, stack]): stacks of 2-8 waves share j halos through LDS; with
# ny = 37 the last stack is partly or wholly past the grid edge
@pytest.mark.parametrize("tiling", [(2, 3), (4, 7), (8, 16), (16, 5), (2, 3, 4), (4, 7, 2), (4, 5, 8), (8, 16, 4)])
def test_fused_step_bitexact(oracle_lib, name, tiling):
    print(f"[fused] variant={name} tiling={tiling}", flush=True)
    shape = (26, 37, 70) if tiling[0] != 16 else (22, 40, 130)
    sc = seeded_scenario(11, shape=shape, **VARIANTS[name])
    n = 30
    st_o, tr_o = cpu.run(oracle_problem(sc, n), n)
    s = make_solver(sc, tiling=tiling)
    for _ in range(n // 2):
        s.step()
    s.step_n(n - n // 2, graph_chunk=4)
    tr = s.get_probe_data()
    for q in range(len(sc.probes)):
        _same(tr[f"probe{q}"], tr_o[:, q], f"probe{q}")
    for k in FIELDS:
        _same(getattr(s, k), st_o[k], k)


def test_config1_probe_and_field_parity(oracle_lib):
    """100^3, 1 mm, AIR_20C, PML 10, 1 kHz pulse at (50,50,50), probe (75,50,50), 1000 steps."""
    sc = Scenario(shape=(100, 100, 100), pml_depth=10, sources=[("point", (50, 50, 50), 1000.0)],
                  probes=[(75, 50, 50)])
    n = 1000
    pb = oracle_problem(sc, n)
    st_o, tr_o = cpu.run(pb, n)
    st_64, tr_64 = cpu.run(pb, n, impl="oracle64")
    s = make_solver(sc)
    s.step_n(n)
    tr = s.get_probe_data()["probe0"]
    e_tr = rel_l2(tr, tr_o[:, 0])
    e_p = rel_l2(s.p, st_o["p"])
    print(f"config1: trace relL2={e_tr:.3e} p relL2={e_p:.3e} "
          f"vs fp64: trace {rel_l2(tr, tr_64[:, 0]):.3e} p {rel_l2(s.p, st_64['p']):.3e}")
    assert e_tr <= 1e-5 and e_p <= 1e-5
    assert np.abs(tr_o).max() > 1e-5  # the pulse reached the probe


def test_graph_replays_match_eager(oracle_lib):
    """Several graph replays with eager work in between equal an eager run."""
    sc = seeded_scenario(3, shape=(30, 33, 66), pml=4, geometry=True)
    a = make_solver(sc)
    b = make_solver(sc)
    for _ in range(3):
        a.step_n(24, graph_chunk=8)   # 3 graph replays per call
        b.step_n(24, graph_chunk=0)   # eager launches
        _ = a.p                       # host round trip between replays
    for k in FIELDS:
        _same(getattr(a, k), getattr(b, k), k)
    _same(a.get_probe_data()["probe0"], b.get_probe_data()["probe0"], "probe0")


def test_host_writes_reach_device(oracle_lib):
    """solver.p[i,j,k] = x before a step is honoured (reference-style initial condition)."""
    sc = Scenario(shape=(20, 22, 24), probes=[(10, 11, 12)])
    s = make_solver(sc)
    s.p[10, 11, 12] = 1.0
    s.step()
    pb = oracle_problem(sc, 1)
    st = cpu.new_state(sc.shape)
    st["p"][10, 11, 12] = 1.0
    st, tr = cpu.run(pb, 1, state=st)
    _same(s.p, st["p"], "p")
    assert s.get_probe_data()["probe0"][0] == tr[0, 0]


def test_reset_restarts(oracle_lib):
    sc = seeded_scenario(5, shape=(20, 21, 40), pml=3)
    s = make_solver(sc)
    s.step_n(10)
    first = s.get_probe_data()["probe3"].copy()
    s.reset()
    assert s.step_count == 0 and s.time == 0.0
    s.step_n(10)
    assert np.array_equal(s.get_probe_data()["probe3"], first)


@pytest.mark.parametrize("order", ["linear", "march"])
def test_stream_ceiling_pass_leaves_the_solve_unchanged(oracle_lib, order):
    """The bandwidth-ceiling pass (16-byte accesses in memory order, or the
    sweep's own strips) writes p, vx, vy, vz + flag byte of every cell into
    the idle buffer set (an odd cell count exercises the 4-cell tail of the
    first, ragged rows and k segments the strips of the second) and leaves
    the solve's state, and so every later step, as it was."""
    torch = _torch()
    sc = seeded_scenario(3, shape=(13, 11, 9), pml=3)
    a, b = make_solver(sc), make_solver(sc)
    a.step_n(7)
    b.step_n(7)
    ms = a.stream_ceiling(3, order=order)
    torch.cuda.synchronize()
    assert ms > 0
    cur = a._cur
    for g in range(4):
        assert torch.equal(a._f[cur ^ 1][g], a._f[cur][g] + a._flags.float())
        assert torch.equal(a._f[cur][g], b._f[cur][g])
    a.step_n(20)
    b.step_n(20)
    for name, tr in b.get_probe_data().items():
        assert np.array_equal(a.get_probe_data()[name], tr), name
    for k in ("p", "vx", "vy", "vz"):
        assert np.array_equal(getattr(a, k), getattr(b, k)), k


def test_material_change_rebuilds_ade_cell_list(oracle_lib):
    """The ADE kernels visit a list of the pole-material cells built when the
    materials are set up; moving the materials after that set-up must rebuild
    it: the run equals the oracle of the moved scene."""
    import dataclasses
    sc = seeded_scenario(21, shape=(24, 20, 28), ade=True)
    moved = np.roll(sc.material, (3, -2, 4), axis=(0, 1, 2))
    if sc.air is not None:
        moved[~sc.air] = 0
    sc2 = dataclasses.replace(sc, material=moved)
    n = 40
    st_o, tr_o = cpu.run(oracle_problem(sc2, n), n)
    s = make_solver(sc)
    s._prepare(1)                      # builds the list for the first placement
    s.set_material_region(np.ones(sc.shape, bool), 0)
    for mid in np.unique(moved[moved > 0]):
        s.set_material_region(moved == mid, int(mid))
    s.step_n(n)
    tr = s.get_probe_data()
    for q in range(len(sc.probes)):
        _same(tr[f"probe{q}"], tr_o[:, q], f"probe{q}")
    for k in FIELDS:
        _same(getattr(s, k), st_o[k], k)


def test_autotune_caches_the_fastest_and_pick_tiling_reads_it(oracle_lib, tmp_path, monkeypatch):
    """The autotuner times its candidates on the device, stores the fastest
    for the grid shape, and pick_tiling returns it; the solve stays bit-exact
    under the chosen launch shape."""
    from strata_fdtd_amd import autotune, tuning
    monkeypatch.setenv("STRATA_FDTD_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sc = seeded_scenario(4, shape=(26, 37, 70), pml=3)
    s = make_solver(sc)
    cands = [(2, 1, 1), (2, 2, 1), (4, 4, 1), (4, 7, 2)]
    best, table = autotune.autotune(s, candidates=cands, steps=5)
    assert len(table) == len(cands) and all(t[3] > 0 for t in table)
    assert min(table, key=lambda t: t[3])[:3] == best
    assert tuning.pick_tiling(sc.shape, False, False) == best
    n = 20
    st_o, tr_o = cpu.run(oracle_problem(sc, n), n)
    t = make_solver(sc, tiling=best)
    t.step_n(n)
    for q in range(len(sc.probes)):
        _same(t.get_probe_data()[f"probe{q}"], tr_o[:, q], f"probe{q}")
    for k in FIELDS:
        _same(getattr(t, k), st_o[k], k)


def test_ade_phase_kernels_bitexact(oracle_lib):
    """The per-phase ADE entry points (update_ade_*, apply_ade_*) equal the
    oracle's ADE phases, which equal the reference's ade.cpp kernels
    (tests/test_oracle.py), on random data with Debye and Lorentz density and
    modulus poles in mixed order."""
    from helpers import ade_phase_case
    from strata_fdtd_amd import _native
    torch = _torch()
    f, mat, poles, rho_inf, K_inf = ade_phase_case(9)
    nx, ny, nz = mat.shape
    n = len(poles)
    st = torch.cuda.current_stream().cuda_stream
    d = {k: _dev(v) for k, v in f.items()}
    dm = _dev(mat)
    h = {k: v.copy() for k, v in f.items()}
    P = lambda k: _native.ptr(d[k])  # noqa: E731
    pl = poles.ctypes.data
    _native.call("fdtd_update_ade_density_debye", P("Jd"), P("p"), _native.ptr(dm), pl, n, nx, ny, nz, st)
    _native.call("fdtd_update_ade_density_lorentz", P("Jl"), P("Jlp"), P("p"), _native.ptr(dm), pl, n, nx, ny, nz, st)
    _native.call("fdtd_apply_ade_velocity_correction", P("vx"), P("vy"), P("vz"), P("Jd"), P("Jl"), _native.ptr(dm),
                 pl, n, rho_inf.ctypes.data, 1.6e-6, 1000.0, nx, ny, nz, st)
    _native.call("fdtd_update_ade_modulus_debye", P("Jd"), P("div"), _native.ptr(dm), pl, n, nx, ny, nz, st)
    _native.call("fdtd_update_ade_modulus_lorentz", P("Jl"), P("Jlp"), P("div"), _native.ptr(dm), pl, n, nx, ny, nz,
                 st)
    _native.call("fdtd_apply_ade_pressure_correction", P("p"), P("Jd"), P("Jl"), _native.ptr(dm), pl, n,
                 K_inf.ctypes.data, 1.6e-6, nx, ny, nz, st)
    cpu.phase("update_ade_density_debye", h["Jd"], h["p"], mat, poles, n, nx, ny, nz)
    cpu.phase("update_ade_density_lorentz", h["Jl"], h["Jlp"], h["p"], mat, poles, n, nx, ny, nz)
    cpu.phase("apply_ade_velocity_correction", h["vx"], h["vy"], h["vz"], h["Jd"], h["Jl"], mat, poles, n, rho_inf,
              1.6e-6, 1000.0, nx, ny, nz)
    cpu.phase("update_ade_modulus_debye", h["Jd"], h["div"], mat, poles, n, nx, ny, nz)
    cpu.phase("update_ade_modulus_lorentz", h["Jl"], h["Jlp"], h["div"], mat, poles, n, nx, ny, nz)
    cpu.phase("apply_ade_pressure_correction", h["p"], h["Jd"], h["Jl"], mat, poles, n, K_inf, 1.6e-6, nx, ny, nz)
    for k in f:
        _same(d[k].cpu().numpy(), h[k], k)


@pytest.mark.slow
def test_wide_offsets_beyond_2_30_cells(oracle_lib):
    """A grid above 2^30 cells (1024 x 1040 x 1009 = 1,074,544,640) takes the
    sweep's 64-bit-offset instantiation; nz = 1009 leaves a partial last
    k segment. Bit-exact against the oracle over 4 steps: traces and fields."""
    shape = (1024, 1040, 1009)
    assert np.prod(shape, dtype=np.int64) > (1 << 30)
    print(f"[wide] {shape}: oracle", flush=True)
    sc = seeded_scenario(13, shape=shape, pml=4, geometry=True)
    sc.sources = [("point", (512, 520, 504), 20000.0)]
    sc.probes = [(512, 520, 505), (511, 519, 504), (513, 521, 503), (1000, 1030, 1000)]
    n = 4
    st_o, tr_o = cpu.run(oracle_problem(sc, n), n)
    print(f"[wide] {shape}: device", flush=True)
    s = make_solver(sc)
    s.step_n(n)
    tr = s.get_probe_data()
    for q in range(len(sc.probes)):
        _same(tr[f"probe{q}"], tr_o[:, q], f"probe{q}")
    for k in FIELDS:
        _same(getattr(s, k), st_o[k], k)
    assert np.abs(tr_o[:, 0]).max() > 0
