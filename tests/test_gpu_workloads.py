"""The three solve workloads at their real sizes, against the CPU oracle.

* 512^3 enclosure (rigid SDF shell + port + ADE sphere, PML 16, 8 probes):
  bit-exact fields and traces vs the oracle after 64 steps at full size; the
  same scene scaled to 128^3 bit-exact over the full 4000-step pulse; then
  size-independent properties over a longer full-size run: a rerun is bit
  identical (no races), p == 0 in every solid cell, no health flag, and the
  traces scale with the source amplitude.
* 320x320x480 nonuniform resonator plate (PML 16, plane source, 16 probes):
  bit-exact vs the oracle after 64 steps at full size, plus the same
  properties.
* Full length at full size (tests/golden/full_*.npz, written by the CPU
  oracle in float32 and float64 by tests/golden/make_full_length.py): the
  100^3 solve over 1000 steps, the 512^3 enclosure over 4000 steps and the
  plate over 3000 steps give the oracle's probe traces bit for bit and final
  fields whose bytes hash to the oracle's; the error against the float64 run
  is printed and bounded.

Amplitude scaling is exact in float32 only while no value is subnormal: once
gradual underflow rounds a doubled value differently, the two runs carry
independent rounding noise (measured on the CPU oracle at 96^3, 1200 steps:
1.3e-6 to 1.9e-6 relative L2 per probe). The bound used is 1e-4.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _oracle_check(solver, n):
    from oracle import cpu
    from oracle.from_solver import problem_from_solver
    pb = problem_from_solver(solver, n)
    st, tr = cpu.run(pb, n)
    solver.step_n(n, graph_chunk=16)
    got = solver.get_probe_data()
    for q, name in enumerate(got):
        assert np.array_equal(got[name], tr[:, q]), name
    for k in FIELDS:
        a = getattr(solver, k)
        assert np.array_equal(a, st[k]), (k, int((a != st[k]).sum()))


def _properties(build, steps):
    a = build()
    a.step_n(steps)
    ta = a.get_probe_data()
    air = a.geometry
    p_solid_max = float(np.abs(a.p[~air]).max()) if (~air).any() else 0.0
    health = a.check_health()
    del a
    again = build()
    again.step_n(steps)
    tr = again.get_probe_data()
    for name in ta:
        assert np.array_equal(ta[name], tr[name]), f"rerun differs at {name}"
    del again
    b = build()
    for s in b._sources:
        s.amplitude = 2.0 * s.amplitude
    b._dirty_sources = True
    b.step_n(steps)
    tb = b.get_probe_data()
    for name in ta:
        x, y = ta[name].astype(np.float64), tb[name].astype(np.float64)
        assert np.all(np.isfinite(x)) and np.abs(x).max() > 0
        err = np.linalg.norm(y - 2 * x) / np.linalg.norm(2 * x)
        assert err < 1e-4, (name, err)
    return p_solid_max, health


@pytest.mark.slow
def test_enclosure_512_full_size():
    from strata_fdtd_amd import workloads as W
    wl = W.enclosure_512()
    print("[workload] enclosure_512 oracle check (64 steps)", flush=True)
    _oracle_check(wl.build(), 64)
    print("[workload] enclosure_512 properties (1200 steps)", flush=True)
    p_solid, health = _properties(wl.build, 1200)
    assert p_solid == 0.0
    assert health == 0


@pytest.mark.slow
def test_enclosure_scaled_full_pulse_bitexact():
    from strata_fdtd_amd import workloads as W
    wl = W.enclosure_512(n=128, steps=4000)
    print("[workload] enclosure_128 oracle check (4000 steps)", flush=True)
    _oracle_check(wl.build(), wl.steps)


@pytest.mark.slow
def test_resonators_nonuniform_full_size():
    from strata_fdtd_amd import workloads as W
    wl = W.resonators_nu()
    print("[workload] resonators_nu oracle check (64 steps)", flush=True)
    _oracle_check(wl.build(), 64)
    print("[workload] resonators_nu properties (600 steps)", flush=True)
    p_solid, health = _properties(wl.build, 600)
    assert p_solid == 0.0
    assert health == 0


@pytest.mark.slow
@pytest.mark.parametrize("name", ["plain_100", "resonators_nu", "enclosure_512"])
def test_full_length_equals_oracle(name):
    if not os.path.exists(os.path.join(GOLDEN, f"full_{name}.npz")):
        pytest.skip(f"no full-length fixture for {name}")
    import bench
    print(f"[workload] {name} full length", flush=True)
    r = bench.full_length_parity(name)
    print(f"{name}: {r}", flush=True)
    assert r["trace_bitexact"] and r["fields_bitexact"], r
    # float32 against the float64 run of the same oracle over the whole run
    # (bit-exact above, so these are the float32 oracle's own errors): the
    # traces (the pulse and its reflections) within 1e-4 in the open and
    # absorbing scenes, 1e-3 in the enclosure, a resonant rigid cavity that
    # keeps its energy for the 4000 steps and accumulates rounding in its
    # ringing (2.6e-4 measured); the final pressure, in the open scenes mostly
    # absorbed and small against the rounding the run accumulated (1.2e-3 at
    # 100^3 after 1000 steps), within 1e-2
    trace_bound = {"enclosure_512": 1e-3}.get(name, 1e-4)
    assert r["trace_rel_l2_vs_f64"] < trace_bound and r["p_sub_rel_l2_vs_f64"] < 1e-2, r
