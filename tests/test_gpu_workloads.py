"""The three solve workloads at their real sizes, against the CPU oracle.

* 512^3 enclosure (rigid SDF shell + port + ADE sphere, PML 16, 8 probes):
  bit-exact fields and traces vs the oracle after 8 steps at full size, then
  size-independent properties over a longer run: exact linearity in the
  source amplitude, p == 0 in every solid cell, all probe samples finite.
* 320x320x480 nonuniform resonator plate (PML 16, plane source, 16 probes):
  bit-exact vs the oracle after 8 steps at full size, plus the same
  properties.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")


def _oracle_check(solver, n):
    from oracle import cpu
    from oracle.from_solver import problem_from_solver
    pb = problem_from_solver(solver, n)
    st, tr = cpu.run(pb, n)
    solver.step_n(n, graph_chunk=4)
    got = solver.get_probe_data()
    for q, name in enumerate(got):
        assert np.array_equal(got[name], tr[:, q]), name
    for k in FIELDS:
        a = getattr(solver, k)
        assert np.array_equal(a, st[k]), (k, int((a != st[k]).sum()))


def _amplitude_linearity(build, steps):
    """Doubling the source amplitude doubles every sample (power-of-two scaling
    commutes with rounding): checked to rel-L2 1e-12 over `steps` steps."""
    a = build()
    a.step_n(steps)
    ta = a.get_probe_data()
    air = a.geometry
    p_solid_max = float(np.abs(a.p[~air]).max()) if (~air).any() else 0.0
    health = a.check_health()
    del a
    b = build()
    for s in b._sources:
        s.amplitude = 2.0 * s.amplitude
    b._dirty_sources = True
    b.step_n(steps)
    tb = b.get_probe_data()
    for name in ta:
        x, y = ta[name].astype(np.float64), tb[name].astype(np.float64)
        assert np.all(np.isfinite(x))
        err = np.linalg.norm(y - 2 * x) / max(np.linalg.norm(2 * x), 1e-30)
        assert err < 1e-12, (name, err)
    return p_solid_max, health


@pytest.mark.slow
def test_enclosure_512_full_size():
    from strata_fdtd_amd import workloads as W
    wl = W.enclosure_512()
    print("[workload] enclosure_512 oracle check (8 steps)", flush=True)
    _oracle_check(wl.build(), 8)
    print("[workload] enclosure_512 linearity (1200 steps)", flush=True)
    p_solid, health = _amplitude_linearity(wl.build, 1200)
    assert p_solid == 0.0
    assert health == 0


@pytest.mark.slow
def test_resonators_nonuniform_full_size():
    from strata_fdtd_amd import workloads as W
    wl = W.resonators_nu()
    print("[workload] resonators_nu oracle check (8 steps)", flush=True)
    _oracle_check(wl.build(), 8)
    print("[workload] resonators_nu linearity (600 steps)", flush=True)
    p_solid, health = _amplitude_linearity(wl.build, 600)
    assert p_solid == 0.0
    assert health == 0
