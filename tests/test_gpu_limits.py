"""Unbounded source, probe and pole lists against the oracle (bit-exact).

The reference keeps plain Python lists of sources and probes and C++ vectors
of poles (core/solver.py:1782-1885, _kernels/include/ade.hpp:44-111). Here the
sweep adds the first 8 point/plane sources itself and the host routes the
rest to the per-cell source tables; the sweep records the first 64 probes and
a small kernel the rest; poles live in a device table. Results must not
depend on where that split falls. Also: table rewrites while steps are still
queued on a solver's non-default stream (probes added between two
asynchronous step_n calls) see every queued step finish first.
"""

import numpy as np
import pytest

from helpers import make_solver, oracle_problem, seeded_scenario
from oracle import cpu

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")


def _same(got, want, what):
    got = np.asarray(got)
    assert got.shape == want.shape, what
    if not np.array_equal(got, want):
        diff = np.abs(got.astype(np.float64) - want.astype(np.float64))
        raise AssertionError(f"{what}: {int((diff > 0).sum())} values differ, max |diff| {diff.max():.3e}")


def _many(seed, shape=(30, 34, 70), n_src=12, n_probe=100):
    sc = seeded_scenario(seed, shape=shape, pml=4, geometry=True, ade=True)
    rng = np.random.default_rng(seed + 100)
    nx, ny, nz = shape
    # a third material with six more poles: ten in all, Debye and Lorentz,
    # density and modulus, interleaved
    sc.material[3:9, 20:27, 40:52] = 3
    sc.material[~sc.air] = 0
    sc.materials[3] = [
        dict(kind="debye", target="density", tau=3e-5, dchi=0.2),
        dict(kind="lorentz", target="modulus", w0=2 * np.pi * 2500, gamma=400.0, dchi=0.3),
        dict(kind="debye", target="modulus", tau=2e-5, dchi=0.25),
        dict(kind="lorentz", target="density", w0=2 * np.pi * 1500, gamma=200.0, dchi=0.15),
        dict(kind="debye", target="density", tau=5e-5, dchi=0.1),
        dict(kind="lorentz", target="modulus", w0=2 * np.pi * 4000, gamma=600.0, dchi=0.2),
    ]
    cells = [tuple(int(v) for v in (rng.integers(5, nx - 5), rng.integers(5, ny - 5), rng.integers(5, nz - 5)))
             for _ in range(n_src - 2)]
    cells += [cells[0], cells[3]]          # two cells get two sources each, in list order
    sc.sources = [("point", c, float(rng.uniform(3000, 9000))) for c in cells]
    sc.probes = [tuple(int(v) for v in (rng.integers(0, nx), rng.integers(0, ny), rng.integers(0, nz)))
                 for _ in range(n_probe - 4)] + [cells[0], cells[9], cells[10], (5, 22, 45)]
    return sc


def test_many_sources_probes_poles_bitexact(oracle_lib):
    sc = _many(31)
    n_pole = sum(len(v) for v in sc.materials.values())
    assert len(sc.sources) == 12 and len(sc.probes) == 100 and n_pole == 10
    n = 40
    print(f"[limits] 12 sources, 100 probes, {n_pole} poles, {n} steps", flush=True)
    st_o, tr_o = cpu.run(oracle_problem(sc, n), n)
    s = make_solver(sc)
    s.step_n(n // 2)
    s.step_n(n - n // 2, graph_chunk=4)
    tr = s.get_probe_data()
    for q in range(len(sc.probes)):
        _same(tr[f"probe{q}"], tr_o[:, q], f"probe{q}")
    for k in FIELDS:
        _same(getattr(s, k), st_o[k], k)
    assert np.abs(tr_o).max() > 0


def test_probe_added_between_queued_steps_on_a_side_stream(oracle_lib):
    """step_n only enqueues; a probe added right after it rewrites the probe
    table, which must wait for the queued steps (a non-blocking side stream,
    as the serving loop uses)."""
    import torch
    sc = seeded_scenario(33, shape=(40, 44, 130), pml=4, geometry=True, ade=True)
    late = (20, 21, 64)
    n1, n2 = 60, 40
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        s = make_solver(sc)
        s.step_n(n1)
        s.add_probe("late", late)
        s.step_n(n2)
    side.synchronize()
    sc_all = seeded_scenario(33, shape=(40, 44, 130), pml=4, geometry=True, ade=True)
    sc_all.probes = list(sc_all.probes) + [late]
    st_o, tr_o = cpu.run(oracle_problem(sc_all, n1 + n2), n1 + n2)
    tr = s.get_probe_data()
    for q in range(len(sc.probes)):
        _same(tr[f"probe{q}"], tr_o[:, q], f"probe{q}")
    _same(tr["late"], tr_o[n1:, len(sc.probes)], "late")
    for k in FIELDS:
        _same(getattr(s, k), st_o[k], k)
