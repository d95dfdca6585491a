"""Mur ABC and radiation-impedance boundaries on the device
(``k_boundary_plane``, after the sweep and the ADE fix-up, before the table
sources and the probes) against the oracle, whose boundary planes are pinned
bit for bit to vectors the reference's own python classes wrote
(tests/test_boundary_planes.py). Whole runs are bit-exact in every field and
probe trace: every axis and face, fixed and pipe-radius reflection, with a
PML before the planes, geometry, ADE materials and a nonuniform grid; sources
and probes on the face planes and on the planes next to them; chunked runs,
graph replay and reset() continue or restart the planes' state like one run.
"""

import numpy as np
import pytest

from helpers import make_solver, oracle_problem, seeded_scenario
from oracle import cpu

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")


def _boundaries(specs):
    import strata_fdtd_amd as sfa
    out = []
    for s in specs:
        if s[0] == "abc":
            out.append(sfa.ABCFirstOrder(axis=s[1]))
        else:
            out.append(sfa.RadiationImpedance(axis=s[1], side=s[2], reflection_coeff=s[3], pipe_radius=s[4]))
    return out


def _on_faces(sc):
    """Sources and probes on the face planes and next to them."""
    nx, ny, nz = sc.shape
    sc.sources = list(sc.sources) + [("point", (0, ny // 2, nz // 3), 7000.0),
                                     ("point", (nx - 2, 2, nz // 2), 5000.0),
                                     ("plane", 1, ny - 1, 6000.0)]
    sc.probes = list(sc.probes) + [(0, ny // 2, nz // 3), (nx - 1, ny // 3, nz // 2), (1, 2, 0),
                                   (nx // 2, 0, nz - 1), (nx - 2, ny - 2, nz - 2)]
    return sc


def _compare(sc, specs, n, chunks=None, graph_chunk=0, label=""):
    s = make_solver(sc)
    bnds = _boundaries(specs)
    for b in bnds:
        s.add_boundary(b)
    print(f"[boundary planes] {label} {sc.shape} {specs} n={n}", flush=True)
    if chunks:
        for c in chunks:
            s.step_n(c)
    else:
        s.step_n(n, graph_chunk=graph_chunk)
    pb = oracle_problem(sc, n)
    pb.boundaries = [pl for b in bnds for pl in b.planes()]
    st, tr = cpu.run(pb, n)
    got = s.get_probe_data()
    for q in range(len(sc.probes)):
        assert np.array_equal(got[f"probe{q}"], tr[:, q]), (label, q)
    for k in FIELDS:
        assert np.array_equal(getattr(s, k), st[k]), (label, k)
    assert np.abs(tr).max() > 0
    return s


CASES = [
    ("mur_all", dict(seed=61, pml=0, geometry=True), [("abc", "all")]),
    ("mur_xz_pml", dict(seed=62, pml=4, geometry=True), [("abc", ("z", "x"))]),
    ("radiation", dict(seed=63, pml=0, geometry=True, plane=True),
     [("rad", "x", "high", 0.7, None), ("rad", "y", "low", None, 0.01), ("rad", "z", "low", 0.0, None)]),
    ("mixed_ade", dict(seed=64, pml=3, geometry=True, ade=True),
     [("abc", "y"), ("rad", "x", "low", 0.45, None), ("rad", "z", "high", 1.0, None)]),
    ("mur_nonuniform", dict(seed=65, pml=0, geometry=True, nonuniform=True), [("abc", "all")]),
]


@pytest.mark.parametrize("name, kw, specs", CASES, ids=[c[0] for c in CASES])
def test_boundary_planes_bitexact(oracle_lib, name, kw, specs):
    kw = dict(kw)
    seed = kw.pop("seed")
    sc = _on_faces(seeded_scenario(seed, shape=(24, 20, 30), **kw))
    _compare(sc, specs, 70, label=name)


def test_boundary_state_across_chunks_graphs_and_reset(oracle_lib):
    """The planes' previous-step state carries across step_n calls and graph
    replays like one run; reset() starts it over."""
    sc = _on_faces(seeded_scenario(66, shape=(22, 18, 26), pml=0, geometry=True))
    specs = [("abc", "all"), ("rad", "y", "high", 0.6, None)]
    _compare(sc, specs, 64, chunks=[5, 17, 1, 41], label="chunks")
    _compare(sc, specs, 64, graph_chunk=8, label="graph")
    s = make_solver(sc)
    for b in _boundaries(specs):
        s.add_boundary(b)
    s.step_n(30)
    first = {k: getattr(s, k).copy() for k in FIELDS}
    s.reset()
    s.step_n(30)
    for k in FIELDS:
        assert np.array_equal(getattr(s, k), first[k]), k


def test_pml_after_boundary_plane_is_refused():
    import strata_fdtd_amd as sfa
    s = sfa.FDTDSolver(grid=sfa.UniformGrid((16, 16, 16), 1e-3), backend="mi355x")
    s.add_boundary(sfa.ABCFirstOrder("x"))
    with pytest.raises(NotImplementedError):
        s.add_boundary(sfa.PML(depth=3))
