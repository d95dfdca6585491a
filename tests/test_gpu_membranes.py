"""Membrane sources and per-cell source tables on the device vs the oracle
(bit for bit): circular and rectangular patches, pressure and normal-velocity
injection, overlapping membranes, solid cells and an ADE material under a
patch, probes on membrane cells, graph-captured chunks; and every source order
(point, x/y/z plane and membrane sources interleaved on shared cells), which
the sweep and the tables together apply in list order."""

import numpy as np
import pytest

from helpers import Scenario, make_solver

pytestmark = pytest.mark.gpu

FIELDS = ("p", "vx", "vy", "vz")


def _scene(uniform=True):
    shape = (30, 34, 28)
    air = np.ones(shape, bool)
    air[14:17, 10:14, 12:16] = False                 # solid cells inside the first patch
    mat = np.zeros(shape, np.uint8)
    mat[10:14, 15:20, 11:15] = 1                     # ADE material under the patch
    sc = Scenario(shape=shape, pml_depth=4, air=air, material=mat,
                  materials={1: [dict(kind="debye", target="modulus", tau=2e-5, dchi=0.4)]},
                  sources=[("point", (8, 20, 12), 12000.0)],
                  probes=[(15, 17, 13), (20, 8, 20), (12, 16, 13)])
    if not uniform:
        xs = [np.cumsum(np.concatenate([[0.5e-3], 1e-3 * 1.03 ** np.abs(np.arange(n - 1) - n // 2)])) for n in shape]
        sc.coords = tuple(xs)
    return sc


def _membranes(s):
    from strata_fdtd_amd import CircularMembraneSource, GaussianPulse, RectangularMembraneSource
    xc, yc, zc = s.grid.x_coords, s.grid.y_coords, s.grid.z_coords
    wf1 = GaussianPulse(position=(0, 0, 0), frequency=9000.0)
    wf2 = GaussianPulse(position=(0, 0, 0), frequency=15000.0, amplitude=0.5)
    wf3 = GaussianPulse(position=(0, 0, 0), frequency=6000.0)
    return [
        # pressure, normal z, on plane k=13 through the solids and the material
        CircularMembraneSource(center=(xc[15], yc[17], zc[13]), radius=5.2e-3, normal_axis="z", waveform=wf1),
        # overlaps the first one: both add, in order
        CircularMembraneSource(center=(xc[12], yc[16], zc[13] + 2e-4), radius=3.1e-3, normal_axis="z",
                               waveform=wf2, mode=(1, 1)),
        # normal-velocity injection (vy), rectangular
        RectangularMembraneSource(center=(xc[20], yc[8], zc[20]), size=(6e-3, 5e-3), normal_axis="y",
                                  waveform=wf3, injection_type="velocity", mode=(1, 2)),
        RectangularMembraneSource(center=(xc[6], yc[24], zc[9]), size=(4e-3, 7e-3), normal_axis="x",
                                  waveform=wf1, mode=(2, 1)),
    ]


def _oracle_check(s, n, chunk):
    from oracle import cpu
    from oracle.from_solver import problem_from_solver
    pb = problem_from_solver(s, n)
    st, tr = cpu.run(pb, n)
    s.step_n(n, graph_chunk=chunk)
    got = s.get_probe_data()
    for q, name in enumerate(got):
        assert np.array_equal(got[name], tr[:, q]), name
    assert np.abs(tr).max() > 0
    for k in FIELDS:
        a = getattr(s, k)
        assert np.array_equal(a, st[k]), (k, int((a != st[k]).sum()))


@pytest.mark.parametrize("uniform,chunk", [(True, 0), (True, 16), (False, 0)])
def test_membranes_equal_oracle(uniform, chunk):
    s = make_solver(_scene(uniform))
    for m in _membranes(s):
        s.add_source(m)
    print(f"[membranes] uniform={uniform} chunk={chunk}", flush=True)
    _oracle_check(s, 70, chunk)


def test_membrane_then_point_source_on_its_cell_equals_oracle():
    """A point source listed after a pressure membrane on one of its cells is
    added after the membrane's weight on that cell, as in the reference."""
    from strata_fdtd_amd import GaussianPulse
    s = make_solver(_scene())
    m = _membranes(s)[0]
    s.add_source(m)
    s.add_source(GaussianPulse(position=(15, 17, 13), frequency=9000.0))
    _oracle_check(s, 50, 0)


@pytest.mark.parametrize("uniform,chunk", [(True, 0), (True, 12), (False, 0)])
def test_interleaved_source_kinds_equal_oracle(uniform, chunk):
    """Point and x/y plane sources before the first z plane are added in the
    sweep, everything from the z plane on from the per-cell tables. The list
    puts sources of both groups on shared cells (a point on the z plane, a
    y plane crossing it and a membrane, probes on those cells), so any
    reordering changes float32 roundings the oracle sees."""
    from strata_fdtd_amd import GaussianPulse
    sc = _scene(uniform)
    sc.sources = []
    s = make_solver(sc)
    mem = _membranes(s)
    plane = lambda ax, idx, f: GaussianPulse(position={"axis": ax, "index": idx}, frequency=f, source_type="plane")
    for src in (GaussianPulse(position=(15, 17, 20), frequency=11000.0),   # sweep
                plane(0, 7, 8000.0),                                      # sweep
                plane(2, 20, 7000.0),                                     # table from here on
                GaussianPulse(position=(15, 17, 20), frequency=5000.0),   # same cell as the first point
                mem[0],
                plane(1, 17, 13000.0),                                    # crosses the z plane and the membrane
                GaussianPulse(position=(15, 17, 13), frequency=9000.0),   # on the membrane
                plane(2, 13, 6000.0)):                                    # the membrane's plane
        s.add_source(src)
    for q, p in enumerate(((15, 17, 20), (15, 17, 13), (7, 5, 20), (3, 17, 2))):
        s.add_probe(f"extra{q}", p)
    print(f"[sources] interleaved uniform={uniform} chunk={chunk}", flush=True)
    _oracle_check(s, 60, chunk)


def test_membrane_weights_follow_geometry_changes():
    s = make_solver(_scene())
    s.add_source(_membranes(s)[0])
    s.step_n(5)
    air = s.geometry.copy()
    air[15:18, 15:19, 13] = False                    # close part of the patch
    s.set_geometry(air)
    s.reset()
    _oracle_check(s, 40, 0)


def test_membrane_tables_are_validated_at_the_c_abi():
    """fdtd_ctx_set_membranes refuses tables that would index outside the grid,
    the entries, the sources or the probes (no kernel runs with them)."""
    import ctypes as C
    from strata_fdtd_amd import _native
    s = make_solver(_scene())
    s._prepare(4)
    ctx = s._ctx
    n = int(np.prod(s.shape))

    def call(cells, estart, src, field, weight, pstart, pidx, n_src=1):
        cells = np.asarray(cells, np.int64)
        estart = np.asarray(estart, np.int32)
        src = np.asarray(src, np.int32)
        field = np.asarray(field, np.int32)
        weight = np.asarray(weight, np.float64)
        pstart = np.asarray(pstart, np.int32)
        pidx = np.asarray(pidx or [0], np.int32)
        wave = s._trace  # any device buffer stands in for the waveform table
        return _native.load().fdtd_ctx_set_membranes(
            C.c_void_p(ctx), len(cells), cells.ctypes.data, estart.ctypes.data, len(src), src.ctypes.data,
            field.ctypes.data, weight.ctypes.data, pstart.ctypes.data, int(pstart[-1]), pidx.ctypes.data, n_src,
            C.c_void_p(wave.data_ptr()), 4)

    assert call([5], [0, 1], [0], [0], [0.5], [0, 0], []) == 0
    assert call([n], [0, 1], [0], [0], [0.5], [0, 0], []) < 0          # cell outside the grid
    assert call([7, 5], [0, 1, 2], [0, 0], [0, 0], [1, 1], [0, 0, 0], []) < 0   # cells not increasing
    assert call([5], [0, 2], [0], [0], [0.5], [0, 0], []) < 0          # entry range beyond the entries
    assert call([5], [0, 1], [1], [0], [0.5], [0, 0], []) < 0          # source row beyond n_src
    assert call([5], [0, 1], [0], [4], [0.5], [0, 0], []) < 0          # field out of range
    assert call([5], [0, 1], [0], [0], [0.5], [0, 1], [99]) < 0        # probe not set
    assert call([], [0], [], [], [], [0], []) == 0                     # removes the tables
