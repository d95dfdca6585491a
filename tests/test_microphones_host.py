"""Microphones without a GPU: the oracle's restatement of the reference's
recording pinned to the reference tests' known answers and to closed forms,
and the product's gather tables / output combination checked against it.
"""

import numpy as np
import pytest

from oracle import microphones as OM
from strata_fdtd_amd.microphones import POLAR_PATTERNS, Microphone


class _Grid:
    """The attributes Microphone._initialize reads."""

    def __init__(self, shape=(20, 20, 20), dx=5e-3, rho=1.2, c=343.0):
        self.shape, self.dx, self.rho, self.c = shape, dx, rho, c
        self.dt = 0.95 * dx / (c * np.sqrt(3))


def _mic(g, pos, **kw):
    m = Microphone(position=pos, **kw)
    m._initialize(g)
    return m


def test_oracle_known_answers_from_reference_tests():
    """tests/test_microphone.py:100-160 of the reference."""
    g = _Grid()
    p = np.zeros(g.shape, np.float32)
    p[5, 5, 5] = 1.0
    assert OM.record(p, p, p, p, (5 * g.dx, 5 * g.dx, 5 * g.dx), g.dx) == pytest.approx(1.0, rel=1e-6)
    p = np.zeros(g.shape, np.float32)
    p[5:7, 5:7, 5:7] = 1.0
    assert OM.record(p, p, p, p, (5.5 * g.dx,) * 3, g.dx) == pytest.approx(1.0, rel=1e-6)
    p = np.zeros(g.shape, np.float32)
    p[6, 5, 5] = 4.0
    assert OM.record(p, p, p, p, (5.25 * g.dx, 5 * g.dx, 5 * g.dx), g.dx) == pytest.approx(1.0, rel=1e-6)


def test_oracle_trilinear_exact_on_linear_fields():
    """Trilinear interpolation reproduces fields linear in the indices; the
    staggered velocity gathers see v_x at i + 1/2 (likewise y, z)."""
    g = _Grid(shape=(16, 18, 20), dx=1e-3)
    i, j, k = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in g.shape], indexing="ij")
    lin = lambda x, y, z: 0.25 + 0.5 * x - 0.125 * y + 0.0625 * z  # noqa: E731
    p = lin(i, j, k).astype(np.float32)
    vx = lin(i + 0.5, j, k).astype(np.float32)
    vy = lin(i, j + 0.5, k).astype(np.float32)
    vz = lin(i, j, k + 0.5).astype(np.float32)
    rng = np.random.default_rng(3)
    for _ in range(50):
        gpos = rng.uniform(1.0, np.array(g.shape) - 2.0)
        pos = tuple(gpos * g.dx)
        gg = OM.grid_position(pos, g.dx)
        want = lin(*gg)
        assert float(OM.interp_pressure(p, gg)) == pytest.approx(want, rel=2e-6)
        for ax, v in enumerate((vx, vy, vz)):
            assert float(OM.interp_velocity(v, gg, ax)) == pytest.approx(want, rel=2e-6)


@pytest.mark.parametrize("pattern", ["cardioid", "supercardioid", "hypercardioid", "figure8", "subcardioid"])
def test_oracle_directional_plane_wave_closed_form(pattern):
    """Uniform p = 1 and v = +-d/Z: on-axis output (1-k) + k = 1, rear
    output 1 - 2k (cardioid rejects the rear)."""
    g = _Grid()
    d = np.array([0.6, 0.0, 0.8])
    Z = g.rho * g.c
    k = POLAR_PATTERNS[pattern]
    p = np.ones(g.shape, np.float32)
    pos = (0.041, 0.047, 0.052)
    for sgn, want in ((1.0, 1.0), (-1.0, 1.0 - 2 * k)):
        v = [np.full(g.shape, sgn * d[a] / Z, np.float32) for a in range(3)]
        out = OM.record(p, *v, pos, g.dx, pattern=pattern, direction=tuple(d), rho=g.rho, c=g.c)
        assert float(out) == pytest.approx(want, abs=1e-6)


def _rand_fields(shape, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(shape).astype(np.float32) for _ in range(4)]


@pytest.mark.parametrize("pattern", ["omni", "cardioid", "figure8", "custom"])
def test_product_manual_record_equals_oracle(pattern):
    g = _Grid(shape=(12, 13, 14), dx=2e-3)
    pat = (lambda th: 0.7 + 0.3 * np.cos(th)) if pattern == "custom" else pattern
    rng = np.random.default_rng(7)
    for t in range(20):
        pos = tuple(rng.uniform(0.0, np.array(g.shape) - 1.0001) * g.dx)
        if t == 0:
            pos = (0.0, 0.0, 0.0)  # velocity gathers clamp at the low edges
        d = tuple(rng.standard_normal(3))
        m = _mic(g, pos, pattern=pat, direction=d)
        f = _rand_fields(g.shape, t)
        m.record(f[0], 0.0, *f[1:])
        want = OM.record(*f, pos, g.dx, pattern=pat, direction=d, rho=g.rho, c=g.c)
        got = m._data[0]
        assert type(got) is type(want) and got == want


@pytest.mark.parametrize("pattern", ["omni", "cardioid", "hypercardioid", "figure8", "custom"])
def test_product_combine_equals_oracle(pattern):
    """The host combination of device records (p, vx, vy, vz) equals the
    reference's per-sample arithmetic, bit for bit, in the reference's type."""
    g = _Grid()
    pat = (lambda th: 0.5 + 0.5 * np.cos(th)) if pattern == "custom" else pattern
    d = (0.3, -0.5, 0.8)
    m = _mic(g, (0.05, 0.05, 0.05), pattern=pat, direction=d)
    rng = np.random.default_rng(11)
    rec = (rng.standard_normal((64, 4)) * 1e-2).astype(np.float32)
    rec[5, 1:] = 0.0  # |v| == 0: theta = 0 branch
    for f64 in ([0, 0, 0, 0], [1, 1, 1, 1], [0, 1, 0, 1]):
        rec_t = [r.astype(np.float64) if f else r for r, f in zip(rec.T, f64)]
        out = m._combine(rec.astype(np.float64), np.array(f64))
        dd = np.asarray(d) / np.linalg.norm(d)
        for q, o in enumerate(out):
            r = [col[q] for col in rec_t]
            if pattern == "omni":
                want = r[0]
            else:
                want = OM.directional(r[0], (r[1], r[2], r[3]), dd, pat, g.rho, g.c)
            assert type(o) is type(want) and float(o) == float(want), (f64, q)


def test_product_gather_tables_match_oracle_corners():
    g = _Grid(shape=(10, 11, 12), dx=1e-3)
    rng = np.random.default_rng(5)
    for t in range(30):
        gp = rng.uniform(0.0, np.array(g.shape) - 1.0001)
        if t < 3:
            gp[t] = 0.0
        pos = tuple(gp * g.dx)
        m = _mic(g, pos)
        idx, w, f64 = m.gather_table(g.shape)
        assert list(f64) == [1, 1, 1, 1]  # numpy float64 positions: float64 weights
        gg = OM.grid_position(pos, g.dx)
        corners, weights = OM.trilinear(*gg)
        assert list(idx[0]) == [np.ravel_multi_index(c, g.shape) for c in corners]
        assert np.array_equal(w[0], np.array(weights, np.float64))
        for ax in range(3):
            h = list(gg)
            h[ax] -= 0.5
            h = [max(0.0, min(v, n - 1.001)) for v, n in zip(h, g.shape)]
            corners, weights = OM.trilinear(*h)
            corners = [tuple(min(c, n - 1) for c, n in zip(cc, g.shape)) for cc in corners]
            assert list(idx[1 + ax]) == [np.ravel_multi_index(c, g.shape) for c in corners]
            assert np.array_equal(w[1 + ax], np.array(weights, np.float64))


def test_microphone_validation():
    g = _Grid()
    with pytest.raises(ValueError):
        _mic(g, (0.15, 0.05, 0.05))
    with pytest.raises(ValueError):
        Microphone(position=(0.05, 0.05, 0.05), pattern="nope")
    with pytest.raises(ValueError):
        Microphone(position=(0.05, 0.05, 0.05), direction=(0, 0, 1), up=(0, 0, 2))
    with pytest.raises(TypeError):
        Microphone(position=(0.05, 0.05, 0.05), pattern=3)


@pytest.mark.parametrize("kind", ["python", "float64", "float32"])
def test_device_gather_arithmetic_equals_oracle(kind):
    """Restates the device gather on the CPU (float32 sum of float32-rounded
    weights, or float64 sum of float64 weights, per the table's flag) and
    compares it with the oracle for positions given as Python floats, numpy
    float64 and numpy float32 scalars."""
    g = _Grid(shape=(24, 26, 28), dx=2e-3)
    rng = np.random.default_rng(2)
    for t in range(200):
        raw = rng.uniform(0.0, np.array(g.shape) - 1.0001) * g.dx
        pos = {"python": tuple(float(v) for v in raw), "float64": tuple(raw),
               "float32": tuple(np.float32(v) for v in raw)}[kind]
        m = _mic(g, pos)
        idx, w, f64 = m.gather_table(g.shape)
        F = [(rng.standard_normal(g.shape) * 1e-4).astype(np.float32) for _ in range(4)]
        gg = OM.grid_position(pos, g.dx)
        for a in range(4):
            flat = F[a].ravel()
            if f64[a]:
                acc = 0.0
                for c in range(8):
                    acc = acc + w[a][c] * float(flat[idx[a][c]])
                got = np.float64(acc)
            else:
                acc = np.float32(0)
                for c in range(8):
                    acc = np.float32(acc + np.float32(np.float32(w[a][c]) * flat[idx[a][c]]))
                got = acc
            want = OM.interp_pressure(F[0], gg) if a == 0 else OM.interp_velocity(F[a], gg, a - 1)
            assert got == want and (f64[a] == 1) == isinstance(want, np.float64), (kind, t, a)
