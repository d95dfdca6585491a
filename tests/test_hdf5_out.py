"""The product's HDF5 writer (strata_fdtd_amd/hdf5_out.py) read back by the
minimal HDF5 parser of strata_fdtd_amd/hdf5_in.py, which reads the
reference's own h5py-written sample file (tests/test_simple_pulse.py pins
what it extracts from that file against the oracle). Every attribute type
the result layout uses, nested and empty groups, a group with more members
than the default symbol-table node holds, and datasets of float32, float64,
int64 and uint8."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))

import hdf5_read as H  # noqa: E402
from strata_fdtd_amd import hdf5_out as W  # noqa: E402


def test_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    root = W.Group()
    meta = root.group("metadata")
    meta.attrs.update(created_at="2026-01-01T00:00:00+00:00", solver_version="0.1.0", backend="mi355x",
                      num_threads=np.int64(1), total_runtime_seconds=0.25, note="ünïcode")
    grid = root.group("grid")
    grid.attrs.update(shape=np.array([5, 6, 7], np.int64), is_uniform=True, resolution=0.002,
                      extent=np.array([0.01, 0.012, 0.014]))
    root.group("empty")
    f32 = rng.standard_normal((3, 5, 6, 7)).astype(np.float32)
    root.group("fields").dataset("pressure", f32, units="Pa", snapshot_interval=np.int64(5))
    geo = (rng.random((5, 6, 7)) > 0.3).astype(np.uint8)
    root.group("materials").dataset("geometry", geo)
    probes = root.group("probes")
    traces = {}
    for q in range(40):     # five symbol-table nodes of 8 names
        traces[f"probe{q:02d}"] = rng.standard_normal(17).astype(np.float32)
        probes.dataset(f"probe{q:02d}", traces[f"probe{q:02d}"], position=np.array([q, 1, 2], np.int64),
                       units="Pa")
    root.group("grid").dataset("x_coords", np.linspace(0, 1, 5))
    root.group("deep").group("er").attrs["flag"] = False
    path = tmp_path / "out.h5"
    W.write(str(path), root)

    data = path.read_bytes()
    assert data[:8] == b"\x89HDF\r\n\x1a\n"
    assert int.from_bytes(data[40:48], "little") == len(data)      # end-of-file address
    f = H.File(str(path))
    r = f.root
    assert sorted(r.keys()) == sorted(["metadata", "grid", "empty", "fields", "materials", "probes", "deep"])
    a = r["metadata"].attrs
    assert a["created_at"] == "2026-01-01T00:00:00+00:00" and a["note"] == "ünïcode"
    assert a["num_threads"] == 1 and a["total_runtime_seconds"] == 0.25 and a["backend"] == "mi355x"
    g = r["grid"].attrs
    assert np.array_equal(g["shape"], [5, 6, 7]) and g["shape"].dtype == np.int64
    assert g["is_uniform"] is np.True_ or g["is_uniform"] == True  # noqa: E712
    assert g["resolution"] == 0.002 and np.array_equal(g["extent"], [0.01, 0.012, 0.014])
    assert r["empty"].keys() == []
    p = r["fields/pressure"]
    assert p.shape == f32.shape and np.array_equal(p.read(), f32) and p.read().dtype == np.float32
    assert p.attrs["units"] == "Pa" and p.attrs["snapshot_interval"] == 5
    assert np.array_equal(r["materials/geometry"].read(), geo)
    assert sorted(r["probes"].keys()) == sorted(traces)
    for name, tr in traces.items():
        d = r[f"probes/{name}"]
        assert np.array_equal(d.read(), tr)
        assert np.array_equal(d.attrs["position"], [int(name[5:]), 1, 2])
    assert np.array_equal(r["grid/x_coords"].read(), np.linspace(0, 1, 5))
    assert bool(r["deep/er"].attrs["flag"]) is False


def test_group_size_limit(tmp_path):
    root = W.Group()
    g = root.group("many")
    for q in range(4 * W.LEAF_K * W.INTERNAL_K + 1):
        g.dataset(f"d{q}", np.zeros(1, np.float32))
    try:
        W.write(str(tmp_path / "x.h5"), root)
    except ValueError as e:
        assert "at most" in str(e)
    else:
        raise AssertionError("a group beyond one B-tree leaf of symbol-table nodes must be refused")


def _sample_storage():
    import json
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "simple_pulse.npz")
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["storage_json"]))


def test_storage_matches_reference_sample(tmp_path):
    """Datasets stored as the reference's writer stores them (h5py, gzip 4):
    the writer's chunk shapes, maximum extents, fill-value and filter-pipeline
    messages equal those of the reference's sample file (simple-pulse.h5,
    40^3, 13 snapshots, 63-step probes), and guess_chunk gives h5py's
    automatic chunk shapes found there."""
    want = _sample_storage()
    assert W.guess_chunk((40, 40, 40), 1) == tuple(want["/materials/geometry"]["chunk"])
    assert W.guess_chunk((0,), 4) == tuple(want["/probes/far_field"]["chunk"])
    root = W.Group()
    root.group("fields").dataset("pressure", np.zeros((13, 40, 40, 40), np.float32), chunks=(1, 40, 40, 40),
                                 maxshape=(None, 40, 40, 40), gzip=W.GZIP_LEVEL)
    root.group("materials").dataset("geometry", np.ones((40, 40, 40), np.uint8),
                                    chunks=W.guess_chunk((40, 40, 40), 1), gzip=W.GZIP_LEVEL)
    for name in ("far_field", "near_field"):
        root.group("probes").dataset(name, np.zeros(63, np.float32), chunks=W.guess_chunk((0,), 4),
                                     maxshape=(None,), gzip=W.GZIP_LEVEL)
    path = tmp_path / "s.h5"
    W.write(str(path), root)
    assert H.storage(H.File(str(path)).root) == want


def test_chunked_round_trip(tmp_path):
    """Chunked datasets: more chunks than one B-tree node indexes (two
    levels), edge chunks padded to the chunk shape, an empty extensible
    dataset, deflated and raw chunks, and booleans."""
    rng = np.random.default_rng(9)
    root = W.Group()
    snaps = rng.standard_normal((150, 4, 5, 6)).astype(np.float32)      # 150 chunks > 64 per node
    root.dataset("snaps", snaps, chunks=(1, 4, 5, 6), maxshape=(None, 4, 5, 6), gzip=4)
    geo = rng.random((45, 41, 43)) > 0.5
    root.dataset("geo", geo.astype(np.uint8), chunks=W.guess_chunk(geo.shape, 1), gzip=4)
    root.dataset("trace", rng.standard_normal(3000).astype(np.float32), chunks=(1024,), maxshape=(None,), gzip=1)
    root.dataset("empty", np.zeros(0, np.float32), chunks=(1024,), maxshape=(None,), gzip=4)
    raw = rng.standard_normal((10, 10)).astype(np.float64)
    root.dataset("raw", raw, chunks=(3, 4))
    flags = rng.random(50) > 0.5
    root.dataset("flags", flags, chunks=(16,), gzip=4)
    path = tmp_path / "c.h5"
    W.write(str(path), root)
    r = H.File(str(path)).root
    assert np.array_equal(r["snaps"].read(), snaps)
    assert np.array_equal(r["geo"].read(), geo.astype(np.uint8))
    assert np.array_equal(r["trace"].read(), root.children["trace"].data)
    assert r["empty"].read().shape == (0,)
    assert np.array_equal(r["raw"].read(), raw)
    assert np.array_equal(r["flags"].read().astype(bool), flags)
    st = H.storage(r)
    assert st["/snaps"]["maxdims"] == [None, 4, 5, 6] and st["/raw"]["filter_hex"] is None
    assert st["/geo"]["chunk"] == list(W.guess_chunk(geo.shape, 1))


def test_result_reader(tmp_path):
    """HDF5ResultReader (io/hdf5.py:243-377) on a file in the result layout."""
    from strata_fdtd_amd import HDF5ResultReader
    rng = np.random.default_rng(4)
    root = W.Group()
    root.group("metadata").attrs.update(solver_version="0.1.0", backend="mi355x")
    root.group("grid").attrs.update(shape=np.array([4, 5, 6], np.int64), is_uniform=False)
    for a, n in zip("xyz", (4, 5, 6)):
        root.group("grid").dataset(f"{a}_coords", np.linspace(0, 1, n))
    root.group("simulation").attrs.update(timestep=1e-6, num_steps=np.int64(9))
    root.group("sources").group("source_0").attrs.update(type="point", frequency=1000.0)
    snaps = rng.standard_normal((3, 4, 5, 6)).astype(np.float32)
    root.group("fields").dataset("pressure", snaps, chunks=(1, 4, 5, 6), maxshape=(None, 4, 5, 6), gzip=4,
                                 units="Pa")
    tr = rng.standard_normal(9).astype(np.float32)
    root.group("probes").dataset("mic", tr, chunks=(1024,), maxshape=(None,), gzip=4,
                                 position=np.array([1, 2, 3], np.int64), units="Pa")
    geo = rng.random((4, 5, 6)) > 0.5
    root.group("materials").dataset("geometry", geo.astype(np.uint8), chunks=(4, 5, 6), gzip=4)
    path = tmp_path / "r.h5"
    W.write(str(path), root)
    with HDF5ResultReader(path) as r:
        md = r.get_metadata()
        assert md["metadata"]["backend"] == "mi355x" and md["simulation"]["num_steps"] == 9
        assert np.array_equal(md["grid"]["y_coords"], np.linspace(0, 1, 5))
        assert md["sources"][0]["type"] == "point" and np.array_equal(md["probes"]["mic"]["position"], [1, 2, 3])
        assert r.get_probe_names() == ["mic"] and r.get_num_snapshots() == 3
        assert np.array_equal(r.load_probe("mic"), tr) and np.array_equal(r.load_timestep(2), snaps[2])
        assert np.array_equal(r.load_geometry(), geo)
        with pytest.raises(KeyError):
            r.load_probe("nope")
    empty = W.Group()
    empty.group("metadata")
    W.write(str(tmp_path / "e.h5"), empty)
    r = HDF5ResultReader(tmp_path / "e.h5")
    assert r.get_num_snapshots() == 0 and r.get_probe_names() == [] and r.load_geometry() is None
    with pytest.raises(ValueError):
        r.load_timestep(0)
