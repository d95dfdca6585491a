"""HDF5 result files in the reference's layout, written without h5py.

``FDTDSolver.run(..., output_file=path, snapshot_interval=n)`` writes the
groups and datasets of the reference's HDF5ResultWriter (io/hdf5.py:24-233):

    /metadata   attrs created_at, solver_version, [script_hash, script_content],
                total_runtime_seconds, backend, num_threads
    /grid       attrs shape, is_uniform, resolution or datasets x/y/z_coords, extent
    /simulation attrs timestep, cfl_number, c, rho, num_steps, total_time
    /sources/source_<i>  attrs type, position, frequency, [bandwidth]
    /fields/pressure     float32 (snapshots, nx, ny, nz), attrs units, snapshot_interval
    /probes/<name>       float32 (steps,), attrs position, units
    /materials/geometry  uint8 (nx, ny, nz)

The file uses the classic format h5py writes by default, which every HDF5
reader accepts: superblock version 0 with 8-byte offsets, version-1 object
headers, groups as symbol tables (a version-1 B-tree leaf over symbol-table
nodes, names in a local heap), attributes as version-1 messages. Datasets are
stored as the reference's writer asks h5py to store them: pressure, probes and
geometry chunked and deflate-compressed at level 4, pressure in chunks of one
snapshot (1, nx, ny, nz) and extensible along time, probes extensible with
h5py's automatic chunk shape, geometry with h5py's automatic chunk shape;
the coordinate tables contiguous. Chunks are indexed by version-1 B-trees
(HDF5's default of 2K = 64 chunks per node, more levels as needed). Booleans
are stored as h5py stores them (an int8 enumeration FALSE/TRUE) and strings
as fixed-length UTF-8. The whole file is built in memory and written once,
when the run ends.
"""

from __future__ import annotations

import itertools
import math
import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
HEAP_FREE_NULL = 1  # a local heap's free-list offset when it has no free block
LEAF_K = 4          # symbol-table nodes of 2K = 8 names (HDF5's default)
INTERNAL_K = 16     # B-tree nodes of 2K = 32 children (HDF5's default)
CHUNK_K = 32        # chunk B-tree nodes of 2K = 64 children (HDF5's default)
GZIP_LEVEL = 4      # the reference writer's compression_level default


def _pad8(n: int) -> int:
    return (n + 7) & ~7


def _padded(b: bytes) -> bytes:
    return b + b"\0" * (_pad8(len(b)) - len(b))


# ---------------------------------------------------------------- datatypes
def _dt_fixed(size: int, signed: bool) -> bytes:
    return struct.pack("<B3sIHH", 0x10, bytes([0x08 if signed else 0, 0, 0]), size, 0, 8 * size)


def _dt_float(size: int) -> bytes:
    if size == 4:
        sign, exp_loc, exp_size, mant_size, bias = 31, 23, 8, 23, 127
    else:
        sign, exp_loc, exp_size, mant_size, bias = 63, 52, 11, 52, 1023
    return struct.pack("<B3sIHHBBBBI", 0x11, bytes([0x20, sign, 0]), size, 0, 8 * size, exp_loc, exp_size, 0,
                       mant_size, bias)


def _dt_string(size: int) -> bytes:
    return struct.pack("<B3sI", 0x13, bytes([0x10, 0, 0]), size)   # null-terminated, UTF-8


def _dt_bool() -> bytes:
    names = b"".join(_padded(n + b"\0") for n in (b"FALSE", b"TRUE"))
    return struct.pack("<B3sI", 0x18, bytes([2, 0, 0]), 1) + _dt_fixed(1, True) + names + bytes([0, 1])


def _encode(value):
    """(datatype message, dims, raw little-endian bytes) of an attribute or dataset value."""
    if isinstance(value, (bool, np.bool_)) or (isinstance(value, np.ndarray) and value.dtype == np.bool_):
        a = np.asarray(value, dtype=np.int8)
        return _dt_bool(), a.shape, a.tobytes()
    if isinstance(value, str):
        raw = value.encode("utf-8") + b"\0"
        return _dt_string(len(raw)), (), raw
    a = np.asarray(value)
    if a.dtype.kind == "U":
        raise TypeError("string arrays are not supported")
    if a.dtype.kind in "iu":
        a = a.astype(a.dtype.newbyteorder("<"))
        return _dt_fixed(a.dtype.itemsize, a.dtype.kind == "i"), a.shape, np.ascontiguousarray(a).tobytes()
    if a.dtype.kind == "f":
        a = a.astype(a.dtype.newbyteorder("<"))
        return _dt_float(a.dtype.itemsize), a.shape, np.ascontiguousarray(a).tobytes()
    raise TypeError(f"cannot store {type(value).__name__} ({a.dtype})")


def _dataspace(dims, maxdims=None) -> bytes:
    """Version-1 dataspace; maxdims (None entries unlimited) adds the maximum sizes."""
    out = struct.pack("<BBBB4x", 1, len(dims), 0 if maxdims is None else 1, 0)
    out += b"".join(struct.pack("<Q", d) for d in dims)
    if maxdims is not None:
        out += b"".join(struct.pack("<Q", UNDEF if m is None else m) for m in maxdims)
    return out


def guess_chunk(shape, typesize: int) -> tuple:
    """h5py's automatic chunk shape (h5py/_hl/filters.py, guess_chunk), which
    the reference's datasets get when they ask for compression without a chunk
    shape: a zero (extensible, empty) extent counts as 1024; the target chunk
    size is 16 KiB x 2^log10(dataset MiB), clamped to [8 KiB, 1 MiB]; axes are
    halved (rounding up) in turn, starting with the first, until the chunk is
    below the target, or within 50% of it, and below 1 MiB."""
    dims = [float(n if n else 1024) for n in shape]
    total = math.prod(dims) * typesize
    target = 16 * 1024 * 2 ** math.log10(total / (1024.0 * 1024))
    target = min(max(target, 8 * 1024), 1024 * 1024)
    axis = 0
    while True:
        nbytes = math.prod(dims) * typesize
        if (nbytes < target or abs(nbytes - target) / target < 0.5) and nbytes < 1024 * 1024:
            break
        if math.prod(dims) == 1:
            break
        dims[axis % len(dims)] = math.ceil(dims[axis % len(dims)] / 2.0)
        axis += 1
    return tuple(int(d) for d in dims)


def _attribute(name: str, value) -> bytes:
    dt, dims, raw = _encode(value)
    nm = name.encode("utf-8") + b"\0"
    ds = _dataspace(dims)
    return (struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds)) + _padded(nm) + _padded(dt) + _padded(ds)
            + raw)


# ---------------------------------------------------------------- the tree
class Group:
    def __init__(self):
        self.attrs: dict = {}
        self.children: dict = {}

    def group(self, name: str) -> "Group":
        g = self.children.get(name)
        if g is None:
            g = self.children[name] = Group()
        return g

    def dataset(self, name: str, data: np.ndarray, chunks=None, maxshape=None, gzip=None, **attrs) -> None:
        ds = Dataset(np.asarray(data), chunks, maxshape, gzip)
        ds.attrs.update(attrs)
        self.children[name] = ds


class Dataset:
    """chunks: chunk shape (None: contiguous); maxshape: maximum extent, None
    entries unlimited (chunked datasets only); gzip: deflate level or None."""

    def __init__(self, data: np.ndarray, chunks=None, maxshape=None, gzip=None):
        self.data = data
        self.attrs: dict = {}
        if chunks is not None:
            chunks = tuple(int(c) for c in chunks)
            if len(chunks) != data.ndim or any(c < 1 for c in chunks):
                raise ValueError(f"chunk shape {chunks} for a dataset of shape {data.shape}")
            if maxshape is None:
                maxshape = data.shape
            if len(maxshape) != data.ndim or any(m is not None and m < n for m, n in zip(maxshape, data.shape)):
                raise ValueError(f"maximum shape {maxshape} for a dataset of shape {data.shape}")
        elif maxshape is not None or gzip is not None:
            raise ValueError("maximum shapes and compression need a chunked dataset")
        self.chunks, self.maxshape, self.gzip = chunks, maxshape, gzip


class _Out:
    def __init__(self):
        self.buf = bytearray()

    def alloc(self, n: int) -> int:
        addr = len(self.buf)
        self.buf += b"\0" * _pad8(n)
        return addr

    def put(self, addr: int, b: bytes) -> None:
        self.buf[addr:addr + len(b)] = b


def _object_header(msgs) -> bytes:
    body = b"".join(struct.pack("<HHB3x", t, len(_padded(d)), flags) + _padded(d) for t, flags, d in msgs)
    return struct.pack("<BBHII4x", 1, 0, len(msgs), 1, len(body)) + body


def _chunk_btree(out: _Out, entries, rank: int) -> int:
    """Version-1 B-tree (node type 1) over the chunks; entries are
    (stored size, chunk offsets, address) in offset order. Each key is
    (size, filter mask, offsets + [0]); a node's last key is the offset just
    past its last chunk (its offsets plus the chunk shape). Returns the root."""
    key_size = 8 + 8 * (rank + 1)
    node_size = 24 + 2 * CHUNK_K * 8 + (2 * CHUNK_K + 1) * key_size

    def key(size, offs):
        return struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in offs)

    # level 0: (first key, last key, child address) per chunk
    items = [(key(size, list(offs) + [0]), key(0, list(end) + [esize]), addr) for size, offs, end, esize, addr in entries]
    level = 0
    while True:
        groups = [items[g:g + 2 * CHUNK_K] for g in range(0, len(items), 2 * CHUNK_K)]
        addrs = [out.alloc(node_size) for _ in groups]
        parents = []
        for n, grp in enumerate(groups):
            left = addrs[n - 1] if n else UNDEF
            right = addrs[n + 1] if n + 1 < len(addrs) else UNDEF
            node = bytearray(b"TREE" + struct.pack("<BBHQQ", 1, level, len(grp), left, right))
            for first, _last, child in grp:
                node += first + struct.pack("<Q", child)
            node += grp[-1][1]
            out.put(addrs[n], bytes(node))
            parents.append((grp[0][0], grp[-1][1], addrs[n]))
        if len(parents) == 1:
            return parents[0][2]
        items = parents
        level += 1


def _write_chunks(out: _Out, ds: Dataset, raw_dtype: np.dtype) -> int:
    """Store every chunk that meets the extent (edge chunks padded with
    zeros to the full chunk shape), deflated when asked; returns the chunk
    B-tree address (UNDEF for an empty dataset)."""
    a = np.ascontiguousarray(np.asarray(ds.data).astype(raw_dtype))
    shape, chunks = a.shape, ds.chunks
    if a.size == 0:
        return UNDEF
    entries = []
    for idx in itertools.product(*(range(-(-n // c)) for n, c in zip(shape, chunks))):
        offs = [i * c for i, c in zip(idx, chunks)]
        block = np.zeros(chunks, dtype=raw_dtype)
        src = tuple(slice(o, min(o + c, n)) for o, c, n in zip(offs, chunks, shape))
        block[tuple(slice(0, s.stop - s.start) for s in src)] = a[src]
        data = block.tobytes()
        if ds.gzip is not None:
            data = zlib.compress(data, ds.gzip)
        addr = out.alloc(len(data))
        out.put(addr, data)
        end = [o + c for o, c in zip(offs, chunks)]
        entries.append((len(data), offs, end, raw_dtype.itemsize, addr))
    return _chunk_btree(out, entries, a.ndim)


def _write_dataset(out: _Out, ds: Dataset) -> int:
    if ds.chunks is not None:
        arr = np.asarray(ds.data)
        dt, _, _ = _encode(arr.reshape(-1)[:1])      # the datatype message only
        dims = arr.shape
        raw_dtype = np.dtype(np.int8) if arr.dtype == np.bool_ else arr.dtype.newbyteorder("<")
        btree = _write_chunks(out, ds, raw_dtype)
        msgs = [
            (0x0001, 0, _dataspace(dims, ds.maxshape)),
            (0x0003, 1, dt),
            # fill value: v2, incremental allocation, written if set, defined as the default (0)
            (0x0005, 1, struct.pack("<BBBBI", 2, 3, 2, 1, 0)),
        ]
        if ds.gzip is not None:
            # filter pipeline v1: deflate (id 1, optional as h5py marks it), its level
            msgs.append((0x000B, 1, struct.pack("<BB6xHHHH", 1, 1, 1, 8, 1, 1) + b"deflate\0"
                         + struct.pack("<I4x", ds.gzip)))
        layout = struct.pack("<BBBQ", 3, 2, len(dims) + 1, btree)
        layout += b"".join(struct.pack("<I", c) for c in ds.chunks) + struct.pack("<I", raw_dtype.itemsize)
        msgs.append((0x0008, 0, layout))
    else:
        dt, dims, raw = _encode(ds.data)
        data_addr = out.alloc(len(raw))
        out.put(data_addr, raw)
        msgs = [
            (0x0001, 0, _dataspace(dims)),
            (0x0003, 1, dt),
            (0x0005, 1, struct.pack("<BBBB", 2, 1, 0, 0)),        # fill value: v2, early, on alloc, undefined
            (0x0008, 0, struct.pack("<BBQQ", 3, 1, data_addr, len(raw))),   # contiguous layout
        ]
    msgs += [(0x000C, 0, _attribute(k, v)) for k, v in ds.attrs.items()]
    hdr = _object_header(msgs)
    addr = out.alloc(len(hdr))
    out.put(addr, hdr)
    return addr


def _write_group(out: _Out, g: Group) -> tuple[int, int, int]:
    """Returns (object header, B-tree, local heap) addresses."""
    names = sorted(g.children)      # byte order of the UTF-8 names, as HDF5 compares them
    if len(names) > 4 * LEAF_K * INTERNAL_K:
        raise ValueError(f"a group holds at most {4 * LEAF_K * INTERNAL_K} members here")
    entries = []
    for name in names:
        child = g.children[name]
        if isinstance(child, Group):
            oh, bt, hp = _write_group(out, child)
            entries.append((name, oh, 1, struct.pack("<QQ", bt, hp)))
        else:
            entries.append((name, _write_dataset(out, child), 0, b"\0" * 16))
    # local heap: offset 0 holds the empty name
    heap_data = bytearray(8)
    offsets = {}
    for name, *_ in entries:
        offsets[name] = len(heap_data)
        heap_data += _padded(name.encode("utf-8") + b"\0")
    seg = out.alloc(len(heap_data))
    out.put(seg, bytes(heap_data))
    heap = out.alloc(32)
    out.put(heap, b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap_data), HEAP_FREE_NULL, seg))
    # symbol-table nodes of up to 2K members each, in name order, under one
    # B-tree leaf whose keys are "" and then the last name of each node
    per = 2 * LEAF_K
    tree = bytearray(b"TREE" + struct.pack("<BBHQQ", 0, 0, (len(entries) + per - 1) // per, UNDEF, UNDEF))
    tree += struct.pack("<Q", 0)
    for first in range(0, len(entries), per):
        chunk = entries[first:first + per]
        snod = out.alloc(8 + per * 40)
        node = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(chunk)))
        for name, oh, cache, scratch in chunk:
            node += struct.pack("<QQII", offsets[name], oh, cache, 0) + scratch
        out.put(snod, bytes(node))
        tree += struct.pack("<QQ", snod, offsets[chunk[-1][0]])
    btree = out.alloc(24 + (2 * INTERNAL_K + 1) * 8 + 2 * INTERNAL_K * 8)
    out.put(btree, bytes(tree))
    msgs = [(0x0011, 0, struct.pack("<QQ", btree, heap))]
    msgs += [(0x000C, 0, _attribute(k, v)) for k, v in g.attrs.items()]
    hdr = _object_header(msgs)
    addr = out.alloc(len(hdr))
    out.put(addr, hdr)
    return addr, btree, heap


def write(path: str, root: Group) -> None:
    """Serialise the tree under `root` into an HDF5 file at `path`."""
    out = _Out()
    out.alloc(96)                     # superblock (with the root's symbol-table entry)
    oh, bt, hp = _write_group(out, root)
    sb = (b"\x89HDF\r\n\x1a\n" + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", LEAF_K, INTERNAL_K, 0)
          + struct.pack("<QQQQ", 0, UNDEF, len(out.buf), UNDEF)
          + struct.pack("<QQII", 0, oh, 1, 0) + struct.pack("<QQ", bt, hp))
    out.put(0, sb)
    with open(path, "wb") as fh:
        fh.write(out.buf)


# ---------------------------------------------------------------- results
def solver_tree(solver, snapshots, snapshot_interval, runtime: float, script_content: str | None,
                created_at: str) -> Group:
    """The reference's result layout (io/hdf5.py:68-233) for a finished run."""
    import hashlib
    root = Group()
    meta = root.group("metadata")
    if script_content:
        meta.attrs["script_hash"] = hashlib.sha256(script_content.encode()).hexdigest()
        meta.attrs["script_content"] = script_content
    meta.attrs["created_at"] = created_at
    meta.attrs["solver_version"] = "0.1.0"
    grid = root.group("grid")
    grid.attrs["shape"] = np.array(solver.shape, np.int64)
    if solver.grid.is_uniform:
        grid.attrs["is_uniform"] = True
        grid.attrs["resolution"] = float(solver.dx)
    else:
        grid.attrs["is_uniform"] = False
        for axis, coords in zip("xyz", (solver.grid.x_coords, solver.grid.y_coords, solver.grid.z_coords)):
            grid.dataset(f"{axis}_coords", np.asarray(coords, np.float64))
    # shape x dx with dx the minimum spacing, on every grid (io/hdf5.py:95-100)
    grid.attrs["extent"] = np.array([n * float(solver.dx) for n in solver.shape], np.float64)
    sim = root.group("simulation")
    sim.attrs["timestep"] = float(solver.dt)
    sim.attrs["cfl_number"] = float(solver.c * solver.dt / float(solver.dx))
    sim.attrs["c"] = float(solver.c)
    sim.attrs["rho"] = float(solver.rho)
    sources = root.group("sources")
    for i, src in enumerate(solver._sources):
        sg = sources.group(f"source_{i}")
        sg.attrs["type"] = str(src.source_type)
        pos = src.position
        if isinstance(pos, dict):
            sg.attrs["position"] = np.array([int(pos.get("axis", 0)), int(pos.get("index", 0)), 0], np.int64)
        elif isinstance(pos, (tuple, list, np.ndarray)):
            sg.attrs["position"] = np.array(pos, dtype=np.int64 if all(
                isinstance(v, (int, np.integer)) for v in pos) else np.float64)
        else:
            sg.attrs["position"] = np.array([pos, pos, pos], np.float64)
        freq = getattr(src, "frequency", None)
        if freq is None and getattr(src, "waveform", None) is not None:
            freq = getattr(src.waveform, "frequency", None)
        if freq is not None:
            sg.attrs["frequency"] = float(freq)
        bw = getattr(src, "bandwidth", None)
        if bw is not None:
            sg.attrs["bandwidth"] = float(bw)
    fields = root.group("fields")
    if snapshots:
        # one chunk per snapshot, extensible along time (io/hdf5.py:186-206)
        fields.dataset("pressure", np.stack([p for _t, p in snapshots]).astype(np.float32),
                       chunks=(1,) + tuple(solver.shape), maxshape=(None,) + tuple(solver.shape), gzip=GZIP_LEVEL,
                       units="Pa", snapshot_interval=np.int64(snapshot_interval))
    probes = root.group("probes")
    for name, data in solver.get_probe_data().items():
        # created empty and extensible, automatic chunks (io/hdf5.py:138-152)
        probes.dataset(name, np.asarray(data, np.float32), chunks=guess_chunk((0,), 4), maxshape=(None,),
                       gzip=GZIP_LEVEL, position=np.array(solver._probes[name].position, np.int64), units="Pa")
    geo = np.asarray(solver.geometry).astype(np.uint8)
    root.group("materials").dataset("geometry", geo, chunks=guess_chunk(geo.shape, 1), gzip=GZIP_LEVEL)
    sim.attrs["num_steps"] = np.int64(solver.step_count)
    sim.attrs["total_time"] = float(solver.time)
    meta.attrs["total_runtime_seconds"] = float(runtime)
    meta.attrs["backend"] = "mi355x"
    meta.attrs["num_threads"] = np.int64(1)
    return root
