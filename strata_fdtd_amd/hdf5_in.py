"""Reading HDF5 result files without h5py: a minimal read-only parser and
``HDF5ResultReader``, the reference's reader interface (io/hdf5.py:243-377:
get_metadata, load_timestep, load_probe, get_probe_names, get_num_snapshots,
load_geometry, close, context manager), for the files ``run(output_file=...)``
writes and for the reference's own result files.

Covers what h5py writes by default ("earliest" file format): superblock
version 0 with 8-byte offsets and lengths, version-1 object headers (with
continuation blocks), groups as symbol tables (version-1 B-trees of SNOD
nodes, names in a local heap), and datasets with contiguous, compact or
chunked storage (version-1 chunk B-trees, deflate and shuffle filters).
Datatypes: little-endian integers and IEEE floats, booleans stored as
h5py's int8 enum, fixed-length and variable-length strings (global heap).
Attributes: message versions 1-3. Anything else raises NotImplementedError,
so an unsupported file cannot be misread silently.

Reference: the HDF5 file format specification (version 0 superblock, v1
object headers, v1 B-trees, layout message version 3, attribute messages
versions 1-3).
"""

from __future__ import annotations

import struct
import zlib

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF


def _u(b, off, n):
    return int.from_bytes(b[off:off + n], "little")


def _pad8(n):
    return (n + 7) & ~7


class Datatype:
    def __init__(self, cls, size, order, signed=False, base=None, vlen_string=False, enum=None):
        self.cls, self.size, self.order, self.signed = cls, size, order, signed
        self.base, self.vlen_string, self.enum = base, vlen_string, enum

    def numpy(self):
        e = ">" if self.order else "<"
        if self.cls == 0:
            return np.dtype(f"{e}{'i' if self.signed else 'u'}{self.size}")
        if self.cls == 1:
            return np.dtype(f"{e}f{self.size}")
        if self.cls == 8:
            return self.base.numpy()
        raise NotImplementedError(f"datatype class {self.cls} as numpy")


class File:
    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.b = fh.read()
        b = self.b
        if b[:8] != b"\x89HDF\r\n\x1a\n":
            raise ValueError("not an HDF5 file")
        if b[8] != 0 or b[13] != 8 or b[14] != 8:
            raise NotImplementedError("only superblock version 0 with 8-byte offsets and lengths")
        self.base = _u(b, 24, 8)
        root_header = _u(b, 56 + 8, 8)
        self.root = Group(self, root_header)

    def addr(self, a):
        return self.base + a

    # ---- object headers
    def messages(self, addr):
        b = self.b
        o = self.addr(addr)
        if b[o] != 1:
            raise NotImplementedError(f"object header version {b[o]}")
        nmsg = _u(b, o + 2, 2)
        size = _u(b, o + 8, 4)
        blocks = [(o + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            p = start
            while p + 8 <= start + length and len(out) < nmsg:
                mtype, msize, flags = _u(b, p, 2), _u(b, p + 2, 2), b[p + 4]
                data = b[p + 8:p + 8 + msize]
                if mtype == 0x10:  # continuation
                    blocks.append((self.addr(_u(data, 0, 8)), _u(data, 8, 8)))
                out.append((mtype, flags, data))
                p += 8 + msize
        return out

    # ---- local and global heaps
    def heap_string(self, heap_addr, offset):
        b = self.b
        o = self.addr(heap_addr)
        if b[o:o + 4] != b"HEAP":
            raise ValueError("bad local heap")
        data = self.addr(_u(b, o + 24, 8))
        end = b.index(b"\0", data + offset)
        return b[data + offset:end].decode()

    def global_object(self, coll_addr, index):
        b = self.b
        o = self.addr(coll_addr)
        if b[o:o + 4] != b"GCOL":
            raise ValueError("bad global heap collection")
        size = _u(b, o + 8, 8)
        p, end = o + 16, o + size
        while p + 16 <= end:
            idx, osize = _u(b, p, 2), _u(b, p + 8, 8)
            if idx == 0:
                break
            if idx == index:
                return b[p + 16:p + 16 + osize]
            p += 16 + _pad8(osize)
        raise KeyError(f"global heap object {index}")

    # ---- B-trees (version 1)
    def btree_children(self, addr, ndims=None):
        """Leaf entries of a v1 B-tree: (key, child address) for chunk trees
        (key = (chunk size, filter mask, offsets)), child addresses for groups."""
        b = self.b
        o = self.addr(addr)
        if b[o:o + 4] != b"TREE":
            raise ValueError("bad B-tree node")
        ntype, level, used = b[o + 4], b[o + 5], _u(b, o + 6, 2)
        p = o + 24
        keysize = 8 if ntype == 0 else 8 + 8 * (ndims + 1)
        out = []
        for e in range(used):
            key = b[p:p + keysize]
            child = _u(b, p + keysize, 8)
            p += keysize + 8
            if level > 0:
                out += self.btree_children(child, ndims)
            elif ntype == 0:
                out.append(child)
            else:
                csize, mask = _u(key, 0, 4), _u(key, 4, 4)
                offs = tuple(_u(key, 8 + 8 * d, 8) for d in range(ndims + 1))
                out.append(((csize, mask, offs), child))
        return out


def _datatype(b):
    cls, ver = b[0] & 0x0F, b[0] >> 4
    bits = b[1] | (b[2] << 8) | (b[3] << 16)
    size = _u(b, 4, 4)
    if cls == 0:
        return Datatype(0, size, bits & 1, signed=bool(bits & 8)), 12
    if cls == 1:
        return Datatype(1, size, bits & 1), 20
    if cls == 3:
        return Datatype(3, size, 0), 8
    if cls == 8:  # enumeration (h5py booleans): base type, then names and values
        base, n = _datatype(b[8:])
        return Datatype(8, size, base.order, base=base), 8 + n
    if cls == 9:  # variable length
        vtype = bits & 0x0F
        base, n = _datatype(b[8:])
        return Datatype(9, size, 0, base=base, vlen_string=(vtype == 1)), 8 + n
    raise NotImplementedError(f"datatype class {cls} version {ver}")


def _dataspace(b):
    ver, nd, flags = b[0], b[1], b[2]
    if ver == 1:
        p = 8
    elif ver == 2:
        if b[3] == 0:  # scalar
            return ()
        p = 4
    else:
        raise NotImplementedError(f"dataspace version {ver}")
    return tuple(_u(b, p + 8 * i, 8) for i in range(nd))


class Node:
    def __init__(self, f: File, addr: int):
        self.f, self.addr = f, addr
        self.msgs = f.messages(addr)

    @property
    def attrs(self) -> dict:
        out = {}
        for mtype, _flags, data in self.msgs:
            if mtype != 0x0C:
                continue
            ver = data[0]
            nlen, tlen, slen = _u(data, 2, 2), _u(data, 4, 2), _u(data, 6, 2)
            if ver == 1:
                p = 8
                name = data[p:p + nlen].split(b"\0")[0].decode()
                p += _pad8(nlen)
                dt, _ = _datatype(data[p:p + tlen])
                p += _pad8(tlen)
                shape = _dataspace(data[p:p + slen])
                p += _pad8(slen)
            elif ver in (2, 3):
                p = 8 + (1 if ver == 3 else 0)
                name = data[p:p + nlen].split(b"\0")[0].decode()
                p += nlen
                dt, _ = _datatype(data[p:p + tlen])
                p += tlen
                shape = _dataspace(data[p:p + slen])
                p += slen
            else:
                raise NotImplementedError(f"attribute message version {ver}")
            out[name] = _values(self.f, dt, shape, data[p:])
        return out


def _values(f: File, dt: Datatype, shape, raw):
    n = int(np.prod(shape)) if shape else 1
    if dt.cls == 9:
        items = []
        for e in range(n):
            q = 16 * e
            length, coll, idx = _u(raw, q, 4), _u(raw, q + 4, 8), _u(raw, q + 12, 4)
            obj = f.global_object(coll, idx) if length else b""
            items.append(obj[:length].decode() if dt.vlen_string else obj[:length])
        return items[0] if not shape else np.array(items, dtype=object).reshape(shape)
    if dt.cls == 3:
        items = [raw[e * dt.size:(e + 1) * dt.size].split(b"\0")[0].decode() for e in range(n)]
        return items[0] if not shape else np.array(items).reshape(shape)
    arr = np.frombuffer(raw[:n * dt.size], dtype=dt.numpy())
    if dt.cls == 8:
        arr = arr.astype(bool)
    return arr[0] if not shape else arr.reshape(shape).copy()


class Group(Node):
    def keys(self):
        return list(self._children())

    def _children(self):
        for mtype, _flags, data in self.msgs:
            if mtype == 0x11:
                btree, heap = _u(data, 0, 8), _u(data, 8, 8)
                out = {}
                for snod in self.f.btree_children(btree):
                    o = self.f.addr(snod)
                    if self.f.b[o:o + 4] != b"SNOD":
                        raise ValueError("bad symbol table node")
                    nsym = _u(self.f.b, o + 6, 2)
                    for e in range(nsym):
                        q = o + 8 + 40 * e
                        name = self.f.heap_string(heap, _u(self.f.b, q, 8))
                        out[name] = _u(self.f.b, q + 8, 8)
                return out
        raise NotImplementedError("group without a symbol table message")

    def __getitem__(self, path: str):
        node = self
        for part in path.strip("/").split("/"):
            addr = node._children()[part]
            msgs = self.f.messages(addr)
            node = (Group if any(m[0] == 0x11 for m in msgs) else Dataset)(self.f, addr)
        return node


class Dataset(Node):
    def _msg(self, mtype):
        for t, _flags, data in self.msgs:
            if t == mtype:
                return data
        return None

    @property
    def shape(self):
        return _dataspace(self._msg(0x01))

    @property
    def dtype(self) -> Datatype:
        return _datatype(self._msg(0x03))[0]

    def _filters(self):
        data = self._msg(0x0B)
        if data is None:
            return []
        ver, nf = data[0], data[1]
        p = 8 if ver == 1 else 2
        out = []
        for _ in range(nf):
            fid = _u(data, p, 2)
            if ver == 1 or fid >= 256:
                nlen = _u(data, p + 2, 2)
                flags, nval = _u(data, p + 4, 2), _u(data, p + 6, 2)
                p += 8 + (_pad8(nlen) if ver == 1 else nlen)
            else:
                flags, nval = _u(data, p + 2, 2), _u(data, p + 4, 2)
                p += 6
            vals = [_u(data, p + 4 * v, 4) for v in range(nval)]
            p += 4 * nval + (4 if ver == 1 and nval % 2 else 0)
            out.append((fid, vals))
        return out

    def read(self) -> np.ndarray:
        f, dt, shape = self.f, self.dtype, self.shape
        lay = self._msg(0x08)
        if lay[0] != 3:
            raise NotImplementedError(f"layout message version {lay[0]}")
        n = int(np.prod(shape)) if shape else 1
        if lay[1] == 0:  # compact
            size = _u(lay, 2, 2)
            return _values(f, dt, shape, lay[4:4 + size])
        if lay[1] == 1:  # contiguous
            addr, size = _u(lay, 2, 8), _u(lay, 10, 8)
            if addr == UNDEF:
                return np.zeros(shape, dtype=dt.numpy())
            raw = f.b[f.addr(addr):f.addr(addr) + size]
            return _values(f, dt, shape, raw)
        if lay[1] != 2:
            raise NotImplementedError(f"layout class {lay[1]}")
        nd = lay[2]
        btree = _u(lay, 3, 8)
        cdims = [_u(lay, 11 + 4 * d, 4) for d in range(nd)]  # last entry: element size
        chunk = tuple(cdims[:-1])
        out = np.zeros(shape, dtype=dt.numpy())
        if btree == UNDEF:
            return out
        filters = self._filters()
        for (csize, mask, offs), caddr in f.btree_children(btree, len(chunk)):
            raw = f.b[f.addr(caddr):f.addr(caddr) + csize]
            for k, (fid, vals) in enumerate(reversed(filters)):
                if mask & (1 << (len(filters) - 1 - k)):
                    continue
                if fid == 1:
                    raw = zlib.decompress(raw)
                elif fid == 2:
                    es = vals[0] if vals else dt.size
                    a = np.frombuffer(raw, np.uint8).reshape(es, -1)
                    raw = a.T.tobytes()
                else:
                    raise NotImplementedError(f"filter {fid}")
            block = np.frombuffer(raw, dtype=dt.numpy()).reshape(chunk)
            sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs[:-1], chunk, shape))
            out[sl] = block[tuple(slice(0, s.stop - s.start) for s in sl)]
        return out


def storage(root) -> dict:
    """How each dataset is stored: layout class, chunk shape, maximum extent
    (None unlimited), and the fill-value and filter-pipeline messages as hex."""
    out = {}
    for opath, node in walk(root):
        if not isinstance(node, Dataset):
            continue
        lay, ds = node._msg(0x08), node._msg(0x01)
        rank, flags = ds[1], ds[2]
        maxdims = None
        if flags & 1:
            maxdims = [None if v == UNDEF else v for v in
                       (_u(ds, 8 + 8 * rank + 8 * d, 8) for d in range(rank))]
        info = {"layout": lay[1], "maxdims": maxdims, "fill_hex": node._msg(0x05).hex()}
        if lay[1] == 2:
            cd = [_u(lay, 11 + 4 * d, 4) for d in range(lay[2])]
            info["chunk"], info["element_size"] = cd[:-1], cd[-1]
        flt = node._msg(0x0B)
        info["filter_hex"] = None if flt is None else flt.hex()
        out[opath] = info
    return out


def walk(g: Group, prefix=""):
    """(path, node) for every object under a group."""
    for name in g.keys():
        node = g[name]
        yield prefix + "/" + name, node
        if isinstance(node, Group):
            yield from walk(node, prefix + "/" + name)





class HDF5ResultReader:
    """A result file's metadata, snapshots, probe traces and geometry
    (io/hdf5.py:243-377), with the reference's method names and errors."""

    def __init__(self, filename):
        from pathlib import Path
        self.filename = Path(filename)
        self.file = File(str(filename))
        self._root = self.file.root

    def _has(self, path: str) -> bool:
        node = self._root
        for part in path.split("/"):
            if not isinstance(node, Group) or part not in node.keys():
                return False
            node = node[part]
        return True

    def get_metadata(self) -> dict:
        r = self._root
        out = {}
        if self._has("metadata"):
            out["metadata"] = dict(r["metadata"].attrs)
        if self._has("grid"):
            grid = dict(r["grid"].attrs)
            if not grid.get("is_uniform", True):
                for a in ("x", "y", "z"):
                    grid[f"{a}_coords"] = r[f"grid/{a}_coords"].read()
            out["grid"] = grid
        if self._has("simulation"):
            out["simulation"] = dict(r["simulation"].attrs)
        if self._has("sources"):
            out["sources"] = [dict(r[f"sources/{n}"].attrs) for n in r["sources"].keys()]
        if self._has("probes"):
            out["probes"] = {n: dict(r[f"probes/{n}"].attrs) for n in r["probes"].keys()}
        return out

    def load_timestep(self, step: int) -> np.ndarray:
        if not self._has("fields/pressure"):
            raise ValueError("No pressure field data in file")
        return self._root["fields/pressure"].read()[step]

    def load_probe(self, probe_name: str) -> np.ndarray:
        if not self._has(f"probes/{probe_name}"):
            raise KeyError(f"Probe '{probe_name}' not found. Available: {self.get_probe_names()}")
        return self._root[f"probes/{probe_name}"].read()

    def get_probe_names(self) -> list:
        return list(self._root["probes"].keys()) if self._has("probes") else []

    def get_num_snapshots(self) -> int:
        return int(self._root["fields/pressure"].shape[0]) if self._has("fields/pressure") else 0

    def load_geometry(self):
        if not self._has("materials/geometry"):
            return None
        return self._root["materials/geometry"].read().astype(bool)

    def close(self) -> None:
        self.file = None

    def __enter__(self):
        return self

    def __exit__(self, exc_type, exc_val, exc_tb):
        self.close()


__all__ = ["File", "Group", "Dataset", "walk", "storage", "HDF5ResultReader"]
