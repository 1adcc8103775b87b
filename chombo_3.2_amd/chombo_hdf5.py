"""Chombo's AMR hierarchy files (HDF5), read and written without an HDF5 library.

Chombo writes and reads AMR data with WriteAMRHierarchyHDF5 / ReadAMRHierarchyHDF5
(lib/src/BoxTools/AMRIO.cpp:54-206, AMRIO.H:56-181) on top of CH_HDF5
(lib/src/BoxTools/CH_HDF5.H:1842-1984, CH_HDF5.cpp).  The file layout is:

  /                      attributes: filetype "VanillaAMRFileType", num_levels,
                         num_components, component_<n> (one string per component)
  /Chombo_global         attributes: SpaceDim (int), testReal (double)
  /level_<l>             attributes: dx, dt, time (double), ref_ratio (int),
                         prob_domain (box compound lo_i.. hi_k, int)
  /level_<l>/boxes       the DisjointBoxLayout: one box compound per box
  /level_<l>/Processors  the owning rank of each box (int)
  /level_<l>/data:datatype=0    every box's FAB (box grown by outputGhost),
                                component after component, Fortran order, double
  /level_<l>/data:offsets=0     start of each box in data:datatype=0 (int64), plus the total
  /level_<l>/data_attributes    attributes: comps (int), objectType "FArrayBox",
                                ghost, outputGhost (intvect compound intvecti..k)

The HDF5 part is the classic on-disk format (superblock version 0, version-1 object
headers, symbol-table groups with a version-1 B-tree, a local heap and symbol nodes,
contiguous datasets), which is what HDF5 1.8 / 1.10 write for Chombo.  The reader also
takes layout messages of versions 1-3 and compact storage, and object header
continuation blocks; chunked or filtered datasets and version-2 object headers are
refused with an error.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass, field
from typing import Dict, List, Tuple

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIG = b"\x89HDF\r\n\x1a\n"


# ===========================================================================
# reading
# ===========================================================================
@dataclass
class H5Dataset:
    name: str
    dtype: np.dtype
    shape: Tuple[int, ...]
    attrs: Dict[str, object] = field(default_factory=dict)
    _raw: bytes = b""

    def read(self) -> np.ndarray:
        n = int(np.prod(self.shape)) if self.shape else 1
        a = np.frombuffer(self._raw, dtype=self.dtype, count=n)
        return a.reshape(self.shape) if self.shape else a


@dataclass
class H5Group:
    name: str
    attrs: Dict[str, object] = field(default_factory=dict)
    members: Dict[str, object] = field(default_factory=dict)

    def __getitem__(self, path):
        node = self
        for part in [p for p in path.split("/") if p]:
            node = node.members[part]
        return node


class _File:
    def __init__(self, data: bytes):
        self.d = data
        if data[:8] != SIG:
            raise ValueError("not an HDF5 file (no signature at offset 0)")
        sbv = data[8]
        if sbv not in (0, 1):
            raise ValueError(f"HDF5 superblock version {sbv} is not supported (0 and 1 are)")
        self.so, self.sl = data[13], data[14]
        if self.so != 8 or self.sl not in (4, 8):
            raise ValueError("only 8-byte offsets and 4- or 8-byte lengths are supported")
        self.leafK, self.internalK = struct.unpack_from("<HH", data, 16)
        p = 24 + (4 if sbv == 1 else 0)
        self.base, _, self.eof, _ = struct.unpack_from("<QQQQ", data, p)
        self.root_entry = p + 32
        self.ent = self.sl + self.so + 24  # symbol table entry size

    def u(self, fmt, off):
        return struct.unpack_from("<" + fmt, self.d, off)

    def L(self, off):
        """a 'size of lengths' field"""
        return struct.unpack_from("<I" if self.sl == 4 else "<Q", self.d, off)[0]

    # ---- symbol table entries, groups ----
    def entry(self, off):
        name_off = self.L(off)
        hdr, cache = self.u("QI", off + self.sl)
        return name_off, hdr, cache

    def heap_data(self, addr):
        if self.d[addr:addr + 4] != b"HEAP":
            raise ValueError("local heap signature missing")
        return self.u("Q", addr + 8 + 2 * self.sl)[0]

    def heap_string(self, data_addr, off):
        s = data_addr + off
        e = self.d.index(b"\0", s)
        return self.d[s:e].decode()

    def btree_group_entries(self, addr, heap_addr, out):
        if self.d[addr:addr + 4] != b"TREE":
            raise ValueError("B-tree signature missing")
        ntype, level, nent = self.d[addr + 4], self.d[addr + 5], self.u("H", addr + 6)[0]
        if ntype != 0:
            raise ValueError("group B-tree expected")
        p = addr + 24
        for c in range(nent):
            child = self.u("Q", p + self.sl + (self.sl + 8) * c)[0]
            if level > 0:
                self.btree_group_entries(child, heap_addr, out)
            else:
                self.snod_entries(child, heap_addr, out)

    def snod_entries(self, addr, heap_addr, out):
        if self.d[addr:addr + 4] != b"SNOD":
            raise ValueError("symbol node signature missing")
        nsym = self.u("H", addr + 6)[0]
        data_addr = self.heap_data(heap_addr)
        for q in range(nsym):
            name_off, hdr, _ = self.entry(addr + 8 + self.ent * q)
            out.append((self.heap_string(data_addr, name_off), hdr))

    # ---- object headers ----
    def messages(self, addr):
        d = self.d
        if d[addr] != 1:
            raise ValueError(f"object header version {d[addr]} is not supported (version 1 is)")
        nmsg, _ref, hsize = self.u("HII", addr + 2)
        blocks = [(addr + 16, hsize)]
        out = []
        while blocks and len(out) < nmsg:
            start, size = blocks.pop(0)
            p = start
            while p + 8 <= start + size and len(out) < nmsg:
                mtype, msize, flags = self.u("HHB", p)
                body = p + 8
                if mtype == 0x10:  # continuation
                    caddr, clen = self.u("Q", body)[0], self.L(body + 8)
                    blocks.append((caddr, clen))
                out.append((mtype, body, msize, flags))
                p = body + msize
        return out

    def obj(self, name, addr):
        msgs = self.messages(addr)
        attrs = {}
        stab = None
        space = dtype = layout = None
        for mtype, body, msize, flags in msgs:
            if mtype == 0x11:
                stab = self.u("QQ", body)
            elif mtype == 0x0C:
                k, v = self.attribute(body)
                attrs[k] = v
            elif mtype == 0x01:
                space = self.dataspace(body)
            elif mtype == 0x03:
                dtype = self.datatype(body)[0]
            elif mtype == 0x08:
                layout = (body, msize)
            elif mtype == 0x0B:
                raise ValueError(f"{name}: filtered datasets are not supported")
        if stab is not None:
            g = H5Group(name, attrs)
            ents = []
            self.btree_group_entries(stab[0], stab[1], ents)
            for n, hdr in ents:
                g.members[n] = self.obj(n, hdr)
            return g
        if space is None or dtype is None or layout is None:
            raise ValueError(f"{name}: neither a group nor a dataset")
        raw = self.layout_data(layout[0], space, dtype)
        return H5Dataset(name, dtype, space, attrs, raw)

    def layout_data(self, body, shape, dtype):
        d = self.d
        ver = d[body]
        nbytes = int(np.prod(shape) if shape else 1) * dtype.itemsize
        if ver == 3:
            cls = d[body + 1]
            if cls == 0:
                size = self.u("H", body + 2)[0]
                return d[body + 4:body + 4 + size]
            if cls == 1:
                addr, size = self.u("Q", body + 2)[0], self.L(body + 10)
                return d[addr:addr + size] if addr != UNDEF else bytes(nbytes)
            raise ValueError("chunked datasets are not supported")
        if ver in (1, 2):
            rank, cls = d[body + 1], d[body + 2]
            if cls == 1:
                addr = self.u("Q", body + 8)[0]
                return d[addr:addr + nbytes] if addr != UNDEF else bytes(nbytes)
            if cls == 0:
                p = body + 8 + 4 * rank
                size = self.u("I", p)[0]
                return d[p + 4:p + 4 + size]
            raise ValueError("chunked datasets are not supported")
        raise ValueError(f"layout message version {ver} is not supported")

    def dataspace(self, body):
        d = self.d
        ver, rank, flags = d[body], d[body + 1], d[body + 2]
        if ver == 1:
            p = body + 8
        elif ver == 2:
            if d[body + 3] == 0:
                return ()
            p = body + 4
        else:
            raise ValueError(f"dataspace version {ver} is not supported")
        return tuple(self.L(p + self.sl * q) for q in range(rank))

    def datatype(self, body):
        """(numpy dtype, bytes of the message)"""
        d = self.d
        cv = d[body]
        cls, ver = cv & 0x0F, cv >> 4
        bits = d[body + 1] | (d[body + 2] << 8) | (d[body + 3] << 16)
        size = self.u("I", body + 4)[0]
        end = "<" if not (bits & 1) else ">"
        if cls == 0:
            signed = bool(bits & 0x08)
            return np.dtype(f"{end}{'i' if signed else 'u'}{size}"), 12
        if cls == 1:
            return np.dtype(f"{end}f{size}"), 20
        if cls == 3:
            return np.dtype(f"S{size}"), 8
        if cls == 6:
            nmem = bits & 0xFFFF
            p = body + 8
            names, offs, fmts = [], [], []
            for _ in range(nmem):
                e = d.index(b"\0", p)
                nm = d[p:e].decode()
                if ver < 3:
                    p += ((e - p + 1) + 7) // 8 * 8
                else:
                    p = e + 1
                if ver < 3:
                    off = self.u("I", p)[0]
                    p += 4
                else:
                    nb = 1 if size < 256 else 2 if size < 65536 else 3 if size < (1 << 24) else 4
                    off = int.from_bytes(d[p:p + nb], "little")
                    p += nb
                if ver == 1:
                    p += 28  # dimensionality, reserved, permutation, reserved, 4 dimension sizes
                mt, mlen = self.datatype(p)
                p += mlen
                names.append(nm)
                offs.append(off)
                fmts.append(mt)
            return np.dtype({"names": names, "formats": fmts, "offsets": offs, "itemsize": size}), p - body
        raise ValueError(f"datatype class {cls} is not supported")

    def attribute(self, body):
        d = self.d
        ver = d[body]
        nlen, tlen, slen = self.u("HHH", body + 2)
        if ver == 1:
            p = body + 8
            pad = lambda n: (n + 7) // 8 * 8
        elif ver in (2, 3):
            p = body + 8 + (1 if ver == 3 else 0)
            pad = lambda n: n
        else:
            raise ValueError(f"attribute message version {ver} is not supported")
        name = d[p:p + nlen].split(b"\0")[0].decode()
        p += pad(nlen)
        dt, _ = self.datatype(p)
        p += pad(tlen)
        shape = self.dataspace(p)
        p += pad(slen)
        n = int(np.prod(shape)) if shape else 1
        val = np.frombuffer(d[p:p + n * dt.itemsize], dtype=dt, count=n)
        if dt.kind == "S":
            v = [x.split(b"\0")[0].decode() for x in val]
            return name, v[0] if not shape else v
        if dt.names:
            return name, val[0] if not shape else val
        return name, val[0].item() if not shape else val.copy()


def read_file(path) -> H5Group:
    """The whole group tree of an HDF5 file (datasets keep their raw bytes)."""
    with open(path, "rb") as f:
        data = f.read()
    F = _File(data)
    _, hdr, _ = F.entry(F.root_entry)
    return F.obj("/", hdr)


# ===========================================================================
# writing
# ===========================================================================
def _pad8(b: bytes) -> bytes:
    return b + b"\0" * ((-len(b)) % 8)


def _dt_int(size=4, signed=True):
    return struct.pack("<BBBBI", 0x10 | 0, 0x08 if signed else 0, 0, 0, size) + struct.pack("<HH", 0, size * 8)


def _dt_double():
    # IEEE little-endian binary64: sign bit 63, exponent 52..62 (bias 1023), mantissa 0..51
    return struct.pack("<BBBBI", 0x10 | 1, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)


def _dt_string(n):
    # fixed-length, null-terminated, ASCII (H5T_C_S1 with H5Tset_size)
    return struct.pack("<BBBBI", 0x10 | 3, 0, 0, 0, n)


def _dt_compound(members, size):
    """members: [(name, offset, member datatype bytes)], compound datatype version 1."""
    body = b""
    for name, off, mt in members:
        body += _pad8(name.encode() + b"\0")
        body += struct.pack("<IB3sI4s4I", off, 0, b"\0\0\0", 0, b"\0\0\0\0", 0, 0, 0, 0)
        body += mt
    return struct.pack("<BBBBI", 0x10 | 6, len(members) & 0xFF, len(members) >> 8, 0, size) + body


_AXES = "ijkl"


def _dt_box(dim):
    m = [(f"lo_{_AXES[d]}", 4 * d, _dt_int()) for d in range(dim)]
    m += [(f"hi_{_AXES[d]}", 4 * (dim + d), _dt_int()) for d in range(dim)]
    return _dt_compound(m, 4 * (2 * dim + 1))  # sizeof(Box): smallEnd, bigEnd and the index type


def _dt_intvect(dim):
    return _dt_compound([(f"intvect{_AXES[d]}", 4 * d, _dt_int()) for d in range(dim)], 4 * dim)


def _space(shape):
    """dataspace message version 1 (rank 0 = scalar)."""
    b = struct.pack("<BBBB4s", 1, len(shape), 0, 0, b"\0" * 4)
    for n in shape:
        b += struct.pack("<Q", n)
    return b


class _Writer:
    """Builds the file in memory: superblock, then objects appended in order."""

    LEAF_K = 32   # symbol node capacity 2K entries
    INTERNAL_K = 16

    def __init__(self):
        self.buf = bytearray(96)

    def alloc(self, data: bytes) -> int:
        while len(self.buf) % 8:
            self.buf.append(0)
        a = len(self.buf)
        self.buf += data
        return a

    def reserve(self, n) -> int:
        return self.alloc(bytes(n))

    def put(self, addr, data):
        self.buf[addr:addr + len(data)] = data

    @staticmethod
    def message(mtype, body, flags=0):
        body = _pad8(body)
        return struct.pack("<HHB3s", mtype, len(body), flags, b"\0\0\0") + body

    def object_header(self, msgs: List[bytes]) -> int:
        body = b"".join(msgs)
        hdr = struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4
        return self.alloc(hdr + body)

    def attribute(self, name, dt: bytes, shape, data: bytes):
        nm = name.encode() + b"\0"
        sp = _space(shape)
        body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(sp)) + _pad8(nm) + _pad8(dt) + _pad8(sp) + data
        return self.message(0x0C, body)

    def dataset(self, array: np.ndarray, dt: bytes, attrs=()) -> int:
        raw = np.ascontiguousarray(array).tobytes()
        addr = self.alloc(raw) if raw else UNDEF
        layout = struct.pack("<BBQQ", 3, 1, addr, len(raw))
        fill = struct.pack("<BBBB", 2, 2, 2, 0)  # fill value message v2: allocate late, write if set, undefined
        msgs = [self.message(0x01, _space(array.shape)), self.message(0x03, dt), self.message(0x05, fill),
                self.message(0x08, layout)] + list(attrs)
        return self.object_header(msgs)

    def group(self, entries: Dict[str, int], attrs=()) -> int:
        """A symbol-table group over entries {name: object header address}."""
        names = sorted(entries)
        if len(names) > 2 * self.LEAF_K:
            raise ValueError("group too large for one symbol node")
        heap = bytearray(b"\0" * 8)  # offset 0: the empty string (the B-tree's first key)
        offs = {}
        for n in names:
            offs[n] = len(heap)
            heap += _pad8(n.encode() + b"\0")
        heap_data = self.alloc(bytes(heap))
        heap_addr = self.alloc(b"HEAP" + struct.pack("<B3sQQQ", 0, b"\0\0\0", len(heap), UNDEF, heap_data))
        snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
        for n in names:
            snod += struct.pack("<QQII16s", offs[n], entries[n], 0, 0, b"\0" * 16)
        snod += b"\0" * 40 * (2 * self.LEAF_K - len(names))
        snod_addr = self.alloc(snod)
        last = offs[names[-1]] if names else 0
        tree = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, UNDEF, UNDEF)
        tree += struct.pack("<QQQ", 0, snod_addr, last)
        tree += b"\0" * (16 * 2 * self.INTERNAL_K - 16)  # room for the rest of a full node
        tree_addr = self.alloc(tree)
        msgs = [self.message(0x11, struct.pack("<QQ", tree_addr, heap_addr))] + list(attrs)
        return self.object_header(msgs)

    def finish(self, root_hdr: int, root_tree_heap: Tuple[int, int]) -> bytes:
        b = self.buf
        b[0:8] = SIG
        b[8:16] = bytes([0, 0, 0, 0, 0, 8, 8, 0])
        struct.pack_into("<HHI", b, 16, self.LEAF_K, self.INTERNAL_K, 0)
        struct.pack_into("<QQQQ", b, 24, 0, UNDEF, len(b), UNDEF)
        struct.pack_into("<QQII", b, 56, 0, root_hdr, 1, 0)
        struct.pack_into("<QQ", b, 80, *root_tree_heap)
        return bytes(b)


def _stab_of(w: _Writer, hdr: int) -> Tuple[int, int]:
    body = hdr + 16 + 8  # the group's first message is its symbol table message
    return struct.unpack_from("<QQ", w.buf, body)


# ===========================================================================
# Chombo AMR hierarchies
# ===========================================================================
@dataclass
class AMRLevelData:
    """One level of a Chombo AMR file: boxes (lo, hi) tuples in layout order, the
    level's problem domain, dx, dt, time, ref_ratio (to the next finer level), and
    data per box as an array [ncomp, nx, ny, (nz)] over the box grown by ghost."""
    boxes: List[Tuple[Tuple[int, ...], Tuple[int, ...]]]
    domain: Tuple[Tuple[int, ...], Tuple[int, ...]]
    dx: float
    dt: float
    time: float
    ref_ratio: int
    data: List[np.ndarray]
    ghost: Tuple[int, ...]
    processors: List[int] = field(default_factory=list)


@dataclass
class AMRFile:
    names: List[str]
    levels: List[AMRLevelData]
    space_dim: int = 3
    test_real: float = 0.0
    filetype: str = "VanillaAMRFileType"


def _box_from(rec, dim):
    return (tuple(int(rec[f"lo_{_AXES[d]}"]) for d in range(dim)),
            tuple(int(rec[f"hi_{_AXES[d]}"]) for d in range(dim)))


def read_amr_hierarchy(path) -> AMRFile:
    """ReadAMRHierarchyHDF5 (AMRIO.H:56-181 over CH_HDF5.H readLevel / read): names,
    levels with their boxes, domain, dx, dt, time, ref_ratio and per-box FAB data."""
    root = read_file(path)
    g = root.members.get("Chombo_global")
    dim = int(g.attrs.get("SpaceDim", 3)) if g is not None else 3
    test_real = float(g.attrs.get("testReal", 0.0)) if g is not None else 0.0
    nlev = int(root.attrs.get("num_levels", 0))
    ncomp = int(root.attrs.get("num_components", 0))
    names = [root.attrs.get(f"component_{c}", "") for c in range(ncomp)]
    levels = []
    for l in range(nlev):
        L = root.members[f"level_{l}"]
        boxes = [_box_from(r, dim) for r in L.members["boxes"].read()] if "boxes" in L.members else []
        procs = [int(v) for v in L.members["Processors"].read()] if "Processors" in L.members else []
        dom = _box_from(L.attrs["prob_domain"], dim)
        ghost = (0,) * dim
        data = []
        if "data:datatype=0" in L.members:
            da = L.members.get("data_attributes")
            if da is not None and "outputGhost" in da.attrs:
                og = da.attrs["outputGhost"]
                ghost = tuple(int(og[f"intvect{_AXES[d]}"]) for d in range(dim))
            comps = int(da.attrs.get("comps", ncomp)) if da is not None else ncomp
            flat = L.members["data:datatype=0"].read()
            offs = L.members["data:offsets=0"].read()
            if len(offs) != len(boxes) + 1 or offs[-1] != flat.size:
                raise ValueError(f"level {l}: offsets do not match the data")
            for b, (lo, hi) in enumerate(boxes):
                shp = tuple(hi[d] - lo[d] + 1 + 2 * ghost[d] for d in range(dim))
                n = int(np.prod(shp))
                chunk = flat[offs[b]:offs[b + 1]]
                if chunk.size != n * comps:
                    raise ValueError(f"level {l} box {b}: {chunk.size} values for {comps} x {n} cells")
                data.append(chunk.reshape((comps,) + shp[::-1]).transpose((0,) + tuple(range(dim, 0, -1))))
        levels.append(AMRLevelData(boxes, dom, float(L.attrs.get("dx", 0.0)), float(L.attrs.get("dt", 0.0)),
                                   float(L.attrs.get("time", 0.0)), int(L.attrs.get("ref_ratio", 1)), data, ghost,
                                   procs))
    return AMRFile(names, levels, dim, test_real, root.attrs.get("filetype", ""))


def write_amr_hierarchy(path, amr: AMRFile):
    """WriteAMRHierarchyHDF5 (AMRIO.cpp:54-206, CH_HDF5.H:1842-1880 writeLevel, write of
    the BoxLayout and of the BoxLayoutData): the layout above."""
    dim = amr.space_dim
    w = _Writer()
    box_t, iv_t = _dt_box(dim), _dt_intvect(dim)
    box_np = np.dtype({"names": [f"lo_{_AXES[d]}" for d in range(dim)] + [f"hi_{_AXES[d]}" for d in range(dim)],
                       "formats": ["<i4"] * (2 * dim), "offsets": [4 * q for q in range(2 * dim)],
                       "itemsize": 4 * (2 * dim + 1)})
    iv_np = np.dtype({"names": [f"intvect{_AXES[d]}" for d in range(dim)], "formats": ["<i4"] * dim})

    def box_bytes(b):
        return np.array([tuple(b[0]) + tuple(b[1])], dtype=box_np).tobytes()

    def str_attr(name, s):
        raw = s.encode()
        return w.attribute(name, _dt_string(max(len(raw), 1)), (), raw if raw else b"\0")

    entries = {}
    glob = w.group({}, [w.attribute("SpaceDim", _dt_int(), (), struct.pack("<i", dim)),
                        w.attribute("testReal", _dt_double(), (), struct.pack("<d", amr.test_real))])
    entries["Chombo_global"] = glob
    for l, L in enumerate(amr.levels):
        members = {}
        members["boxes"] = w.dataset(np.array([tuple(b[0]) + tuple(b[1]) for b in L.boxes], dtype=box_np), box_t)
        procs = L.processors or [0] * len(L.boxes)
        members["Processors"] = w.dataset(np.asarray(procs, dtype="<i4"), _dt_int())
        if L.data:
            comps = L.data[0].shape[0]
            parts, offs = [], [0]
            for b, (lo, hi) in enumerate(L.boxes):
                a = np.asarray(L.data[b], dtype="<f8")
                shp = tuple(hi[d] - lo[d] + 1 + 2 * L.ghost[d] for d in range(dim))
                if a.shape != (comps,) + shp:
                    raise ValueError(f"level {l} box {b}: data shape {a.shape}, expected {(comps,) + shp}")
                flat = a.transpose((0,) + tuple(range(dim, 0, -1))).reshape(-1)
                parts.append(flat)
                offs.append(offs[-1] + flat.size)
            members["data:datatype=0"] = w.dataset(np.concatenate(parts) if parts else np.zeros(0), _dt_double())
            members["data:offsets=0"] = w.dataset(np.asarray(offs, dtype="<i8"), _dt_int(8))
            ogb = np.array([tuple(L.ghost)], dtype=iv_np).tobytes()
            members["data_attributes"] = w.group({}, [
                w.attribute("comps", _dt_int(), (), struct.pack("<i", comps)),
                w.attribute("ghost", iv_t, (), ogb),
                str_attr("objectType", "FArrayBox"),
                w.attribute("outputGhost", iv_t, (), ogb)])
        attrs = [w.attribute("dt", _dt_double(), (), struct.pack("<d", L.dt)),
                 w.attribute("dx", _dt_double(), (), struct.pack("<d", L.dx)),
                 w.attribute("prob_domain", box_t, (), box_bytes(L.domain)),
                 w.attribute("ref_ratio", _dt_int(), (), struct.pack("<i", L.ref_ratio)),
                 w.attribute("time", _dt_double(), (), struct.pack("<d", L.time))]
        entries[f"level_{l}"] = w.group(members, attrs)
    rattrs = [w.attribute("num_components", _dt_int(), (), struct.pack("<i", len(amr.names))),
              w.attribute("num_levels", _dt_int(), (), struct.pack("<i", len(amr.levels)))]
    for c, nm in enumerate(amr.names):
        rattrs.append(str_attr(f"component_{c}", nm))
    rattrs.append(str_attr("filetype", amr.filetype))
    root = w.group(entries, rattrs)
    data = w.finish(root, _stab_of(w, root))
    with open(path, "wb") as f:
        f.write(data)


# ---- between Chombo files and this package's hierarchies and valid-packed fields ----
def hierarchy_from_file(amr: AMRFile, **kw):
    """A hierarchy.Hierarchy of the file's levels (3-D, refinement ratio 2 throughout):
    level-0 domain and boxes from the file, dx0 from level 0; kw go to Hierarchy."""
    import hierarchy as H
    if amr.space_dim != 3:
        raise ValueError("the solver takes 3-D hierarchies")
    for L in amr.levels[:-1]:
        if L.ref_ratio != 2:
            raise ValueError("the solver takes refinement ratio 2")
    levels = [[H.box(lo, hi) for (lo, hi) in L.boxes] for L in amr.levels]
    return H.Hierarchy(H.box(*amr.levels[0].domain), levels, amr.levels[0].dx, **kw)


def field_from_file(amr: AMRFile, comp: int = 0) -> np.ndarray:
    """Component comp of every level as one valid-packed array (level by level, box
    by box in the file's layout order, Fortran order), ghosts dropped."""
    parts = []
    for L in amr.levels:
        for b, (lo, hi) in enumerate(L.boxes):
            a = L.data[b][comp]
            g = L.ghost
            sl = tuple(slice(g[d], a.shape[d] - g[d]) for d in range(len(g)))
            parts.append(np.asarray(a[sl]).ravel(order="F"))
    return np.concatenate(parts) if parts else np.zeros(0)


def file_from_hierarchy(hier, fields: Dict[str, np.ndarray], dt=0.0, time=0.0) -> AMRFile:
    """An AMRFile of valid-packed fields (name -> array over all levels) on a
    hierarchy.Hierarchy, ghost width 0, as WriteAMRHierarchyHDF5 writes them."""
    names = list(fields)
    levels = []
    offs = hier.level_offsets()
    for l in range(hier.nlev):
        data = []
        o = offs[l]
        for (lo, hi) in hier.levels[l]:
            shp = tuple(hi[d] - lo[d] + 1 for d in range(3))
            n = int(np.prod(shp))
            data.append(np.stack([np.asarray(fields[k][o:o + n]).reshape(shp, order="F") for k in names]))
            o += n
        dom = hier.level_domain(l)
        levels.append(AMRLevelData([tuple(map(tuple, b)) for b in hier.levels[l]], dom, hier.dx(l),
                                   dt / (2 ** l), time, 2 if l + 1 < hier.nlev else 1, data, (0, 0, 0)))
    return AMRFile(names, levels, 3, 0.0)
