"""Reader of the HDF5 particle domains of the reference (runMPMYDW.py:19-23,
70-81: datasets "x" (3, N), "particle_volume" (N,) and optional per-particle
properties) without h5py, which is not part of this build.

It implements the part of the published HDF5 file format that these files
use, as h5py writes them by default: superblock version 0 or 1, version-1
object headers (with continuation blocks), the root group as a symbol table
(version-1 B-tree of "SNOD" nodes, names in a local heap), and datasets of
fixed-point or IEEE floating-point elements (either byte order) stored
contiguously or compactly.  Anything else (version-2 object headers, link
messages, chunked or filtered storage) raises HDF5FormatError rather than
being guessed at.
"""
from __future__ import annotations

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"


class HDF5FormatError(ValueError):
    pass


class _File:
    def __init__(self, data: bytes):
        self.d = data
        at = 0
        while data[at:at + 8] != SIGNATURE:  # the superblock may sit at 0, 512, 1024, ...
            at = 512 if at == 0 else at * 2
            if at >= len(data):
                raise HDF5FormatError("no HDF5 signature")
        self.base = at
        ver = data[at + 8]
        if ver not in (0, 1):
            raise HDF5FormatError(f"superblock version {ver} is not supported")
        self.so, self.sl = data[at + 13], data[at + 14]
        p = at + 24 + (4 if ver == 1 else 0)
        self.base_addr = self.addr(p)
        p += 4 * self.so  # base, free-space, end-of-file and driver-info addresses
        # root group symbol table entry: name offset, then the object header address
        self.root = self.addr(p + self.so)
        self.undef = (1 << (8 * self.so)) - 1

    # ---- primitives ----
    def uint(self, p: int, n: int) -> int:
        return int.from_bytes(self.d[p:p + n], "little")

    def addr(self, p: int) -> int:
        return self.uint(p, self.so)

    def length(self, p: int) -> int:
        return self.uint(p, self.sl)

    def abs(self, a: int) -> int:
        return self.base_addr + a

    # ---- object headers ----
    def messages(self, oh: int) -> list:
        """[(type, flags, body bytes)] of the object header at address oh."""
        d = self.d
        p = self.abs(oh)
        out = []
        ver = d[p]
        if ver != 1:
            raise HDF5FormatError(f"object header version {ver} is not supported")
        nmsg = self.uint(p + 2, 2)
        hsize = self.uint(p + 8, 4)
        blocks = [(p + 16, p + 16 + hsize)]  # header prefix is 12 bytes, padded to 16
        while blocks and len(out) < nmsg + 64:
            s, e = blocks.pop(0)
            while s + 8 <= e:
                mtype, msize, mflags = self.uint(s, 2), self.uint(s + 2, 2), d[s + 4]
                body = d[s + 8:s + 8 + msize]
                s += 8 + msize
                if mtype == 0x10:
                    ca = int.from_bytes(body[:self.so], "little")
                    cl = int.from_bytes(body[self.so:self.so + self.sl], "little")
                    blocks.append((self.abs(ca), self.abs(ca) + cl))
                elif mtype != 0:
                    out.append((mtype, mflags, body))
        return out

    # ---- groups ----
    def links(self, oh: int) -> dict:
        """name -> object header address of the members of the group at oh."""
        out = {}
        for mtype, _, body in self.messages(oh):
            if mtype == 0x11:  # symbol table: v1 B-tree + local heap
                btree = int.from_bytes(body[:self.so], "little")
                heap = int.from_bytes(body[self.so:2 * self.so], "little")
                self._walk_group_btree(btree, self._heap_data(heap), out)
            elif mtype in (0x02, 0x06, 0x0A):  # link info / link / group info: newer group storage
                raise HDF5FormatError("groups stored as links are not supported")
        return out

    def _heap_data(self, heap: int) -> int:
        p = self.abs(heap)
        if self.d[p:p + 4] != b"HEAP":
            raise HDF5FormatError("bad local heap")
        return self.abs(self.addr(p + 8 + 2 * self.sl))

    def _walk_group_btree(self, node: int, heap_data: int, out: dict) -> None:
        d = self.d
        p = self.abs(node)
        if d[p:p + 4] != b"TREE":
            raise HDF5FormatError("bad group B-tree node")
        level, used = d[p + 5], self.uint(p + 6, 2)
        q = p + 8 + 2 * self.so  # past the sibling addresses
        q += self.sl  # key 0
        for _ in range(used):
            child = self.addr(q)
            q += self.so + self.sl
            if level > 0:
                self._walk_group_btree(child, heap_data, out)
                continue
            s = self.abs(child)
            if d[s:s + 4] != b"SNOD":
                raise HDF5FormatError("bad symbol table node")
            nsym = self.uint(s + 6, 2)
            e = s + 8
            for _ in range(nsym):
                name_off, oh = self.length(e), self.addr(e + self.sl)
                nm = d[heap_data + name_off:d.index(b"\0", heap_data + name_off)].decode()
                out[nm] = oh
                e += self.sl + self.so + 4 + 4 + 16

    # ---- datasets ----
    def dataset(self, oh: int) -> np.ndarray:
        shape = dtype = layout = None
        for mtype, _, body in self.messages(oh):
            if mtype == 0x01:
                shape = self._dataspace(body)
            elif mtype == 0x03:
                dtype = self._datatype(body)
            elif mtype == 0x08:
                layout = body
            elif mtype == 0x0B:
                raise HDF5FormatError("filtered datasets are not supported")
        if shape is None or dtype is None or layout is None:
            raise HDF5FormatError("object is not a dataset")
        n = int(np.prod(shape)) if shape else 1
        ver = layout[0]
        if ver != 3:
            raise HDF5FormatError(f"data layout version {ver} is not supported")
        cls = layout[1]
        if cls == 0:  # compact: the data sits in the message
            size = int.from_bytes(layout[2:4], "little")
            raw = layout[4:4 + size]
        elif cls == 1:  # contiguous
            a = int.from_bytes(layout[2:2 + self.so], "little")
            size = int.from_bytes(layout[2 + self.so:2 + self.so + self.sl], "little")
            if a == self.undef:
                raw = bytes(n * dtype.itemsize)
            else:
                raw = self.d[self.abs(a):self.abs(a) + size]
        else:
            raise HDF5FormatError(f"layout class {cls} is not supported")
        return np.frombuffer(raw[:n * dtype.itemsize], dtype).reshape(shape).copy()

    def _dataspace(self, b: bytes) -> tuple:
        ver, rank, flags = b[0], b[1], b[2]
        q = 8 if ver == 1 else 4
        return tuple(int.from_bytes(b[q + k * self.sl:q + (k + 1) * self.sl], "little") for k in range(rank))

    @staticmethod
    def _datatype(b: bytes) -> np.dtype:
        cls = b[0] & 0x0F
        bits0 = b[1]
        size = int.from_bytes(b[4:8], "little")
        order = ">" if bits0 & 1 else "<"
        if cls == 0:
            kind = "i" if bits0 & 0x08 else "u"
        elif cls == 1:
            kind = "f"
        else:
            raise HDF5FormatError(f"datatype class {cls} is not supported")
        return np.dtype(f"{order}{kind}{size}")


def read_datasets(path: str, names=None) -> dict:
    """name -> numpy array for the datasets in the root group (all of them, or
    ``names`` that exist)."""
    with open(path, "rb") as f:
        data = f.read()
    F = _File(data)
    root = F.links(F.root)
    want = root.keys() if names is None else [n for n in names if n in root]
    out = {}
    for name in want:
        try:
            out[name] = F.dataset(root[name])
        except HDF5FormatError:
            if names is not None:
                raise
    return out


def load_domain_h5(path: str) -> dict:
    """The domain dict of driver.setup from a reference HDF5 domain:
    x = dataset "x" transposed to (N, 3), particle_volume, and the optional
    per-particle properties (runMPMYDW.py:19-23, 70-81)."""
    from .fs5PlotUtils import DOMAIN_PROPS
    ds = read_datasets(path, ["x", "particle_volume", *DOMAIN_PROPS])
    if "x" not in ds or "particle_volume" not in ds:
        raise HDF5FormatError(f"{path}: needs datasets 'x' and 'particle_volume'")
    dom = {"x": np.asarray(ds["x"], np.float64).T.copy(), "particle_volume": np.asarray(ds["particle_volume"],
                                                                                       np.float64)}
    for name in DOMAIN_PROPS:
        if name in ds:
            dom[name] = np.asarray(ds[name], np.float32)
    return dom
