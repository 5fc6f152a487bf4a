"""VTK XML files (PolyData .vtp, ImageData .vti) without the vtk package.

The reference writes its particle and grid snapshots, and the restart data of
a checkpoint, with vtkXMLPolyDataWriter / vtkXMLImageDataWriter and reads the
restart back with vtkXMLPolyDataReader (fs5PlotUtils.py:253-646).  vtk is not
part of this build, so this module implements the published VTK XML format:

* DataArray formats "ascii", "binary" (base64 inline) and "appended" (with
  AppendedData encoding "base64" or "raw"), header_type UInt32 or UInt64,
  byte_order LittleEndian, with or without vtkZLibDataCompressor;
* compressed data is a header [#blocks, block size, size of the last block
  (0 when it is full), compressed size of each block] followed by the
  zlib-compressed blocks; in base64 form the header and the blocks are
  encoded separately, uncompressed data carries a one-integer byte count.

The writer emits what VTK's writers emit in binary mode with the zlib
compressor and 32 KiB blocks, including the RangeMin/RangeMax attributes
(component range for scalars, L2-norm range for vectors).
"""
from __future__ import annotations

import base64
import re
import xml.etree.ElementTree as ET
import zlib

import numpy as np

_DTYPES = {"Int8": "<i1", "UInt8": "<u1", "Int16": "<i2", "UInt16": "<u2", "Int32": "<i4", "UInt32": "<u4",
           "Int64": "<i8", "UInt64": "<u8", "Float32": "<f4", "Float64": "<f8"}
_NAMES = {np.dtype(v): k for k, v in _DTYPES.items()}
BLOCK = 32768


class VTKFormatError(ValueError):
    pass


# ---------------------------------------------------------------------------
# DataArray payloads
# ---------------------------------------------------------------------------
def _b64len(nbytes: int) -> int:
    return 4 * ((nbytes + 2) // 3)


def _decode_compressed_b64(text: str, hdt: np.dtype) -> bytes:
    hs = hdt.itemsize
    first = base64.b64decode(text[:_b64len(hs)])[:hs]
    nblocks = int(np.frombuffer(first, hdt)[0])
    hlen = _b64len(hs * (3 + nblocks))
    header = np.frombuffer(base64.b64decode(text[:hlen]), hdt)[:3 + nblocks]
    data = base64.b64decode(text[hlen:])
    return _inflate(header, data)


def _inflate(header: np.ndarray, data: bytes) -> bytes:
    nblocks, bsize, last = (int(v) for v in header[:3])
    sizes = [int(v) for v in header[3:3 + nblocks]]
    out = []
    pos = 0
    for k, cs in enumerate(sizes):
        raw = zlib.decompress(data[pos:pos + cs])
        want = (last if last else bsize) if k == nblocks - 1 else bsize
        if len(raw) != want:
            raise VTKFormatError(f"zlib block {k} holds {len(raw)} bytes, header says {want}")
        out.append(raw)
        pos += cs
    return b"".join(out)


def _decode_plain_b64(text: str, hdt: np.dtype) -> bytes:
    hs = hdt.itemsize
    # VTK encodes the byte count and the data either as two base64 units or as one
    n = int(np.frombuffer(base64.b64decode(text[:_b64len(hs)])[:hs], hdt)[0])
    rest = base64.b64decode(text[_b64len(hs):]) if len(text) > _b64len(hs) else b""
    if len(rest) == n:
        return rest
    both = base64.b64decode(text)
    if len(both) < hs + n:
        raise VTKFormatError("binary DataArray shorter than its byte count")
    return both[hs:hs + n]


def decode_array(el: ET.Element, hdt: np.dtype, compressed: bool, appended: bytes | None, raw_appended: bool):
    """The values of one DataArray element as a numpy array (tuples x components)."""
    dt = np.dtype(_DTYPES[el.get("type")])
    ncomp = int(el.get("NumberOfComponents", "1"))
    fmt = el.get("format", "ascii")
    if fmt == "ascii":
        vals = np.array((el.text or "").split(), dtype=dt)
    elif fmt == "binary":
        text = re.sub(r"\s+", "", el.text or "")
        buf = _decode_compressed_b64(text, hdt) if compressed else _decode_plain_b64(text, hdt)
        vals = np.frombuffer(buf, dt).copy()
    elif fmt == "appended":
        if appended is None:
            raise VTKFormatError("appended DataArray without an AppendedData section")
        off = int(el.get("offset"))
        if raw_appended:
            hs = hdt.itemsize
            if compressed:
                nb = int(np.frombuffer(appended[off:off + hs], hdt)[0])
                header = np.frombuffer(appended[off:off + hs * (3 + nb)], hdt)
                start = off + hs * (3 + nb)
                buf = _inflate(header, appended[start:start + int(header[3:].sum())])
            else:
                n = int(np.frombuffer(appended[off:off + hs], hdt)[0])
                buf = appended[off + hs:off + hs + n]
        else:
            text = appended[off:].decode("ascii")
            buf = _decode_compressed_b64(text, hdt) if compressed else _decode_plain_b64(text, hdt)
        vals = np.frombuffer(buf, dt).copy()
    else:
        raise VTKFormatError(f"unknown DataArray format {fmt!r}")
    if ncomp > 1:
        vals = vals.reshape(-1, ncomp)
    return vals


def encode_binary(values: np.ndarray, header_type: str = "UInt32", compress: bool = True, level: int = 6) -> str:
    """base64 payload of a binary DataArray (compressed as VTK's zlib compressor does)."""
    raw = np.ascontiguousarray(values).tobytes()
    hdt = np.dtype(_DTYPES[header_type])
    if not compress:
        return base64.b64encode(np.array([len(raw)], hdt).tobytes()).decode() + base64.b64encode(raw).decode()
    blocks = [raw[i:i + BLOCK] for i in range(0, len(raw), BLOCK)]  # none for an empty array
    comp = [zlib.compress(b, level) for b in blocks]
    last = len(blocks[-1]) if blocks and len(blocks[-1]) != BLOCK else 0
    header = np.array([len(blocks), BLOCK, last] + [len(c) for c in comp], hdt)
    return base64.b64encode(header.tobytes()).decode() + base64.b64encode(b"".join(comp)).decode()


def _range(values: np.ndarray) -> tuple[float, float]:
    v = np.asarray(values)
    if v.size == 0:
        return 1e299, -1e299
    if v.ndim == 2 and v.shape[1] > 1:
        v = np.sqrt(np.sum(v.astype(np.float64) ** 2, axis=1))
    return float(np.min(v)), float(np.max(v))


def _fmt_num(x: float, integer: bool) -> str:
    """As VTK prints a range: integers plainly, reals with 17 significant digits."""
    if abs(x) < 1e18 and (integer or float(x).is_integer()):
        return str(int(x))
    return f"{x:.17g}"


def data_array_xml(name: str | None, values: np.ndarray, indent: str, header_type: str = "UInt32",
                   compress: bool = True, extra: str = "") -> str:
    v = np.asarray(values)
    dt = np.dtype(v.dtype).newbyteorder("<")
    if dt not in _NAMES:
        raise VTKFormatError(f"no VTK type for {v.dtype}")
    ncomp = v.shape[1] if v.ndim == 2 else 1
    lo, hi = _range(v)
    integer = np.issubdtype(dt, np.integer)
    attrs = f'type="{_NAMES[dt]}"'
    if name is not None:
        attrs += f' Name="{name}"'
    if ncomp > 1:
        attrs += f' NumberOfComponents="{ncomp}"'
    attrs += extra
    attrs += f' format="binary" RangeMin="{_fmt_num(lo, integer)}" RangeMax="{_fmt_num(hi, integer)}"'
    payload = encode_binary(v.astype(dt, copy=False), header_type, compress)
    return f"{indent}<DataArray {attrs}>\n{indent}  {payload}\n{indent}</DataArray>\n"


# ---------------------------------------------------------------------------
# files
# ---------------------------------------------------------------------------
def _split_appended(blob: bytes):
    """(XML text with the AppendedData payload removed, payload bytes, raw?)."""
    k = blob.find(b"<AppendedData")
    if k < 0:
        return blob.decode("utf-8"), None, False
    head_end = blob.index(b">", k)
    tag = blob[k:head_end + 1].decode("ascii")
    raw = 'encoding="raw"' in tag
    start = blob.index(b"_", head_end) + 1
    end = blob.rindex(b"</AppendedData>")
    payload = blob[start:end]
    if not raw:
        payload = re.sub(rb"\s+", b"", payload)
    else:
        payload = payload.rstrip(b"\n\r\t ")
    text = blob[:k].decode("utf-8") + "<AppendedData/>" + blob[end + len(b"</AppendedData>"):].decode("utf-8")
    return text, payload, raw


def read(path: str) -> dict:
    """Read a .vtp or .vti file: {"type", "points" (n,3) or None, "point_data",
    "cell_data", "field_data" (name -> array), "verts" (connectivity, offsets),
    and for ImageData "extent", "origin", "spacing"}."""
    with open(path, "rb") as f:
        blob = f.read()
    text, appended, raw = _split_appended(blob)
    root = ET.fromstring(text)
    if root.tag != "VTKFile":
        raise VTKFormatError(f"{path}: not a VTK XML file")
    if root.get("byte_order", "LittleEndian") != "LittleEndian":
        raise VTKFormatError(f"{path}: big-endian VTK files are not supported")
    hdt = np.dtype(_DTYPES[root.get("header_type", "UInt32")])
    compressed = root.get("compressor") is not None
    if compressed and root.get("compressor") != "vtkZLibDataCompressor":
        raise VTKFormatError(f"{path}: compressor {root.get('compressor')} is not supported")
    kind = root.get("type")
    ds = root.find(kind)
    piece = ds.find("Piece")

    def arrays(parent):
        out = {}
        if parent is None:
            return out
        for el in parent.findall("DataArray"):
            out[el.get("Name")] = decode_array(el, hdt, compressed, appended, raw)
        return out

    res = {"type": kind, "point_data": arrays(piece.find("PointData")), "cell_data": arrays(piece.find("CellData")),
           "field_data": arrays(ds.find("FieldData")), "points": None, "verts": None}
    pts = piece.find("Points")
    if pts is not None:
        res["points"] = decode_array(pts.find("DataArray"), hdt, compressed, appended, raw).reshape(-1, 3)
        n = int(piece.get("NumberOfPoints"))
        if res["points"].shape[0] != n:
            raise VTKFormatError(f"{path}: {res['points'].shape[0]} points, header says {n}")
    verts = piece.find("Verts")
    if verts is not None:
        a = arrays(verts)
        res["verts"] = (a.get("connectivity"), a.get("offsets"))
    if kind == "ImageData":
        res["extent"] = [int(v) for v in ds.get("WholeExtent").split()]
        res["origin"] = [float(v) for v in ds.get("Origin").split()]
        res["spacing"] = [float(v) for v in ds.get("Spacing").split()]
    return res


def _header(kind: str, header_type: str, compress: bool) -> str:
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    version = "0.1" if header_type == "UInt32" else "1.0"  # VTK's own numbering of the two header types
    return (f'<?xml version="1.0"?>\n<VTKFile type="{kind}" version="{version}" byte_order="LittleEndian" '
            f'header_type="{header_type}"{comp}>\n')


def write_vtp(path: str, points: np.ndarray, point_data: dict | None = None, field_data: dict | None = None,
              verts: bool = False, header_type: str = "UInt32", compress: bool = True) -> None:
    """PolyData with the given points, point arrays (n,) or (n, k) and field
    arrays; with ``verts`` one vertex cell per point."""
    pts = np.asarray(points)
    n = pts.shape[0]
    point_data = point_data or {}
    kw = {"header_type": header_type, "compress": compress}
    out = [_header("PolyData", header_type, compress), "  <PolyData>\n"]
    if field_data:
        out.append("    <FieldData>\n")
        for name, v in field_data.items():
            v = np.atleast_1d(np.asarray(v))
            out.append(data_array_xml(name, v, "      ", extra=f' NumberOfTuples="{v.shape[0]}"', **kw))
        out.append("    </FieldData>\n")
    nv = n if verts else 0
    out.append(f'    <Piece NumberOfPoints="{n}" NumberOfVerts="{nv}" NumberOfLines="0" NumberOfStrips="0" '
               f'NumberOfPolys="0">\n')
    out.append("      <PointData>\n")
    for name, v in point_data.items():
        v = np.asarray(v)
        if v.shape[0] != n:
            raise VTKFormatError(f"point array {name} has {v.shape[0]} tuples for {n} points")
        out.append(data_array_xml(name, v.reshape(n, -1) if v.ndim > 1 else v, "        ", **kw))
    out.append("      </PointData>\n      <CellData>\n      </CellData>\n      <Points>\n")
    out.append(data_array_xml("Points", pts.reshape(n, 3), "        ", **kw))
    out.append("      </Points>\n")
    for cell in ("Verts", "Lines", "Strips", "Polys"):
        out.append(f"      <{cell}>\n")
        m = nv if cell == "Verts" else 0
        out.append(data_array_xml("connectivity", np.arange(m, dtype=np.int64), "        ", **kw))
        out.append(data_array_xml("offsets", np.arange(1, m + 1, dtype=np.int64), "        ", **kw))
        out.append(f"      </{cell}>\n")
    out.append("    </Piece>\n  </PolyData>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(out))


def write_vti(path: str, dims, origin, spacing, point_data: dict, scalars: str | None = None,
              header_type: str = "UInt32", compress: bool = True) -> None:
    """ImageData of dims = (nx, ny, nz) points; each point array is given in
    VTK's point order (x fastest)."""
    nx, ny, nz = (int(d) for d in dims)
    ext = f"0 {nx - 1} 0 {ny - 1} 0 {nz - 1}"
    org = " ".join(repr(float(v)) for v in origin)
    sp = " ".join(repr(float(v)) for v in spacing)
    kw = {"header_type": header_type, "compress": compress}
    out = [_header("ImageData", header_type, compress),
           f'  <ImageData WholeExtent="{ext}" Origin="{org}" Spacing="{sp}" Direction="1 0 0 0 1 0 0 0 1">\n',
           f'    <Piece Extent="{ext}">\n']
    sc = f' Scalars="{scalars}"' if scalars else ""
    out.append(f"      <PointData{sc}>\n")
    for name, v in point_data.items():
        out.append(data_array_xml(name, np.asarray(v), "        ", **kw))
    out.append("      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n")
    with open(path, "w") as f:
        f.write("".join(out))
