"""Device arrays, launches and the hash grid that the hot path runs on.

This replaces the slice of NVIDIA Warp that MPMWarpYDW's hot path uses
(``wp.array``/``wp.zeros``/``.numpy()``/``.zero_()``/``.fill_()``/``.assign``,
``wp.copy``, ``wp.launch``, ``wp.HashGrid``), with the same names so that the
reference's launch sites read the same:

    runtime.launch(kernel=mpmRoutines.g2p, dim=sim.nPoints, inputs=[...], device=device)

Storage is HBM owned by PyTorch (plumbing only); the compute is libmpmx's HIP
kernels.  Layout is structure-of-arrays: a length-n ``vec3`` array is stored as
three float planes ``[3, n]``, a ``mat33`` array as nine planes ``[9, n]``, grid
arrays of shape (gx, gy, gz) flatten to G = gx*gy*gz nodes (z fastest).  The
host views (``numpy()``, ``assign``) use the reference's AoS shapes: ``(n, 3)``,
``(n, 3, 3)``, ``(gx, gy, gz)``, ``(gx, gy, gz, 3)``.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Sequence

import numpy as np
import torch

from ._lib import MpmxError, lib

float32 = "float32"
float64 = "float64"
int32 = "int32"
vec3 = "vec3"
mat33 = "mat33"
_NCOMP = {float32: 1, float64: 1, int32: 1, vec3: 3, mat33: 9}
_TORCH = {float32: torch.float32, float64: torch.float64, int32: torch.int32, vec3: torch.float32,
          mat33: torch.float32}
_NUMPY = {float32: np.float32, float64: np.float64, int32: np.int32, vec3: np.float32, mat33: np.float32}
_DEVICE = "cuda:0"


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise MpmxError("no GPU: the MPM hot path runs only on the device (there is no CPU fallback)")
    return torch.device(_DEVICE)


def stream_handle() -> C.c_void_p:
    """The current torch stream as a hipStream_t for libmpmx."""
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Array:
    """A device array with the reference's shape and an SoA storage tensor."""

    def __init__(self, shape, dtype: str, storage: torch.Tensor | None = None):
        if isinstance(shape, (int, np.integer)):
            shape = (int(shape),)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = dtype
        self.size = int(np.prod(self.shape)) if self.shape else 1
        nc = _NCOMP[dtype]
        tdt = _TORCH[dtype]
        if storage is None:
            storage = torch.zeros((nc, self.size) if nc > 1 else (self.size,), dtype=tdt, device=device())
        self.t = storage
        # bumped by every write issued from the host (assign from numpy, fill_):
        # the fused substep rebuilds its particle bins when x or the labels change
        self.version = 0
        # set by permute_: the items sit in HBM in a permuted order, and
        # numpy()/assign() translate from and to the input order, so the
        # permutation stays invisible to the host (``_order[i]`` is the input
        # index of item i, ``_words`` the 32-bit words per item)
        self._order = None
        self._words = 1

    # --- pointers ---------------------------------------------------------
    @property
    def ptr(self) -> C.c_void_p:
        return C.c_void_p(self.t.data_ptr())

    def __len__(self):
        return self.shape[0]

    # --- wp.array API used by the reference -------------------------------
    def _to_input_order(self, h: np.ndarray) -> np.ndarray:
        """host copy of the storage (SoA planes) -> same layout in input order"""
        if self._order is None:
            return h
        order = self._order.t.detach().cpu().numpy()
        w = self._words
        if h.ndim == 2:  # planes
            out = np.empty_like(h)
            out[:, order] = h
            return out
        out = np.empty_like(h).reshape(-1, w)
        out[order] = h.reshape(-1, w)
        return out.reshape(h.shape)

    def _from_input_order(self, h: np.ndarray) -> np.ndarray:
        if self._order is None:
            return h
        order = self._order.t.detach().cpu().numpy()
        w = self._words
        if h.ndim == 2:
            return np.ascontiguousarray(h[:, order])
        return np.ascontiguousarray(h.reshape(-1, w)[order].reshape(h.shape))

    def numpy(self) -> np.ndarray:
        """Synchronising download in the reference's AoS shape (input order)."""
        h = self._to_input_order(self.t.detach().cpu().numpy())
        nc = _NCOMP[self.dtype]
        if nc == 1:
            return h.reshape(self.shape)
        h = np.ascontiguousarray(h.T)
        if self.dtype == vec3:
            return h.reshape(self.shape + (3,))
        return h.reshape(self.shape + (3, 3))

    def assign(self, src) -> "Array":
        """Upload from numpy (AoS) or copy from another Array (wp.array.assign)."""
        if isinstance(src, Array):
            copy(self, src)
            return self
        a = np.asarray(src)
        nc = _NCOMP[self.dtype]
        npdt = _NUMPY[self.dtype]
        a = np.ascontiguousarray(a, dtype=npdt).reshape(self.size, nc) if nc > 1 else \
            np.ascontiguousarray(a, dtype=npdt).reshape(self.size)
        host = self._from_input_order(np.ascontiguousarray(a.T) if nc > 1 else a)
        self.t.copy_(torch.from_numpy(host).to(self.t.device, non_blocking=False))
        self.version += 1
        return self

    def zero_(self) -> "Array":
        return self.fill_(0)

    def fill_(self, value) -> "Array":
        self.version += 1
        nc = _NCOMP[self.dtype]
        if nc > 1 and np.ndim(value) > 0:
            vals = np.asarray(value, dtype=np.float32).reshape(-1)
            for c in range(nc):
                lib("mpmx_fill_f32", C.c_void_p(self.t[c].data_ptr()), self.size, float(vals[c]), stream_handle())
        elif self.dtype == int32:
            lib("mpmx_fill_i32", self.ptr, self.size, int(value), stream_handle())
        elif self.dtype == float64:
            self.t.fill_(float(value))
        else:
            lib("mpmx_fill_f32", self.ptr, self.size * nc, float(value), stream_handle())
        return self

    def permute_(self, perm: "Array", item_words: int = 1, order: "Array | None" = None) -> "Array":
        """Reorder the items by a device permutation (new[i] = old[perm[i]]):
        SoA components as planes, ``item_words`` 32-bit words per item.  With
        ``order`` (the input index of every item after the permutation) the
        host views keep showing the input order."""
        if self.dtype == float64:
            raise MpmxError("permute_: float64 arrays are not per-particle state")
        nc = _NCOMP[self.dtype]
        n = self.size // item_words
        if perm.size != n:
            raise MpmxError(f"permute_: permutation of {perm.size} items for an array of {n}")
        new = torch.empty_like(self.t)
        lib("mpmx_gather_words", C.c_void_p(new.data_ptr()), self.ptr, perm.ptr, n, nc, int(item_words),
            stream_handle())
        self.t = new
        self.version += 1
        self._order = order
        self._words = int(item_words)
        return self

    def __repr__(self):
        return f"Array(shape={self.shape}, dtype={self.dtype})"


def zeros(shape, dtype: str = float32, device=None) -> Array:
    return Array(shape, dtype)


def empty(shape, dtype: str = float32, device=None) -> Array:
    a = Array(shape, dtype, storage=None)
    return a


def array(data, dtype: str = float32, device=None) -> Array:
    data = np.asarray(data)
    nc = _NCOMP[dtype]
    shape = data.shape if nc == 1 else data.shape[: data.ndim - (1 if dtype == vec3 else 2)]
    a = Array(shape, dtype)
    a.assign(data)
    return a


def copy(dst: Array, src: Array) -> None:
    """wp.copy(dst, src) for equally sized float or int arrays."""
    if dst.size != src.size or dst.dtype != src.dtype:
        raise MpmxError(f"copy: shape/dtype mismatch {dst} <- {src}")
    n = dst.t.numel()
    if dst.dtype in (int32, float64):
        dst.t.copy_(src.t)
    else:
        lib("mpmx_copy_f32", dst.ptr, src.ptr, n, stream_handle())


def synchronize() -> None:
    lib("mpmx_stream_sync", stream_handle())


class Workspace:
    """Device scratch owned by a kernel wrapper, grown on demand outside hot loops."""

    def __init__(self):
        self.t = None

    def get(self, nbytes: int) -> tuple[C.c_void_p, int]:
        nbytes = max(int(nbytes), 256)
        if self.t is None or self.t.numel() < nbytes:
            self.t = torch.empty(nbytes, dtype=torch.uint8, device=device())
        return C.c_void_p(self.t.data_ptr()), self.t.numel()


class Kernel:
    """A launchable kernel with the reference's argument list.

    ``fn(dim, *args)`` receives the reference's inputs+outputs in order; it
    converts them to the C ABI call.
    """

    def __init__(self, name: str, fn, reference: str):
        self.name = name
        self.fn = fn
        self.reference = reference
        self.__doc__ = f"{name}: replaces the Warp kernel at {reference}"

    def __call__(self, dim, *args):
        return self.fn(dim, *args)

    def __repr__(self):
        return f"<mpmx kernel {self.name} ({self.reference})>"


def launch(kernel: Kernel, dim, inputs: Sequence = (), outputs: Iterable = (), device=None):
    """wp.launch(kernel, dim, inputs, outputs, device) for libmpmx kernels."""
    return kernel(dim, *list(inputs), *list(outputs))


class HashGrid:
    """wp.HashGrid(dim_x, dim_y, dim_z): a counting-sorted cell table on the device.

    After ``build(points, cell_width)``: ``point_ids[tid]`` lists the points in
    cell order (ids ascending inside a cell) and cell c holds
    ``point_ids[cell_start[c]:cell_start[c+1]]``.
    """

    def __init__(self, dim_x: int, dim_y: int, dim_z: int, device=None):
        self.dims = (int(dim_x), int(dim_y), int(dim_z))
        ncell = self.dims[0] * self.dims[1] * self.dims[2]
        self.cell_start = zeros(ncell + 1, int32)
        self.point_ids = None
        self.point_pos = None
        self.cell_width = None
        self.n = 0
        self._ws = Workspace()

    @property
    def id(self):
        return self

    def build(self, points: Array, cell_width: float, labels: Array | None = None, label: int = 2,
              order: Array | None = None) -> None:
        """Bin all points, or (with ``labels``) only those whose label equals ``label``.
        With ``order`` the ids inside a cell ascend in order[id] instead of id."""
        n = points.shape[0]
        if self.point_ids is None or self.point_ids.shape[0] != n:
            self.point_ids = zeros(n, int32)
        self.n = n
        self.cell_width = float(cell_width)
        dx, dy, dz = self.dims
        ws, nb = self._ws.get(lib.raw("mpmx_hashgrid_workspace_bytes")(n, dx, dy, dz))
        if order is not None:
            lib("mpmx_hashgrid_build_ordered", n, points.ptr, self.cell_width, dx, dy, dz,
                None if labels is None else labels.ptr, int(label), order.ptr, self.point_ids.ptr,
                self.cell_start.ptr, ws, nb, stream_handle())
        elif labels is None:
            lib("mpmx_hashgrid_build", n, points.ptr, self.cell_width, dx, dy, dz, self.point_ids.ptr,
                self.cell_start.ptr, ws, nb, stream_handle())
        else:
            lib("mpmx_hashgrid_build_label", n, points.ptr, self.cell_width, dx, dy, dz, labels.ptr, int(label),
                self.point_ids.ptr, self.cell_start.ptr, ws, nb, stream_handle())

    def cell_starts_ends(self):
        cs = self.cell_start.numpy()
        return cs[:-1], cs[1:]

    def near_cells(self) -> Array:
        """The grid's neighbourhood bitmap (mpmx_hashgrid_near_cells): bit c set
        when cell c or one of its 26 neighbours holds a point.  Kept on the grid
        (64-bit words stored as pairs of int32) and returned."""
        words = int(lib.raw("mpmx_hashgrid_near_words")(*self.dims))
        if getattr(self, "near", None) is None or self.near.shape[0] != 2 * words:
            self.near = zeros(2 * words, int32)
            self._occ = zeros(2 * words, int32)
        lib("mpmx_hashgrid_near_cells", self.cell_start.ptr, *self.dims, self._occ.ptr, self.near.ptr,
            stream_handle())
        return self.near

    def build_pair(self, other: "HashGrid", points: Array, cell_width: float, labels: Array, label: int = 2,
                   order: Array | None = None, lazy: bool = False) -> None:
        """``self.build(points, cell_width, order=order)`` and
        ``other.build(points, cell_width, labels=labels, label=label, order=order)``
        in one set of launches (mpmx_hashgrid_build_pair); each grid keeps its own dims.
        Each grid's ``point_pos`` then holds every point's position in its
        ``point_ids`` (-1 where a point is not in ``other``), when ``other``
        holds a point (else it is left as it was).  With ``lazy``, while no point
        has the label, ``self`` is left as it was too (the fused XPBD step's
        contact solve reads neither grid then)."""
        n = points.shape[0]
        for g in (self, other):
            if g.point_ids is None or g.point_ids.shape[0] != n:
                g.point_ids = zeros(n, int32)
            if getattr(g, "point_pos", None) is None or g.point_pos.shape[0] != n:
                g.point_pos = zeros(n, int32)
            g.n = n
            g.cell_width = float(cell_width)
        dx, dy, dz = self.dims
        lx, ly, lz = other.dims
        ws, nb = self._ws.get(lib.raw("mpmx_hashgrid_pair_workspace_bytes")(n, dx, dy, dz, lx, ly, lz))
        lib("mpmx_hashgrid_build_pair", n, points.ptr, self.cell_width, dx, dy, dz, labels.ptr, int(label), lx, ly, lz,
            None if order is None else order.ptr, self.point_ids.ptr, self.cell_start.ptr, other.point_ids.ptr,
            other.cell_start.ptr, self.point_pos.ptr, other.point_pos.ptr, 1 if lazy else 0, ws, nb, stream_handle())
