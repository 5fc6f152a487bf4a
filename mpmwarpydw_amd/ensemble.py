"""Ensembles: many independent scenes advanced by one launch of every kernel of
the fused MPM substep (mpmx_ensemble_*), captured in a HIP graph by the
caller.  No counterpart in the reference, which runs one scene per process.

Each scene is set up alone by driver.setup (its grid, walls, material and
XPBD box), then the per-particle arrays of all scenes are concatenated into
one state (scene after scene), grid arrays likewise, and a scene table tells
the kernels where every scene's tiles and nodes start.  Scenes must share dx,
dt, gravity and the boundary settings; materials are per particle anyway.
A scene's results equal running it alone through the fused substep, bit for
bit (same kernels, same summation order: particles inside a cell are summed in
ascending input index, and every tile and node belongs to one scene).

The XPBD contact step is per scene (each has its own contact box); the
ensemble advances the MPM substep, which is where the time goes.
"""
from __future__ import annotations

import copy
import ctypes as C

import numpy as np
import torch

from . import runtime as rt
from ._lib import Scene, lib
from .simulationRoutines import step_structs

SHARED = ("dt", "dx", "invdx", "gravity")
SHARED_MPM = ("rpic_damping", "eff", "grid_v_damping_scale", "boundFriction", "boundaryPadding", "boundaryCondition",
              "boundRestitution", "update_cov", "volumetric_locking_correction")


class Ensemble:
    def __init__(self, sims: list):
        if not sims:
            raise ValueError("an ensemble needs at least one scene")
        s0 = sims[0]
        for S in sims[1:]:
            for k in SHARED:
                if getattr(S.sim, k) != getattr(s0.sim, k):
                    raise ValueError(f"scenes differ in sim.{k}")
            for k in SHARED_MPM:
                if getattr(S.mpm, k) != getattr(s0.mpm, k):
                    raise ValueError(f"scenes differ in mpm.{k}")
        self.scenes_sims = sims
        self.counts = [S.sim.nPoints for S in sims]
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        n = int(self.offsets[-1])
        self.n = n
        # combined state objects: scene 0's scalars, concatenated per-particle arrays
        self.sim = copy.copy(s0.sim)
        self.mpm = copy.copy(s0.mpm)
        self.xpbd = copy.copy(s0.xpbd)
        for name, obj in (("sim", self.sim), ("mpm", self.mpm), ("xpbd", self.xpbd)):
            for attr, a0 in list(vars(getattr(s0, name)).items()):
                if not isinstance(a0, rt.Array) or len(a0.shape) != 1 or a0.dtype == rt.float64:
                    continue
                n0 = s0.sim.nPoints
                if a0.shape[0] == n0:
                    per = 1
                elif a0.shape[0] == 6 * n0:
                    per = 6
                else:
                    continue
                parts = [getattr(getattr(S, name), attr).t for S in sims]
                cat = torch.cat(parts, dim=-1)
                setattr(obj, attr, rt.Array((per * n,), a0.dtype, storage=cat))
        # input index over the whole ensemble, and the scene of every particle
        ids = np.concatenate([np.arange(c, dtype=np.int32) + o for c, o in zip(self.counts, self.offsets[:-1])])
        self.sim.particle_id = rt.array(ids.astype(np.int32), rt.int32)
        self.scene = rt.array(np.repeat(np.arange(len(sims), dtype=np.int32), self.counts), rt.int32)
        self.sim.nPoints = n
        # grids scene after scene
        self.table = (Scene * len(sims))()
        self.node_offsets = [0]
        for k, S in enumerate(sims):
            gx, gy, gz = S.mpm.gridDims
            mb = S.sim.minBounds
            self.table[k] = Scene(gx=gx, gy=gy, gz=gz, minx=mb[0], miny=mb[1], minz=mb[2])
            self.node_offsets.append(self.node_offsets[-1] + gx * gy * gz)
        G = self.node_offsets[-1]
        self.grid_m = rt.zeros(G)
        self.grid_v_in = rt.zeros(G, rt.vec3)
        self.grid_v_out = rt.zeros(G, rt.vec3)
        self._ws = rt.Workspace()
        self._binned = False

    # ---- launches -----------------------------------------------------------
    def _structs(self):
        P, _, prm = step_structs(self.sim, self.mpm, self.xpbd)
        P.scene = self.scene.t.data_ptr()
        return P, prm

    def _workspace(self):
        need = lib.raw("mpmx_ensemble_workspace_bytes")(self.n, self.table, len(self.table))
        if need == 0:
            raise ValueError("ensemble: bad scene table")
        before = self._ws.t
        ws, nb = self._ws.get(need)
        if self._ws.t is not before:
            self._binned = False
        return ws, nb

    def bin(self) -> None:
        P, prm = self._structs()
        ws, nb = self._workspace()
        lib("mpmx_ensemble_bin", C.byref(P), self.table, len(self.table), C.byref(prm), ws, nb, rt.stream_handle())
        self._binned = True

    def substep(self, learned=None, tol: float = 1e-4, counters: torch.Tensor | None = None,
                use_net: bool = True) -> None:
        """One fused MPM substep of every scene (one launch per kernel).  With
        ``learned`` (a learned.LearnedReturnMap) the constitutive update uses
        its projection, judged by the exact map at ``tol`` (``use_net`` False:
        the same kernel with the exact projection for every state);
        ``counters`` (int64 device tensor of 3, optional) accumulates
        {plastic, accepted, updated}."""
        ws, nb = self._workspace()
        if not self._binned:
            self.bin()
        P, prm = self._structs()
        if learned is None:
            lib("mpmx_ensemble_substep", C.byref(P), self.table, len(self.table), self.grid_m.ptr,
                self.grid_v_in.ptr, self.grid_v_out.ptr, C.byref(prm), ws, nb, rt.stream_handle())
            return
        cp = None
        if counters is not None:
            if counters.dtype != torch.int64 or counters.numel() < 3 or not counters.is_cuda:
                raise ValueError("counters: an int64 device tensor of 3 elements")
            cp = C.c_void_p(counters.data_ptr())
        lib("mpmx_ensemble_substep_learned", C.byref(P), self.table, len(self.table), self.grid_m.ptr,
            self.grid_v_in.ptr, self.grid_v_out.ptr, C.byref(prm),
            C.c_void_p(learned.params.data_ptr()) if use_net else None, float(tol), cp, ws, nb, rt.stream_handle())

    def reorder(self) -> None:
        """Permute every per-particle array into tile order (as driver.Simulation.reorder)."""
        P, prm = self._structs()
        ws, nb = self._workspace()
        perm = rt.zeros(self.n, rt.int32)
        lib("mpmx_ensemble_reorder_permutation", C.byref(P), self.table, len(self.table), C.byref(prm), perm.ptr, ws,
            nb, rt.stream_handle())
        for a in self._state_arrays():
            words = 6 if a.shape[0] == 6 * self.n else 1
            a.permute_(perm, words, None if a is self.sim.particle_id else self.sim.particle_id)
        self._binned = False

    def _state_arrays(self):
        arrays, seen = [self.scene], {id(self.scene)}
        for obj in (self.sim, self.mpm, self.xpbd):
            for name, a in vars(obj).items():
                if isinstance(a, rt.Array) and not name.startswith("_") and len(a.shape) == 1 \
                        and a.dtype != rt.float64 and a.shape[0] in (self.n, 6 * self.n) and id(a) not in seen:
                    seen.add(id(a))
                    arrays.append(a)
        return arrays

    def restore_input_order(self) -> None:
        """Undo reorder(): every per-particle array back in input order."""
        back = rt.zeros(self.n, rt.int32)  # slot of every input index
        lib("mpmx_invert_permutation", self.sim.particle_id.ptr, back.ptr, self.n, rt.stream_handle())
        for a in self._state_arrays():
            a.permute_(back, 6 if a.shape[0] == 6 * self.n else 1)
        self._binned = False

    # ---- views ---------------------------------------------------------------
    def scene_array(self, k: int, name: str, obj: str = "sim") -> np.ndarray:
        """Scene k's part of a per-particle array (input order)."""
        a = getattr(getattr(self, obj), name).numpy()
        per = a.shape[0] // self.n
        lo, hi = int(self.offsets[k]), int(self.offsets[k + 1])
        if per == 6 and a.ndim == 1:
            return a[6 * lo:6 * hi]
        return a[lo:hi]

    def scene_grid(self, k: int, name: str) -> np.ndarray:
        a = getattr(self, name).numpy()
        return a[self.node_offsets[k]:self.node_offsets[k + 1]]


def block_scene(n_target: int, E: float, ys: float, phi_deg: float, seed: int, dt: float = 1e-4):
    """A block-collapse style scene of about n_target particles (cube lattice at
    0.25 m with +-0.05 m jitter) with one material: (domain, overrides)."""
    from .scenes import VOL, _common_overrides, _lattice, alpha_from_friction_angle
    side = max(4, int(round(n_target ** (1.0 / 3.0))))
    rng = np.random.default_rng(seed)
    x, _ = _lattice(side, side, side, rng)
    dom = {"x": x, "particle_volume": np.full(len(x), VOL)}
    ov = _common_overrides(dt, 1000, 10)
    ov.update({"boundary_padding_mask": "111101", "grid_padding": 4.0, "E": float(E), "ys": float(ys),
               "alpha": float(alpha_from_friction_angle(phi_deg))})
    return dom, ov


def c4_scene_list(count: int = 64, seed: int = 4, n_lo: float = 2e4, n_hi: float = 5e5):
    """The batched-ensemble workload: particle counts log-uniform in [n_lo, n_hi],
    friction angle U(20, 40) degrees, cohesion log-uniform in [1e4, 1e6] Pa,
    E log-uniform in [1e7, 1e9] Pa, all seeded."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = float(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi))))
        phi = float(rng.uniform(20.0, 40.0))
        c = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        E = float(np.exp(rng.uniform(np.log(1e7), np.log(1e9))))
        out.append(block_scene(int(n), E, c, phi, seed=1000 + k))
    return out
