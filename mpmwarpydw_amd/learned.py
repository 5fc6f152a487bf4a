"""Learned return map: a 10-64-64-3 GELU network that predicts the projection
step of the Drucker-Prager return map (drucker_prager_return_mapping,
mpmRoutines.py:360-413), trained on the GPU by libmpmx's own forward/backward
(bfloat16 matrix cores, FP32 accumulation) and Adam kernels on pairs that the
simulator generates, and used inside the constitutive update with the exact
map as judge and fallback (mpmx_compute_stress_learned).

No counterpart in the reference.  Device tensors are torch's (plumbing); every
computation is a libmpmx kernel.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import runtime as rt
from ._lib import lib
from .simulationRoutines import step_structs

NI, NH, NO = 10, 64, 3
# features that the exact targets do not depend on (4 nu - 1, log10 E - 8,
# damage) and the ranges they are redrawn from in training
INVARIANT_COLS = [7, 8, 9]
INVARIANT_LO = [-0.6, -2.0, 0.0]
INVARIANT_HI = [0.6, 2.0, 1.0]
LAYERS = ((NH, NI), (NH, NH), (NO, NH))  # (out, in)


def _p(t: torch.Tensor) -> C.c_void_p:
    return C.c_void_p(t.data_ptr())


def init_params(seed: int = 0) -> np.ndarray:
    """Glorot-uniform weights, zero biases, in the flat layout of the C ABI."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_out, fan_in in LAYERS:
        a = np.sqrt(6.0 / (fan_in + fan_out))
        parts.append(rng.uniform(-a, a, fan_out * fan_in))
        parts.append(np.zeros(fan_out))
    return np.concatenate(parts).astype(np.float32)


class LearnedReturnMap:
    """Parameters (FP32 master copy), Adam state and the kernels that use them."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        self.nparams = int(lib.raw("mpmx_mlp_nparams")())
        p = init_params(seed) if params is None else np.asarray(params, np.float32)
        if p.shape != (self.nparams,):
            raise ValueError(f"expected {self.nparams} parameters, got {p.shape}")
        dev = rt.device()
        self.params = torch.from_numpy(p.copy()).to(dev)
        self.m = torch.zeros(self.nparams, dtype=torch.float32, device=dev)
        self.v = torch.zeros(self.nparams, dtype=torch.float32, device=dev)
        self.grad = torch.zeros(self.nparams, dtype=torch.float32, device=dev)
        self.loss = torch.zeros(1, dtype=torch.float64, device=dev)
        self.t = 0
        self._ws = rt.Workspace()

    # ---- network ---------------------------------------------------------
    def forward(self, X: torch.Tensor) -> torch.Tensor:
        X = X.contiguous().float()
        Y = torch.empty((X.shape[0], NO), dtype=torch.float32, device=X.device)
        lib("mpmx_mlp_forward", X.shape[0], _p(self.params), _p(X), _p(Y), rt.stream_handle())
        return Y

    def loss_grad(self, X: torch.Tensor, T: torch.Tensor) -> tuple[torch.Tensor, torch.Tensor]:
        """Mean squared error and its gradient (device tensors, no synchronisation)."""
        X = X.contiguous().float()
        T = T.contiguous().float()
        B = X.shape[0]
        ws, nb = self._ws.get(lib.raw("mpmx_mlp_workspace_bytes")(B))
        lib("mpmx_mlp_loss_grad", B, _p(self.params), _p(X), _p(T), _p(self.grad), _p(self.loss), ws, nb,
            rt.stream_handle())
        return self.loss, self.grad

    def adam(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.t += 1
        lib("mpmx_mlp_adam_step", self.nparams, _p(self.params), _p(self.grad), _p(self.m), _p(self.v), float(lr),
            float(beta1), float(beta2), float(eps), self.t, rt.stream_handle())

    def train(self, X: torch.Tensor, T: torch.Tensor, steps: int, batch: int = 4096, lr: float = 1e-3,
              lr_final: float | None = None, seed: int = 0, log_every: int = 0,
              redraw_invariant: bool = True) -> list[float]:
        """Adam on random minibatches of (X, T) (device tensors); returns the
        loss of every ``log_every``-th step (one synchronisation each).

        The exact targets (s, alpha) are functions of the overshoot, yield
        ratio, kap and friction features alone: Poisson's ratio, E and damage
        enter them only through those normalised features.  With
        ``redraw_invariant`` those three columns are redrawn in every minibatch
        (4 nu - 1 in [-0.6, 0.6], log10 E - 8 in [-2, 2], damage in [0, 1]), so
        that the network learns that they do not matter instead of meeting
        combinations the training runs never produced."""
        g = torch.Generator(device=X.device)
        g.manual_seed(seed)
        n = X.shape[0]
        hist = []
        lo = torch.tensor(INVARIANT_LO, device=X.device)
        span = torch.tensor(INVARIANT_HI, device=X.device) - lo
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=g)
            xb = X[idx]
            if redraw_invariant:
                xb[:, INVARIANT_COLS] = lo + span * torch.rand((xb.shape[0], len(INVARIANT_COLS)), device=X.device,
                                                              generator=g)
            self.loss_grad(xb, T[idx])
            frac = k / max(1, steps - 1)
            cur = lr if lr_final is None else lr * (lr_final / lr) ** frac
            self.adam(cur)
            if log_every and (k % log_every == 0 or k == steps - 1):
                hist.append(float(self.loss.item()))
        return hist

    # ---- simulator side ---------------------------------------------------
    @staticmethod
    def harvest(sim, mpm, xpbd, plastic_only: bool = True):
        """(X, T, e_trial) of the active MPM particles' current trial states
        (after G2P, before the constitutive update), plastic ones by default."""
        n = sim.nPoints
        dev = rt.device()
        X = torch.empty((n, NI), dtype=torch.float32, device=dev)
        T = torch.empty((n, NO), dtype=torch.float32, device=dev)
        E = torch.empty((n, 3), dtype=torch.float32, device=dev)
        flag = torch.empty(n, dtype=torch.int32, device=dev)
        P, _, _ = step_structs(sim, mpm, xpbd)
        lib("mpmx_lrm_harvest", C.byref(P), _p(X), _p(T), _p(E), _p(flag), rt.stream_handle())
        sel = flag == 1 if plastic_only else flag >= 0
        return X[sel], T[sel], E[sel]

    def stress_update(self, sim, mpm, xpbd, tol: float = 1e-4, use_net: bool = True) -> torch.Tensor:
        """compute_stress_from_F_trial with the learned projection judged by the
        exact map; returns the device counters {plastic, accepted, updated}."""
        P, _, _ = step_structs(sim, mpm, xpbd)
        counters = torch.zeros(3, dtype=torch.int64, device=rt.device())
        ws, nb = self._ws.get(lib.raw("mpmx_lrm_workspace_bytes")(sim.nPoints))
        lib("mpmx_compute_stress_learned", C.byref(P), float(mpm.eff), _p(self.params) if use_net else None,
            float(tol), _p(counters), ws, nb, rt.stream_handle())
        return counters

    # ---- persistence ------------------------------------------------------
    def save(self, path: str) -> None:
        np.savez(path, params=self.params.cpu().numpy(), t=self.t)

    @classmethod
    def load(cls, path: str) -> "LearnedReturnMap":
        with np.load(path, allow_pickle=False) as f:
            obj = cls(params=f["params"])
            obj.t = int(f["t"])
        return obj


def harvest_run(sim_obj, substeps: int, start: int = 1, every: int = 1, max_pairs: int | None = None,
                sample: int | None = None, seed: int = 0):
    """Advance a driver.Simulation ``substeps`` substeps (the reference's loop
    order, XPBD steps included) and collect the plastic trial states the return
    map meets after each ``every``-th substep (at most ``sample`` of them per
    substep, drawn at random): (X, T) device tensors."""
    xs, ts = [], []
    total = 0
    g = None
    if sample is not None:
        g = torch.Generator(device=rt.device())
        g.manual_seed(seed)
    for k in range(substeps):
        counter = start + k
        sim_obj.substep(counter)
        if k % every:
            continue
        X, T, _ = LearnedReturnMap.harvest(sim_obj.sim, sim_obj.mpm, sim_obj.xpbd)
        if sample is not None and X.shape[0] > sample:
            keep = torch.randperm(X.shape[0], device=X.device, generator=g)[:sample]
            X, T = X[keep], T[keep]
        if X.shape[0]:
            xs.append(X)
            ts.append(T)
            total += X.shape[0]
        if max_pairs is not None and total >= max_pairs:
            break
    dev = rt.device()
    if not xs:
        return torch.zeros((0, NI), device=dev), torch.zeros((0, NO), device=dev)
    return torch.cat(xs), torch.cat(ts)


def draw_block_material(rng: np.random.Generator) -> tuple[float, float, float]:
    """(E, cohesion, friction angle in degrees) of a block-collapse run:
    E log-uniform in [1e7, 1e9] Pa, cohesion log-uniform in [1e4, 1e6] Pa,
    friction angle U(20, 40) degrees (the batched-ensemble ranges)."""
    E = float(np.exp(rng.uniform(np.log(1e7), np.log(1e9))))
    c = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
    phi = float(rng.uniform(20.0, 40.0))
    return E, c, phi


def harvest_blocks(target_pairs: int, max_runs: int = 64, substeps: int = 500, side_particles: int = 32768,
                   seed: int = 0, per_run: int | None = None, every: int = 1, log=None):
    """Training pairs from block-collapse runs (a cube of ``side_particles``
    particles on the 0.25 m lattice, dt 1e-4 s, ``substeps`` substeps with the
    XPBD steps at their ratio), one material per run drawn by
    draw_block_material from default_rng(seed), the plastic trial states of
    every ``every``-th substep kept, at most ``per_run`` from one run (spread
    evenly over its substeps, drawn at random within each): runs are added
    until ``target_pairs`` pairs are collected or ``max_runs`` ran.
    Returns (X, T, stats) with X, T device tensors."""
    from . import driver, getArgs
    from .ensemble import block_scene
    rng = np.random.default_rng(seed)
    xs, ts, runs = [], [], []
    total = 0
    while total < target_pairs and len(runs) < max_runs:
        E, c, phi = draw_block_material(rng)
        dom, ov = block_scene(side_particles, E, c, phi, seed=seed * 1000 + len(runs))
        S = driver.setup(getArgs.get_args([], ov), dom)
        cap = target_pairs - total if per_run is None else min(per_run, target_pairs - total)
        sample = None if per_run is None else max(1, per_run // max(1, substeps // every))
        X, T = harvest_run(S, substeps=substeps, start=1, every=every, max_pairs=cap, sample=sample,
                           seed=seed * 1000 + len(runs))
        X, T = X[:cap], T[:cap]
        runs.append({"E": E, "cohesion": c, "friction_deg": phi, "pairs": int(X.shape[0])})
        if log:
            log(f"run {len(runs)}: E {E:.3g} cohesion {c:.3g} friction {phi:.1f} deg -> {X.shape[0]} pairs")
        if X.shape[0]:
            xs.append(X)
            ts.append(T)
            total += X.shape[0]
        del S
    dev = rt.device()
    X = torch.cat(xs) if xs else torch.zeros((0, NI), device=dev)
    T = torch.cat(ts) if ts else torch.zeros((0, NO), device=dev)
    return X[:target_pairs], T[:target_pairs], {"runs": runs, "pairs": int(min(total, target_pairs))}
