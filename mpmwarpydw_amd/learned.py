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
              lr_final: float | None = None, seed: int = 0, log_every: int = 0) -> list[float]:
        """Adam on random minibatches of (X, T) (device tensors); returns the
        loss of every ``log_every``-th step (one synchronisation each)."""
        g = torch.Generator(device=X.device)
        g.manual_seed(seed)
        n = X.shape[0]
        hist = []
        for k in range(steps):
            idx = torch.randint(0, n, (min(batch, n),), device=X.device, generator=g)
            self.loss_grad(X[idx], T[idx])
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
        counters = torch.zeros(3, dtype=torch.int32, device=rt.device())
        lib("mpmx_compute_stress_learned", C.byref(P), float(mpm.eff), _p(self.params) if use_net else None,
            float(tol), _p(counters), rt.stream_handle())
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


def harvest_run(sim_obj, substeps: int, start: int = 1, every: int = 1, max_pairs: int | None = None):
    """Advance a driver.Simulation ``substeps`` substeps (the reference's loop
    order, XPBD steps included) and collect the plastic trial states the return
    map meets after each ``every``-th substep: (X, T) device tensors."""
    xs, ts = [], []
    total = 0
    for k in range(substeps):
        counter = start + k
        sim_obj.substep(counter)
        if k % every:
            continue
        X, T, _ = LearnedReturnMap.harvest(sim_obj.sim, sim_obj.mpm, sim_obj.xpbd)
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
