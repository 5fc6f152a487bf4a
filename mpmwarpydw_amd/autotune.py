"""Autotuner of the fused substep's launch shape and particle order.

Knobs (none changes results, only speed):
  xcd_tiles   deal tiles to the chip's 8 XCDs in contiguous ranges (1), in
              runs of 16 tiles in turn (2) or in block order (0)
                                                   (mpmx_set_tuning)
  split_stress run the constitutive update as its own particle kernel before
              a P2G-only tile kernel (1) or inside the tile kernel (0)
                                                   (mpmx_set_tuning)
  g2p_tiles   G2P per tile with the grid velocities in LDS, one block per
              tile (1) or two blocks sharing a tile's particles (2), or one
              thread per particle reading the grid through the caches (0)
                                                   (mpmx_set_tuning)
  p2g_chunk   particles staged in LDS per pass of the P2G tile kernel that
              follows a split constitutive update: 512 (0), 384 (1) or 256
              (2), three instantiations of the kernel template with their own
              occupancy                            (mpmx_set_tuning)
  reorder     keep the particle arrays in tile order in HBM (the sort-bin
              layout of the tile kernels) or in input order  (Simulation.reorder,
              Ensemble.reorder, and restore_input_order to go back)
Every candidate is timed on the caller's own state with HIP events over a few
substeps, best of ``trials``; candidates are visited in a fixed order and ties
keep the earlier one, so a tuning run is reproducible given the timings, and
the log records every measurement.
"""
from __future__ import annotations

import time

import torch

from ._lib import lib

MPMX_TUNE_XCD_TILES = 0
MPMX_TUNE_SPLIT_STRESS = 1
MPMX_TUNE_G2P_TILES = 2
MPMX_TUNE_P2G_CHUNK = 3
P2G_CHUNKS = (512, 384, 256)  # knob value -> particles per LDS pass


def set_xcd_tiles(mode: int) -> None:
    """0: block order; 1: 8 contiguous ranges; 2: runs of 16 tiles in turn
    (True and False stand for 1 and 0)."""
    lib("mpmx_set_tuning", MPMX_TUNE_XCD_TILES, int(mode))


def set_split_stress(on: bool) -> None:
    lib("mpmx_set_tuning", MPMX_TUNE_SPLIT_STRESS, 1 if on else 0)


def set_g2p_tiles(mode: int) -> None:
    """0: particle-parallel G2P; 1, 2: tile G2P with 1 or 2 blocks per tile
    (True and False stand for 1 and 0)."""
    mode = int(mode)
    if mode not in (0, 1, 2):
        raise ValueError(f"g2p_tiles takes 0, 1 or 2, not {mode}")
    lib("mpmx_set_tuning", MPMX_TUNE_G2P_TILES, mode)


def set_p2g_chunk(mode: int) -> None:
    """0, 1, 2: 512, 384 or 256 particles per LDS pass of the split P2G kernel."""
    mode = int(mode)
    if mode not in (0, 1, 2):
        raise ValueError(f"p2g_chunk takes 0, 1 or 2, not {mode}")
    lib("mpmx_set_tuning", MPMX_TUNE_P2G_CHUNK, mode)


def current() -> dict:
    """The knob settings in force."""
    g = lambda k: int(lib.raw("mpmx_get_tuning")(k))  # noqa: E731
    return {"xcd_tiles": g(MPMX_TUNE_XCD_TILES), "split_stress": g(MPMX_TUNE_SPLIT_STRESS),
            "g2p_tiles": g(MPMX_TUNE_G2P_TILES), "p2g_chunk": P2G_CHUNKS[g(MPMX_TUNE_P2G_CHUNK)]}


def apply(cfg: dict) -> None:
    """Set the knobs of a candidate (as tune() logs them)."""
    set_xcd_tiles(int(cfg["xcd_tiles"]))
    set_split_stress(bool(cfg["split_stress"]))
    set_g2p_tiles(cfg["g2p_tiles"])
    set_p2g_chunk(P2G_CHUNKS.index(cfg.get("p2g_chunk", 512)))


def _time(step, steps: int, warmup: int, trials: int) -> float:
    for _ in range(warmup):
        step()
    best = float("inf")
    for _ in range(trials):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        a.record()
        for _ in range(steps):
            step()
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b) / steps)
    return best


def _state_arrays(state):
    """Every device array of a driver.Simulation or an ensemble.Ensemble."""
    from . import runtime as rt
    if hasattr(state, "_state_arrays"):
        return state._state_arrays() + [state.grid_m, state.grid_v_in, state.grid_v_out]
    out, seen = [], set()
    for obj in (state.sim, state.mpm, state.xpbd):
        for a in vars(obj).values():
            if isinstance(a, rt.Array) and id(a) not in seen:
                seen.add(id(a))
                out.append(a)
    return out


def _restore(state, arrays, snap):
    for a, t in zip(arrays, snap):
        a.t.copy_(t)
    if hasattr(state, "_binned"):
        state._binned = False
    else:
        state.mpm._bins = None


def tune(state, step, steps: int = 10, warmup: int = 3, trials: int = 2, orders=(False, True),
         split_forced: bool = False) -> dict:
    """Tune on ``state`` (a driver.Simulation or an ensemble.Ensemble, with
    ``reorder()`` and ``restore_input_order()``) advanced by ``step()``, which
    must start from input order.  Every knob combination is timed from the
    same starting state (a device copy, restored before each candidate), first
    in input order, then in tile order; the best one is applied and the state
    is left as it started, in the order of the best candidate.
    ``split_forced``: the step always runs the constitutive update as its own
    kernel (the learned update does), so only the split candidates are timed.
    Returns the log."""
    t0 = time.perf_counter()
    log = []
    snaps = {}
    for reorder in orders:
        if reorder:
            if snaps:  # back to the starting state before it is permuted
                _restore(state, *snaps[False])
            state.reorder()
        arrays = _state_arrays(state)
        snap = [a.t.clone() for a in arrays]
        snaps[reorder] = (arrays, snap)
        for split in ((1,) if split_forced else (1, 0)):
            set_split_stress(bool(split))
            for chunk in (range(len(P2G_CHUNKS)) if split else (0,)):
                set_p2g_chunk(chunk)
                for g2p in (2, 1, 0):
                    set_g2p_tiles(g2p)
                    for xcd in (1, 2, 0):
                        set_xcd_tiles(xcd)
                        _restore(state, arrays, snap)
                        ms = _time(step, steps, warmup, trials)
                        log.append({"reorder": reorder, "split_stress": split,
                                    "p2g_chunk": P2G_CHUNKS[chunk] if split else None, "g2p_tiles": g2p,
                                    "xcd_tiles": xcd, "ms_per_substep": round(ms, 4)})
    best = min(log, key=lambda r: r["ms_per_substep"])  # min keeps the first of equal entries
    apply({**best, "p2g_chunk": best["p2g_chunk"] or 512})
    if orders[-1] and not best["reorder"]:
        state.restore_input_order()
        arrays = _state_arrays(state)
        _restore(state, arrays, snaps[False][1])
    else:
        _restore(state, *snaps[orders[-1]])
    return {"best": best, "candidates": log, "kept_reordered": bool(best["reorder"]),
            "tuning_s": round(time.perf_counter() - t0, 2)}
