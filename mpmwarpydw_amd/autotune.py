"""Autotuner of the fused substep's launch shape and particle order.

Knobs (none changes results, only speed):
  xcd_tiles   deal tiles to the chip's 8 XCDs in contiguous ranges (1) or in
              block order (0)                      (mpmx_set_tuning)
  split_stress run the constitutive update as its own particle kernel before
              a P2G-only tile kernel (1) or inside the tile kernel (0)
                                                   (mpmx_set_tuning)
  g2p_tiles   G2P per tile with the grid velocities in LDS, one block per
              tile (1) or two blocks sharing a tile's particles (2), or one
              thread per particle reading the grid through the caches (0)
                                                   (mpmx_set_tuning)
  reorder     keep the particle arrays in tile order in HBM (the sort-bin
              layout of the tile kernels) or in input order  (Simulation.reorder,
              Ensemble.reorder)
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


def set_xcd_tiles(on: bool) -> None:
    lib("mpmx_set_tuning", MPMX_TUNE_XCD_TILES, 1 if on else 0)


def set_split_stress(on: bool) -> None:
    lib("mpmx_set_tuning", MPMX_TUNE_SPLIT_STRESS, 1 if on else 0)


def set_g2p_tiles(mode: int) -> None:
    """0: particle-parallel G2P; 1, 2: tile G2P with 1 or 2 blocks per tile
    (True and False stand for 1 and 0)."""
    mode = int(mode)
    if mode not in (0, 1, 2):
        raise ValueError(f"g2p_tiles takes 0, 1 or 2, not {mode}")
    lib("mpmx_set_tuning", MPMX_TUNE_G2P_TILES, mode)


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


def tune(state, step, steps: int = 10, warmup: int = 3, trials: int = 2) -> dict:
    """Tune on ``state`` (a driver.Simulation or an ensemble.Ensemble with
    a ``reorder()`` method) advanced by ``step()``.  Leaves the state reordered
    and the best knob settings applied; returns the log."""
    t0 = time.perf_counter()
    log = []
    for reorder in (False, True):
        if reorder:
            state.reorder()
        for split in (1, 0):
            set_split_stress(bool(split))
            for g2p in (2, 1, 0):
                set_g2p_tiles(g2p)
                for xcd in (1, 0):
                    set_xcd_tiles(bool(xcd))
                    ms = _time(step, steps, warmup, trials)
                    log.append({"reorder": reorder, "split_stress": split, "g2p_tiles": g2p, "xcd_tiles": xcd,
                                "ms_per_substep": round(ms, 4)})
    best = min(log, key=lambda r: r["ms_per_substep"])  # min keeps the first of equal entries
    set_xcd_tiles(bool(best["xcd_tiles"]))
    set_split_stress(bool(best["split_stress"]))
    set_g2p_tiles(best["g2p_tiles"])
    # the state is reordered at this point; input order is kept only when it was faster
    return {"best": best, "candidates": log, "kept_reordered": True, "tuning_s": round(time.perf_counter() - t0, 2)}
