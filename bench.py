"""Benchmark of the MPM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scene undercut_slab]

A step is one MPM substep of the plain method over the whole scene (stress +
return map, deterministic P2G with the grid update, G2P), with the XPBD
contact step every ``mpmStepsPerXpbdStep`` substeps as the reference's loop
runs it (runMPMYDW.py:677-682).  The default workload is the undercut slab:
960,000 particles, grid 104x54x31, dt = 5e-5 s, XPBD every 10 substeps.
Inputs are resident in HBM before the timed region; W untimed substeps, then
K timed substeps between two device synchronisations.

Prints one JSON line with the metric (MPM particle-substeps/s), per-phase
device times from HIP events recorded on the launch stream inside the timed
region, the roofline entry of the dominant kernel and a CPU baseline (the
FP32 restatement under oracle/, one core, on a bounded sample).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=400)
    p.add_argument("--warmup", type=int, default=40)
    p.add_argument("--scene", default="undercut_slab")
    p.add_argument("--scale", type=float, default=1.0)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    return p.parse_args()


def cpu_baseline(S, budget_s):
    """The FP32 restatement (oracle/, reference semantics, P2G sequential in
    particle order) timed on one host core over a bounded number of substeps of
    the same state."""
    from oracle import Oracle
    o = Oracle(32)
    sim, mpm = S.sim, S.mpm
    f = lambda a: np.ascontiguousarray(a.numpy().astype(np.float32))
    H = {k: f(getattr(sim, k)) for k in ("particle_x", "particle_v", "particle_mass", "particle_vol",
                                         "particle_density")}
    for k in ("particle_F", "particle_F_trial", "particle_stress", "particle_accumulated_strain", "particle_damage",
              "particle_C", "particle_cov", "mu", "lam", "ys", "alpha", "hardening", "softening", "eta_shear",
              "eta_bulk", "strainCriteria"):
        H[k] = f(getattr(mpm, k))
    xi = np.ascontiguousarray(S.xpbd.particle_x_initial_xpbd.numpy().astype(np.float32))
    vi = np.ascontiguousarray(S.xpbd.particle_v_initial_xpbd.numpy().astype(np.float32))
    act = np.ascontiguousarray(sim.activeLabel.numpy())
    lab = np.ascontiguousarray(sim.materialLabel.numpy())
    gm = np.zeros(mpm.gridDims, np.float32)
    gv = np.zeros(mpm.gridDims + (3,), np.float32)
    go = np.zeros(mpm.gridDims + (3,), np.float32)
    n_mpm = int(((act == 1) & (lab <= 1)).sum())
    steps, t0 = 0, time.perf_counter()
    while True:
        o.compute_stress_from_F_trial(act, lab, H["particle_x"], H["particle_v"], xi, vi, H["particle_F"],
                                      H["particle_F_trial"], H["mu"], H["lam"], H["ys"], H["hardening"],
                                      H["softening"], H["particle_density"], H["strainCriteria"], mpm.eff,
                                      H["eta_shear"], H["eta_bulk"], H["particle_C"], H["particle_stress"],
                                      H["particle_accumulated_strain"], H["particle_damage"], 1, H["alpha"])
        gm[:] = 0
        gv[:] = 0
        go[:] = 0
        o.p2g_apic_with_stress(act, lab, H["particle_stress"], H["particle_x"], H["particle_v"], H["particle_C"],
                               H["particle_vol"], H["particle_mass"], sim.dx, sim.invdx, sim.minBounds,
                               mpm.rpic_damping, gm, gv, sim.dt)
        o.grid_normalization_and_gravity(gm, gv, go, sim.gravity, sim.dt)
        o.collideBounds(go, *mpm.gridDims, mpm.boundFriction, sim.dt, abs(sim.gravity[2]), mpm.boundaryPadding)
        o.g2p(sim.dt, act, lab, H["particle_x"], H["particle_v"], H["particle_C"], H["particle_F"],
              H["particle_F_trial"], H["particle_cov"], sim.invdx, go, bool(mpm.update_cov), sim.minBounds)
        steps += 1
        el = time.perf_counter() - t0
        if el > budget_s or steps >= 50:
            break
    return {"value": n_mpm * steps / el, "unit": "particle-substeps/s", "cores": 1, "kind": "port",
            "sample": f"{steps} MPM substeps (stress+P2G+grid+G2P, FP32, no XPBD) of the same {n_mpm}-particle "
                      f"state, {el:.1f} s on one host core"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (every box here has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    from mpmwarpydw_amd import driver, getArgs, scenes, simulationRoutines as SR

    dom, ov = scenes.SCENES[a.scene](scale=a.scale)
    args = getArgs.get_args([], ov)
    S = driver.setup(args, dom)
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    n = sim.nPoints
    ratio = sim.mpmStepsPerXpbdStep
    print(f"[bench] scene={a.scene} n={n} grid={mpm.gridDims} ratio={ratio} steps={a.steps} warmup={a.warmup}",
          file=sys.stderr, flush=True)

    counter = 0
    for _ in range(a.warmup):
        S.substep(counter)
        counter += 1
    torch.cuda.synchronize()

    # per-phase events on the launch stream (torch's current stream is the one libmpmx launches on)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    phases = {"stress": 0.0, "p2g_grid": 0.0, "g2p": 0.0, "xpbd": 0.0}
    marks = []
    n_mpm_total = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        e = [ev() for _ in range(5)]
        e[0].record()
        SR.stress_phase(sim, mpm, xpbd)
        e[1].record()
        SR.p2g_grid_phase(sim, mpm, xpbd)
        e[2].record()
        SR.g2p_phase(sim, mpm, xpbd)
        e[3].record()
        if counter % ratio == 0 and counter > 0:
            SR.xpbdSimulationStep(sim, mpm, xpbd)
        e[4].record()
        marks.append(e)
        counter += 1
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for e in marks:
        phases["stress"] += e[0].elapsed_time(e[1])
        phases["p2g_grid"] += e[1].elapsed_time(e[2])
        phases["g2p"] += e[2].elapsed_time(e[3])
        phases["xpbd"] += e[3].elapsed_time(e[4])
    lab = sim.materialLabel.numpy()
    act = sim.activeLabel.numpy()
    n_mpm = int(((act == 1) & (lab <= 1)).sum())
    value = n_mpm * a.steps / wall
    G = int(np.prod(mpm.gridDims))
    per_phase_ms = {k: v / a.steps for k, v in phases.items()}
    # stress kernel: 124 B read + 84 B written per MPM particle (F_trial, C, 9 material floats, acc, damage, labels
    # in; F, stress, acc, damage, ys out)
    stress_bytes = 208.0 * n_mpm
    achieved = stress_bytes / (per_phase_ms["stress"] * 1e-3) / 1e9
    out = {
        "metric": "MPM particle-substeps/s (plain method)",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{a.scene}: {n} particles ({n_mpm} MPM), grid {mpm.gridDims[0]}x{mpm.gridDims[1]}x"
                               f"{mpm.gridDims[2]}, dt {sim.dt:g} s, XPBD every {ratio} substeps",
                   "particles": n, "grid_nodes": G},
        "phase_ms": per_phase_ms,
        "roofline": {"bound": "hbm", "kernel": "stress_kernel", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None},
        "algorithmic_bytes_per_substep": 300.0 * n_mpm + 28.0 * G,
        "substep_hbm_frac": (300.0 * n_mpm + 28.0 * G) / (wall / a.steps) / 1e9 / HBM_PEAK_GBS,
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(S, a.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
