"""Benchmark of the MPM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scene undercut_slab]

A step is one MPM substep of the plain method over the whole scene: the fused
substep (constitutive update + deterministic P2G, grid update, G2P with
re-binning), plus the XPBD contact step every ``mpmStepsPerXpbdStep`` substeps
as the reference's loop runs it (runMPMYDW.py:677-682).  The default workload
is the undercut slab: 960,000 particles, grid 104x54x31, dt = 5e-5 s, XPBD
every 10 substeps.  Inputs are resident in HBM before the timed region; at
setup the particle arrays are permuted into the substep's tile order
(Simulation.reorder, a layout change that leaves the results bit for bit the
same; --no-reorder keeps the input order).

One XPBD period (10 substeps, the XPBD step included) is captured once in a
HIP graph; W untimed substeps run first, then the graph is replayed K/10
times between two device synchronisations.  Per-kernel durations come from
the library's launch profiler (HIP events just before and after every kernel,
on the launch stream) over one more period launched eagerly right after the
timed region.

Prints one JSON line: the metric (MPM particle-substeps/s), per-kernel device
times, the roofline entry of the dominant kernel and a CPU baseline (the FP32
restatement under oracle/, one core, on a bounded sample).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per MPM particle of the substep's particle kernels:
#   substep_tile(stress+p2g), constitutive update inside the P2G tile pass: reads
#     F_trial 36, C 36, nine material floats 36, accumulated strain 4, damage 4,
#     label 4, x 12, v 12, vol 4, mass 4, bin id 4; writes F 36, stress 36,
#     accumulated strain / damage / ys 12
#   substep_stress, the constitutive update alone: reads F_trial 36, nine
#     material floats 36, accumulated strain 4, damage 4, two labels 8 (and C 36
#     when a viscosity is nonzero; the scenes here have none); writes F 36,
#     stress 36, accumulated strain / damage / ys 12
#   substep_tile(p2g), the P2G tile pass after it: reads bin id 4, label 4, x 12,
#     v 12, C 36, vol 4, mass 4, stress 24 (symmetric: 6 of its 9 planes)
#   substep_g2p: reads bin id 4, label 4, x 12, F 36; writes x 12, v 12, C 36,
#     F_trial 36, next bin key 4, rank 4
KERNEL_BYTES_PER_PARTICLE = {
    "substep_tile(stress+p2g)": 156.0 + 84.0,
    "substep_stress": 88.0 + 84.0,
    "substep_tile(p2g)": 100.0,
    "substep_g2p": 56.0 + 104.0,
}
PROFILE_HOLD_US = 50000  # long enough for the host to queue a whole profiled period


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=400)
    p.add_argument("--warmup", type=int, default=40)
    p.add_argument("--scene", default="undercut_slab", help="a scene of scenes.SCENES, or 'ensemble'")
    p.add_argument("--scale", type=float, default=1.0)
    p.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a HIP graph")
    p.add_argument("--no-reorder", action="store_true",
                   help="keep the particle arrays in input order (no spatial reorder at setup)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU baseline sample")
    p.add_argument("--ensemble-scenes", type=int, default=64, help="scenes of the batched ensemble (--scene ensemble)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run the 256 MiB copy that calibrates the HBM byte counters")
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
            "sample": f"{steps} MPM substeps (stress+P2G+grid+G2P, FP32, no XPBD step) of the same "
                      f"{n_mpm}-particle state, {el:.1f} s on one host core"}


def profile_report():
    from mpmwarpydw_amd._lib import lib
    buf = C.create_string_buffer(1 << 16)
    lib("mpmx_profile_report", buf, len(buf))
    out = {}
    for line in buf.value.decode().splitlines():
        name, cnt, ms = line.rsplit(" ", 2)
        out[name] = {"launches": int(cnt), "total_ms": float(ms), "avg_us": 1e3 * float(ms) / max(1, int(cnt))}
    return out


def roofline_kernel(kernels, n_mpm):
    """(name, algorithmic bytes per launch, achieved GB/s) of the particle kernel
    with the most device time."""
    have = {k: v for k, v in kernels.items() if k in KERNEL_BYTES_PER_PARTICLE}
    if not have:
        return None, None, None
    name = max(have, key=lambda k: have[k]["total_ms"])
    nbytes = KERNEL_BYTES_PER_PARTICLE[name] * n_mpm
    return name, nbytes, nbytes / (have[name]["avg_us"] * 1e-6) / 1e9


def measured_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 counter
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on a copy kernel), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_substep_kernels.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except OSError:
        return None, None
    if kernel not in d.get("bytes_per_launch", {}):
        return None, None
    return float(d["bytes_per_launch"][kernel]), d.get("source")


def copy_calibration():
    """A plain copy of 256 MiB through the library's copy kernel, three times:
    the reference point for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    from mpmwarpydw_amd._lib import lib
    n = 64 << 20
    src = torch.ones(n, device="cuda")
    dst = torch.empty(n, device="cuda")
    for _ in range(3):
        lib("mpmx_copy_f32", dst.data_ptr(), src.data_ptr(), n, None)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 3}), flush=True)


def main():
    a = parse()
    if a.copy_calibration:
        return copy_calibration()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {a.gpus} is not supported (every box here has one GPU).",
              file=sys.stderr)
        sys.exit(2)
    if a.scene == "ensemble":
        return bench_ensemble(a)
    import torch
    from mpmwarpydw_amd import driver, getArgs, scenes
    from mpmwarpydw_amd._lib import lib
    from mpmwarpydw_amd.runtime import stream_handle

    dom, ov = scenes.SCENES[a.scene](scale=a.scale)
    args = getArgs.get_args([], ov)
    S = driver.setup(args, dom)
    sim, mpm = S.sim, S.mpm
    if not a.no_reorder:
        S.reorder()  # part of setup: tile order in HBM (results are the same bits either way)
    n = sim.nPoints
    ratio = sim.mpmStepsPerXpbdStep
    periods = max(1, a.steps // ratio)
    steps = periods * ratio
    print(f"[bench] scene={a.scene} n={n} grid={mpm.gridDims} ratio={ratio} steps={steps} warmup={a.warmup}",
          file=sys.stderr, flush=True)

    counter = 1  # start inside a period so that warmup + replays keep the reference's XPBD phase
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for _ in range(max(a.warmup, ratio)):
            S.substep(counter)
            counter += 1
        while counter % ratio != 1:
            S.substep(counter)
            counter += 1
    torch.cuda.synchronize()

    period = [counter + k for k in range(ratio)]  # the XPBD step falls on the multiple of ratio inside
    graph = None
    if not a.no_graph:
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph, stream=stream):
            for c in period:
                S.substep(c)
        graph.replay()  # the captured period runs once before timing (untimed)
        torch.cuda.synchronize()

    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(periods):
            if graph is not None:
                graph.replay()
            else:
                for c in period:
                    S.substep(c)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lab = sim.materialLabel.numpy()
    act = sim.activeLabel.numpy()
    n_mpm = int(((act == 1) & (lab <= 1)).sum())  # MPM particles at the end of the timed region

    # per-kernel device time: one more period launched eagerly on the same stream, every
    # library kernel carrying dispatch-stamped events (graph launches cannot carry them)
    with torch.cuda.stream(stream):
        # the periods queue up behind the hold and then run back to back; the
        # first brings the clocks back up after the idle hold, the second is timed
        lib("mpmx_profile_hold", stream_handle(), PROFILE_HOLD_US)
        for c in period:
            S.substep(c)
        lib("mpmx_profile_begin", 2 * (64 * ratio + 64))
        for c in period:
            S.substep(c)
    torch.cuda.synchronize()
    lib("mpmx_profile_end")
    kernels = profile_report()
    how = ("HIP start/stop events stamped by each kernel's own dispatch (hipExtLaunchKernelGGL) over one eager "
           "period after the timed region; it and a warm-up period are queued behind a hold kernel")

    value = n_mpm * steps / wall
    G = int(np.prod(mpm.gridDims))
    dom_name, dom_k = max(kernels.items(), key=lambda kv: kv[1]["total_ms"])
    rk, rbytes, achieved = roofline_kernel(kernels, n_mpm)
    sub_bytes = 300.0 * n_mpm + 28.0 * G
    traffic, traffic_src = measured_traffic(rk) if a.scene == "undercut_slab" else (None, None)
    out = {
        "metric": "MPM particle-substeps/s (plain method)",
        "value": value,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{a.scene}: {n} particles ({n_mpm} MPM), grid {mpm.gridDims[0]}x{mpm.gridDims[1]}x"
                               f"{mpm.gridDims[2]}, dt {sim.dt:g} s, XPBD every {ratio} substeps",
                   "particles": n, "grid_nodes": G, "hip_graph": graph is not None,
                   "particle_order": "input order" if a.no_reorder else "tile order (reordered once at setup)"},
        "roofline": {"bound": "hbm", "kernel": rk, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": traffic, "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": rbytes},
        "substep_algorithmic_bytes": sub_bytes,
        "substep_hbm_frac": sub_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        "dominant_kernel": dom_name,
        "kernel_timing": how,
        "kernels_us_per_launch": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
    }
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(S, a.cpu_seconds)
    print(json.dumps(out), flush=True)


def bench_ensemble(a):
    """The batched ensemble: 64 block-collapse scenes of uneven size (log-uniform
    in [20k, 500k] particles, about 8M in total, seeded materials), one launch
    per kernel for all scenes, ten substeps captured per HIP graph replay,
    launch shape and particle order chosen by the autotuner on the same state.
    The XPBD step is per scene and not part of this measurement."""
    import torch
    from mpmwarpydw_amd import autotune, driver, getArgs
    from mpmwarpydw_amd._lib import lib
    from mpmwarpydw_amd.ensemble import Ensemble, c4_scene_list
    from mpmwarpydw_amd.runtime import stream_handle

    t0 = time.perf_counter()
    specs = c4_scene_list(count=a.ensemble_scenes)
    sims = [driver.setup(getArgs.get_args([], ov), dom) for dom, ov in specs]
    E = Ensemble(sims)
    del sims
    n = E.n
    G = E.node_offsets[-1]
    print(f"[bench] ensemble: {len(E.table)} scenes, {n} particles, {G} grid nodes, setup "
          f"{time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        E.substep()
        tuning = autotune.tune(E, E.substep)
        for _ in range(a.warmup):
            E.substep()
    torch.cuda.synchronize()
    print(f"[bench] autotune: {tuning['best']}", file=sys.stderr, flush=True)
    per_graph = 10
    reps = max(1, a.steps // per_graph)
    steps = reps * per_graph
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for _ in range(per_graph):
            E.substep()
    graph.replay()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    with torch.cuda.stream(stream):
        for _ in range(reps):
            graph.replay()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    lab = E.sim.materialLabel.numpy()
    act = E.sim.activeLabel.numpy()
    n_mpm = int(((act == 1) & (lab <= 1)).sum())  # MPM particles at the end of the timed region
    with torch.cuda.stream(stream):
        lib("mpmx_profile_hold", stream_handle(), PROFILE_HOLD_US)
        for _ in range(per_graph):
            E.substep()
        lib("mpmx_profile_begin", 4096)
        for _ in range(per_graph):
            E.substep()
    torch.cuda.synchronize()
    lib("mpmx_profile_end")
    kernels = profile_report()
    rk, rbytes, achieved = roofline_kernel(kernels, n_mpm)
    sub_bytes = 300.0 * n_mpm + 28.0 * G
    out = {
        "metric": "MPM particle-substeps/s (batched ensemble)",
        "value": n_mpm * steps / wall,
        "unit": "particle-substeps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"batched ensemble: {len(E.table)} block-collapse scenes, {n} particles "
                               f"(log-uniform 20k-500k per scene), {G} grid nodes, dt 1e-4 s, MPM substep only",
                   "particles": n, "grid_nodes": G, "scenes": len(E.table), "hip_graph": True,
                   "substeps_per_graph": per_graph},
        "autotune": tuning,
        "roofline": {"bound": "hbm", "kernel": rk, "achieved": achieved,
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
                     "traffic": None},
        "substep_algorithmic_bytes": sub_bytes,
        "substep_hbm_frac": sub_bytes / (wall / steps) / 1e9 / HBM_PEAK_GBS,
        "kernels_us_per_launch": {k: round(v["avg_us"], 2) for k, v in kernels.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
