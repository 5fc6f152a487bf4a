"""Benchmark of the MPM hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--scene undercut_slab]

A step is one MPM substep of the plain method over the whole scene: the fused
substep (constitutive update + deterministic P2G, grid update, G2P with
re-binning), plus the XPBD contact step every ``mpmStepsPerXpbdStep`` substeps
as the reference's loop runs it (runMPMYDW.py:677-682).  The default workload
is the undercut slab as configured: 960,000 particles, grid 104x54x31,
dt = 5e-5 s, XPBD every 10 substeps, 2,000 substeps from rest with damage on
(K = 2000 times the whole run).  Inputs are resident in HBM before the timed
region; at setup the particle arrays are permuted into the substep's tile
order (Simulation.reorder, a layout change that leaves the results bit for bit
the same; --no-reorder keeps the input order).

One XPBD period (10 substeps, the XPBD step included) is captured once in a
HIP graph.  W untimed substeps run first on the same state (kernel loading,
clocks, graph capture); then the device state is restored to t = 0 from a
device copy taken at setup and binned, and the graph is replayed K/10 times
between two device synchronisations: substeps 1..K of the run.

After the timed region:
  * the configured run: when K is shorter than the scene's configured run
    (2,000 substeps on the slab), substeps 1..2000 from t = 0 replayed the
    same way, reported in configured_run beside the K-substep value;
  * matched error: one more substep whose constitutive update is checked
    against the FP64 oracle on every MPM particle (per particle, relative to
    the larger of its largest stress component and its yield stress);
  * per-kernel durations from the library's launch profiler (HIP events
    stamped by each kernel's dispatch) over one eagerly launched period;
  * the driver loop's rate: the same substeps launched one by one with the
    per-substep residual and damage readback of runMPMYDW.py:684-709;
  * the CPU baseline: the FP32 restatement under oracle/ with OpenMP on all
    host threads from the same t = 0 state over a bounded number of substeps
    (XPBD steps at their ratio included), and on one core over fewer.

Prints one JSON line.
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
    p.add_argument("--steps", type=int, default=2000)
    p.add_argument("--warmup", type=int, default=40)
    p.add_argument("--scene", default="undercut_slab", help="a scene of scenes.SCENES, or 'ensemble'")
    p.add_argument("--scale", type=float, default=1.0)
    p.add_argument("--no-graph", action="store_true", help="launch eagerly instead of replaying a HIP graph")
    p.add_argument("--no-reorder", action="store_true",
                   help="keep the particle arrays in input order (no spatial reorder at setup)")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=12.0, help="budget of each CPU baseline sample")
    p.add_argument("--no-error-sample", action="store_true",
                   help="skip the FP64 check of the constitutive update after the timed region")
    p.add_argument("--ensemble-scenes", type=int, default=64, help="scenes of the batched ensemble")
    p.add_argument("--ensemble-steps", type=int, default=1000, help="substeps of each batched-ensemble run")
    p.add_argument("--no-learned", action="store_true", help="skip the learned+batched measurement")
    p.add_argument("--learned-only", action="store_true", help="only the learned+batched measurement")
    p.add_argument("--learned-tol", type=float, default=5e-5,
                   help="judge tolerance of the learned map (residuals against tol * cohesion)")
    p.add_argument("--no-autotune", action="store_true", help="default knobs instead of the autotuner's")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run the 256 MiB copy that calibrates the HBM byte counters")
    return p.parse_args()


def _host_state(S):
    """The t = 0 state of a scene as FP32 host arrays in the reference's AoS
    layout, for the CPU baseline."""
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    f = lambda a: np.ascontiguousarray(a.numpy().astype(np.float32))  # noqa: E731
    H = {k: f(getattr(sim, k)) for k in ("particle_x", "particle_v", "particle_mass", "particle_vol",
                                         "particle_density", "particle_radius")}
    for k in ("particle_F", "particle_F_trial", "particle_stress", "particle_accumulated_strain", "particle_damage",
              "particle_C", "particle_cov", "mu", "lam", "ys", "alpha", "hardening", "softening", "eta_shear",
              "eta_bulk", "strainCriteria"):
        H[k] = f(getattr(mpm, k))
    for k in ("particle_x_initial_xpbd", "particle_v_initial_xpbd", "particle_cumDist_xpbd"):
        H[k] = f(getattr(xpbd, k))
    H["activeLabel"] = np.ascontiguousarray(sim.activeLabel.numpy())
    H["materialLabel"] = np.ascontiguousarray(sim.materialLabel.numpy())
    return H


def _cpu_substep(o, H, S, counter, grids):
    """One substep of the reference's loop on the oracle's kernels: stress, P2G,
    grid update, G2P, and every ratio-th substep the XPBD step
    (simulationRoutines.py:6-526) with the reference's full candidate walk."""
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    act, lab = H["activeLabel"], H["materialLabel"]
    gm, gv, go = grids
    o.compute_stress_from_F_trial(act, lab, H["particle_x"], H["particle_v"], H["particle_x_initial_xpbd"],
                                  H["particle_v_initial_xpbd"], H["particle_F"], H["particle_F_trial"], H["mu"],
                                  H["lam"], H["ys"], H["hardening"], H["softening"], H["particle_density"],
                                  H["strainCriteria"], mpm.eff, H["eta_shear"], H["eta_bulk"], H["particle_C"],
                                  H["particle_stress"], H["particle_accumulated_strain"], H["particle_damage"], 1,
                                  H["alpha"])
    gm[:] = 0
    gv[:] = 0
    go[:] = 0
    o.p2g_apic_with_stress(act, lab, H["particle_stress"], H["particle_x"], H["particle_v"], H["particle_C"],
                           H["particle_vol"], H["particle_mass"], sim.dx, sim.invdx, sim.minBounds, mpm.rpic_damping,
                           gm, gv, sim.dt)
    o.grid_normalization_and_gravity(gm, gv, go, sim.gravity, sim.dt)
    o.collideBounds(go, *mpm.gridDims, mpm.boundFriction, sim.dt, abs(sim.gravity[2]), mpm.boundaryPadding)
    o.g2p(sim.dt, act, lab, H["particle_x"], H["particle_v"], H["particle_C"], H["particle_F"],
          H["particle_F_trial"], H["particle_cov"], sim.invdx, go, bool(mpm.update_cov), sim.minBounds)
    if counter % sim.mpmStepsPerXpbdStep or counter <= 0:
        return
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    n = sim.nPoints
    x0, v0 = H["particle_x"].copy(), H["particle_v"].copy()
    xi, vi = x0.copy(), v0.copy()
    dims = xpbd.particle_grid.dims
    ids, cs, ce = o.hashgrid_build(x0, np.float32(sim.dx), dims)
    o.integrateParticlesXPBD(act, lab, xi, vi, np.float32(sim.gravity), f32(sim.dtxpbd), xi, vi,
                             f32(xpbd.particle_v_max))
    if xpbd.mpm_contact_transition_lock:
        o.reset_mpm_contact_labels(act, lab)
    for _ in range(xpbd.xpbd_iterations):
        d = np.zeros((n, 3), np.float32)
        o.my_solve_particle_bound_contacts(act, lab, xi, vi, H["particle_mass"], H["particle_radius"],
                                           f32(xpbd.dynamicParticleFriction), f32(xpbd.staticVelocityThreshold),
                                           f32(xpbd.staticParticleFriction), np.float32(xpbd.minBoundsXPBD),
                                           np.float32(xpbd.maxBoundsXPBD), f32(sim.dtxpbd),
                                           f32(xpbd.xpbd_relaxation), d)
        o.my_solve_particle_particle_contacts(act, lab, (ids, cs, ce, np.float32(sim.dx), dims), xi, vi,
                                              H["particle_mass"], H["particle_radius"],
                                              f32(xpbd.dynamicParticleFriction), f32(xpbd.staticVelocityThreshold),
                                              f32(xpbd.staticParticleFriction), f32(xpbd.particle_cohesion),
                                              f32(xpbd.max_radius), f32(sim.dtxpbd), f32(xpbd.xpbd_relaxation),
                                              int(xpbd.mpm_contact_transition_lock), d)
        xo, vo = xi.copy(), vi.copy()
        o.my_apply_particle_deltas(act, lab, x0, xi, d, f32(sim.dtxpbd), f32(xpbd.particle_v_max), xo, vo)
        xi, vi = xo, vo
    o.sleepParticles(act, lab, f32(xpbd.sleepThreshold), H["particle_radius"], x0, xi, H["particle_cumDist_xpbd"],
                     f32(sim.dtxpbd))
    o.apply_mpm_contact_impulses(act, lab, x0, v0, xi, vi, f32(sim.dtxpbd), f32(xpbd.particle_v_max),
                                 f32(xpbd.xpbd_mpm_coupling_strength), 0, vi)
    H["particle_x"], H["particle_v"] = xi, vi


def cpu_baseline(S, H0, budget_s):
    """The FP32 restatement under oracle/ (reference semantics) from the t = 0
    state: with OpenMP on every host thread for up to ``budget_s`` seconds of
    substeps (whole XPBD periods when the budget allows), then on one core for
    a shorter sample.  Returns the cpu_baseline object of the JSON line."""
    from oracle import Oracle
    sim, mpm = S.sim, S.mpm
    ratio = sim.mpmStepsPerXpbdStep
    act, lab = H0["activeLabel"], H0["materialLabel"]
    n_mpm = int(((act == 1) & (lab <= 1)).sum())
    res = {}
    for name, o, budget, cap in (("all", Oracle(32, omp=True), budget_s, 10 * ratio),
                                 ("one", Oracle(32), budget_s / 3, ratio)):
        H = {k: v.copy() for k, v in H0.items()}
        grids = (np.zeros(mpm.gridDims, np.float32), np.zeros(mpm.gridDims + (3,), np.float32),
                 np.zeros(mpm.gridDims + (3,), np.float32))
        steps, t0 = 0, time.perf_counter()
        while True:
            steps += 1
            _cpu_substep(o, H, S, steps, grids)
            el = time.perf_counter() - t0
            if el > budget or steps >= cap:
                break
        res[name] = (n_mpm * steps / el, steps, el, o.threads())
    v, steps, el, threads = res["all"]
    v1, steps1, el1, _ = res["one"]
    xp = sum(1 for c in range(1, steps + 1) if c % ratio == 0)
    return {"value": v, "unit": "particle-substeps/s", "cores": threads, "kind": "port",
            "sample": f"substeps 1-{steps} of the same run from t = 0 ({xp} XPBD steps with the reference's full "
                      f"candidate walk), FP32 restatement (oracle/, OpenMP) on {threads} host threads, {el:.1f} s",
            "single_core_value": v1,
            "single_core_sample": f"substeps 1-{steps1}, serial FP32 build on one core, {el1:.1f} s"}


def stress_error_sample(S, eff, step=None, sample=None, seed=0, what="constitutive update of one substep after "
                        "the timed run, every active MPM particle"):
    """The constitutive update of one substep (``step()``, by default the fused
    substep of S) against the FP64 oracle on every active MPM particle (or on
    ``sample`` particles drawn at random), from the device state it starts
    from: per particle |tau_gpu - tau_64| / max(|tau_64|, ys_64, 1 Pa) (largest
    component)."""
    from mpmwarpydw_amd import simulationRoutines as SR
    from oracle import Oracle
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    idx = None
    if sample is not None and sample < sim.nPoints:
        idx = np.sort(np.random.default_rng(seed).choice(sim.nPoints, sample, replace=False))
    pick = (lambda a: a) if idx is None else (lambda a: a[idx])  # noqa: E731
    g = lambda a: np.ascontiguousarray(pick(a.numpy()).astype(np.float64))  # noqa: E731
    H = {k: g(getattr(sim, k)) for k in ("particle_x", "particle_v", "particle_density")}
    for k in ("particle_F", "particle_F_trial", "particle_stress", "particle_accumulated_strain", "particle_damage",
              "particle_C", "mu", "lam", "ys", "alpha", "hardening", "softening", "eta_shear", "eta_bulk",
              "strainCriteria"):
        H[k] = g(getattr(mpm, k))
    for k in ("particle_x_initial_xpbd", "particle_v_initial_xpbd"):
        H[k] = g(getattr(xpbd, k))
    act = np.ascontiguousarray(pick(sim.activeLabel.numpy()))
    lab = np.ascontiguousarray(pick(sim.materialLabel.numpy()))
    sel = (act == 1) & (lab <= 1)
    H0_Ft = H["particle_F_trial"].copy()
    mu0, lam0, ys0, al0, D0 = (H[k].copy() for k in ("mu", "lam", "ys", "alpha", "particle_damage"))
    if step is None:
        SR.mpmSimulationStepFused(sim, mpm, xpbd)
    else:
        step()
    tau_gpu = pick(mpm.particle_stress.numpy()).astype(np.float64)
    lab_gpu = pick(sim.materialLabel.numpy())
    t0 = time.perf_counter()
    acc0 = H["particle_accumulated_strain"].copy()
    Oracle(64).compute_stress_from_F_trial(act, lab, H["particle_x"], H["particle_v"], H["particle_x_initial_xpbd"],
                                           H["particle_v_initial_xpbd"], H["particle_F"], H["particle_F_trial"],
                                           H["mu"], H["lam"], H["ys"], H["hardening"], H["softening"],
                                           H["particle_density"], H["strainCriteria"], eff, H["eta_shear"],
                                           H["eta_bulk"], H["particle_C"], H["particle_stress"],
                                           H["particle_accumulated_strain"], H["particle_damage"], 1, H["alpha"])
    el = time.perf_counter() - t0
    tau = H["particle_stress"][sel]
    scale = np.maximum(np.abs(tau).max(axis=(1, 2)), np.maximum(np.abs(H["ys"][sel]), 1.0))
    err = np.abs(tau_gpu[sel] - tau).max(axis=(1, 2)) / scale
    plastic = (H["particle_accumulated_strain"] != acc0)[sel]
    margin = yield_margin(H0_Ft, mu0, lam0, ys0, al0, D0)[sel]
    # the reference's projection jumps at the yield surface (dgam = |dev e| + 1e-6 - Y / (2 mu) is 1e-6,
    # not 0, there): where the FP64 margin is below FP32 resolution, FP32 may decide either way
    marginal = np.abs(margin) < MARGIN
    ok = err[~marginal]
    return {"max": float(ok.max()) if ok.size else 0.0, "p99_9": float(np.quantile(ok, 0.999)) if ok.size else 0.0,
            "mean": float(ok.mean()) if ok.size else 0.0, "particles": int(sel.sum()),
            "plastic_in_sample": int(plastic.sum()), "labels_identical": bool(np.array_equal(lab_gpu, lab)),
            "marginal_particles": int(marginal.sum()),
            "marginal_max": float(err[marginal].max()) if marginal.any() else 0.0,
            "definition": f"{what}, "
                          "against the FP64 oracle: max over particles of |tau - tau_fp64|_max / "
                          "max(|tau_fp64|_max, ys, 1 Pa); 'marginal' particles, whose FP64 yield margin "
                          f"|sigma_eq - Y| / |Y| is below {MARGIN:g}, are counted apart: there the FP32 yield test "
                          "can decide either way and the reference's map jumps by 2 mu 1e-6 in strain",
            "oracle_seconds": round(el, 1)}


MARGIN = 1e-6


def yield_margin(Ft, mu, lam, ys, alpha, D):
    """(sigma_eq - Y) / |Y| of the return map's yield test (mpmRoutines.py:341-361)
    in FP64, from the principal stretches of F_trial clamped at 0.01 (the
    negative one of an inverted F included)."""
    s = np.linalg.svd(Ft, compute_uv=False)
    s[:, 2] *= np.sign(np.linalg.det(Ft))
    e = np.log(np.maximum(s, 0.01))
    dev = e - e.mean(axis=1, keepdims=True)
    seq = 2 * mu * np.linalg.norm(dev, axis=1)
    Y = (1 - D) * (ys - alpha * (2 * mu / 3 + lam) * e.sum(axis=1))
    with np.errstate(divide="ignore", invalid="ignore"):
        return (seq - Y) / np.abs(Y)


def driver_loop_rate(S, counter0, substeps):
    """Particle-substeps/s of the driver loop as runMPMYDW.py:677-709 runs it:
    substeps launched one by one, each followed by the FP64 residual and mean
    damage reductions and their readback (a host synchronisation)."""
    import torch
    sim = S.sim
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(substeps):
        S.substep(counter0 + k)
        S.loop_scalars()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    lab = sim.materialLabel.numpy()
    act = sim.activeLabel.numpy()
    n_mpm = int(((act == 1) & (lab <= 1)).sum())
    return {"value": n_mpm * substeps / el, "substeps": substeps, "ms_per_step": el * 1e3 / substeps,
            "what": "substeps launched one by one, each with the residual and mean-damage readback"}


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
    if a.learned_only:
        print(json.dumps(bench_learned_batched(a)), flush=True)
        return
    import torch
    from mpmwarpydw_amd import driver, getArgs, scenes
    from mpmwarpydw_amd import simulationRoutines as SR
    from mpmwarpydw_amd._lib import lib
    from mpmwarpydw_amd.runtime import stream_handle

    from mpmwarpydw_amd import autotune
    dom, ov = scenes.SCENES[a.scene](scale=a.scale)
    args = getArgs.get_args([], ov)
    S = driver.setup(args, dom)
    sim, mpm = S.sim, S.mpm
    tuning = None
    if not a.no_autotune:
        # launch shape and particle order (part of setup): every candidate timed over substeps 1-23 from the
        # t = 0 state, which the tuner restores before each candidate and leaves in the order it keeps
        tc = [1]

        def tstep():
            S.substep(tc[0])
            tc[0] += 1

        with torch.cuda.stream(torch.cuda.Stream()):
            tuning = autotune.tune(S, tstep, orders=(False,) if a.no_reorder else (False, True))
        torch.cuda.synchronize()
        print(f"[bench] autotune: {tuning['best']} ({len(tuning['candidates'])} candidates, "
              f"{tuning['tuning_s']} s)", file=sys.stderr, flush=True)
        reorder = tuning["kept_reordered"]
    else:
        reorder = not a.no_reorder
        if reorder:
            S.reorder()  # part of setup: tile order in HBM (results are the same bits either way)
    n = sim.nPoints
    ratio = sim.mpmStepsPerXpbdStep
    periods = max(1, a.steps // ratio)
    steps = periods * ratio
    print(f"[bench] scene={a.scene} n={n} grid={mpm.gridDims} ratio={ratio} steps={steps} warmup={a.warmup}",
          file=sys.stderr, flush=True)
    snap = snapshot(S)  # the t = 0 device state, restored before the timed run
    H0 = None if a.no_cpu_baseline else _host_state(S)

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

    # back to t = 0: the device arrays from the copy taken at setup, then the bins of that state
    with torch.cuda.stream(stream):
        restore(S, snap)
        SR.mpmSimulationBin(sim, mpm, S.xpbd)
    torch.cuda.synchronize()

    t0 = time.perf_counter()
    with torch.cuda.stream(stream):
        for p in range(periods):
            if graph is not None:
                graph.replay()
            else:
                for k in range(ratio):
                    S.substep(p * ratio + k + 1)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lab = sim.materialLabel.numpy()
    act = sim.activeLabel.numpy()
    n_mpm = int(((act == 1) & (lab <= 1)).sum())  # MPM particles at the end of the timed region
    state_end = {"mpm_particles": n_mpm, "xpbd_particles": int(((act == 1) & (lab == 2)).sum()),
                 "plastic_particles": int((mpm.particle_accumulated_strain.numpy() > 0).sum()),
                 "mean_damage": float(mpm.particle_damage.numpy()[act == 1].mean())}

    # the configured run as a whole when the timed window is shorter: back to t = 0 and substeps
    # 1..nSteps replayed the same way; reported beside the window, which stays the line's value
    configured = None
    total = int(sim.nSteps) // ratio * ratio
    if graph is not None and total > steps:
        with torch.cuda.stream(stream):
            restore(S, snap)
            SR.mpmSimulationBin(sim, mpm, S.xpbd)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            for p in range(total // ratio):
                graph.replay()
        torch.cuda.synchronize()
        wall_full = time.perf_counter() - t1
        lab = sim.materialLabel.numpy()
        act = sim.activeLabel.numpy()
        n_full = int(((act == 1) & (lab <= 1)).sum())
        configured = {
            "substeps": total, "value": n_full * total / wall_full, "ms_per_step": wall_full * 1e3 / total,
            "state_at_end": {"mpm_particles": n_full, "xpbd_particles": int(((act == 1) & (lab == 2)).sum()),
                             "plastic_particles": int((mpm.particle_accumulated_strain.numpy() > 0).sum()),
                             "mean_damage": float(mpm.particle_damage.numpy()[act == 1].mean())},
            "what": f"substeps 1-{total} of the configured run from t = 0, replayed from the same HIP graph "
                    "after the timed region (MPM particles at the end x substeps / wall time)"}

    # per-kernel device time over the same substeps: back to t = 0 and substeps 1..K again, launched
    # eagerly with every library kernel carrying dispatch-stamped events (graph launches cannot carry
    # them), in chunks of two periods queued behind a hold kernel so that each chunk runs back to back
    # at full clocks; the run repeats the timed one bit for bit
    with torch.cuda.stream(stream):
        restore(S, snap)
        SR.mpmSimulationBin(sim, mpm, S.xpbd)
    torch.cuda.synchronize()
    del snap
    chunk = 2 * ratio
    lib("mpmx_profile_begin", min(1 << 22, 2 * 64 * (steps + ratio)))
    for c0 in range(1, steps + 1, chunk):
        with torch.cuda.stream(stream):
            lib("mpmx_profile_hold", stream_handle(), PROFILE_HOLD_US)
            for c in range(c0, min(c0 + chunk, steps + 1)):
                S.substep(c)
        torch.cuda.synchronize()
    lib("mpmx_profile_end")
    kernels = profile_report()
    kernels.pop("hold", None)
    how = ("HIP start/stop events stamped by each kernel's own dispatch (hipExtLaunchKernelGGL) over substeps "
           f"1-{steps} run again from t = 0 after the timed region, eagerly, in chunks of {chunk} substeps queued "
           "behind a hold kernel")
    counter = steps + 1

    err = None
    if not a.no_error_sample:  # one more substep (counter steps + 1: no XPBD step), checked against FP64
        with torch.cuda.stream(stream):
            err = stress_error_sample(S, mpm.eff)
        counter += 1
        while counter % ratio != 1:
            S.substep(counter)
            counter += 1
    period = [counter + k for k in range(ratio)]
    counter = period[-1] + 1
    with torch.cuda.stream(stream):
        loop = driver_loop_rate(S, counter, 5 * ratio)

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
                               f"{mpm.gridDims[2]}, dt {sim.dt:g} s, XPBD every {ratio} substeps, substeps 1-{steps} "
                               f"of the run from t = 0",
                   "particles": n, "grid_nodes": G, "hip_graph": graph is not None,
                   "particle_order": "tile order (reordered once at setup)" if reorder else "input order",
                   "knobs": autotune.current()},
        "scope": ("the timed region replays the substeps (MPM substep + XPBD step at its ratio) as HIP graphs with "
                  "no host synchronisation; the driver loop's per-substep readback is measured separately in "
                  "driver_loop"),
        "state_at_end": state_end,
        "configured_run": configured,
        "stress_error_vs_fp64": err,
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
        "driver_loop": loop,
    }
    if tuning is not None:
        out["autotune"] = {"best": tuning["best"], "kept_reordered": tuning["kept_reordered"],
                           "candidates": tuning["candidates"], "tuning_s": tuning["tuning_s"],
                           "tuned_on": "the scene itself at setup: each candidate over substeps 1-23 from the "
                                       "t = 0 state (XPBD steps at their ratio), restored before each candidate"}
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(S, H0, a.cpu_seconds)
    if not a.no_learned:
        del S
        torch.cuda.empty_cache()
        out["learned_batched"] = bench_learned_batched(a)
    print(json.dumps(out), flush=True)


def snapshot(S):
    """Device copies of every array of the simulation state."""
    from mpmwarpydw_amd import runtime as rt
    snap, seen = [], set()
    for obj in (S.sim, S.mpm, S.xpbd):
        for name, arr in vars(obj).items():
            if isinstance(arr, rt.Array) and id(arr) not in seen:
                seen.add(id(arr))
                snap.append((arr, arr.t.clone()))
    return snap


def restore(S, snap):
    for arr, t in snap:
        arr.t.copy_(t)


def train_block_network(log=None):
    """The learned return map used by the batched measurement: trained on the
    GPU (the library's own forward, backward and Adam kernels) on 10M plastic
    trial states harvested from 32,768-particle block-collapse runs of 500
    substeps whose materials span the ensemble's ranges."""
    import torch
    from mpmwarpydw_amd.learned import LearnedReturnMap, harvest_blocks
    t0 = time.perf_counter()
    X, T, stats = harvest_blocks(10_000_000, max_runs=100, substeps=500, seed=0, per_run=400_000, every=5)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    net = LearnedReturnMap(seed=5)
    hist = net.train(X, T, steps=30000, batch=8192, lr=3e-3, lr_final=2e-5, seed=1, log_every=10000)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    info = {"pairs": stats["pairs"], "runs": len(stats["runs"]),
            "runs_with_plastic_states": sum(1 for r in stats["runs"] if r["pairs"]), "harvest_s": round(t1 - t0, 2),
            "train_steps": 30000, "batch": 8192, "train_s": round(t2 - t1, 2), "final_loss": hist[-1],
            "what": "10M plastic trial states from 32,768-particle block-collapse runs of 500 substeps, materials "
                    "drawn from the ensemble's ranges; Adam on the library's bfloat16 MFMA forward/backward"}
    if log:
        log(f"[bench] network: {info}")
    del X, T
    return net, info


def _ensemble_graph(E, stream, per_graph, learned=None, counters=None, tol=1e-4):
    """A HIP graph of ``per_graph`` ensemble substeps (plain, or with the
    learned map), captured after one eager substep (workspace and bins
    exist).  Replays use the arrays as they are now: nothing may reallocate
    them afterwards."""
    import torch

    def sub():
        if learned is None:
            E.substep()
        else:
            E.substep(learned=learned, counters=counters, tol=tol)

    with torch.cuda.stream(stream):
        sub()
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=stream):
        for _ in range(per_graph):
            sub()
    return graph


def plain_per_scene(E, stream, steps, per_graph=10):
    """The plain GPU path over the ensemble's scenes one at a time, as the
    reference runs a scene per process: each scene alone, in tile order, from
    its t = 0 state, one eager substep, then ``steps`` MPM substeps replayed
    from a HIP graph of ``per_graph`` fused substeps.  Returns the aggregate
    particle-substeps/s (the scenes' particle-substeps over the sum of their
    wall times) and the per-scene figures of the smallest and largest scene."""
    import torch
    from mpmwarpydw_amd import simulationRoutines as SR
    reps = max(1, steps // per_graph)
    walls, work, rows = 0.0, 0, []
    for S in E.scenes_sims:
        S.reorder()
        sub = lambda: SR.mpmSimulationStepFused(S.sim, S.mpm, S.xpbd)  # noqa: E731
        with torch.cuda.stream(stream):
            sub()
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=stream):
            for _ in range(per_graph):
                sub()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            for _ in range(reps):
                g.replay()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        del g
        lab, act = S.sim.materialLabel.numpy(), S.sim.activeLabel.numpy()
        n_mpm = int(((act == 1) & (lab <= 1)).sum())
        walls += wall
        work += n_mpm * reps * per_graph
        rows.append((S.sim.nPoints, wall * 1e3 / (reps * per_graph)))
    rows.sort()
    return {"value": work / walls, "ms_per_step": walls * 1e3 / (reps * per_graph), "steps": reps * per_graph,
            "smallest_scene": {"particles": rows[0][0], "ms_per_step": round(rows[0][1], 4)},
            "largest_scene": {"particles": rows[-1][0], "ms_per_step": round(rows[-1][1], 4)},
            "what": "each scene alone through the plain fused substep (substeps 2 to steps + 1 from t = 0, HIP "
                    f"graphs of {per_graph} substeps), one scene after another; ms_per_step sums the scenes"}


def bench_learned_batched(a, net=None, net_info=None):
    """The batched ensemble (64 block scenes of 20k-500k particles, seeded
    materials) for ``a.ensemble_steps`` substeps from t = 0, once with the plain
    constitutive update and once with the learned map (judged by the exact
    map), each replayed from HIP graphs of 10
    substeps; launch shape and particle order from the autotuner on the
    ensemble's own state.  Then one learned substep is checked against the
    FP64 oracle on 2M particles drawn at random.  The judge's tolerance
    (a.learned_tol, 5e-5 of the cohesion on each residual) keeps the learned
    stresses within 1e-4 of the FP64 ones, the plain method's bound."""
    import torch
    from mpmwarpydw_amd import autotune, driver, getArgs
    from mpmwarpydw_amd.ensemble import Ensemble, c4_scene_list
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if net is None:
        net, net_info = train_block_network(log)
    t0 = time.perf_counter()
    specs = c4_scene_list(count=a.ensemble_scenes)
    E = Ensemble([driver.setup(getArgs.get_args([], ov), dom) for dom, ov in specs])
    setup_s = time.perf_counter() - t0
    n, G = E.n, E.node_offsets[-1]
    log(f"[bench] ensemble: {len(E.table)} scenes, {n} particles, {G} nodes, setup {setup_s:.1f} s")
    stream = torch.cuda.Stream()
    knobs_before = autotune.current()
    state = lambda: E._state_arrays() + [E.grid_m, E.grid_v_in, E.grid_v_out]  # noqa: E731
    # the tuner times every candidate from the t = 0 state and leaves it there, in the order it keeps
    # (reordering allocates new arrays: it happens before any graph capture)
    with torch.cuda.stream(stream):
        tuning = autotune.tune(E, lambda: E.substep(learned=net, tol=a.learned_tol), split_forced=True)
    torch.cuda.synchronize()
    log(f"[bench] ensemble autotune: {tuning['best']}")
    with torch.cuda.stream(stream):
        E.bin()
    torch.cuda.synchronize()
    snap = [(arr, arr.t.clone()) for arr in state()]
    per_graph = 10
    reps = max(1, a.ensemble_steps // per_graph)
    steps = reps * per_graph
    out = {}
    cnt = torch.zeros(3, dtype=torch.int64, device="cuda")
    for form in ("plain", "learned"):
        g = _ensemble_graph(E, stream, per_graph, learned=None if form == "plain" else net, counters=cnt,
                            tol=a.learned_tol)
        for arr, t in snap:  # back to t = 0 in place: the graph's pointers stay valid
            arr.t.copy_(t)
        with torch.cuda.stream(stream):
            E.bin()
        cnt.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        with torch.cuda.stream(stream):
            for _ in range(reps):
                g.replay()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t1
        lab, act = E.sim.materialLabel.numpy(), E.sim.activeLabel.numpy()
        n_mpm = int(((act == 1) & (lab <= 1)).sum())
        out[form] = {"value": n_mpm * steps / wall, "ms_per_step": wall * 1e3 / steps, "mpm_particles_at_end": n_mpm}
        if form == "learned":
            c = cnt.cpu().numpy()  # plastic states and accepted predictions over the run
            out[form].update({"plastic_states": int(c[0]), "accepted": int(c[1]),
                              "accepted_fraction": float(c[1] / max(c[0], 1))})
        log(f"[bench] ensemble {form}: {out[form]}")
        del g
    del snap
    # per-kernel device time of one eager substep of each form (dispatch-stamped events)
    from mpmwarpydw_amd._lib import lib
    from mpmwarpydw_amd.runtime import stream_handle
    kernels = {}
    for form in ("plain", "learned"):
        with torch.cuda.stream(stream):
            lib("mpmx_profile_hold", stream_handle(), PROFILE_HOLD_US)
            for _ in range(2):
                E.substep() if form == "plain" else E.substep(learned=net, tol=a.learned_tol)
            lib("mpmx_profile_begin", 4096)
            for _ in range(2):
                E.substep() if form == "plain" else E.substep(learned=net, tol=a.learned_tol)
        torch.cuda.synchronize()
        lib("mpmx_profile_end")
        kernels[form] = {k: round(v["avg_us"], 2) for k, v in profile_report().items()}
    log(f"[bench] ensemble kernels: {kernels}")
    # matched error: one more learned substep against the FP64 oracle on a random sample
    err = stress_error_sample(E, E.mpm.eff, step=lambda: E.substep(learned=net, tol=a.learned_tol),
                              sample=2_000_000, seed=7,
                              what=f"constitutive update (learned map, judge tolerance {a.learned_tol:g} of the "
                                   f"cohesion) of substep {steps + 1} of the batched ensemble, 2,000,000 particles "
                                   "drawn at random")
    autotune.apply(knobs_before)
    # the plain GPU path without batching: every scene alone, one after another
    per_scene = plain_per_scene(E, stream, steps)
    log(f"[bench] plain, scene by scene: {per_scene}")
    res = {
        "metric": "MPM particle-substeps/s (learned+batched)",
        "value": out["learned"]["value"],
        "unit": "particle-substeps/s",
        "ms_per_step": out["learned"]["ms_per_step"],
        "steps": steps,
        "workload": f"batched ensemble: {len(E.table)} block-collapse scenes, {n} particles (log-uniform 20k-500k "
                    f"per scene, friction 20-40 deg, cohesion 1e4-1e6 Pa, E 1e7-1e9 Pa, seeded), {G} grid nodes, "
                    f"dt 1e-4 s, substeps 1-{steps} from t = 0, MPM substep (the XPBD step is per scene and not "
                    "part of this measurement), HIP graphs of 10 substeps",
        "learned": out["learned"],
        "plain_batched": out["plain"],
        "plain_per_scene": per_scene,
        "learned_vs_plain_batched": out["learned"]["value"] / out["plain"]["value"],
        "learned_batched_vs_plain_per_scene": out["learned"]["value"] / per_scene["value"],
        "plain_batched_vs_plain_per_scene": out["plain"]["value"] / per_scene["value"],
        "judge_tolerance": a.learned_tol,
        "stress_error_vs_fp64": err,
        "kernels_us_per_launch": kernels,
        "network": net_info,
        "autotune": {"best": tuning["best"], "kept_reordered": tuning["kept_reordered"],
                     "candidates": len(tuning["candidates"]), "tuning_s": tuning["tuning_s"],
                     "tuned_on": "the learned ensemble substep, each candidate over substeps 1-23 from the "
                                 "t = 0 state"},
        "setup_s": round(setup_s, 2),
    }
    del E
    torch.cuda.empty_cache()
    return res


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
