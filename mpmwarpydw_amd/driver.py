"""Setup and substep loop of the reference's runMPMYDW.py around the device
hot path.

``setup(args, domain)`` reproduces the driver's setup arithmetic
(runMPMYDW.py:22-67 grid spacing and bounds, :110-169 CFL dt and the integer
XPBD ratio, :278-312 material upload, Lame constants and mass, :413-470
geostatic/identity F, damage and radius, :472-495 the XPBD box) on a domain
given as arrays; ``load_domain`` reads the reference's HDF5 domain files (with
this package's own reader, h5py is not part of this build) or .vtp domains.

``run`` is the big-step loop of runMPMYDW.py:638-877: the combined MPM+XPBD
phase with its gravity pulse and termination rules (minSteps = 1000, velocity
residual, damage stall), the XPBD-only phase with its sleep test, and the
deactivation below the datum; its reductions are FP64 on the device with one
readback per substep.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import mpmRoutines, simulationRoutines
from . import runtime as rt
from ._lib import lib
from .simStates import MPMState, SimState, XPBDState

MPM_PADDING = 2  # runMPMYDW.py:36

PROPERTY_NAMES = ["density", "E", "nu", "ys", "alpha", "hardening", "softening", "eta_shear", "eta_bulk",
                  "strainCriteria", "damage"]


@dataclass
class Geometry:
    dx: float
    invdx: float
    minBounds: np.ndarray
    maxBounds: np.ndarray
    gridDims: tuple
    dt: float
    mpmStepsPerXpbdStep: int


def grid_geometry(x: np.ndarray, particle_volume: np.ndarray, args) -> tuple[float, np.ndarray, np.ndarray, tuple]:
    """dx, padded bounds and gridDims exactly as runMPMYDW.py:24-65."""
    particleDiameter = np.mean(particle_volume) ** 0.33
    dx = particleDiameter * args.grid_particle_spacing_scale
    pmin, pmax = np.min(x, 0), np.max(x, 0)
    mask = args.boundary_padding_mask
    if len(mask) != 6 or not all(c in "01" for c in mask):
        raise ValueError(f"boundary_padding_mask must be 6-digit binary string, got: {mask}")
    f = [int(c) for c in mask]
    loose, tight = args.grid_padding * dx, MPM_PADDING * dx
    minB = np.array([pmin[0] - (loose if f[0] else tight), pmin[1] - (loose if f[2] else tight),
                     pmin[2] - (loose if f[4] else tight)])
    maxB = np.array([pmax[0] + (loose if f[1] else tight), pmax[1] + (loose if f[3] else tight),
                     pmax[2] + (loose if f[5] else tight)])
    dims = tuple(int(d) for d in np.ceil((maxB - minB) / dx).astype(int))
    return float(dx), minB, maxB, dims


def time_step(args, dx, E, nu, density):
    """CFL auto-dt when dt <= 0 and the integer XPBD ratio (runMPMYDW.py:110-169)."""
    E_mean, nu_mean, rho_mean = np.mean(E), np.mean(nu), np.mean(density)
    K = E_mean / (3 * (1 - 2 * nu_mean))
    G = E_mean / (2 * (1 + nu_mean))
    c_p = np.sqrt((K + 4 * G / 3) / rho_mean)
    dt = args.dt
    if dt <= 0:
        sf = 0.1
        if nu_mean > 0.45:
            sf *= np.exp(-10 * (nu_mean - 0.45))
        if E_mean < 1e9:
            sf *= np.clip((E_mean / 1e9) ** 0.3, 0.1, 1.0)
        sf = np.clip(sf, 0.005, 0.2)
        dt = sf * dx / c_p
    ratio = int(np.round(args.dtxpbd / dt))
    ratio = max(ratio, 1)
    return float(args.dtxpbd / ratio), ratio


@dataclass
class Simulation:
    args: object
    sim: SimState
    mpm: MPMState
    xpbd: XPBDState
    geom: Geometry
    fused: bool = True
    log: list = field(default_factory=list)

    # ---- one substep of the combined phase (runMPMYDW.py:677-682) ----
    def substep(self, counter: int) -> None:
        SR = simulationRoutines
        (SR.mpmSimulationStepFused if self.fused else SR.mpmSimulationStep)(self.sim, self.mpm, self.xpbd)
        if counter % self.sim.mpmStepsPerXpbdStep == 0 and counter > 0:
            (SR.xpbdSimulationStepFused if self.fused else SR.xpbdSimulationStep)(self.sim, self.mpm, self.xpbd)

    def reorder(self) -> None:
        """Permute every per-particle array in HBM into the fused substep's tile
        order (tile by tile, cell by cell, input index ascending inside a cell),
        so that each tile's particles are contiguous and the substep's loads
        coalesce.  Active XPBD particles follow in the same tile order (the
        contact walk gathers the particles of neighbouring cells), then the
        rest in input order.  ``sim.particle_id`` holds each slot's input index; the bins
        and hash tables order particles by it, so a reordered run gives the same
        bits as one that never reorders (the FP64 convergence sums aside, whose
        order follows the slots), and the host views (numpy, assign) keep the
        input order: the permutation is invisible outside the device."""
        SR = simulationRoutines
        sim, mpm, xpbd = self.sim, self.mpm, self.xpbd
        n = sim.nPoints
        P, G, prm = SR.step_structs(sim, mpm, xpbd)
        ws, nb = SR.substep_workspace(sim, mpm)
        perm = rt.zeros(n, rt.int32)
        lib("mpmx_reorder_permutation", C.byref(P), C.byref(G), C.byref(prm), perm.ptr, ws, nb, rt.stream_handle())
        seen = set()
        for obj in (sim, mpm, xpbd):
            for name, a in list(vars(obj).items()):
                if (not isinstance(a, rt.Array) or name.startswith("_") or id(a) in seen or a.dtype == rt.float64
                        or len(a.shape) != 1):
                    continue
                if a.shape[0] == n:
                    a.permute_(perm, 1, None if a is sim.particle_id else sim.particle_id)
                elif a.shape[0] == 6 * n:  # the covariance arrays hold 6 floats per particle
                    a.permute_(perm, 6, sim.particle_id)
                else:
                    continue
                seen.add(id(a))
        # array index of every input index, for the reference's index-based tests
        # in the contact solve (not a per-particle array itself: reorder skips it)
        if getattr(sim, "_slot_of_id", None) is None:
            sim._slot_of_id = rt.zeros(n, rt.int32)
        lib("mpmx_invert_permutation", sim.particle_id.ptr, sim._slot_of_id.ptr, n, rt.stream_handle())
        mpm._bins = None

    def restore_input_order(self) -> None:
        """Undo reorder(): permute every per-particle array back to input order
        (slot i holds input index i again)."""
        SR = simulationRoutines
        sim, mpm, xpbd = self.sim, self.mpm, self.xpbd
        n = sim.nPoints
        back = rt.zeros(n, rt.int32)  # slot of every input index
        lib("mpmx_invert_permutation", sim.particle_id.ptr, back.ptr, n, rt.stream_handle())
        seen = set()
        for obj in (sim, mpm, xpbd):
            for name, a in list(vars(obj).items()):
                if (not isinstance(a, rt.Array) or name.startswith("_") or id(a) in seen or a.dtype == rt.float64
                        or len(a.shape) != 1):
                    continue
                if a.shape[0] == n:
                    a.permute_(back, 1)
                elif a.shape[0] == 6 * n:
                    a.permute_(back, 6)
                else:
                    continue
                seen.add(id(a))
        sim._slot_of_id = None
        mpm._bins = None
        SR.mpmSimulationBin(sim, mpm, xpbd)

    # ---- what the loop reads back (runMPMYDW.py:686-722) ----
    def loop_scalars(self) -> tuple[float, float]:
        """(velocity residual, mean damage) after a substep: sum |v|/r over
        active particles / active count (runMPMYDW.py:686-708) and the mean of
        particle_damage over active particles (compute_mean_gpu, filtered),
        both FP64 on the device, in one readback."""
        import torch
        sim = self.sim
        if getattr(self, "_scal", None) is None:
            dev = rt.device()
            self._scal = (torch.zeros(2, dtype=torch.float64, device=dev), torch.zeros(2, dtype=torch.int32,
                                                                                        device=dev),
                          rt.Workspace())
        s, c, ws_ = self._scal
        n = sim.nPoints
        ws, nb = ws_.get(lib.raw("mpmx_reduce_workspace_bytes")(int(n)))
        p = lambda t, k: C.c_void_p(t.data_ptr() + k * t.element_size())  # noqa: E731
        lib("mpmx_velocity_convergence", n, sim.activeLabel.ptr, sim.particle_v.ptr, sim.particle_radius.ptr,
            p(s, 0), p(c, 0), ws, nb, rt.stream_handle())
        lib("mpmx_sum_float_array_active", n, self.mpm.particle_damage.ptr, sim.activeLabel.ptr, p(s, 1), p(c, 1), ws,
            nb, rt.stream_handle())
        h = torch.cat([s, c.to(torch.float64)]).cpu().numpy()
        residual = float(h[0] / h[2]) if h[2] > 0 else float("inf")
        damage = float(h[1] / h[3]) if h[3] > 0 else 0.0
        return residual, damage

    def residual(self) -> float:
        return self.loop_scalars()[0]

    def mean_damage(self) -> float:
        return simulationRoutines.compute_mean_gpu(self.mpm.particle_damage, self.sim)

    def xpbd_only_step(self) -> None:
        """One step of the XPBD-only phase (runMPMYDW.py:784): MPM frozen."""
        SR = simulationRoutines
        (SR.xpbdSimulationStepFused if self.fused else SR.xpbdSimulationStep)(self.sim, self.mpm, self.xpbd,
                                                                               xpbd_only=True)

    def sleep_counts(self) -> tuple[int, int]:
        """(XPBD particles, sleeping ones) from the last XPBD step's sleepParticles."""
        return int(self.sim.numTotalXPBD.numpy()[0]), int(self.sim.numSleepingXPBD.numpy()[0])

    def deactivate_below_datum(self) -> None:
        """End of a big step (runMPMYDW.py:849-877): XPBD particles below the datum
        (xpbd_deactivation_z_datum, else the grid's lower z bound) are switched
        off and moved down by twice the domain height."""
        sim, g = self.sim, self.geom
        zd = getattr(self.args, "xpbd_deactivation_z_datum", None)
        z_datum = float(g.minBounds[2]) if zd is None else float(zd)
        teleport = 2.0 * float(g.maxBounds[2] - g.minBounds[2])
        rt.launch(mpmRoutines.deactivate_and_teleport_xpbd_below_datum, sim.nPoints,
                  inputs=[sim.particle_x, sim.particle_v, sim.activeLabel, sim.materialLabel, z_datum, teleport])
        self.mpm._bins = None  # activity changed

    # ---- the loop ----
    def run_phase1(self, max_steps: int | None = None, verbose: bool = False, base_gravity=None) -> tuple:
        return run_phase1(self, max_steps=max_steps, verbose=verbose, base_gravity=base_gravity)

    def run(self, start_big_step: int = 0, big_steps: int | None = None, max_phase1_steps: int | None = None,
            checkpoint=None, verbose: bool = False) -> list:
        return run(self, start_big_step=start_big_step, big_steps=big_steps, max_phase1_steps=max_phase1_steps,
                   checkpoint=checkpoint, verbose=verbose)


def apply_restart(S, restart: dict) -> int:
    """Restore a checkpoint (fs5PlotUtils.load_restart_from_vtp) into a state
    set up from the same domain, as runMPMYDW.py:496-576 does: time, positions,
    velocities, labels, F (and F_trial = F), C, accumulated strain, damage,
    stress, yield stresses, material and physical properties, radius and the
    XPBD state that the file holds.  Returns the big step to start from."""
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    n = sim.nPoints
    if len(restart["particle_x"]) != n:
        raise ValueError(f"Restart file has {len(restart['particle_x'])} particles, but domain has {n}. "
                         "Cannot restart with different particle count.")
    sim.t = float(restart["sim_time"])
    sim.particle_x.assign(restart["particle_x"])
    sim.particle_v.assign(restart["particle_v"])
    for key, arr in (("materialLabel", sim.materialLabel), ("activeLabel", sim.activeLabel)):
        if restart.get(key) is not None:
            arr.assign(restart[key])
    mpm.particle_F.assign(restart["particle_F"])
    mpm.particle_C.assign(restart["particle_C"])
    rt.launch(mpmRoutines.set_mat33_to_copy, n, inputs=[mpm.particle_F, mpm.particle_F_trial])
    for key, arr in (("particle_accumulated_strain", mpm.particle_accumulated_strain),
                     ("particle_damage", mpm.particle_damage), ("particle_stress", mpm.particle_stress),
                     ("ys", mpm.ys), ("ys_base", mpm.ys_base)):
        if restart.get(key) is not None:
            arr.assign(restart[key])
    for prop in ("mu", "lam", "bulk", "alpha", "hardening", "softening", "eta_shear", "eta_bulk", "strainCriteria"):
        if restart.get(prop) is not None:
            getattr(mpm, prop).assign(restart[prop])
    for key, arr in (("particle_vol", sim.particle_vol), ("particle_mass", sim.particle_mass),
                     ("particle_density", sim.particle_density), ("particle_radius", sim.particle_radius)):
        if restart.get(key) is not None:
            arr.assign(restart[key])
    if restart.get("particle_x_initial_xpbd") is not None:
        xpbd.particle_x_initial_xpbd.assign(restart["particle_x_initial_xpbd"])
    if restart.get("particle_cumDist_xpbd") is not None:
        xpbd.particle_cumDist_xpbd.assign(restart["particle_cumDist_xpbd"])
    mpm._bins = None
    return int(restart["bigStep"]) + 1


MIN_STEPS = 1000  # runMPMYDW.py:662


def run_phase1(S, max_steps: int | None = None, verbose: bool = False, base_gravity=None, after_step=None) -> tuple:
    """The combined MPM+XPBD phase of one big step (runMPMYDW.py:643-770) with
    its gravity pulse and termination rules; ``S`` is a Simulation (or any
    object with its substep / loop_scalars / mean_damage methods).  Returns
    (substeps, reason) with reason "nSteps", "residual" or "damage_stall".

    The loop runs while counter < nSteps and (counter < 1000 or residual >
    residualThreshold); after each substep the mean damage is compared with
    the previous one and a run of damage_stall_duration/dt substeps past the
    first 1000 with a change below damage_stall_threshold ends it early.  The
    gravity pulse (gravity_pulse_factor for gravity_pulse_duration) is set at
    the start and restored before the substep whose counter equals its length,
    or on an early stop inside it; a phase that ends otherwise inside the pulse
    keeps the pulsed gravity, as the reference does.  ``after_step(counter)``
    runs at the end of every substep that did not stop the loop early (where
    the reference renders and saves, :765-769)."""
    sim, args = S.sim, S.args
    base = tuple(sim.gravity) if base_gravity is None else tuple(base_gravity)
    factor = float(getattr(args, "gravity_pulse_factor", 1.0))
    dur = float(getattr(args, "gravity_pulse_duration", 0.0))
    pulse_steps = int(dur / sim.dt) if dur > 0 else 0
    pulsing = factor != 1.0 and pulse_steps > 0
    if pulsing:
        sim.gravity = tuple(g * factor for g in base)
    nSteps = sim.nSteps if max_steps is None else min(sim.nSteps, max_steps)
    counter = 0
    residual = 1e10
    prev = S.mean_damage()
    stall = 0
    stall_steps = int(sim.damage_stall_duration / sim.dt)
    reason = None
    t0 = time.time()
    while counter < nSteps and (counter < MIN_STEPS or residual > sim.residualThreshold):
        if pulsing and counter == pulse_steps:
            sim.gravity = base
        S.substep(counter)
        residual, dmg = S.loop_scalars()
        counter += 1
        sim.t += sim.dt
        if dmg - prev < sim.damage_stall_threshold and counter > MIN_STEPS:
            stall += 1
        else:
            stall = 0
        prev = dmg
        if stall >= stall_steps and counter > MIN_STEPS:
            if pulsing and counter < pulse_steps:
                sim.gravity = base
            reason = "damage_stall"
            break
        if verbose and counter % 100 == 0:
            print(f"  Step: {counter}, t: {sim.t:.4f}s, residual: {residual:.4e}, mean_damage: {dmg:.6f}, "
                  f"t_real: {time.time() - t0:.2f}s", flush=True)
        if after_step is not None:
            after_step(counter)
    if reason is None:
        reason = "nSteps" if counter >= nSteps else "residual"
    return counter, reason


def run_phase2(S, verbose: bool = False, after_step=None) -> tuple:
    """The XPBD-only debris phase (runMPMYDW.py:775-839): up to xpbdOnlySteps
    steps of dtxpbd, ending early once the sleeping fraction of the XPBD
    particles reaches xpbd_sleep_termination_ratio (when that is below 1).
    Returns (steps, reason) with reason "steps", "asleep" or "off";
    ``after_step(k)`` runs after every step that did not end the phase."""
    sim = S.sim
    if sim.xpbdOnlySteps <= 0:
        return 0, "off"
    steps = 0
    for k in range(1, sim.xpbdOnlySteps + 1):
        S.xpbd_only_step()
        sim.t += sim.dtxpbd
        steps = k
        if sim.xpbd_sleep_termination_ratio < 1.0:
            total, asleep = S.sleep_counts()
            if total > 0 and asleep / total >= sim.xpbd_sleep_termination_ratio:
                return steps, "asleep"
        if verbose and k % 100 == 0:
            print(f"    XPBD Step: {k}/{sim.xpbdOnlySteps}, t: {sim.t:.4f}s", flush=True)
        if after_step is not None:
            after_step(k)
    return steps, "steps"


def run(S, start_big_step: int = 0, big_steps: int | None = None, max_phase1_steps: int | None = None,
        checkpoint=None, verbose: bool = False) -> list:
    """The big-step loop of runMPMYDW.py:638-877: for each big step the combined
    phase, the XPBD-only phase, ``checkpoint(S, bigStep)`` (where the reference
    saves its restart file) and the deactivation below the datum.  Returns one
    record per big step.

    With sim.saveFlag set, snapshots follow check_render_save (:609-628): one
    whenever sim.t passes the next multiple of render_interval (within half a
    step), one forced at the end of the combined phase, and a checkpoint with
    the restart data at the end of every big step (fs5PlotUtils.save_mpm)."""
    sim = S.sim
    base = tuple(sim.gravity)  # the pulse scales the gravity the run started with
    end = sim.bigSteps if big_steps is None else min(sim.bigSteps, start_big_step + big_steps)
    saving = bool(getattr(sim, "saveFlag", 0))
    interval = float(getattr(S.args, "render_interval", 0.1)) if saving else 0.0
    state = {"next": sim.t + interval, "counter": 0, "big": start_big_step}

    def save(dt_tol, force=False, restart=False):
        if not saving:
            return
        if force or sim.t >= state["next"] - 0.5 * dt_tol:
            from . import fs5PlotUtils
            fs5PlotUtils.save_mpm(sim, S.mpm, state["big"], state["counter"], include_restart_data=restart,
                                  xpbd=S.xpbd if restart else None)
            if not force:
                state["next"] += interval

    def after_mpm(counter):
        state["counter"] = counter
        save(sim.dt)

    out = []
    for big in range(start_big_step, end):
        state["big"] = big
        n1, why1 = run_phase1(S, max_steps=max_phase1_steps, verbose=verbose, base_gravity=base,
                              after_step=after_mpm if saving else None)
        state["counter"] = n1
        save(sim.dt, force=True)

        def after_xpbd(k, n1=n1):
            state["counter"] = n1 + k
            save(sim.dtxpbd)

        n2, why2 = run_phase2(S, verbose=verbose, after_step=after_xpbd if saving else None)
        state["counter"] = n1 + n2
        save(sim.dt, force=True, restart=True)
        if checkpoint is not None:
            checkpoint(S, big)
        S.deactivate_below_datum()
        out.append({"bigStep": big, "phase1_steps": n1, "phase1_end": why1, "phase2_steps": n2, "phase2_end": why2,
                    "t": sim.t})
    return out


def setup(args, domain: dict, fused: bool = True) -> Simulation:
    """Build the device state from a domain dict with ``x`` (N,3) and
    ``particle_volume`` (N,), optional per-particle property arrays named as the
    HDF5 datasets (PROPERTY_NAMES), and optional ``materialLabel``/``radius``
    overrides (for scenes that start partly in the XPBD phase)."""
    x = np.asarray(domain["x"], dtype=np.float64)
    vol = np.asarray(domain["particle_volume"], dtype=np.float64)
    n = x.shape[0]
    dx, minB, maxB, dims = grid_geometry(x, vol, args)

    def prop(name, default):
        if name in domain and domain[name] is not None:
            return np.asarray(domain[name], dtype=np.float32)
        return np.full(n, default, dtype=np.float32)

    density = prop("density", args.density)
    E = prop("E", args.E)
    nu = prop("nu", args.nu)
    dt, ratio = time_step(args, dx, E, nu, density)
    args.dt = dt

    sim = SimState(args, n)
    mpm = MPMState(args, n, dims)
    xpbd = XPBDState(args, n)
    sim.minBounds = tuple(float(v) for v in minB)
    sim.maxBounds = tuple(float(v) for v in maxB)
    sim.dx = dx
    sim.invdx = 1.0 / dx
    mpm.boundaryPadding = MPM_PADDING

    yMod = rt.array(E)
    pr = rt.array(nu)
    mpm.mu = rt.zeros(n)
    mpm.lam = rt.zeros(n)
    mpm.bulk = rt.zeros(n)
    rt.launch(mpmRoutines.compute_mu_lam_bulk_from_E_nu, n, inputs=[yMod, pr], outputs=[mpm.mu, mpm.lam, mpm.bulk])
    ys = prop("ys", args.ys)
    mpm.alpha = rt.array(prop("alpha", args.alpha))
    mpm.strainCriteria = rt.array(prop("strainCriteria", args.strainCriteria))
    mpm.ys_base = rt.array(ys.copy())
    mpm.ys = rt.array(ys)
    mpm.hardening = rt.array(prop("hardening", args.hardening))
    mpm.softening = rt.array(prop("softening", args.softening))
    mpm.eta_shear = rt.array(prop("eta_shear", args.eta_shear))
    mpm.eta_bulk = rt.array(prop("eta_bulk", args.eta_bulk))

    labels = np.asarray(domain.get("materialLabel", np.ones(n)), dtype=np.int32)
    sim.materialLabel = rt.array(labels, rt.int32)
    sim.activeLabel = rt.array(np.ones(n, np.int32), rt.int32)
    sim.particle_density = rt.array(density)
    sim.particle_x = rt.array(x.astype(np.float32), rt.vec3)
    sim.particle_v = rt.zeros(n, rt.vec3)
    sim.particle_vol = rt.array(vol.astype(np.float32))
    sim.particle_mass = rt.zeros(n)
    sim.particleBCMask = rt.zeros(n, rt.int32)
    sim.particle_id = rt.array(np.arange(n, dtype=np.int32), rt.int32)  # input index (Simulation.reorder)
    rt.launch(mpmRoutines.get_float_array_product, n, inputs=[sim.particle_density, sim.particle_vol,
                                                              sim.particle_mass])

    # geostatic or identity F (runMPMYDW.py:407-445); z_top defaults to the highest particle
    z_top = float(np.max(x[:, 2])) if getattr(args, "z_top", None) is None else float(args.z_top)
    if args.initialise_geostatic:
        k = (mpmRoutines.initialize_geostatic_F_planestrain if args.boundaryCondition == "plane_strain_y_global"
             else mpmRoutines.initialize_geostatic_F)
        rt.launch(k, n, inputs=[sim.particle_x, mpm.particle_F, mpm.mu, mpm.lam, sim.particle_density,
                                abs(float(args.gravity)), z_top, mpm.K0])
    else:
        rt.launch(mpmRoutines.set_mat33_to_identity, n, inputs=[mpm.particle_F])
    rt.launch(mpmRoutines.set_mat33_to_copy, n, inputs=[mpm.particle_F, mpm.particle_F_trial])
    damage = prop("damage", 0.0)
    mpm.particle_damage = rt.array(damage)

    radius = np.asarray(vol ** (1 / 3) * np.pi / 6, dtype=np.float64)
    if "radius" in domain and domain["radius"] is not None:
        radius = np.asarray(domain["radius"], dtype=np.float64)
    sim.particle_radius = rt.array(radius.astype(np.float32))
    xpbd.particleBaseRadius = rt.array(radius.astype(np.float32))
    max_r = radius * (1 + xpbd.swellingRatio) ** (1 / 3)
    xpbd.particleMaxRadius = rt.array(max_r.astype(np.float32))
    xpbd.max_radius = float(np.max(max_r))
    pad = -xpbd.max_radius * 1.5 - (mpm.boundaryPadding - 1) * sim.dx
    xpbd.minBoundsXPBD = tuple(float(v - pad) for v in minB)
    xpbd.maxBoundsXPBD = tuple(float(v + pad) for v in maxB)

    # particles already in the XPBD phase carry the transition's zeroed material (mpmRoutines.py:404-407)
    xp = labels == 2
    if np.any(xp):
        for arr in (mpm.mu, mpm.lam, mpm.alpha, mpm.ys):
            h = arr.numpy()
            h[xp] = 0.0
            arr.assign(h)
    geom = Geometry(dx, 1.0 / dx, minB, maxB, dims, dt, ratio)
    return Simulation(args, sim, mpm, xpbd, geom, fused=fused)


def load_domain(path: str) -> dict:
    """A particle domain from the reference's HDF5 format (runMPMYDW.py:19-23,
    70-81) or from a .vtp holding the same arrays."""
    if path.lower().endswith(".vtp"):
        from .fs5PlotUtils import load_domain_vtp
        return load_domain_vtp(path)
    from .hdf5 import load_domain_h5
    return load_domain_h5(path)


def main(argv=None, verbose: bool = True) -> list:
    """runMPMYDW.py end to end: flags and JSON config, the domain file, setup,
    an optional restart (--restart a checkpoint .vtp), then the big-step loop
    with snapshots and checkpoints under outputFolder when saveFlag is set."""
    from . import getArgs
    args = getArgs.get_args(argv)
    S = setup(args, load_domain(args.domainFile))
    start = 0
    if args.restart is not None:
        from .fs5PlotUtils import load_restart_from_vtp
        start = apply_restart(S, load_restart_from_vtp(args.restart))
    return S.run(start_big_step=start, verbose=verbose)


if __name__ == "__main__":
    import sys
    for rec in main(sys.argv[1:]):
        print(rec)
