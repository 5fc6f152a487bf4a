"""Setup and substep loop of the reference's runMPMYDW.py around the device
hot path.

``setup(args, domain)`` reproduces the driver's setup arithmetic
(runMPMYDW.py:22-67 grid spacing and bounds, :110-169 CFL dt and the integer
XPBD ratio, :278-312 material upload, Lame constants and mass, :413-470
geostatic/identity F, damage and radius, :472-495 the XPBD box) on a domain
given as arrays instead of an HDF5 file (h5py is not part of this build).

``Simulation.run_phase1`` is the combined MPM+XPBD loop of
runMPMYDW.py:638-770 with its termination rules (minSteps = 1000, velocity
residual, damage stall); its reductions are FP64 on the device.
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
        coalesce.  ``sim.particle_id`` holds each slot's input index; the bins
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

    def residual(self) -> float:
        """sum |v|/r / N_active over active particles (runMPMYDW.py:686-708): one readback."""
        s, c = simulationRoutines.device_velocity_convergence(self.sim.nPoints, self.sim.activeLabel,
                                                              self.sim.particle_v, self.sim.particle_radius)
        import torch
        both = torch.cat([s, c.to(torch.float64)]).cpu().numpy()
        return float(both[0] / both[1]) if both[1] > 0 else float("inf")

    def run_phase1(self, max_steps: int | None = None, verbose: bool = False) -> int:
        """The combined loop with the reference's termination rules; returns the substep count."""
        sim, mpm = self.sim, self.mpm
        nSteps = sim.nSteps if max_steps is None else min(sim.nSteps, max_steps)
        minSteps = 1000
        counter = 0
        residual = 1e10
        prev_damage = simulationRoutines.compute_mean_gpu(mpm.particle_damage, sim)
        stall = 0
        stall_steps = int(sim.damage_stall_duration / sim.dt)
        t0 = time.time()
        while counter < nSteps and (counter < minSteps or residual > sim.residualThreshold):
            self.substep(counter)
            residual = self.residual()
            counter += 1
            sim.t += sim.dt
            dmg = simulationRoutines.compute_mean_gpu(mpm.particle_damage, sim)
            if dmg - prev_damage < sim.damage_stall_threshold and counter > minSteps:
                stall += 1
            else:
                stall = 0
            prev_damage = dmg
            if stall >= stall_steps and counter > minSteps:
                break
            if verbose and counter % 100 == 0:
                print(f"  Step: {counter}, t: {sim.t:.4f}s, residual: {residual:.4e}, mean_damage: {dmg:.6f}, "
                      f"t_real: {time.time() - t0:.2f}s", flush=True)
        return counter


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

    if args.initialise_geostatic:
        raise NotImplementedError("initialise_geostatic=1 (mpmRoutines.py:51-124) is not on the device path yet; "
                                  "set it to 0")
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
