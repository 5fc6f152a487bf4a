"""State containers of the hot path: the field names, shapes and initial
values of the reference's utils/simStates.py (SimState :10, MPMState :74,
XPBDState :133), stored as libmpmx device arrays (runtime.Array, SoA in HBM).

Differences in storage only (values and names are the reference's):
  * the per-substep reductions accumulate in FP64 on the device, so
    ``residual``/``scratchFloat`` are float64 one-element arrays;
  * the int counters are views into one int32 block so that the kernels that
    write several of them (sleepParticles, countParticlesByType) get them
    adjacent.
"""
from __future__ import annotations

import numpy as np
import torch

from . import runtime as rt


class SimState:
    """simStates.py:10-71"""

    def __init__(self, args, nPoints, device="cuda:0"):
        self.nPoints = int(nPoints)
        self.dt = args.dt
        self.dtxpbd = args.dtxpbd
        self.bigStepDuration = args.bigStepDuration
        self.xpbdOnlyDuration = args.xpbdOnlyDuration
        self.bigSteps = args.bigSteps
        self.nSteps = int(self.bigStepDuration / self.dt)
        self.xpbdOnlySteps = int(self.xpbdOnlyDuration / self.dtxpbd) if self.xpbdOnlyDuration > 0 else 0
        self.mpmStepsPerXpbdStep = int(self.dtxpbd / self.dt)
        self.t = 0.0
        self.residualThreshold = args.residualThreshold
        self.damage_stall_threshold = args.damage_stall_threshold
        self.damage_stall_duration = args.damage_stall_duration
        self.xpbd_sleep_termination_ratio = args.xpbd_sleep_termination_ratio
        self.render = args.render
        self.saveFlag = args.saveFlag
        self.color_mode = args.color_mode
        self.outputFolder = args.outputFolder
        self.gravity = (0.0, 0.0, float(args.gravity))

        self.particle_x = None
        self.particle_v = None
        self.particle_mass = None
        self.particle_vol = None
        self.particle_radius = None
        self.particle_density = None
        self.materialLabel = None
        self.activeLabel = None
        self.particleBCMask = None

        self.minBounds = None
        self.maxBounds = None
        self.dx = None
        self.invdx = None

        self.residual = rt.array(np.array([1e10]), rt.float64)
        block = torch.zeros(8, dtype=torch.int32, device=rt.device())
        self._counters = block

        def view(i, k=1):
            return rt.Array(k, rt.int32, storage=block[i:i + k])

        self.numActiveParticles = view(0)
        self.numTotalXPBD = view(1)      # {total, sleeping} adjacent for sleepParticles
        self.numSleepingXPBD = view(2)
        self.numActiveMPM = view(3)      # the four countParticlesByType outputs adjacent
        self.numActiveXPBD = view(4)
        self.numInactiveMPM = view(5)
        self.numInactiveXPBD = view(6)
        self.scratchInt = view(7)
        self.scratchFloat = rt.zeros(1, rt.float64)


class MPMState:
    """simStates.py:74-130"""

    def __init__(self, args, nPoints, gridDims, device="cuda:0"):
        self.device = device
        self.nPoints = int(nPoints)
        self.gridDims = tuple(int(g) for g in gridDims)
        self.rpic_damping = args.rpic_damping
        self.grid_v_damping_scale = args.grid_v_damping_scale
        self.update_cov = args.update_cov
        self.constitutive_model = args.constitutive_model
        self.K0 = args.K0
        self.boundaryCondition = args.boundaryCondition
        self.boundFriction = args.boundFriction
        self.boundRestitution = args.boundRestitution
        self.boundaryPadding = None

        self.mu = None
        self.lam = None
        self.bulk = None
        self.ys = None
        self.ys_base = None
        self.alpha = None
        self.hardening = None
        self.softening = None
        self.eta_shear = None
        self.eta_bulk = None
        self.strainCriteria = None

        n = self.nPoints
        self.particle_F = rt.zeros(n, rt.mat33)
        self.particle_F_trial = rt.zeros(n, rt.mat33)
        self.particle_stress = rt.zeros(n, rt.mat33)
        self.particle_accumulated_strain = rt.zeros(n, rt.float32)
        self.particle_damage = rt.zeros(n, rt.float32)
        self.particle_C = rt.zeros(n, rt.mat33)
        self.particle_init_cov = rt.zeros(n * 6, rt.float32)
        self.particle_cov = rt.zeros(n * 6, rt.float32)
        # particle_cov starts at zero and its update (mpmRoutines.py:779-802) is
        # homogeneous in cov, so it stays zero until someone stores a non-zero
        # value; the substep skips the update while this holds (same values).
        self.cov_is_zero = True

        self.grid_m = rt.zeros(self.gridDims, rt.float32)
        self.grid_v_in = rt.zeros(self.gridDims, rt.vec3)
        self.grid_v_out = rt.zeros(self.gridDims, rt.vec3)
        self.gridBCIndex = rt.zeros(self.gridDims, rt.int32)

        # volumetric locking correction (simStates.py:123-126): cell sums of the
        # reference's launch-by-launch step; the fused substep keeps its own in
        # its workspace
        self.volumetric_locking_correction = args.volumetric_locking_correction
        self.cell_div_v_weighted = rt.zeros(self.gridDims, rt.float32)
        self.cell_vol_sum = rt.zeros(self.gridDims, rt.float32)

        self.particle_x_initial = rt.zeros(n, rt.vec3)
        self.eff = args.eff


class XPBDState:
    """simStates.py:133-179"""

    def __init__(self, args, nPoints, device="cuda:0"):
        self.device = device
        self.nPoints = int(nPoints)
        self.xpbd_iterations = args.xpbd_iterations
        self.xpbd_relaxation = args.xpbd_relaxation
        self.xpbd_mpm_coupling_strength = args.xpbd_mpm_coupling_strength
        self.mpm_contact_transition_lock = args.mpm_contact_transition_lock
        self.dynamicParticleFriction = args.dynamicParticleFriction
        self.staticParticleFriction = args.staticParticleFriction
        self.staticVelocityThreshold = args.staticVelocityThreshold
        self.particle_cohesion = args.particle_cohesion
        self.sleepThreshold = args.sleepThreshold
        self.particle_v_max = args.particle_v_max
        self.swellingRatio = args.swellingRatio
        self.swellingActivationFactor = args.swellingActivationFactor
        self.swellingMaxFactor = args.swellingMaxFactor
        self.particleBaseRadius = None
        self.particleMaxRadius = None

        n = self.nPoints
        self.particle_x_initial_xpbd = rt.zeros(n, rt.vec3)
        self.particle_x_initial_xpbd.fill_([1e6, 1e6, 1e6])
        self.particle_v_initial_xpbd = rt.zeros(n, rt.vec3)
        self.particle_cumDist_xpbd = rt.zeros(n, rt.float32)
        self.particle_distance_total = rt.zeros(n, rt.float32)
        self.particle_x_integrated = rt.zeros(n, rt.vec3)
        self.particle_v_integrated = rt.zeros(n, rt.vec3)
        self.particle_x_deltaInt = rt.zeros(n, rt.vec3)
        self.particle_v_deltaInt = rt.zeros(n, rt.vec3)
        self.particle_delta = rt.zeros(n, rt.vec3)
        self.particle_v_before_contact = rt.zeros(n, rt.vec3)
        self.particle_grid = rt.HashGrid(128, 128, 128)
        # the label-2 points only, same cell width and cell count as the full table (the
        # MPM side of the contact solve walks it): with fewer cells, XPBD cells 64 cells
        # apart would wrap onto one another, and the neighbourhood bitmap would then list
        # the MPM particles beside every alias of an XPBD cell for a walk
        self.particle_grid_xpbd = rt.HashGrid(128, 128, 128)
        self.minBoundsXPBD = None
        self.maxBoundsXPBD = None
        self.max_radius = None
