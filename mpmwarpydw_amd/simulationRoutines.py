"""Substep orchestration of the hot path: the call surface of the reference's
utils/simulationRoutines.py (mpmSimulationStep :6, xpbdSimulationStep :295,
reductions :530-675) driving libmpmx's HIP kernels.

``mpmSimulationStep`` issues the reference's launches in the reference's
order.  ``mpmSimulationStepFused`` computes the same substep with the grid
update folded into the P2G combine pass (no grid zeroing, no separate
normalisation/boundary launches); it is what the driver and the benchmark run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import mpmRoutines, xpbdRoutines
from . import runtime as rt
from ._lib import MpmxError, XpbdBounds, lib
from .runtime import Kernel, Workspace, launch, stream_handle

def _bc_code(mpm):
    """The MPMX_BC_* code of mpm.boundaryCondition (simulationRoutines.py:124-228).
    An unknown name raises: the reference silently applies no boundary then."""
    if mpm.boundaryCondition not in mpmRoutines.BC_CODES:
        raise MpmxError(f"boundaryCondition {mpm.boundaryCondition!r} is not one of "
                        f"{sorted(mpmRoutines.BC_CODES)} (simulationRoutines.py:124-228)")
    return mpmRoutines.BC_CODES[mpm.boundaryCondition]


def _collide(sim, mpm):
    """Step 6 of mpmSimulationStep (simulationRoutines.py:122-228): the boundary
    kernel that mpm.boundaryCondition names."""
    bc = mpm.boundaryCondition
    _bc_code(mpm)
    dims = list(mpm.gridDims)
    g = abs(sim.gravity[2])
    if bc == "friction":
        launch(kernel=mpmRoutines.collideBounds, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundFriction, sim.dt, g, mpm.boundaryPadding])
    elif bc == "friction_gradual":
        launch(kernel=mpmRoutines.collideBoundsGradualFriction, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundFriction, sim.dt, g, mpm.boundaryPadding])
    elif bc == "restitution":
        launch(kernel=mpmRoutines.collideBoundsRestitution, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundRestitution, mpm.boundFriction, sim.dt, g,
                       mpm.boundaryPadding])
    elif bc == "absorbing":
        launch(kernel=mpmRoutines.collideBoundsAbsorbing, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundaryPadding])
    elif bc == "plane_strain_y":
        launch(kernel=mpmRoutines.collideBoundsPlaneStrainY, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundFriction, sim.dt, g, mpm.boundaryPadding])
    else:
        launch(kernel=mpmRoutines.collideBoundsPlaneStrainYGlobal, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, *dims, mpm.boundFriction, sim.dt, g, mpm.boundaryPadding])


def volumetric_locking_phase(sim, mpm, xpbd=None):
    """Step 8 of mpmSimulationStep (simulationRoutines.py:252-293): the cell sums
    (both arrays are overwritten, which is the reference's zero_() + atomics)
    and the correction of C."""
    launch(kernel=mpmRoutines.accumulate_cell_divergence, dim=sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, mpm.particle_C, sim.particle_vol, sim.invdx,
                   sim.minBounds, mpm.cell_div_v_weighted, mpm.cell_vol_sum])
    launch(kernel=mpmRoutines.apply_volumetric_locking_correction, dim=sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, mpm.particle_C, sim.invdx, sim.minBounds,
                   mpm.cell_div_v_weighted, mpm.cell_vol_sum])


def _cov_flag(mpm):
    return bool(mpm.update_cov) and not getattr(mpm, "cov_is_zero", False)


def mpmSimulationStep(sim, mpm, xpbd, device=None):
    """One MPM substep in the reference's launch order (simulationRoutines.py:6-293)."""
    _bc_code(mpm)
    mpm.grid_m.zero_()
    mpm.grid_v_in.zero_()
    mpm.grid_v_out.zero_()
    launch(kernel=mpmRoutines.compute_stress_from_F_trial, dim=sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, xpbd.particle_x_initial_xpbd,
                   xpbd.particle_v_initial_xpbd, mpm.particle_F, mpm.particle_F_trial, mpm.mu, mpm.lam, mpm.ys,
                   mpm.hardening, mpm.softening, sim.particle_density, mpm.strainCriteria, mpm.eff, mpm.eta_shear,
                   mpm.eta_bulk, mpm.particle_C, mpm.particle_stress, mpm.particle_accumulated_strain,
                   mpm.particle_damage, mpm.constitutive_model, mpm.alpha])
    launch(kernel=mpmRoutines.p2g_apic_with_stress, dim=sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, mpm.particle_stress, sim.particle_x, sim.particle_v,
                   mpm.particle_C, sim.particle_vol, sim.particle_mass, sim.dx, sim.invdx, sim.minBounds,
                   mpm.rpic_damping, mpm.grid_m, mpm.grid_v_in, sim.dt])
    launch(kernel=mpmRoutines.grid_normalization_and_gravity, dim=mpm.gridDims,
           inputs=[mpm.grid_m, mpm.grid_v_in, mpm.grid_v_out, sim.gravity, sim.dt])
    if mpm.grid_v_damping_scale < 1.0:
        launch(kernel=mpmRoutines.add_damping_via_grid, dim=mpm.gridDims,
               inputs=[mpm.grid_v_out, mpm.grid_v_damping_scale])
    _collide(sim, mpm)
    launch(kernel=mpmRoutines.g2p, dim=sim.nPoints,
           inputs=[sim.dt, sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, mpm.particle_C,
                   mpm.particle_F, mpm.particle_F_trial, mpm.particle_cov, sim.invdx, mpm.grid_v_out,
                   _cov_flag(mpm), sim.minBounds])
    if mpm.volumetric_locking_correction:
        volumetric_locking_phase(sim, mpm)


def stress_phase(sim, mpm, xpbd):
    """compute_stress_from_F_trial over all particles (simulationRoutines.py:15-45)."""
    launch(kernel=mpmRoutines.compute_stress_from_F_trial, dim=sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, xpbd.particle_x_initial_xpbd,
                   xpbd.particle_v_initial_xpbd, mpm.particle_F, mpm.particle_F_trial, mpm.mu, mpm.lam, mpm.ys,
                   mpm.hardening, mpm.softening, sim.particle_density, mpm.strainCriteria, mpm.eff, mpm.eta_shear,
                   mpm.eta_bulk, mpm.particle_C, mpm.particle_stress, mpm.particle_accumulated_strain,
                   mpm.particle_damage, mpm.constitutive_model, mpm.alpha])


def p2g_grid_phase(sim, mpm, xpbd):
    """P2G (simulationRoutines.py:48-69) with the grid update of :94-140 in its combine pass."""
    bc = _bc_code(mpm)
    n = sim.nPoints
    gx, gy, gz = mpm.gridDims
    ws, nb = mpmRoutines.p2g_workspace(n, mpm.gridDims)
    g = sim.gravity
    lib("mpmx_p2g_grid_update", n, sim.activeLabel.ptr, sim.materialLabel.ptr, mpm.particle_stress.ptr,
        sim.particle_x.ptr, sim.particle_v.ptr, mpm.particle_C.ptr, sim.particle_vol.ptr, sim.particle_mass.ptr,
        float(sim.dx), float(sim.invdx), *[float(v) for v in sim.minBounds], float(mpm.rpic_damping), gx, gy, gz,
        mpm.grid_m.ptr, mpm.grid_v_in.ptr, mpm.grid_v_out.ptr, float(g[0]), float(g[1]), float(g[2]), float(sim.dt),
        float(mpm.grid_v_damping_scale), bc, float(mpm.boundFriction), float(abs(g[2])), int(mpm.boundaryPadding),
        float(mpm.boundRestitution), ws, nb, stream_handle())


def g2p_phase(sim, mpm, xpbd):
    """g2p over all particles (simulationRoutines.py:231-250)."""
    launch(kernel=mpmRoutines.g2p, dim=sim.nPoints,
           inputs=[sim.dt, sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, mpm.particle_C,
                   mpm.particle_F, mpm.particle_F_trial, mpm.particle_cov, sim.invdx, mpm.grid_v_out,
                   _cov_flag(mpm), sim.minBounds])


def mpmSimulationStepPhased(sim, mpm, xpbd, device=None):
    """The substep as three calls: stress, P2G with the grid update, G2P (and the
    volumetric locking correction when it is on)."""
    stress_phase(sim, mpm, xpbd)
    p2g_grid_phase(sim, mpm, xpbd)
    g2p_phase(sim, mpm, xpbd)
    if mpm.volumetric_locking_correction:
        volumetric_locking_phase(sim, mpm)


def _aptr(a):
    return None if a is None else a.t.data_ptr()


def step_structs(sim, mpm, xpbd):
    """The mpmx_particles / mpmx_grid / mpmx_step_params views of the state."""
    from ._lib import Grid, Particles, StepParams
    P = Particles(n=sim.nPoints, activeLabel=_aptr(sim.activeLabel), materialLabel=_aptr(sim.materialLabel),
                  x=_aptr(sim.particle_x), v=_aptr(sim.particle_v), x_init_xpbd=_aptr(xpbd.particle_x_initial_xpbd),
                  v_init_xpbd=_aptr(xpbd.particle_v_initial_xpbd), F=_aptr(mpm.particle_F),
                  F_trial=_aptr(mpm.particle_F_trial), C=_aptr(mpm.particle_C), stress=_aptr(mpm.particle_stress),
                  cov=_aptr(mpm.particle_cov), mu=_aptr(mpm.mu), lam=_aptr(mpm.lam), ys=_aptr(mpm.ys),
                  alpha=_aptr(mpm.alpha), hardening=_aptr(mpm.hardening), softening=_aptr(mpm.softening),
                  density=_aptr(sim.particle_density), strainCriteria=_aptr(mpm.strainCriteria),
                  eta_shear=_aptr(mpm.eta_shear), eta_bulk=_aptr(mpm.eta_bulk),
                  acc_strain=_aptr(mpm.particle_accumulated_strain), damage=_aptr(mpm.particle_damage),
                  vol=_aptr(sim.particle_vol), mass=_aptr(sim.particle_mass),
                  order=_aptr(getattr(sim, "particle_id", None)))
    gx, gy, gz = mpm.gridDims
    G = Grid(gx=gx, gy=gy, gz=gz, m=_aptr(mpm.grid_m), v_in=_aptr(mpm.grid_v_in), v_out=_aptr(mpm.grid_v_out))
    g = sim.gravity
    mb = sim.minBounds
    prm = StepParams(dt=sim.dt, dx=sim.dx, inv_dx=sim.invdx, minx=mb[0], miny=mb[1], minz=mb[2], gravity_x=g[0],
                     gravity_y=g[1], gravity_z=g[2], rpic_damping=mpm.rpic_damping, eff=mpm.eff,
                     damping_scale=mpm.grid_v_damping_scale, friction=mpm.boundFriction,
                     gravity_magnitude=abs(g[2]), bc=_bc_code(mpm), boundary_padding=int(mpm.boundaryPadding),
                     update_cov=int(_cov_flag(mpm)), constitutive_model=int(mpm.constitutive_model),
                     restitution=float(mpm.boundRestitution), vlc=int(bool(mpm.volumetric_locking_correction)))
    return P, G, prm


def _bin_key(sim, mpm):
    return (sim.nPoints, mpm.gridDims, id(sim.particle_x), sim.particle_x.version, id(sim.materialLabel),
            sim.materialLabel.version, id(sim.activeLabel), sim.activeLabel.version, tuple(sim.minBounds), sim.invdx)


def substep_workspace(sim, mpm):
    if getattr(mpm, "_substep_ws", None) is None:
        mpm._substep_ws = Workspace()
        mpm._bins = None
    gx, gy, gz = mpm.gridDims
    before = mpm._substep_ws.t
    ws, nb = mpm._substep_ws.get(lib.raw("mpmx_substep_workspace_bytes")(sim.nPoints, gx, gy, gz))
    if mpm._substep_ws.t is not before:  # new storage: the bins it held are gone
        mpm._bins = None
    return ws, nb


def mpmSimulationStepFused(sim, mpm, xpbd, device=None, *, learned=None, tol=1e-4, counters=None):
    """mpmSimulationStep (simulationRoutines.py:6-293) as libmpmx's fused substep:
    a tile pass (constitutive update + deterministic P2G), a node pass (grid
    update) and a tile pass (G2P, which also bins the particles for the next
    substep), then the volumetric locking correction when it is on.  The bins
    are rebuilt first whenever x, the labels or the activity were written from
    the host since the last substep.  With ``learned`` (a
    learned.LearnedReturnMap) the constitutive update uses the network's
    projection, judged by the exact map at ``tol``; ``counters`` (int64 device
    tensor of 3) accumulates {plastic, accepted, updated}."""
    P, G, prm = step_structs(sim, mpm, xpbd)
    ws, nb = substep_workspace(sim, mpm)
    key = _bin_key(sim, mpm)
    if mpm._bins != key:
        lib("mpmx_substep_bin", C.byref(P), C.byref(G), C.byref(prm), ws, nb, stream_handle())
    if learned is None:
        lib("mpmx_mpm_substep", C.byref(P), C.byref(G), C.byref(prm), ws, nb, stream_handle())
    else:
        cp = None if counters is None else C.c_void_p(counters.data_ptr())
        lib("mpmx_mpm_substep_learned", C.byref(P), C.byref(G), C.byref(prm), C.c_void_p(learned.params.data_ptr()),
            float(tol), cp, ws, nb, stream_handle())
    mpm._bins = key


def mpmSimulationBin(sim, mpm, xpbd, device=None):
    """Bin the current state for the fused substep (what its first call does
    for a fresh state), e.g. after the device arrays were overwritten in place."""
    P, G, prm = step_structs(sim, mpm, xpbd)
    ws, nb = substep_workspace(sim, mpm)
    lib("mpmx_substep_bin", C.byref(P), C.byref(G), C.byref(prm), ws, nb, stream_handle())
    mpm._bins = _bin_key(sim, mpm)


def misbinned_count(sim, mpm):
    """Particles found outside their bin since the last rebin (must be 0)."""
    ws, nb = substep_workspace(sim, mpm)
    out = C.c_int(-1)
    gx, gy, gz = mpm.gridDims
    lib("mpmx_substep_misbinned", ws, nb, gx, gy, gz, sim.nPoints, C.byref(out), stream_handle())
    return out.value


def xpbdSimulationStep(sim, mpm, xpbd, device=None, xpbd_only=False, split_grid=True):
    """The XPBD contact step (simulationRoutines.py:295-526), launch for launch.

    With ``split_grid`` (default) a second table of the label-2 points is built
    next to the full one and the particle-contact solve lets MPM particles walk
    it (same candidates in the same order, so the same results; see
    mpmx_solve_particle_particle_contacts_split)."""
    n = sim.nPoints
    order = getattr(sim, "particle_id", None)
    xpbd.particle_grid.build(sim.particle_x, sim.dx, order=order)
    if split_grid:
        xpbd.particle_grid_xpbd.build(sim.particle_x, sim.dx, labels=sim.materialLabel, label=2, order=order)
    rt.copy(xpbd.particle_x_integrated, sim.particle_x)
    rt.copy(xpbd.particle_v_integrated, sim.particle_v)
    launch(kernel=xpbdRoutines.integrateParticlesXPBD, dim=n,
           inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_x_integrated, xpbd.particle_v_integrated,
                   sim.gravity, sim.dtxpbd, xpbd.particle_x_integrated, xpbd.particle_v_integrated,
                   xpbd.particle_v_max])
    if xpbd.mpm_contact_transition_lock and not xpbd_only:
        launch(kernel=xpbdRoutines.reset_mpm_contact_labels, dim=n, inputs=[sim.activeLabel, sim.materialLabel])
    contact_lock = xpbd.mpm_contact_transition_lock if not xpbd_only else 0
    for _ in range(xpbd.xpbd_iterations):
        xpbd.particle_delta.zero_()
        launch(kernel=xpbdRoutines.my_solve_particle_bound_contacts, dim=n,
               inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_x_integrated, xpbd.particle_v_integrated,
                       sim.particle_mass, sim.particle_radius, xpbd.dynamicParticleFriction,
                       xpbd.staticVelocityThreshold, xpbd.staticParticleFriction, xpbd.minBoundsXPBD,
                       xpbd.maxBoundsXPBD, sim.dtxpbd, xpbd.xpbd_relaxation],
               outputs=[xpbd.particle_delta])
        contact_args = [xpbd.particle_x_integrated, xpbd.particle_v_integrated, sim.particle_mass,
                        sim.particle_radius, xpbd.dynamicParticleFriction, xpbd.staticVelocityThreshold,
                        xpbd.staticParticleFriction, xpbd.particle_cohesion, xpbd.max_radius, sim.dtxpbd,
                        xpbd.xpbd_relaxation, contact_lock]
        if split_grid:
            launch(kernel=xpbdRoutines.solve_particle_particle_contacts_split, dim=n,
                   inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_grid.id, xpbd.particle_grid_xpbd.id,
                           *contact_args], outputs=[xpbd.particle_delta, getattr(sim, "_slot_of_id", None)])
        else:
            launch(kernel=xpbdRoutines.my_solve_particle_particle_contacts, dim=n,
                   inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_grid.id, *contact_args],
                   outputs=[xpbd.particle_delta])
        rt.copy(xpbd.particle_v_deltaInt, xpbd.particle_v_integrated)
        rt.copy(xpbd.particle_x_deltaInt, xpbd.particle_x_integrated)
        launch(kernel=xpbdRoutines.my_apply_particle_deltas, dim=n,
               inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, xpbd.particle_x_integrated,
                       xpbd.particle_delta, sim.dtxpbd, xpbd.particle_v_max],
               outputs=[xpbd.particle_x_deltaInt, xpbd.particle_v_deltaInt])
        rt.copy(xpbd.particle_v_integrated, xpbd.particle_v_deltaInt)
        rt.copy(xpbd.particle_x_integrated, xpbd.particle_x_deltaInt)
    sim.numTotalXPBD.zero_()
    sim.numSleepingXPBD.zero_()
    launch(kernel=xpbdRoutines.sleepParticles, dim=n,
           inputs=[sim.activeLabel, sim.materialLabel, xpbd.sleepThreshold, sim.particle_radius, sim.particle_x,
                   xpbd.particle_x_integrated, xpbd.particle_cumDist_xpbd, sim.dtxpbd, sim.numTotalXPBD,
                   sim.numSleepingXPBD])
    rt.copy(xpbd.particle_v_before_contact, sim.particle_v)
    launch(kernel=xpbdRoutines.apply_mpm_contact_impulses, dim=n,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, xpbd.particle_x_integrated,
                   xpbd.particle_v_integrated, sim.dtxpbd, xpbd.particle_v_max, xpbd.xpbd_mpm_coupling_strength,
                   1 if xpbd_only else 0],
           outputs=[xpbd.particle_v_integrated])
    sim.particle_x.assign(xpbd.particle_x_integrated)
    sim.particle_v.assign(xpbd.particle_v_integrated)
    if xpbd.swellingRatio > 0:
        launch(kernel=xpbdRoutines.swellParticlesType2, dim=n,
               inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_x_initial_xpbd, sim.particle_x,
                       sim.particle_radius, xpbd.swellingActivationFactor, xpbd.swellingMaxFactor,
                       xpbd.particleMaxRadius, xpbd.particleBaseRadius])


def _contact_records(xpbd, n):
    """The candidate records of the packed contact solve in table order (2n
    records, the layout of mpmx_xpbd_contact_pack), kept on the XPBD state
    between steps; begin and each delta application write them."""
    rec = getattr(xpbd, "_contact_records", None)
    if rec is None or rec[2].shape[0] != 2 * n:
        rec = (rt.empty(8 * n, rt.float32), rt.empty(8 * n, rt.float32), rt.empty(2 * n, rt.int32))
        xpbd._contact_records = rec
    return rec


def _contact_list(xpbd, n):
    """The fused step's contact list (mpmx_xpbd_contact_list) and its count."""
    cl = getattr(xpbd, "_contact_list", None)
    if cl is None or cl[0].shape[0] != n:
        cl = (rt.empty(n, rt.int32), rt.zeros(1, rt.int32))
        xpbd._contact_list = cl
    return cl


def xpbdSimulationStepFused(sim, mpm, xpbd, device=None, xpbd_only=False):
    """xpbdSimulationStep with the copies folded into its kernels: x, v, the
    labels, particle_cumDist_xpbd and the sleep counters end exactly as after
    the launch sequence; the scratch arrays particle_x_deltaInt,
    particle_v_deltaInt and particle_v_before_contact are not written (nothing
    outside the step reads them)."""
    n = sim.nPoints
    order = getattr(sim, "particle_id", None)
    # lazy: with no XPBD particle the contact solve reads neither table, and the full one is not built
    xpbd.particle_grid.build_pair(xpbd.particle_grid_xpbd, sim.particle_x, sim.dx, sim.materialLabel, 2, order,
                                  lazy=True)
    g = sim.gravity
    s = stream_handle()
    reset = 1 if (xpbd.mpm_contact_transition_lock and not xpbd_only) else 0
    xi, vi = xpbd.particle_x_integrated, xpbd.particle_v_integrated
    mn, mx = xpbd.minBoundsXPBD, xpbd.maxBoundsXPBD
    iters = int(xpbd.xpbd_iterations)
    # each iteration's particle-wall deltas are folded into the kernel before it (begin, or
    # the previous iteration's delta application) and the last application into finish
    bounds = XpbdBounds(mass=sim.particle_mass.ptr.value, radius=sim.particle_radius.ptr.value,
                        mu=float(xpbd.dynamicParticleFriction), static_threshold=float(xpbd.staticVelocityThreshold),
                        static_friction=float(xpbd.staticParticleFriction), minx=float(mn[0]), miny=float(mn[1]),
                        minz=float(mn[2]), maxx=float(mx[0]), maxy=float(mx[1]), maxz=float(mx[2]),
                        dt=float(sim.dtxpbd), relaxation=float(xpbd.xpbd_relaxation),
                        delta=xpbd.particle_delta.ptr.value)
    grid, xgrid = xpbd.particle_grid, xpbd.particle_grid_xpbd
    near = xgrid.near_cells()  # MPM particles with no XPBD point beside their cell skip their walk
    xr, vw, flags = _contact_records(xpbd, n)
    clist, ccount = _contact_list(xpbd, n)
    bounds.xr, bounds.vw, bounds.flags = xr.ptr.value, vw.ptr.value, flags.ptr.value
    bounds.pos, bounds.xpbd_pos = grid.point_pos.ptr.value, xgrid.point_pos.ptr.value
    bounds.xpbd_count = xgrid.cell_start.ptr.value + 4 * int(np.prod(xgrid.dims))  # no records without XPBD points
    bref = C.byref(bounds)
    dims = grid.dims
    lib("mpmx_xpbd_contact_list", n, sim.materialLabel.ptr, grid.point_pos.ptr, xgrid.cell_start.ptr,
        float(grid.cell_width), *xgrid.dims, sim.particle_x.ptr, sim.particle_radius.ptr,
        float(xpbd.particle_cohesion), float(xpbd.max_radius), near.ptr, clist.ptr, ccount.ptr, s)
    lib("mpmx_xpbd_begin", n, sim.activeLabel.ptr, sim.materialLabel.ptr, sim.particle_x.ptr, sim.particle_v.ptr,
        float(g[0]), float(g[1]), float(g[2]), float(sim.dtxpbd), float(xpbd.particle_v_max), reset, xi.ptr, vi.ptr,
        sim.numTotalXPBD.ptr, bref if iters > 0 else None, s)  # also zeroes numTotalXPBD and numSleepingXPBD
    contact_lock = xpbd.mpm_contact_transition_lock if not xpbd_only else 0
    dims = grid.dims
    for it in range(iters):
        # the records are this iteration's: begin or the last delta application wrote them
        lib("mpmx_solve_particle_particle_contacts_packed", n, sim.activeLabel.ptr, sim.materialLabel.ptr,
            grid.point_ids.ptr, grid.cell_start.ptr, xgrid.point_ids.ptr, xgrid.cell_start.ptr,
            float(grid.cell_width), dims[0], dims[1], dims[2], *xgrid.dims, xi.ptr, vi.ptr, sim.particle_mass.ptr,
            sim.particle_radius.ptr, xr.ptr, vw.ptr, flags.ptr, float(xpbd.dynamicParticleFriction),
            float(xpbd.staticVelocityThreshold), float(xpbd.staticParticleFriction), float(xpbd.particle_cohesion),
            float(xpbd.max_radius), float(sim.dtxpbd), float(xpbd.xpbd_relaxation), int(contact_lock),
            xpbd.particle_delta.ptr, _aptr(getattr(sim, "_slot_of_id", None)), near.ptr, clist.ptr, ccount.ptr, s)
        if it + 1 < iters:
            lib("mpmx_xpbd_apply_deltas_inplace", n, sim.activeLabel.ptr, sim.particle_x.ptr, xpbd.particle_delta.ptr,
                float(sim.dtxpbd), float(xpbd.particle_v_max), xi.ptr, vi.ptr, sim.materialLabel.ptr, bref, s)
    lib("mpmx_xpbd_finish", n, sim.activeLabel.ptr, sim.materialLabel.ptr, float(xpbd.sleepThreshold),
        sim.particle_radius.ptr, xpbd.particle_cumDist_xpbd.ptr, float(sim.dtxpbd), float(xpbd.particle_v_max),
        float(xpbd.xpbd_mpm_coupling_strength), 1 if xpbd_only else 0, sim.particle_x.ptr, sim.particle_v.ptr,
        xi.ptr, vi.ptr, sim.numTotalXPBD.ptr, xpbd.particle_delta.ptr if iters > 0 else None, s)
    if xpbd.swellingRatio > 0:
        launch(kernel=xpbdRoutines.swellParticlesType2, dim=n,
               inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_x_initial_xpbd, sim.particle_x,
                       sim.particle_radius, xpbd.swellingActivationFactor, xpbd.swellingMaxFactor,
                       xpbd.particleMaxRadius, xpbd.particleBaseRadius])


# ---------------------------------------------------------------------------
# reductions (simulationRoutines.py:530-675), FP64 on the device
# ---------------------------------------------------------------------------
_red_ws = Workspace()
_red_out = {}


def _red_buffers():
    if not _red_out:
        _red_out["sum"] = torch.zeros(1, dtype=torch.float64, device=rt.device())
        _red_out["cnt"] = torch.zeros(4, dtype=torch.int32, device=rt.device())
    return _red_out["sum"], _red_out["cnt"]


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def device_velocity_convergence(n, activeLabel, particle_v, particle_radius):
    """(sum over active of |v|/r as a float64 device scalar, active count as an int32 device scalar)."""
    s, c = _red_buffers()
    ws, nb = _red_ws.get(lib.raw("mpmx_reduce_workspace_bytes")(int(n)))
    lib("mpmx_velocity_convergence", int(n), activeLabel.ptr, particle_v.ptr, particle_radius.ptr, _ptr(s), _ptr(c),
        ws, nb, stream_handle())
    return s, c[:1]


def _velocityConvergence(dim, activeLabel, materialLabel, particle_v, particle_radius, residual):
    s, _ = device_velocity_convergence(dim, activeLabel, particle_v, particle_radius)
    residual.t.add_(s.to(residual.t.dtype))


velocityConvergence = Kernel("velocityConvergence", _velocityConvergence, "simulationRoutines.py:538-549")


def _countActiveParticles(dim, activeLabel, out):
    _, c = device_velocity_convergence_count(dim, activeLabel)
    out.t.add_(c.to(out.t.dtype))


def device_velocity_convergence_count(n, activeLabel):
    s, c = _red_buffers()
    ws, nb = _red_ws.get(lib.raw("mpmx_reduce_workspace_bytes")(int(n)))
    lib("mpmx_count_particles_by_type", int(n), activeLabel.ptr, activeLabel.ptr, _ptr(c), ws, nb, stream_handle())
    # with materialLabel := activeLabel every particle counts as "MPM": slot 0 = active
    return s, c[:1]


countActiveParticles = Kernel("countActiveParticles", _countActiveParticles, "simulationRoutines.py:551-558")


def _countParticlesByType(dim, activeLabel, materialLabel, numActiveMPM, numActiveXPBD, numInactiveMPM,
                          numInactiveXPBD):
    _, c = _red_buffers()
    ws, nb = _red_ws.get(lib.raw("mpmx_reduce_workspace_bytes")(int(dim)))
    lib("mpmx_count_particles_by_type", int(dim), activeLabel.ptr, materialLabel.ptr, _ptr(c), ws, nb,
        stream_handle())
    for k, arr in enumerate((numActiveMPM, numActiveXPBD, numInactiveMPM, numInactiveXPBD)):
        arr.t.add_(c[k:k + 1].to(arr.t.dtype))


countParticlesByType = Kernel("countParticlesByType", _countParticlesByType, "simulationRoutines.py:560-585")


def device_sum_active(n, arr, activeLabel):
    s, c = _red_buffers()
    ws, nb = _red_ws.get(lib.raw("mpmx_reduce_workspace_bytes")(int(n)))
    lib("mpmx_sum_float_array_active", int(n), arr.ptr, activeLabel.ptr, _ptr(s), _ptr(c), ws, nb, stream_handle())
    return s, c[:1]


def compute_mean_gpu(arr, sim, device=None, filtered=True):
    """compute_mean_gpu (simulationRoutines.py:633-675): FP64 device sum, one readback."""
    if not filtered:
        ones = rt.Array(sim.nPoints, rt.int32, storage=torch.ones(sim.nPoints, dtype=torch.int32,
                                                                  device=rt.device()))
        s, c = device_sum_active(sim.nPoints, arr, ones)
        return float(s.item()) / sim.nPoints if sim.nPoints > 0 else 0.0
    s, c = device_sum_active(sim.nPoints, arr, sim.activeLabel)
    both = torch.cat([s, c.to(torch.float64)]).cpu().numpy()
    total, count = both[0], int(both[1])
    return total / count if count > 0 else 0.0
