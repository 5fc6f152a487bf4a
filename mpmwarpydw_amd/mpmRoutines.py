"""MPM kernels of the hot path, with the names and argument lists of the
reference's utils/mpmRoutines.py.

Each object is launched exactly like the Warp kernel it replaces::

    runtime.launch(kernel=mpmRoutines.p2g_apic_with_stress, dim=sim.nPoints,
                   inputs=[sim.activeLabel, sim.materialLabel, mpm.particle_stress, ...])

and runs the matching libmpmx entry point on the GPU.  The docstring of each
object cites the reference kernel.
"""
from __future__ import annotations

import numpy as np

from ._lib import lib
from .runtime import Kernel, Workspace, stream_handle


def _v3(v):
    a = np.asarray(v, dtype=np.float64).reshape(3)
    return float(a[0]), float(a[1]), float(a[2])


def _compute_mu_lam_bulk_from_E_nu(dim, E, nu, mu, lam, bulk):
    lib("mpmx_compute_mu_lam_bulk", int(dim), E.ptr, nu.ptr, mu.ptr, lam.ptr, bulk.ptr, stream_handle())


compute_mu_lam_bulk_from_E_nu = Kernel("compute_mu_lam_bulk_from_E_nu", _compute_mu_lam_bulk_from_E_nu,
                                       "mpmRoutines.py:206-226")


def _geostatic(plane_strain):
    def fn(dim, particle_x, particle_F, mu, lam, density, gravity, z_top, K0):
        lib("mpmx_initialize_geostatic_F", int(dim), particle_x.ptr, particle_F.ptr, mu.ptr, lam.ptr, density.ptr,
            float(gravity), float(z_top), float(K0), int(plane_strain), stream_handle())
    return fn


initialize_geostatic_F = Kernel("initialize_geostatic_F", _geostatic(False), "mpmRoutines.py:52-124")
initialize_geostatic_F_planestrain = Kernel("initialize_geostatic_F_planestrain", _geostatic(True),
                                            "mpmRoutines.py:128-203")


def _deactivate_and_teleport_xpbd_below_datum(dim, particle_x, particle_v, activeLabel, materialLabel, z_datum,
                                              teleport_distance):
    lib("mpmx_deactivate_and_teleport_xpbd_below_datum", int(dim), particle_x.ptr, particle_v.ptr, activeLabel.ptr,
        materialLabel.ptr, float(z_datum), float(teleport_distance), stream_handle())


deactivate_and_teleport_xpbd_below_datum = Kernel("deactivate_and_teleport_xpbd_below_datum",
                                                  _deactivate_and_teleport_xpbd_below_datum,
                                                  "mpmRoutines.py:1446-1475")


def _compute_stress_from_F_trial(dim, activeLabel, materialLabel, particlePosition, particleVelocity,
                                 initialPhaseChangePosition, initialPhaseChangeVelocity, particle_F, particle_F_trial,
                                 mu, lam, yield_stress, hardening, softening, particle_density, strainCriteria,
                                 efficiency, eta_shear, eta_bulk, particle_C, particle_stress,
                                 particle_accumulated_strain, particle_damage, constitutive_model, alpha):
    lib("mpmx_compute_stress_from_F_trial", int(dim), activeLabel.ptr, materialLabel.ptr, particlePosition.ptr,
        particleVelocity.ptr, initialPhaseChangePosition.ptr, initialPhaseChangeVelocity.ptr, particle_F.ptr,
        particle_F_trial.ptr, mu.ptr, lam.ptr, yield_stress.ptr, hardening.ptr, softening.ptr, particle_density.ptr,
        strainCriteria.ptr, float(efficiency), eta_shear.ptr, eta_bulk.ptr, particle_C.ptr, particle_stress.ptr,
        particle_accumulated_strain.ptr, particle_damage.ptr, int(constitutive_model), alpha.ptr, stream_handle())


compute_stress_from_F_trial = Kernel("compute_stress_from_F_trial", _compute_stress_from_F_trial,
                                     "mpmRoutines.py:229-471 (return map :314, kirchoff_stress :456)")

_p2g_ws = Workspace()


def p2g_workspace(n, dims):
    ws, nb = _p2g_ws.get(lib.raw("mpmx_p2g_workspace_bytes")(int(n), *[int(d) for d in dims]))
    return ws, nb


def _p2g_apic_with_stress(dim, activeLabel, materialLabel, particle_stress, particle_x, particle_v, particle_C,
                          particle_vol, particle_mass, dx, inv_dx, minBounds, rpic_damping, grid_m, grid_v_in, dt):
    gx, gy, gz = grid_m.shape
    ws, nb = p2g_workspace(dim, grid_m.shape)
    lib("mpmx_p2g_apic_with_stress", int(dim), activeLabel.ptr, materialLabel.ptr, particle_stress.ptr,
        particle_x.ptr, particle_v.ptr, particle_C.ptr, particle_vol.ptr, particle_mass.ptr, float(dx), float(inv_dx),
        *_v3(minBounds), float(rpic_damping), gx, gy, gz, grid_m.ptr, grid_v_in.ptr, float(dt), ws, nb, None,
        stream_handle())


p2g_apic_with_stress = Kernel("p2g_apic_with_stress", _p2g_apic_with_stress, "mpmRoutines.py:485-569")


def p2g_out_of_grid(dim, activeLabel, materialLabel, particle_stress, particle_x, particle_v, particle_C,
                    particle_vol, particle_mass, dx, inv_dx, minBounds, rpic_damping, grid_m, grid_v_in, dt):
    """p2g_apic_with_stress that also returns the number of particles whose stencil left the grid (syncs)."""
    import ctypes as C
    gx, gy, gz = grid_m.shape
    ws, nb = p2g_workspace(dim, grid_m.shape)
    out = C.c_int(0)
    lib("mpmx_p2g_apic_with_stress", int(dim), activeLabel.ptr, materialLabel.ptr, particle_stress.ptr,
        particle_x.ptr, particle_v.ptr, particle_C.ptr, particle_vol.ptr, particle_mass.ptr, float(dx), float(inv_dx),
        *_v3(minBounds), float(rpic_damping), gx, gy, gz, grid_m.ptr, grid_v_in.ptr, float(dt), ws, nb,
        C.byref(out), stream_handle())
    return out.value


def _grid_normalization_and_gravity(dim, grid_m, grid_v_in, grid_v_out, gravity, dt):
    gx, gy, gz = grid_m.shape
    lib("mpmx_grid_normalization_and_gravity", gx, gy, gz, grid_m.ptr, grid_v_in.ptr, grid_v_out.ptr, *_v3(gravity),
        float(dt), stream_handle())


grid_normalization_and_gravity = Kernel("grid_normalization_and_gravity", _grid_normalization_and_gravity,
                                        "mpmRoutines.py:691-704")


def _add_damping_via_grid(dim, grid_v_out, scale):
    gx, gy, gz = grid_v_out.shape
    lib("mpmx_add_damping_via_grid", gx, gy, gz, grid_v_out.ptr, float(scale), stream_handle())


add_damping_via_grid = Kernel("add_damping_via_grid", _add_damping_via_grid, "mpmRoutines.py:706-711")


def _collideBounds(dim, grid_v_out, grid_dim_x, grid_dim_y, grid_dim_z, friction_coefficient, dt,
                   gravity_magnitude, boundary_padding):
    lib("mpmx_collide_bounds", grid_v_out.ptr, int(grid_dim_x), int(grid_dim_y), int(grid_dim_z),
        float(friction_coefficient), float(dt), float(gravity_magnitude), int(boundary_padding), stream_handle())


collideBounds = Kernel("collideBounds", _collideBounds, "mpmRoutines.py:942-993 (apply_coulomb_friction :915)")

# boundary-condition selectors of the C ABI (include/mpmx.h MPMX_BC_*)
BC_CODES = {"friction": 1, "friction_gradual": 2, "restitution": 3, "absorbing": 4, "plane_strain_y": 5,
            "plane_strain_y_global": 6}


def _variant(bc, grid_v_out, gx, gy, gz, mu=0.0, dt=0.0, g=0.0, pad=0, restitution=0.0):
    lib("mpmx_collide_bounds_variant", grid_v_out.ptr, int(gx), int(gy), int(gz), BC_CODES[bc], float(mu), float(dt),
        float(g), int(pad), float(restitution), stream_handle())


def _collideBoundsGradualFriction(dim, grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude,
                                  boundary_padding):
    _variant("friction_gradual", grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude,
             boundary_padding)


collideBoundsGradualFriction = Kernel("collideBoundsGradualFriction", _collideBoundsGradualFriction,
                                      "mpmRoutines.py:1264-1340")


def _collideBoundsRestitution(dim, grid_v_out, gx, gy, gz, restitution, friction_coefficient, dt, gravity_magnitude,
                              boundary_padding):
    _variant("restitution", grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude, boundary_padding,
             restitution)


collideBoundsRestitution = Kernel("collideBoundsRestitution", _collideBoundsRestitution, "mpmRoutines.py:1182-1261")


def _collideBoundsAbsorbing(dim, grid_v_out, gx, gy, gz, boundary_padding):
    _variant("absorbing", grid_v_out, gx, gy, gz, pad=boundary_padding)


collideBoundsAbsorbing = Kernel("collideBoundsAbsorbing", _collideBoundsAbsorbing, "mpmRoutines.py:1129-1179")


def _collideBoundsPlaneStrainY(dim, grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude,
                               boundary_padding):
    _variant("plane_strain_y", grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude, boundary_padding)


collideBoundsPlaneStrainY = Kernel("collideBoundsPlaneStrainY", _collideBoundsPlaneStrainY, "mpmRoutines.py:997-1064")


def _collideBoundsPlaneStrainYGlobal(dim, grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude,
                                     boundary_padding):
    _variant("plane_strain_y_global", grid_v_out, gx, gy, gz, friction_coefficient, dt, gravity_magnitude,
             boundary_padding)


collideBoundsPlaneStrainYGlobal = Kernel("collideBoundsPlaneStrainYGlobal", _collideBoundsPlaneStrainYGlobal,
                                         "mpmRoutines.py:1067-1126")

_vlc_ws = Workspace()


def _accumulate_cell_divergence(dim, activeLabel, materialLabel, particle_x, particle_C, particle_vol, inv_dx,
                                minBounds, cell_div_v_weighted, cell_vol_sum):
    gx, gy, gz = cell_vol_sum.shape
    ws, nb = _vlc_ws.get(lib.raw("mpmx_vlc_workspace_bytes")(gx, gy, gz))
    lib("mpmx_accumulate_cell_divergence", int(dim), activeLabel.ptr, materialLabel.ptr, particle_x.ptr,
        particle_C.ptr, particle_vol.ptr, float(inv_dx), *_v3(minBounds), gx, gy, gz, cell_div_v_weighted.ptr,
        cell_vol_sum.ptr, ws, nb, stream_handle())


accumulate_cell_divergence = Kernel("accumulate_cell_divergence", _accumulate_cell_divergence,
                                    "mpmRoutines.py:818-857 (cell sums in 64-bit fixed point, order-independent)")


def _apply_volumetric_locking_correction(dim, activeLabel, materialLabel, particle_x, particle_C, inv_dx, minBounds,
                                         cell_div_v_weighted, cell_vol_sum):
    gx, gy, gz = cell_vol_sum.shape
    lib("mpmx_apply_volumetric_locking_correction", int(dim), activeLabel.ptr, materialLabel.ptr, particle_x.ptr,
        particle_C.ptr, float(inv_dx), *_v3(minBounds), gx, gy, gz, cell_div_v_weighted.ptr, cell_vol_sum.ptr,
        stream_handle())


apply_volumetric_locking_correction = Kernel("apply_volumetric_locking_correction",
                                             _apply_volumetric_locking_correction, "mpmRoutines.py:860-911")


def _g2p(dim, dt, activeLabel, materialLabel, particle_x, particle_v, particle_C, particle_F, particle_F_trial,
         particle_cov, inv_dx, grid_v_out, update_cov_with_F, minBounds):
    gx, gy, gz = grid_v_out.shape
    lib("mpmx_g2p", float(dt), int(dim), activeLabel.ptr, materialLabel.ptr, particle_x.ptr, particle_v.ptr,
        particle_C.ptr, particle_F.ptr, particle_F_trial.ptr, particle_cov.ptr, float(inv_dx), grid_v_out.ptr, gx, gy,
        gz, int(bool(update_cov_with_F)), *_v3(minBounds), stream_handle())


g2p = Kernel("g2p", _g2p, "mpmRoutines.py:715-802 (update_cov :779)")


# ---- setup kernels used by the driver (runMPMYDW.py:309-312, 440-451) ----------------------------
def _get_float_array_product(dim, arrayA, arrayB, arrayC):
    arrayC.t.copy_(arrayA.t * arrayB.t)


get_float_array_product = Kernel("get_float_array_product", _get_float_array_product, "mpmRoutines.py:681-689")


def _set_mat33_to_identity(dim, target_array):
    target_array.fill_([1, 0, 0, 0, 1, 0, 0, 0, 1])


set_mat33_to_identity = Kernel("set_mat33_to_identity", _set_mat33_to_identity, "mpmRoutines.py:635-639")


def _set_mat33_to_copy(dim, source_array, target_array):
    target_array.assign(source_array)


set_mat33_to_copy = Kernel("set_mat33_to_copy", _set_mat33_to_copy, "mpmRoutines.py:641-645")
