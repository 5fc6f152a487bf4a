"""XPBD contact kernels of the hot path, with the names and argument lists of
the reference's utils/xpbdRoutines.py, launched like the Warp kernels they
replace (``runtime.launch(kernel=xpbdRoutines.X, dim=n, inputs=[...], outputs=[...])``).
The hash grid argument of the particle-contact kernel is a ``runtime.HashGrid``
(passed as ``grid.id``, like ``wp.HashGrid.id``).
"""
from __future__ import annotations

import numpy as np

from ._lib import lib
from .runtime import Kernel, stream_handle


def _v3(v):
    a = np.asarray(v, dtype=np.float64).reshape(3)
    return float(a[0]), float(a[1]), float(a[2])


def _reset_mpm_contact_labels(dim, activeLabel, materialLabel):
    lib("mpmx_reset_mpm_contact_labels", int(dim), activeLabel.ptr, materialLabel.ptr, stream_handle())


reset_mpm_contact_labels = Kernel("reset_mpm_contact_labels", _reset_mpm_contact_labels, "xpbdRoutines.py:3-21")


def _integrateParticlesXPBD(dim, activeLabel, materialLabel, x, v, gravity, dt, x_new, v_new, v_max):
    lib("mpmx_integrate_particles_xpbd", int(dim), activeLabel.ptr, materialLabel.ptr, x.ptr, v.ptr, *_v3(gravity),
        float(dt), x_new.ptr, v_new.ptr, float(v_max), stream_handle())


integrateParticlesXPBD = Kernel("integrateParticlesXPBD", _integrateParticlesXPBD, "xpbdRoutines.py:23-48")


def _my_solve_particle_bound_contacts(dim, activeLabel, materialLabel, particle_x, particle_v, mass, particle_radius,
                                      mu, staticGroundVelocityThresholdRatio, staticGroundFriction, minbounds,
                                      maxbounds, dt, relaxation, delta):
    lib("mpmx_solve_particle_bound_contacts", int(dim), activeLabel.ptr, materialLabel.ptr, particle_x.ptr,
        particle_v.ptr, mass.ptr, particle_radius.ptr, float(mu), float(staticGroundVelocityThresholdRatio),
        float(staticGroundFriction), *_v3(minbounds), *_v3(maxbounds), float(dt), float(relaxation), delta.ptr,
        stream_handle())


my_solve_particle_bound_contacts = Kernel("my_solve_particle_bound_contacts", _my_solve_particle_bound_contacts,
                                          "xpbdRoutines.py:49-136")


def _my_solve_particle_particle_contacts(dim, activeLabel, materialLabel, grid, particle_x, particle_v,
                                         particle_mass, particle_radius, k_mu, staticParticleVelocityThresholdRatio,
                                         staticParticleFriction, k_cohesion, max_radius, dt, relaxation,
                                         mpm_contact_transition_lock, deltas):
    dx, dy, dz = grid.dims
    lib("mpmx_solve_particle_particle_contacts", int(dim), activeLabel.ptr, materialLabel.ptr, grid.point_ids.ptr,
        grid.cell_start.ptr, float(grid.cell_width), dx, dy, dz, particle_x.ptr, particle_v.ptr, particle_mass.ptr,
        particle_radius.ptr, float(k_mu), float(staticParticleVelocityThresholdRatio), float(staticParticleFriction),
        float(k_cohesion), float(max_radius), float(dt), float(relaxation), int(mpm_contact_transition_lock),
        deltas.ptr, stream_handle())


my_solve_particle_particle_contacts = Kernel("my_solve_particle_particle_contacts",
                                             _my_solve_particle_particle_contacts,
                                             "xpbdRoutines.py:138-219 (hash_grid_query order of wp.HashGrid)")


def _solve_particle_particle_contacts_split(dim, activeLabel, materialLabel, grid, xpbd_grid, particle_x, particle_v,
                                            particle_mass, particle_radius, k_mu,
                                            staticParticleVelocityThresholdRatio, staticParticleFriction, k_cohesion,
                                            max_radius, dt, relaxation, mpm_contact_transition_lock, deltas,
                                            slot_of_id=None):
    dx, dy, dz = grid.dims
    if xpbd_grid.cell_width != grid.cell_width:
        raise ValueError("the XPBD-only table needs the full table's cell width")
    lib("mpmx_solve_particle_particle_contacts_split", int(dim), activeLabel.ptr, materialLabel.ptr,
        grid.point_ids.ptr, grid.cell_start.ptr, xpbd_grid.point_ids.ptr, xpbd_grid.cell_start.ptr,
        float(grid.cell_width), dx, dy, dz, *xpbd_grid.dims, particle_x.ptr, particle_v.ptr, particle_mass.ptr,
        particle_radius.ptr,
        float(k_mu), float(staticParticleVelocityThresholdRatio), float(staticParticleFriction), float(k_cohesion),
        float(max_radius), float(dt), float(relaxation), int(mpm_contact_transition_lock), deltas.ptr,
        None if slot_of_id is None else slot_of_id.ptr, stream_handle())


solve_particle_particle_contacts_split = Kernel(
    "solve_particle_particle_contacts_split", _solve_particle_particle_contacts_split,
    "xpbdRoutines.py:138-219 with MPM particles walking an XPBD-only table (same results)")


def _my_apply_particle_deltas(dim, activeLabel, materialLabel, x_orig, x_pred, delta, dt, v_max, x_out, v_out):
    lib("mpmx_apply_particle_deltas", int(dim), activeLabel.ptr, materialLabel.ptr, x_orig.ptr, x_pred.ptr, delta.ptr,
        float(dt), float(v_max), x_out.ptr, v_out.ptr, stream_handle())


my_apply_particle_deltas = Kernel("my_apply_particle_deltas", _my_apply_particle_deltas, "xpbdRoutines.py:221-257")


def _apply_mpm_contact_impulses(dim, activeLabel, materialLabel, x_orig, v_orig, x_final, v_final, dt, v_max,
                                coupling_strength, xpbd_only, v_out):
    lib("mpmx_apply_mpm_contact_impulses", int(dim), activeLabel.ptr, materialLabel.ptr, x_orig.ptr, v_orig.ptr,
        x_final.ptr, v_final.ptr, float(dt), float(v_max), float(coupling_strength), int(xpbd_only), v_out.ptr,
        stream_handle())


apply_mpm_contact_impulses = Kernel("apply_mpm_contact_impulses", _apply_mpm_contact_impulses,
                                    "xpbdRoutines.py:259-310")


def _sleepParticles(dim, activeLabel, materialLabel, sleepThreshold, radius, particle_positions_prev,
                    particle_positions_after, particle_distance_total, dt, total_xpbd, sleeping_xpbd):
    # the C ABI increments a {total, sleeping} pair; SimState allocates the two
    # one-element counters adjacently, other layouts go through a pair buffer
    import ctypes as C
    adjacent = sleeping_xpbd.t.data_ptr() == total_xpbd.t.data_ptr() + 4
    if adjacent:
        pair_ptr, pair = total_xpbd.ptr, None
    else:
        import torch
        pair = torch.cat([total_xpbd.t.reshape(-1)[:1], sleeping_xpbd.t.reshape(-1)[:1]]).contiguous()
        pair_ptr = C.c_void_p(pair.data_ptr())
    lib("mpmx_sleep_particles", int(dim), activeLabel.ptr, materialLabel.ptr, float(sleepThreshold), radius.ptr,
        particle_positions_prev.ptr, particle_positions_after.ptr, particle_distance_total.ptr, float(dt), pair_ptr,
        stream_handle())
    if pair is not None:
        total_xpbd.t.reshape(-1)[:1].copy_(pair[:1])
        sleeping_xpbd.t.reshape(-1)[:1].copy_(pair[1:])


sleepParticles = Kernel("sleepParticles", _sleepParticles, "xpbdRoutines.py:313-334")


def _swellParticlesType2(dim, activeLabel, materialLabel, particle_initial_position, particle_position,
                         particle_radius, factor, factor_max, swellRadius, baseRadius):
    lib("mpmx_swell_particles_type2", int(dim), activeLabel.ptr, materialLabel.ptr, particle_initial_position.ptr,
        particle_position.ptr, particle_radius.ptr, float(factor), float(factor_max), swellRadius.ptr, baseRadius.ptr,
        stream_handle())


swellParticlesType2 = Kernel("swellParticlesType2", _swellParticlesType2, "xpbdRoutines.py:354-376")
