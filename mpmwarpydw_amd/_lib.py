"""ctypes binding of libmpmx.so (the C ABI declared in include/mpmx.h).

The library is loaded once.  Every binding declares its argument types, so a
call with the wrong number or kind of arguments raises instead of corrupting
the stack, and every non-zero return code raises ``MpmxError`` with the
library's own message.  There is no fallback: without the built library, or
without a GPU, the hot path fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpmx.so")

P = C.c_void_p
I = C.c_int
L = C.c_long
F = C.c_float
S = C.c_size_t
IP = C.POINTER(C.c_int)

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "mpmx_last_error": [],
    "mpmx_version": [],
    "mpmx_struct_layout": [P, I],
    "mpmx_set_tuning": [I, I],
    "mpmx_get_tuning": [I],
    "mpmx_ctx_create": [I, P],
    "mpmx_ctx_destroy": [P],
    "mpmx_ctx_stream": [P],
    "mpmx_buf_alloc": [P, S, P],
    "mpmx_buf_free": [P, P],
    "mpmx_buf_upload": [P, P, P, S],
    "mpmx_buf_download": [P, P, P, S],
    "mpmx_graph_begin": [P],
    "mpmx_graph_end": [P, P],
    "mpmx_graph_launch": [P, P],
    "mpmx_graph_destroy": [P],
    "mpmx_profile_begin": [I],
    "mpmx_profile_mark": [P],
    "mpmx_profile_end": [],
    "mpmx_profile_hold": [P, I],
    "mpmx_profile_report": [P, S],
    "mpmx_substep_workspace_bytes": [I, I, I, I],
    "mpmx_substep_bin": [P, P, P, P, S, P],
    "mpmx_mpm_substep": [P, P, P, P, S, P],
    "mpmx_substep_misbinned": [P, S, I, I, I, I, IP, P],
    "mpmx_reorder_permutation": [P, P, P, P, P, S, P],
    "mpmx_ensemble_workspace_bytes": [I, P, I],
    "mpmx_ensemble_bin": [P, P, I, P, P, S, P],
    "mpmx_ensemble_substep": [P, P, I, P, P, P, P, P, S, P],
    "mpmx_ensemble_reorder_permutation": [P, P, I, P, P, P, S, P],
    "mpmx_gather_words": [P, P, P, L, I, I, P],
    "mpmx_invert_permutation": [P, P, L, P],
    "mpmx_stream_sync": [P],
    "mpmx_fill_i32": [P, L, I, P],
    "mpmx_fill_f32": [P, L, F, P],
    "mpmx_copy_f32": [P, P, L, P],
    "mpmx_compute_mu_lam_bulk": [I, P, P, P, P, P, P],
    "mpmx_initialize_geostatic_F": [I, P, P, P, P, P, F, F, F, I, P],
    "mpmx_deactivate_and_teleport_xpbd_below_datum": [I, P, P, P, P, F, F, P],
    "mpmx_compute_stress_from_F_trial": [I, P, P, P, P, P, P, P, P, P, P, P, P, P, P, P, F, P, P, P, P, P, P, I, P,
                                         P],
    "mpmx_p2g_workspace_bytes": [I, I, I, I],
    "mpmx_p2g_apic_with_stress": [I, P, P, P, P, P, P, P, P, F, F, F, F, F, F, I, I, I, P, P, F, P, S, IP, P],
    "mpmx_p2g_grid_update": [I, P, P, P, P, P, P, P, P, F, F, F, F, F, F, I, I, I, P, P, P, F, F, F, F, F, I, F, F,
                             I, F, P, S, P],
    "mpmx_grid_normalization_and_gravity": [I, I, I, P, P, P, F, F, F, F, P],
    "mpmx_add_damping_via_grid": [I, I, I, P, F, P],
    "mpmx_collide_bounds": [P, I, I, I, F, F, F, I, P],
    "mpmx_grid_update": [I, I, I, P, P, P, F, F, F, F, F, I, F, F, I, F, P],
    "mpmx_collide_bounds_variant": [P, I, I, I, I, F, F, F, I, F, P],
    "mpmx_vlc_workspace_bytes": [I, I, I],
    "mpmx_accumulate_cell_divergence": [I, P, P, P, P, P, F, F, F, F, I, I, I, P, P, P, S, P],
    "mpmx_apply_volumetric_locking_correction": [I, P, P, P, P, F, F, F, F, I, I, I, P, P, P],
    "mpmx_g2p": [F, I, P, P, P, P, P, P, P, P, F, P, I, I, I, I, F, F, F, P],
    "mpmx_reduce_workspace_bytes": [I],
    "mpmx_velocity_convergence": [I, P, P, P, P, P, P, S, P],
    "mpmx_velocity_convergence_noalloc": [I, P, P, P, P, P, P],
    "mpmx_sum_float_array_active": [I, P, P, P, P, P, S, P],
    "mpmx_count_particles_by_type": [I, P, P, P, P, S, P],
    "mpmx_hashgrid_workspace_bytes": [I, I, I, I],
    "mpmx_hashgrid_build": [I, P, F, I, I, I, P, P, P, S, P],
    "mpmx_hashgrid_build_label": [I, P, F, I, I, I, P, I, P, P, P, S, P],
    "mpmx_hashgrid_build_ordered": [I, P, F, I, I, I, P, I, P, P, P, P, S, P],
    "mpmx_hashgrid_pair_workspace_bytes": [I, I, I, I, I, I, I],
    "mpmx_hashgrid_build_pair": [I, P, F, I, I, I, P, I, I, I, I, P, P, P, P, P, P, P, I, P, S, P],
    "mpmx_reset_mpm_contact_labels": [I, P, P, P],
    "mpmx_integrate_particles_xpbd": [I, P, P, P, P, F, F, F, F, P, P, F, P],
    "mpmx_solve_particle_bound_contacts": [I, P, P, P, P, P, P, F, F, F, F, F, F, F, F, F, F, F, P, P],
    "mpmx_solve_particle_particle_contacts": [I, P, P, P, P, F, I, I, I, P, P, P, P, F, F, F, F, F, F, F, I, P, P],
    "mpmx_solve_particle_particle_contacts_split": [I, P, P, P, P, P, P, F, I, I, I, I, I, I, P, P, P, P, F, F, F,
                                                    F, F, F, F, I, P, P, P],
    "mpmx_xpbd_contact_pack": [I, P, P, P, P, P, P, P, P, P, P, P, P, I, I, I, P, P, P, P],
    "mpmx_solve_particle_particle_contacts_packed": [I, P, P, P, P, P, P, F, I, I, I, I, I, I, P, P, P, P, P, P, P,
                                                     F, F, F, F, F, F, F, I, P, P, P, P, P, P],
    "mpmx_xpbd_contact_list": [I, P, P, P, F, I, I, I, P, P, F, F, P, P, P, P],
    "mpmx_hashgrid_near_words": [I, I, I],
    "mpmx_hashgrid_near_cells": [P, I, I, I, P, P, P],
    "mpmx_apply_particle_deltas": [I, P, P, P, P, P, F, F, P, P, P],
    "mpmx_sleep_particles": [I, P, P, F, P, P, P, P, F, P, P],
    "mpmx_apply_mpm_contact_impulses": [I, P, P, P, P, P, P, F, F, F, I, P, P],
    "mpmx_swell_particles_type2": [I, P, P, P, P, P, F, F, P, P, P],
    "mpmx_xpbd_begin": [I, P, P, P, P, F, F, F, F, F, I, P, P, P, P, P],
    "mpmx_xpbd_bound_contacts_write": [I, P, P, P, P, P, P, F, F, F, F, F, F, F, F, F, F, F, P, P],
    "mpmx_xpbd_apply_deltas_inplace": [I, P, P, P, F, F, P, P, P, P, P],
    "mpmx_xpbd_finish": [I, P, P, F, P, P, F, F, F, I, P, P, P, P, P, P, P],
    "mpmx_mlp_nparams": [],
    "mpmx_mlp_forward": [I, P, P, P, P],
    "mpmx_mlp_workspace_bytes": [I],
    "mpmx_mlp_loss_grad": [I, P, P, P, P, P, P, S, P],
    "mpmx_mlp_adam_step": [I, P, P, P, P, F, F, F, F, I, P],
    "mpmx_lrm_harvest": [P, P, P, P, P, P],
    "mpmx_lrm_workspace_bytes": [I],
    "mpmx_compute_stress_learned": [P, F, P, F, P, P, S, P],
    "mpmx_mpm_substep_learned": [P, P, P, P, F, P, P, S, P],
    "mpmx_ensemble_substep_learned": [P, P, I, P, P, P, P, P, F, P, P, S, P],
}
_RESTYPE = {
    "mpmx_substep_workspace_bytes": S,
    "mpmx_vlc_workspace_bytes": S,
    "mpmx_last_error": C.c_char_p,
    "mpmx_p2g_workspace_bytes": S,
    "mpmx_reduce_workspace_bytes": S,
    "mpmx_hashgrid_workspace_bytes": S,
    "mpmx_hashgrid_pair_workspace_bytes": S,
    "mpmx_mlp_workspace_bytes": S,
    "mpmx_lrm_workspace_bytes": S,
    "mpmx_ctx_stream": P,
    "mpmx_ensemble_workspace_bytes": S,
}


class _Strict(C.Structure):
    """A ctypes structure whose constructor rejects names that are not fields
    (plain ctypes would keep them as Python attributes and pass a struct
    without them)."""

    def __init__(self, *args, **kw):
        names = {f[0] for f in self._fields_}
        bad = [k for k in kw if k not in names]
        if bad:
            raise TypeError(f"{type(self).__name__}: no field(s) {bad}")
        super().__init__(*args, **kw)


class Particles(_Strict):
    """mpmx_particles (include/mpmx.h)"""
    _fields_ = [("n", C.c_int)] + [(name, C.c_void_p) for name in (
        "activeLabel", "materialLabel", "x", "v", "x_init_xpbd", "v_init_xpbd", "F", "F_trial", "C", "stress", "cov",
        "mu", "lam", "ys", "alpha", "hardening", "softening", "density", "strainCriteria", "eta_shear", "eta_bulk",
        "acc_strain", "damage", "vol", "mass", "order", "scene")]


class Scene(_Strict):
    """mpmx_scene (include/mpmx.h)"""
    _fields_ = [("gx", C.c_int), ("gy", C.c_int), ("gz", C.c_int), ("minx", C.c_float), ("miny", C.c_float),
                ("minz", C.c_float)]


class Grid(_Strict):
    """mpmx_grid (include/mpmx.h)"""
    _fields_ = [("gx", C.c_int), ("gy", C.c_int), ("gz", C.c_int), ("m", C.c_void_p), ("v_in", C.c_void_p),
                ("v_out", C.c_void_p)]


class StepParams(_Strict):
    """mpmx_step_params (include/mpmx.h)"""
    _fields_ = [(name, C.c_float) for name in (
        "dt", "dx", "inv_dx", "minx", "miny", "minz", "gravity_x", "gravity_y", "gravity_z", "rpic_damping", "eff",
        "damping_scale", "friction", "gravity_magnitude")] + [(name, C.c_int) for name in (
        "bc", "boundary_padding", "update_cov", "constitutive_model")] + [("restitution", C.c_float), ("vlc", C.c_int)]


class XpbdBounds(_Strict):
    """mpmx_xpbd_bounds (include/mpmx.h)"""
    _fields_ = [("mass", C.c_void_p), ("radius", C.c_void_p)] + [(name, C.c_float) for name in (
        "mu", "static_threshold", "static_friction", "minx", "miny", "minz", "maxx", "maxy", "maxz", "dt",
        "relaxation")] + [(name, C.c_void_p) for name in ("delta", "xr", "vw", "flags", "pos", "xpbd_pos",
                                                            "xpbd_count")]


class MpmxError(RuntimeError):
    pass


class _Lib:
    def __init__(self):
        self._dll = None

    def load(self):
        if self._dll is None:
            if not os.path.exists(LIB_PATH):
                raise MpmxError(
                    f"{LIB_PATH} is missing: build it with `python -m mpmwarpydw_amd.build` "
                    "(there is no CPU fallback for the hot path)")
            dll = C.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(dll, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, C.c_int)
            self._dll = dll
        return self._dll

    def raw(self, name):
        return getattr(self.load(), name)

    def __call__(self, name, *args):
        rc = getattr(self.load(), name)(*args)
        if rc != 0:
            msg = self.load().mpmx_last_error().decode(errors="replace")
            raise MpmxError(f"{name} returned {rc}: {msg}")
        return rc


lib = _Lib()
