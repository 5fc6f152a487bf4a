"""Snapshots, checkpoints and restart: the output half of the reference's
utils/fs5PlotUtils.py (save_grid_and_particles_vti_vtp :433-646, save_mpm
:785-881, load_restart_from_vtp :253-430), with the same file names, array
names, layouts and field data, written and read by vtkxml (the vtk package is
not part of this build).  Rendering, ParaView and XDMF helpers are not here.

Files written here open in ParaView and in the reference's own
load_restart_from_vtp; checkpoints written by the reference load here.  Two
additions, ignored by VTK readers that do not know them: the field arrays
``sim_time_f64`` and ``dx_f64`` keep the time and the spacing in FP64 (the
reference stores its field data as Float32, :585-592), so that a restart here
resumes at the exact time.

``load_domain_vtp`` reads a particle domain from a .vtp file holding the
arrays of the reference's HDF5 domains ("volume" and the optional per-particle
properties), as the domain .vtp files next to the reference's benchmark .h5
files do.
"""
from __future__ import annotations

import os

import numpy as np

from . import vtkxml

MATERIAL_PROPS = ["mu", "lam", "bulk", "alpha", "hardening", "softening", "eta_shear", "eta_bulk", "strainCriteria"]
DOMAIN_PROPS = ["density", "E", "nu", "ys", "alpha", "hardening", "softening", "eta_shear", "eta_bulk",
                "strainCriteria", "damage"]


def _f32(a):
    return np.ascontiguousarray(np.asarray(a), dtype=np.float32)


def save_grid_and_particles_vti_vtp(output_prefix, grid_mass, minBounds, dx, particle_positions, particle_radius=0.5,
                                    particle_velocity=None, particle_ys=None, particle_effective_ys=None,
                                    particle_damage=None, particle_mean_stress=None, particle_von_mises=None,
                                    particle_stress_tensor=None, particle_active_label=None,
                                    particle_material_label=None, restart_data=None):
    """``{prefix}_grid.vti`` (grid mass, Fortran order) and ``{prefix}_particles.vtp``
    with the visualisation arrays and, given ``restart_data``, the checkpoint
    arrays and field data (fs5PlotUtils.py:433-646)."""
    folder = os.path.dirname(output_prefix)
    if folder:
        os.makedirs(folder, exist_ok=True)
    gm = np.asarray(grid_mass)
    vtkxml.write_vti(f"{output_prefix}_grid.vti", gm.shape, minBounds, (dx, dx, dx),
                     {"grid_mass": _f32(gm.ravel(order="F"))}, scalars="grid_mass")

    pos = _f32(particle_positions).reshape(-1, 3)
    n = pos.shape[0]
    pd = {}
    pd["radius"] = np.full(n, particle_radius, np.float32) if np.isscalar(particle_radius) else _f32(particle_radius)
    if particle_velocity is not None:
        pd["velocity"] = _f32(particle_velocity).reshape(n, 3)
    for name, v in (("yield_stress", particle_ys), ("effective_yield_stress", particle_effective_ys),
                    ("damage", particle_damage), ("mean_stress", particle_mean_stress),
                    ("von_mises_stress", particle_von_mises)):
        if v is not None:
            pd[name] = _f32(v)
    if particle_stress_tensor is not None:
        s = np.asarray(particle_stress_tensor)
        if s.ndim == 3:  # xx, yy, zz, xy, xz, yz (:520-534)
            pd["stress_tensor"] = _f32(np.stack([s[:, 0, 0], s[:, 1, 1], s[:, 2, 2], s[:, 0, 1], s[:, 0, 2],
                                                 s[:, 1, 2]], axis=1))
    if particle_active_label is not None:
        pd["active_label"] = np.ascontiguousarray(particle_active_label, dtype=np.int32)
    if particle_material_label is not None:
        pd["material_label"] = np.ascontiguousarray(particle_material_label, dtype=np.int32)

    fd = None
    if restart_data is not None:
        r = restart_data
        fd = {}
        for key in ("sim_time", "bigStep", "dx"):
            if key in r:
                fd[key] = np.array([r[key]], np.float32)
        for key in ("sim_time", "dx"):
            if key in r:
                fd[f"{key}_f64"] = np.array([r[key]], np.float64)
        if "particle_F" in r:
            pd["deformation_gradient_F"] = _f32(r["particle_F"]).reshape(n, 9)
        if "particle_C" in r:
            pd["apic_C"] = _f32(r["particle_C"]).reshape(n, 9)
        if "particle_accumulated_strain" in r:
            pd["accumulated_strain"] = _f32(r["particle_accumulated_strain"])
        if "ys_base" in r:
            pd["yield_stress_base"] = _f32(r["ys_base"])
        for prop in MATERIAL_PROPS:
            if prop in r:
                pd[f"mpm_{prop}"] = _f32(r[prop])
        for key, name in (("particle_vol", "particle_volume"), ("particle_mass", "particle_mass"),
                          ("particle_density", "particle_density")):
            if key in r:
                pd[name] = _f32(r[key])
        if "particle_x_initial_xpbd" in r:
            pd["xpbd_initial_position"] = _f32(r["particle_x_initial_xpbd"]).reshape(n, 3)
        if "particle_cumDist_xpbd" in r:
            pd["xpbd_cumulative_distance"] = _f32(r["particle_cumDist_xpbd"])
    vtkxml.write_vtp(f"{output_prefix}_particles.vtp", pos, pd, fd, verts=False)


def snapshot_name(sim, bigStep: int, counter: int) -> str:
    """sim_step_BBBBCCCCCCTTTTTTTTTTTTTTTT (save_mpm, fs5PlotUtils.py:863)."""
    return f"{sim.outputFolder}/sim_step_{bigStep:04d}{counter:06d}{int(np.ceil(sim.t * 1e8)):016d}"


def save_mpm(sim, mpm, bigStep, counter, include_restart_data=False, xpbd=None):
    """Snapshot (and with include_restart_data, checkpoint) of the state
    (fs5PlotUtils.py:785-881): mean stress, von Mises and effective yield
    stress are derived on the host in FP64 as there.  Returns the prefix."""
    sigma = mpm.particle_stress.numpy().astype(np.float64)
    mean_stress = np.trace(sigma, axis1=1, axis2=2) / 3.0
    s = sigma - mean_stress[:, None, None] * np.eye(3)
    von_mises = np.sqrt(1.5 * np.sum(s ** 2, axis=(1, 2)))
    ys = mpm.ys.numpy()
    alpha = mpm.alpha.numpy()
    damage = mpm.particle_damage.numpy()
    effective_ys = (ys - alpha * mean_stress) * (1.0 - damage)
    restart = None
    if include_restart_data:
        restart = {"sim_time": sim.t, "bigStep": bigStep, "dx": sim.dx,
                   "particle_F": mpm.particle_F.numpy(), "particle_C": mpm.particle_C.numpy(),
                   "particle_accumulated_strain": mpm.particle_accumulated_strain.numpy(),
                   "ys_base": mpm.ys_base.numpy(),
                   "particle_vol": sim.particle_vol.numpy(), "particle_mass": sim.particle_mass.numpy(),
                   "particle_density": sim.particle_density.numpy()}
        for prop in MATERIAL_PROPS:
            restart[prop] = getattr(mpm, prop).numpy()
        if xpbd is not None:
            restart["particle_x_initial_xpbd"] = xpbd.particle_x_initial_xpbd.numpy()
            restart["particle_cumDist_xpbd"] = xpbd.particle_cumDist_xpbd.numpy()
    prefix = snapshot_name(sim, bigStep, counter)
    save_grid_and_particles_vti_vtp(prefix, mpm.grid_m.numpy(), sim.minBounds, sim.dx, sim.particle_x.numpy(),
                                    particle_radius=sim.particle_radius.numpy(),
                                    particle_velocity=sim.particle_v.numpy(), particle_ys=ys,
                                    particle_effective_ys=effective_ys, particle_damage=damage,
                                    particle_mean_stress=mean_stress, particle_von_mises=von_mises,
                                    particle_stress_tensor=sigma, particle_active_label=sim.activeLabel.numpy(),
                                    particle_material_label=sim.materialLabel.numpy(), restart_data=restart)
    return prefix


def load_restart_from_vtp(vtp_path):
    """The restart dictionary of fs5PlotUtils.py:253-430 (same keys, defaults
    and shapes): F defaults to I, C to 0, accumulated strain to 0, ys_base to
    ys; stress is rebuilt symmetric from its 6 components."""
    d = vtkxml.read(vtp_path)
    x = d["points"]
    if x is None or x.shape[0] == 0:
        raise ValueError(f"Failed to read VTP file or file is empty: {vtp_path}")
    n = x.shape[0]
    pdat, fdat = d["point_data"], d["field_data"]

    def get(name, comps=1, required=False):
        v = pdat.get(name)
        if v is None:
            if required:
                raise ValueError(f"Required array '{name}' not found in VTP file")
            return None
        return v.reshape(-1, comps) if comps > 1 else v

    r = {"particle_x": x.astype(np.float32), "particle_radius": get("radius", required=True)}
    v = get("velocity", 3)
    r["particle_v"] = v if v is not None else np.zeros((n, 3), np.float32)
    r["ys"] = get("yield_stress")
    r["particle_damage"] = get("damage")
    r["materialLabel"] = get("material_label")
    r["activeLabel"] = get("active_label")
    s6 = get("stress_tensor", 6)
    if s6 is not None:
        st = np.zeros((n, 3, 3), np.float32)
        st[:, 0, 0], st[:, 1, 1], st[:, 2, 2] = s6[:, 0], s6[:, 1], s6[:, 2]
        st[:, 0, 1] = st[:, 1, 0] = s6[:, 3]
        st[:, 0, 2] = st[:, 2, 0] = s6[:, 4]
        st[:, 1, 2] = st[:, 2, 1] = s6[:, 5]
        r["particle_stress"] = st
    F = get("deformation_gradient_F", 9)
    r["particle_F"] = F.reshape(-1, 3, 3) if F is not None else np.tile(np.eye(3, dtype=np.float32), (n, 1, 1))
    Cm = get("apic_C", 9)
    r["particle_C"] = Cm.reshape(-1, 3, 3) if Cm is not None else np.zeros((n, 3, 3), np.float32)
    acc = get("accumulated_strain")
    r["particle_accumulated_strain"] = acc if acc is not None else np.zeros(n, np.float32)
    r["ys_base"] = get("yield_stress_base")
    if r["ys_base"] is None and r["ys"] is not None:
        r["ys_base"] = r["ys"].copy()
    for prop in MATERIAL_PROPS:
        a = get(f"mpm_{prop}")
        if a is not None:
            r[prop] = a
    r["particle_vol"] = get("particle_volume")
    r["particle_mass"] = get("particle_mass")
    r["particle_density"] = get("particle_density")
    xi = get("xpbd_initial_position", 3)
    if xi is not None:
        r["particle_x_initial_xpbd"] = xi
    r["particle_cumDist_xpbd"] = get("xpbd_cumulative_distance")

    def field(name, default=None):
        a = fdat.get(name)
        return default if a is None else a.ravel()[0].item()

    r["sim_time"] = field("sim_time_f64", field("sim_time", 0.0))
    r["bigStep"] = int(field("bigStep", 0))
    r["dx"] = field("dx_f64", field("dx", None))
    return r


def load_domain_vtp(path: str) -> dict:
    """A particle domain for driver.setup from a .vtp: positions, "volume" as
    particle_volume, and whichever per-particle property arrays it holds."""
    d = vtkxml.read(path)
    pdat = d["point_data"]
    vol = pdat.get("particle_volume", pdat.get("volume"))
    if vol is None:
        raise ValueError(f"{path}: no 'volume' array")
    dom = {"x": d["points"].astype(np.float64), "particle_volume": np.asarray(vol, np.float64)}
    for name in DOMAIN_PROPS:
        if name in pdat:
            dom[name] = np.asarray(pdat[name], np.float32)
    return dom
