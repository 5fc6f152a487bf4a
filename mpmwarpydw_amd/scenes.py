"""Synthetic, seeded scenes of the measured workloads (no input files).

Each builder returns ``(domain, overrides)``: ``domain`` holds the particle
arrays that runMPMYDW.py would read from its HDF5 file (x, particle_volume and
per-particle properties), ``overrides`` the JSON keys of the run.

  block_collapse   32^3 = 32,768 particles on a 0.25 m lattice with +-0.05 m
                   jitter, Drucker-Prager block collapsing under gravity,
                   500 substeps at dt = 1e-4 s
  undercut_slab    200x100x50 lattice minus a 40x100x10 undercut = 960,000
                   particles, lognormal E (median 5e8 Pa, sigma_ln 0.3) and yield
                   stress (median 2e5 Pa, sigma_ln 0.4), 2,000 substeps at 5e-5 s
  contact_heavy    200x100x100 = 2,000,000 particles, a 100x100x100 sub-block
                   already in the XPBD phase (radius 0.25 m), the rest as in
                   undercut_slab, XPBD every 10 substeps with 4 iterations
Sizes can be scaled down with ``scale`` for tests (lattice edges multiplied).
"""
from __future__ import annotations

import numpy as np

SPACING = 0.25
VOL = SPACING ** 3


def alpha_from_friction_angle(phi_deg):
    """Drucker-Prager alpha from the friction angle (createRandomRockDomain.py:164)."""
    s = np.sin(np.deg2rad(phi_deg))
    return 2.0 * s / (np.sqrt(3.0) * (3.0 - s))


def _lattice(nx, ny, nz, rng=None, jitter=0.05):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    x = np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1).astype(np.float64)
    x = (x + 0.5) * SPACING
    if rng is not None and jitter > 0:
        x += rng.uniform(-jitter, jitter, size=x.shape)
    return x, np.stack([i.ravel(), j.ravel(), k.ravel()], axis=1)


def _common_overrides(dt, n_sub, ratio):
    return {
        "dt": dt, "dtxpbd": dt * ratio, "bigStepDuration": dt * n_sub, "bigSteps": 1,
        "grid_particle_spacing_scale": 0.5 / VOL ** 0.33,  # dx = 0.5 m exactly
        "density": 2500.0, "E": 1e8, "nu": 0.25, "ys": 1e5, "alpha": float(alpha_from_friction_angle(30.0)),
        "hardening": 0.0, "softening": 0.0, "strainCriteria": 0.05, "eff": 0.0, "gravity": -9.81,
        "initialise_geostatic": 0, "rpic_damping": 0.0, "volumetric_locking_correction": 0,
        "eta_shear": 0.0, "eta_bulk": 0.0, "grid_v_damping_scale": 1.0, "update_cov": 1, "boundFriction": 0.5,
        "boundaryCondition": "friction", "xpbd_iterations": 4, "staticParticleFriction": 0.5,
        "dynamicParticleFriction": 0.5, "particle_cohesion": 0.0, "swellingRatio": 0.0,
        "mpm_contact_transition_lock": 0,
    }


def block_collapse(scale: float = 1.0, seed: int = 0):
    n = max(4, int(round(32 * scale)))
    rng = np.random.default_rng(seed)
    x, _ = _lattice(n, n, n, rng)
    dom = {"x": x, "particle_volume": np.full(len(x), VOL)}
    ov = _common_overrides(1e-4, 500, 10)
    ov.update({"boundary_padding_mask": "111101", "grid_padding": 8.0})
    return dom, ov


def undercut_slab(scale: float = 1.0, seed: int = 1):
    nx, ny, nz = (max(10, int(round(v * scale))) for v in (200, 100, 50))
    rng = np.random.default_rng(seed)
    x, ijk = _lattice(nx, ny, nz, None)
    cut_x0, cut_x1, cut_z = int(round(0.4 * nx)), int(round(0.6 * nx)), int(round(0.2 * nz))
    keep = ~((ijk[:, 0] >= cut_x0) & (ijk[:, 0] < cut_x1) & (ijk[:, 2] < cut_z))
    x = x[keep]
    n = len(x)
    E = 5e8 * np.exp(0.3 * rng.standard_normal(n))
    ys = 2e5 * np.exp(0.4 * rng.standard_normal(n))
    x = x + rng.uniform(-0.05, 0.05, size=x.shape)
    dom = {"x": x, "particle_volume": np.full(n, VOL), "E": E, "ys": ys}
    ov = _common_overrides(5e-5, 2000, 10)
    ov.update({"boundary_padding_mask": "000001", "grid_padding": 4.0})
    return dom, ov


def contact_heavy(scale: float = 1.0, seed: int = 2):
    nx, ny, nz = (max(10, int(round(v * scale))) for v in (200, 100, 100))
    rng = np.random.default_rng(seed)
    x, ijk = _lattice(nx, ny, nz, None)
    n = len(x)
    half = nx // 2
    s = int(rng.integers(0, nx - half + 1))
    xp = (ijk[:, 0] >= s) & (ijk[:, 0] < s + half)
    E = 5e8 * np.exp(0.3 * rng.standard_normal(n))
    ys = 2e5 * np.exp(0.4 * rng.standard_normal(n))
    labels = np.where(xp, 2, 1).astype(np.int32)
    damage = np.where(xp, 1.0, 0.0)
    dom = {"x": x, "particle_volume": np.full(n, VOL), "E": E, "ys": ys, "materialLabel": labels,
           "damage": damage, "radius": np.full(n, 0.25)}
    ov = _common_overrides(5e-5, 2000, 10)
    ov.update({"boundary_padding_mask": "000001", "grid_padding": 4.0, "staticParticleFriction": 0.6,
               "dynamicParticleFriction": 0.4})
    return dom, ov


SCENES = {"block_collapse": block_collapse, "undercut_slab": undercut_slab, "contact_heavy": contact_heavy}
