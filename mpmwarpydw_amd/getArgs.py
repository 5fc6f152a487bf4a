"""Configuration of a run: the flags, defaults and JSON overlay rules of the
reference's utils/getArgs.py (get_args :13-153), so that the reference's JSON
configs load unchanged.

Precedence is the reference's: command line > JSON file > defaults, and JSON
keys that are not flags are ignored (getArgs.py:137-148).  ``get_args`` takes
an explicit argv and an optional dict overlay so that tests and benchmarks can
build configurations without touching sys.argv.
"""
from __future__ import annotations

import argparse
import json
import os

# (flag, type, default) -- getArgs.py:22-123, in the reference's order
FLAGS = [
    ("dt", float, 0.0), ("dtxpbd", float, 1e-2), ("bigStepDuration", float, 10.0), ("xpbdOnlyDuration", float, 0.0),
    ("bigSteps", int, 100), ("residualThreshold", float, 1e-8), ("damage_stall_threshold", float, 1e-9),
    ("damage_stall_duration", float, 0.1), ("xpbd_sleep_termination_ratio", float, 0.95),
    ("rpic_damping", float, 0.2), ("grid_v_damping_scale", float, 1.1), ("update_cov", int, 1),
    ("render", int, 0), ("render_interval", float, 0.1), ("color_mode", str, "effective_ys"), ("saveFlag", int, 0),
    ("outputFolder", str, "./output/"), ("restart", str, None), ("open_paraview", int, 0),
    ("domainFile", str, "./exampleDomains/annular_arch_particles.h5"), ("grid_padding", float, 25.0),
    ("grid_particle_spacing_scale", float, 2.0), ("boundary_padding_mask", str, "111101"),
    ("density", float, 3000.0), ("E", float, 1e8), ("nu", float, 0.2), ("constitutive_model", int, 1),
    ("ys", float, 3e8), ("alpha", float, 0.3), ("hardening", float, 0.0), ("softening", float, 0.0),
    ("eta_shear", float, 1e5), ("eta_bulk", float, 1e5), ("volumetric_locking_correction", int, 1),
    ("gravity", float, -9.81), ("K0", float, 0.5), ("z_top", float, None), ("initialise_geostatic", int, 1),
    ("gravity_pulse_factor", float, 1.0), ("gravity_pulse_duration", float, 0.0), ("boundFriction", float, 0.2),
    ("boundaryCondition", str, "friction"), ("boundRestitution", float, 0.0), ("eff", float, 0.05),
    ("strainCriteria", float, 0.05), ("xpbd_relaxation", float, 1.0), ("xpbd_mpm_coupling_strength", float, 0.25),
    ("xpbd_deactivation_z_datum", float, None), ("mpm_contact_transition_lock", int, 0),
    ("dynamicParticleFriction", float, 0.05), ("staticVelocityThreshold", float, 1e-5),
    ("staticParticleFriction", float, 0.1), ("xpbd_iterations", int, 4), ("particle_cohesion", float, 0.0),
    ("sleepThreshold", float, 0.5), ("swellingRatio", float, 0.2), ("swellingActivationFactor", float, 1.0),
    ("swellingMaxFactor", float, 2.0), ("particle_v_max", float, 1000.0),
]
# the reference's --boundaryCondition choices (getArgs.py:96); its JSON configs
# also select the two plane-strain modes (simulationRoutines.py:190-228)
BOUNDARY_CONDITIONS = ["friction", "friction_gradual", "restitution", "absorbing", "plane_strain_y",
                       "plane_strain_y_global"]


def make_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="MPM-XPBD simulation parameters (MI355X hot path)")
    p.add_argument("--config", type=str, help="Path to JSON config file")
    for name, typ, default in FLAGS:
        kw = {"type": typ, "default": default}
        if name == "boundaryCondition":
            kw["choices"] = BOUNDARY_CONDITIONS
        p.add_argument(f"--{name}", **kw)
    return p


def get_args(argv=None, overrides: dict | None = None, quiet: bool = True):
    """Parse ``argv`` (default: no flags), overlay the JSON of --config and then
    ``overrides``, keeping every flag that was given on the command line."""
    argv = list(argv or [])
    parser = make_parser()
    args, _unknown = parser.parse_known_args(argv)
    cli = {a.lstrip("-").replace("-", "_") for a in argv if a.startswith("--")}
    cfg = {}
    if args.config and os.path.exists(args.config):
        with open(args.config) as f:
            cfg = json.load(f)
    cfg.update(overrides or {})
    for k, v in cfg.items():
        if k not in cli and hasattr(args, k):
            setattr(args, k, v)
    if not quiet:
        print("\n===== Loaded Simulation Parameters =====")
        for name, _, _ in FLAGS:
            print(f"{name}: {getattr(args, name)}")
        print("=======================================\n")
    return args
