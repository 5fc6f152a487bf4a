"""CPU checks of the boundary and the host logic: libmpmx.so loads and exports
every entry point include/mpmx.h declares, the ctypes signatures cover them,
the config flags/defaults and JSON precedence follow getArgs.py, and the
driver's setup arithmetic (runMPMYDW.py:22-169) gives the documented grids."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mpmx.h")).read()
    return sorted(set(re.findall(r"\b(mpmx_[A-Za-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mpmwarpydw_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from mpmwarpydw_amd import build
        build.build()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines()}
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert set(declared_symbols()) == set(_lib.SIGNATURES), "ctypes table out of sync with the header"
    _lib.lib.load()  # dlopen resolves without a GPU


def test_ctypes_structs_match_the_library_layout():
    """The ctypes mirrors of mpmx_particles / mpmx_grid / mpmx_step_params have
    the size and field offsets the library was compiled with (a struct passed
    by pointer that is shorter than the C one makes the library read past it),
    and unknown field names are refused."""
    import ctypes as C
    from mpmwarpydw_amd._lib import Grid, Particles, Scene, StepParams, XpbdBounds, lib
    out = (C.c_long * 12)()
    k = lib.raw("mpmx_struct_layout")(out, 12)
    assert k == 10
    assert list(out[:10]) == [C.sizeof(Particles), Particles.order.offset, C.sizeof(Grid), C.sizeof(StepParams),
                              StepParams.bc.offset, Particles.scene.offset, C.sizeof(Scene), C.sizeof(XpbdBounds),
                              XpbdBounds.delta.offset, XpbdBounds.xpbd_count.offset]
    with pytest.raises(TypeError):
        Particles(n=1, not_a_field=1)


def test_workspace_queries_without_gpu():
    from mpmwarpydw_amd._lib import lib
    assert lib.raw("mpmx_p2g_workspace_bytes")(1000, 20, 20, 20) > 0
    assert lib.raw("mpmx_p2g_workspace_bytes")(10, 2, 5, 5) == 0  # grids need 3 nodes per axis
    assert lib.raw("mpmx_hashgrid_workspace_bytes")(10, 128, 128, 128) > 128 ** 3 * 8
    assert lib.raw("mpmx_reduce_workspace_bytes")(10) > 0


def test_argument_errors_are_reported():
    from mpmwarpydw_amd._lib import MpmxError, lib
    with pytest.raises(MpmxError, match="n < 0"):
        lib("mpmx_compute_mu_lam_bulk", -1, None, None, None, None, None, None)
    with pytest.raises(MpmxError, match="unsupported boundary"):
        lib("mpmx_grid_update", 4, 4, 4, 1, 1, 1, 0.0, 0.0, 0.0, 1.0, 1.0, 7, 0.0, 0.0, 2, 0.0, None)
    with pytest.raises(MpmxError, match="unsupported boundary"):
        lib("mpmx_collide_bounds_variant", 1, 4, 4, 4, 0, 0.0, 0.0, 0.0, 2, 0.0, None)
    assert lib.raw("mpmx_vlc_workspace_bytes")(4, 4, 4) >= 2 * 64 * 8


def test_get_args_defaults_and_precedence(tmp_path):
    from mpmwarpydw_amd.getArgs import get_args
    a = get_args([])
    assert (a.rpic_damping, a.grid_v_damping_scale, a.update_cov, a.volumetric_locking_correction) == (0.2, 1.1, 1, 1)
    assert (a.eta_shear, a.swellingRatio, a.xpbd_iterations, a.boundFriction, a.grid_padding) == (1e5, 0.2, 4, 0.2,
                                                                                                  25.0)
    assert a.boundary_padding_mask == "111101" and a.grid_particle_spacing_scale == 2.0
    cfg = tmp_path / "c.json"
    cfg.write_text(json.dumps({"dt": 1e-4, "boundFriction": 0.7, "render_backend": "x", "E": 2e9}))
    b = get_args(["--config", str(cfg), "--E", "3e9"])
    assert b.dt == 1e-4 and b.boundFriction == 0.7 and b.E == 3e9  # CLI beats JSON
    assert not hasattr(b, "render_backend")  # unknown JSON keys are dropped


def test_grid_geometry_and_time_step():
    from mpmwarpydw_amd import driver, getArgs, scenes
    for name, n, dims in (("block_collapse", 32768, (32, 32, 26)), ("undercut_slab", 960000, (104, 54, 31)),
                          ("contact_heavy", 2000000, (104, 54, 56))):
        dom, ov = scenes.SCENES[name]()
        a = getArgs.get_args([], ov)
        dx, mn, mx, d = driver.grid_geometry(dom["x"], dom["particle_volume"], a)
        assert len(dom["x"]) == n and d == dims and abs(dx - 0.5) < 1e-12
    # dx follows mean(vol)**0.33 * scale (runMPMYDW.py:24-25), not the cube root
    a = getArgs.get_args([])
    dx, *_ = driver.grid_geometry(np.zeros((2, 3)) + [[0, 0, 0], [1, 1, 1]], np.full(2, 0.25 ** 3), a)
    assert np.isclose(dx, (0.25 ** 3) ** 0.33 * 2.0)
    # CFL auto dt and the integer XPBD ratio (runMPMYDW.py:110-169)
    a = getArgs.get_args([], {"dt": 0, "dtxpbd": 1e-3})
    dt, ratio = driver.time_step(a, 0.5, np.full(4, 1e9), np.full(4, 0.25), np.full(4, 2500.0))
    c_p = np.sqrt((1e9 / 1.5 + 4 * 1e9 / 2.5 / 3) / 2500.0)
    assert ratio == int(np.round(1e-3 / (0.1 * 0.5 / c_p))) and np.isclose(dt * ratio, 1e-3)


def test_alpha_mapping():
    from mpmwarpydw_amd.scenes import alpha_from_friction_angle
    assert np.isclose(alpha_from_friction_angle(30.0), 1.0 / (np.sqrt(3) * 2.5))


def test_ctypes_signatures_have_the_declared_arity():
    """Every ctypes signature lists one type per parameter of its declaration in
    include/mpmx.h (a missing or extra entry shifts every later argument)."""
    from mpmwarpydw_amd._lib import SIGNATURES
    src = open(os.path.join(ROOT, "include", "mpmx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    bad = []
    for m in re.finditer(r"\b(mpmx_[A-Za-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src):
        name, params = m.group(1), m.group(2).strip()
        if name not in SIGNATURES:
            continue
        arity = 0 if params in ("", "void") else params.count(",") + 1
        if len(SIGNATURES[name]) != arity:
            bad.append((name, arity, len(SIGNATURES[name])))
    assert not bad, bad
