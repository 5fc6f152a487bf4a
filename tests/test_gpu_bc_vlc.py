"""GPU parity of the grid boundary conditions other than "friction"
(mpmRoutines.py:997-1340) and of the volumetric locking correction
(mpmRoutines.py:818-911) against the FP64 oracle, kernel by kernel.

Tolerances: grid velocities 1e-5 relative to the array's scale (FP32
arithmetic); the correction's cell sums are 64-bit fixed point on the device,
so C agrees to 1e-5 of its scale and the cell arrays to 1e-6, and a run with
the particle arrays permuted gives the same bits."""
import numpy as np
import pytest

from conftest import rel_err
from state_utils import host_state, make_sim

pytestmark = pytest.mark.gpu
TOL = 1e-5

VARIANTS = ["friction_gradual", "restitution", "absorbing", "plane_strain_y", "plane_strain_y_global"]


def _grid_state(S, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(S.mpm.gridDims + (3,))
    v[rng.random(S.mpm.gridDims) < 0.1] = 0.0  # nodes at rest
    S.mpm.grid_v_out.assign(v)
    return v.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("kind", VARIANTS)
@pytest.mark.parametrize("pad", [0, 2, 3])
def test_bc_variant_kernel_parity(o64, kind, pad):
    from mpmwarpydw_amd import mpmRoutines as M
    from mpmwarpydw_amd.runtime import launch
    S = make_sim(seed=30, perturb=False)
    sim, mpm = S.sim, S.mpm
    dims = list(mpm.gridDims)
    v = _grid_state(S, 30 + pad)
    mu, dt, g, e = 0.4, float(sim.dt), 9.81, 0.6
    print(f"[bc variant] {kind} pad={pad} dims={dims}", flush=True)
    if kind == "friction_gradual":
        launch(M.collideBoundsGradualFriction, mpm.gridDims, inputs=[mpm.grid_v_out, *dims, mu, dt, g, pad])
    elif kind == "restitution":
        launch(M.collideBoundsRestitution, mpm.gridDims, inputs=[mpm.grid_v_out, *dims, e, mu, dt, g, pad])
    elif kind == "absorbing":
        launch(M.collideBoundsAbsorbing, mpm.gridDims, inputs=[mpm.grid_v_out, *dims, pad])
    elif kind == "plane_strain_y":
        launch(M.collideBoundsPlaneStrainY, mpm.gridDims, inputs=[mpm.grid_v_out, *dims, mu, dt, g, pad])
    else:
        launch(M.collideBoundsPlaneStrainYGlobal, mpm.gridDims, inputs=[mpm.grid_v_out, *dims, mu, dt, g, pad])
    want = v.copy()
    o64.collide_bounds_variant(kind, want, *dims, np.float32(mu), np.float32(dt), np.float32(g), pad,
                               restitution=np.float32(e))
    got = mpm.grid_v_out.numpy()
    # absorbing divides by the padding: with 0 its boundary layers are NaN in
    # the reference too (mpmRoutines.py:1147), and must be NaN at the same nodes
    nan = np.isnan(want)
    assert np.array_equal(np.isnan(got), nan)
    assert nan.any() == (kind == "absorbing" and pad == 0)
    assert rel_err(np.where(nan, 0.0, got), np.where(nan, 0.0, want)) < TOL
    if pad > 0 and kind != "friction_gradual":
        assert not np.array_equal(want, v)  # the boundary did something


def _vlc_state(seed):
    return make_sim(seed=seed)  # about 8 particles per cell, random C, mixed labels


def test_vlc_kernels_parity(o64):
    from mpmwarpydw_amd import simulationRoutines as SR
    S = _vlc_state(31)
    sim, mpm = S.sim, S.mpm
    H = host_state(S)
    print(f"[vlc kernels] n={sim.nPoints} dims={mpm.gridDims}", flush=True)
    SR.volumetric_locking_phase(sim, mpm)
    cd, cv = o64.volumetric_locking_correction(H["activeLabel"], H["materialLabel"], H["particle_x"],
                                               H["particle_C"], H["particle_vol"], np.float32(sim.invdx),
                                               np.float32(sim.minBounds), *mpm.gridDims)
    assert rel_err(mpm.cell_vol_sum.numpy(), cv) < 1e-6
    assert rel_err(mpm.cell_div_v_weighted.numpy(), cd) < 1e-6
    C_gpu = mpm.particle_C.numpy()
    assert rel_err(C_gpu, H["particle_C"]) < TOL
    # the correction moved the diagonals of the active MPM particles only
    C0 = host_state(_vlc_state(31))["particle_C"].reshape(-1, 3, 3)
    moved = np.abs(C_gpu.reshape(-1, 3, 3) - C0).max(axis=(1, 2)) > 0
    mpm_active = (H["activeLabel"] == 1) & (H["materialLabel"] <= 1)
    assert not moved[~mpm_active].any()
    assert moved[mpm_active].mean() > 0.5
    off = ~np.eye(3, dtype=bool)
    assert np.array_equal(C_gpu.reshape(-1, 3, 3)[:, off], C0[:, off].astype(np.float32))
    # the cell-averaged divergence is what every corrected particle now carries
    tr = np.trace(C_gpu.reshape(-1, 3, 3).astype(np.float64), axis1=1, axis2=2)
    x = H["particle_x"].reshape(-1, 3).astype(np.float32)  # the device's float32 cell arithmetic
    cell = ((x - np.asarray(sim.minBounds, np.float32)) * np.float32(sim.invdx)).astype(np.int64)
    gx, gy, gz = mpm.gridDims
    k = (cell[:, 0] * gy + cell[:, 1]) * gz + cell[:, 2]
    dc = (cd.ravel() / np.where(cv.ravel() > 1e-20, cv.ravel(), 1.0))[k]
    assert np.abs(tr - dc)[mpm_active].max() < 1e-5 * np.abs(C0).max()


def test_vlc_is_order_independent():
    """The correction's cell sums are exact integers: permuting the particle
    arrays (Simulation.reorder) leaves every result bit-identical."""
    A = make_sim(seed=32, volumetric_locking_correction=1)
    B = make_sim(seed=32, volumetric_locking_correction=1)
    B.reorder()
    print(f"[vlc order] n={A.sim.nPoints}", flush=True)
    for counter in range(1, 6):
        A.substep(counter)
        B.substep(counter)
    for name in ("particle_C", "particle_x", "particle_v", "particle_F"):
        obj = A.sim if name in ("particle_x", "particle_v") else A.mpm
        objb = B.sim if name in ("particle_x", "particle_v") else B.mpm
        assert np.array_equal(getattr(objb, name).numpy(), getattr(obj, name).numpy()), name
