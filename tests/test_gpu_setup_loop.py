"""GPU parity of the setup and end-of-big-step kernels (geostatic F,
mpmRoutines.py:52-203; deactivation below the datum, :1446-1475) against the
FP64 oracle, and one short run of the big-step loop on the device.

Tolerance for F: FP32 stores exp(eps) near 1, so |F_gpu - F_oracle| may reach
one float32 ulp of 1 (1.2e-7) on top of 1e-5 of |F - I|."""
import os

import numpy as np
import pytest

from state_utils import host_state, make_sim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("plane_strain", [False, True])
def test_geostatic_F_kernel_parity(o64, plane_strain):
    from mpmwarpydw_amd import mpmRoutines as M
    from mpmwarpydw_amd.runtime import launch
    S = make_sim(seed=40, perturb=False)
    sim, mpm = S.sim, S.mpm
    n = sim.nPoints
    rng = np.random.default_rng(40)
    mu = rng.uniform(5e6, 5e8, n)
    lam = rng.uniform(5e6, 9e8, n)
    mpm.mu.assign(mu)
    mpm.lam.assign(lam)
    H = host_state(S)
    x = H["particle_x"]
    z_top = float(np.float32(np.median(x[:, 2])))  # particles above and below it
    g, K0 = 9.81, 0.55
    print(f"[geostatic] plane_strain={plane_strain} n={n}", flush=True)
    k = M.initialize_geostatic_F_planestrain if plane_strain else M.initialize_geostatic_F
    launch(k, n, inputs=[sim.particle_x, mpm.particle_F, mpm.mu, mpm.lam, sim.particle_density, g, z_top, mpm.K0])
    F = np.zeros((n, 9))
    o64.initialize_geostatic_F(x, F, H["mu"], H["lam"], H["particle_density"], np.float32(g), np.float32(z_top),
                               np.float32(mpm.K0), plane_strain)
    got = mpm.particle_F.numpy().reshape(n, 9)
    dev = np.abs(F - np.eye(3).ravel()).max()
    assert dev > 1e-6  # a real prestrain
    assert np.abs(got - F).max() <= 1.2e-7 + 1e-5 * dev


def test_setup_with_geostatic_initialisation(o64):
    S = make_sim(seed=41, perturb=False, initialise_geostatic=1, K0=0.5)
    sim, mpm = S.sim, S.mpm
    H = host_state(S)
    n = sim.nPoints
    z_top = float(np.float32(H["particle_x"][:, 2].max()))
    F = np.zeros((n, 9))
    o64.initialize_geostatic_F(H["particle_x"], F, H["mu"], H["lam"], H["particle_density"],
                               np.float32(abs(S.args.gravity)), np.float32(z_top), np.float32(0.5))
    dev = np.abs(F - np.eye(3).ravel()).max()
    assert np.abs(H["particle_F"].reshape(n, 9) - F).max() <= 1.2e-7 + 1e-5 * dev
    assert np.array_equal(H["particle_F_trial"], H["particle_F"])  # F_trial copies F (runMPMYDW.py:448-451)


def test_deactivate_and_teleport_parity(o64):
    S = make_sim(seed=42)  # mixed labels and activity
    sim = S.sim
    H = host_state(S)
    z_datum = float(np.float32(np.quantile(H["particle_x"][:, 2], 0.4)))
    print(f"[teleport] n={sim.nPoints} z_datum={z_datum}", flush=True)
    S.args.xpbd_deactivation_z_datum = z_datum
    S.deactivate_below_datum()
    teleport = 2.0 * float(S.geom.maxBounds[2] - S.geom.minBounds[2])
    act = H["activeLabel"].copy()
    o64.deactivate_and_teleport(H["particle_x"], H["particle_v"], act, H["materialLabel"], np.float32(z_datum),
                                np.float32(teleport))
    G = host_state(S)
    assert np.array_equal(G["activeLabel"], act)
    assert (act != H["activeLabel"]).sum() > 10
    assert np.abs(G["particle_x"] - H["particle_x"]).max() <= 1e-6 * teleport
    assert np.array_equal(G["particle_v"] == 0, H["particle_v"] == 0)


def test_big_step_loop_on_device():
    """Two big steps of a small scene: phase 1 to nSteps, the XPBD-only phase,
    the checkpoint hook and the deactivation, with the loop's FP64 scalars."""
    S = make_sim(seed=43, bigSteps=2, xpbdOnlyDuration=3.5e-3, xpbd_sleep_termination_ratio=1.0)
    sim = S.sim
    sim.nSteps = 25
    print(f"[loop] n={sim.nPoints} nSteps={sim.nSteps} xpbdOnlySteps={sim.xpbdOnlySteps}", flush=True)
    seen = []
    rec = S.run(checkpoint=lambda s, b: seen.append((b, s.loop_scalars())))
    assert [r["phase1_steps"] for r in rec] == [25, 25]
    assert [r["phase2_steps"] for r in rec] == [sim.xpbdOnlySteps] * 2
    assert sim.xpbdOnlySteps == 3
    assert [b for b, _ in seen] == [0, 1]
    res, dmg = seen[-1][1]
    assert np.isfinite(res) and res > 0 and 0 <= dmg <= 1
    x = sim.particle_x.numpy()
    assert np.isfinite(x[sim.activeLabel.numpy() == 1]).all()


def test_checkpoint_restart_continues_bit_for_bit(tmp_path):
    """save_mpm with the restart data, load_restart_from_vtp and apply_restart
    into a state set up from the same domain: every restored array equals the
    saved one, and the next substep gives the same bits as the run that was
    saved (after applying the restart's F_trial = F to it too)."""
    from mpmwarpydw_amd import driver, fs5PlotUtils
    from state_utils import F_FIELDS_MPM, F_FIELDS_SIM, F_FIELDS_XPBD
    A = make_sim(seed=44, outputFolder=str(tmp_path))
    for counter in range(1, 4):
        A.substep(counter)
    A.sim.t = 0.0123456789
    print(f"[restart] n={A.sim.nPoints}", flush=True)
    prefix = fs5PlotUtils.save_mpm(A.sim, A.mpm, 3, 7, include_restart_data=True, xpbd=A.xpbd)
    assert os.path.exists(prefix + "_grid.vti")
    r = fs5PlotUtils.load_restart_from_vtp(prefix + "_particles.vtp")
    B = make_sim(seed=44)
    assert driver.apply_restart(B, r) == 4
    assert B.sim.t == A.sim.t
    # what the reference's checkpoint leaves out comes over by hand
    B.xpbd.particle_v_initial_xpbd.assign(A.xpbd.particle_v_initial_xpbd.numpy())
    B.mpm.particle_cov.assign(A.mpm.particle_cov.numpy())
    A.mpm.particle_F_trial.assign(A.mpm.particle_F.numpy())
    grid = {"grid_m", "grid_v_in", "grid_v_out"}
    for objs, names in (((A.sim, B.sim), F_FIELDS_SIM), ((A.mpm, B.mpm), F_FIELDS_MPM),
                        ((A.xpbd, B.xpbd), F_FIELDS_XPBD)):
        for k in names:
            if k in grid or k == "particle_stress":
                continue
            assert np.array_equal(getattr(objs[1], k).numpy(), getattr(objs[0], k).numpy()), k
    A.substep(4)
    B.substep(4)
    a, b = host_state(A), host_state(B)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
