"""A block collapse run for 20 substeps of the reference's loop
(mpmSimulationStep every substep, xpbdSimulationStep at counters 10 and 20,
runMPMYDW.py:677-682), checked two ways against the oracle.

* Step by step: every substep and XPBD step of the GPU run is compared with the
  oracle's step from the GPU's own state before it (downloaded in the
  reference's particle order).  The GPU carries state that the download does
  not show -- the particle order after the reorder, the bins built by the
  previous G2P, the hash tables and contact records of the XPBD step -- so a
  step that reads stale hidden state shows here as a step error, however
  small the single-step tests' states were.  Bound: 1e-5 of each array's scale,
  labels identical.
* Free running: the oracle advanced from the starting state alone, nothing fed
  back.  The two runs round differently, and the grid nodes at the rim of the
  block, whose masses are the cube of a small weight, turn a 1e-7 difference of
  momentum into a 1e-4 difference of node velocity and of the velocity gradient
  of the particles next to them (the reference's FP32 run has the same
  sensitivity).  Bound: 1e-3 of each array's scale against FP64, labels
  identical."""
import numpy as np
import pytest

from conftest import rel_err
from state_utils import host_state, oracle_substep, oracle_xpbd_step

pytestmark = pytest.mark.gpu
STEP_TOL = 1e-5
DRIFT = 1e-3
KEYS = ("particle_x", "particle_v", "particle_C", "particle_F", "particle_F_trial", "particle_stress",
        "particle_damage", "particle_accumulated_strain", "ys", "grid_m", "grid_mv_out", "particle_cumDist_xpbd")


def _with_momentum(H):
    """Grid velocity compared as momentum m v: a node at the rim of the block
    holds a mass of order 1e-6 of an interior node's, which turns the rounding
    of its momentum into a large velocity error (v = momentum / m) that no
    particle feels (its weight on that node is as small)."""
    H = dict(H)
    H["grid_mv_out"] = H["grid_v_out"] * H["grid_m"][..., None]
    return H


def _compare(G, H, tol, what):
    G, H = _with_momentum(G), _with_momentum(H)
    diff = np.flatnonzero(G["materialLabel"] != H["materialLabel"])
    assert diff.size == 0, (what, diff[:8])
    worst = max(((rel_err(G[k], H[k]), k) for k in KEYS))
    assert worst[0] < tol, (what, worst)
    return worst


def _compare_xpbd(G, H, dt, what):
    """The XPBD step's bounds (test_gpu_xpbd.py): x to 16 FP32 ulps of the
    domain's largest coordinate, v = (x_new - x_orig) / dt to 16 ulps / dt,
    the rest to 1e-5 of its scale."""
    G, H = _with_momentum(G), _with_momentum(H)
    diff = np.flatnonzero(G["materialLabel"] != H["materialLabel"])
    assert diff.size == 0, (what, diff[:8])
    ulp = float(np.spacing(np.float32(np.max(np.abs(H["particle_x"])))))
    assert np.max(np.abs(G["particle_x"] - H["particle_x"])) <= 16 * ulp, what
    assert np.max(np.abs(G["particle_v"] - H["particle_v"])) <= 16 * ulp / dt + STEP_TOL * np.max(
        np.abs(H["particle_v"])), what
    assert np.max(np.abs(G["particle_cumDist_xpbd"] - H["particle_cumDist_xpbd"])) <= 16 * ulp, what
    rest = [k for k in KEYS if k not in ("particle_x", "particle_v", "particle_cumDist_xpbd")]
    worst = max(((rel_err(G[k], H[k]), k) for k in rest))
    assert worst[0] < STEP_TOL, (what, worst)
    return worst


@pytest.mark.parametrize("fused", [True, False])
def test_block_collapse_run(o64, fused):
    from mpmwarpydw_amd import driver, getArgs, scenes
    from mpmwarpydw_amd import simulationRoutines as SR
    from state_utils import perturb_state
    dom, ov = scenes.block_collapse(scale=0.5, seed=0)
    S = driver.setup(getArgs.get_args([], ov), dom, fused=fused)
    perturb_state(S, 14)
    # damage grows 2000 times slower except on the brittle particles, so that
    # no label decision of the free run comes within rounding of its threshold
    S.mpm.particle_F_trial.assign(S.mpm.particle_F.numpy())
    sc = S.mpm.strainCriteria.numpy()
    S.mpm.strainCriteria.assign(np.where(sc < 1e-3, sc, 100.0))
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    free = host_state(S)
    start_labels = free["materialLabel"].copy()
    mpm_step = SR.mpmSimulationStepFused if fused else SR.mpmSimulationStep
    xpbd_step = SR.xpbdSimulationStepFused if fused else SR.xpbdSimulationStep
    print(f"[block collapse run] fused={fused} n={sim.nPoints} xpbd every {sim.mpmStepsPerXpbdStep}", flush=True)
    worst_step = (0.0, "")
    for counter in range(1, 21):
        H = host_state(S)
        mpm_step(sim, mpm, xpbd)
        oracle_substep(o64, H, S)
        G = host_state(S)
        worst_step = max(worst_step, _compare(G, H, STEP_TOL, f"substep {counter}"))
        oracle_substep(o64, free, S)
        if counter % sim.mpmStepsPerXpbdStep == 0:
            xpbd_step(sim, mpm, xpbd)
            oracle_xpbd_step(o64, G, S)
            X = host_state(S)
            _compare_xpbd(X, G, float(np.float32(sim.dtxpbd)), f"xpbd step {counter}")
            oracle_xpbd_step(o64, free, S)
    G, free = _with_momentum(host_state(S)), _with_momentum(free)
    drift = {k: rel_err(G[k], free[k]) for k in KEYS if k != "particle_v"}
    # XPBD velocities are position differences over dtxpbd: 16 ulps of x / dt as in the step bound
    ulp = float(np.spacing(np.float32(np.max(np.abs(free["particle_x"])))))
    vscale = np.max(np.abs(free["particle_v"]))
    drift["particle_v"] = max(0.0, np.max(np.abs(G["particle_v"] - free["particle_v"])) -
                              16 * ulp / float(np.float32(sim.dtxpbd))) / vscale
    print(f"  worst step error {worst_step[0]:.1e} ({worst_step[1]}); free-run drift:",
          " ".join(f"{k}={v:.1e}" for k, v in drift.items()))
    assert np.array_equal(G["materialLabel"], free["materialLabel"])
    assert max(drift.values()) < DRIFT, drift
    assert (G["materialLabel"] == 2).sum() > (start_labels == 2).sum()  # transitions happened
