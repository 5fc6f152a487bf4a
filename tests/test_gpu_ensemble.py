"""An ensemble of uneven scenes advanced by one launch per kernel gives every
scene the bits it gets when run alone through the fused substep."""
import numpy as np
import pytest

from state_utils import perturb_state

pytestmark = pytest.mark.gpu

SPECS = [(3000, 1e8, 2e4, 30.0, 1), (9000, 3e8, 5e4, 25.0, 2), (1500, 5e7, 1e4, 38.0, 3), (600, 1e9, 1e6, 21.0, 4)]
FIELDS = [("sim", "particle_x"), ("sim", "particle_v"), ("sim", "materialLabel"), ("mpm", "particle_F"),
          ("mpm", "particle_F_trial"), ("mpm", "particle_C"), ("mpm", "particle_stress"),
          ("mpm", "particle_accumulated_strain"), ("mpm", "particle_damage"), ("mpm", "ys"), ("mpm", "mu")]


def _scenes(perturb, **extra):
    from mpmwarpydw_amd import driver, getArgs
    from mpmwarpydw_amd.ensemble import block_scene
    out = []
    for k, spec in enumerate(SPECS):
        dom, ov = block_scene(*spec)
        ov.update(extra)
        S = driver.setup(getArgs.get_args([], ov), dom)
        if perturb:
            perturb_state(S, 60 + k)
        out.append(S)
    return out


VARIANTS = {"friction": {}, "restitution+vlc": {"boundaryCondition": "restitution", "boundRestitution": 0.4,
                                                "volumetric_locking_correction": 1}}


@pytest.mark.parametrize("reorder,variant", [(False, "friction"), (True, "friction"), (True, "restitution+vlc")])
def test_ensemble_equals_scenes_alone(reorder, variant):
    from mpmwarpydw_amd import simulationRoutines as SR
    from mpmwarpydw_amd.ensemble import Ensemble
    alone = _scenes(True, **VARIANTS[variant])
    ens = Ensemble(_scenes(True, **VARIANTS[variant]))
    print(f"[ensemble] {variant} reorder={reorder}: {len(SPECS)} scenes, {ens.n} particles, "
          f"{ens.node_offsets[-1]} nodes", flush=True)
    if reorder:
        ens.reorder()
    for _ in range(6):
        ens.substep()
        for S in alone:
            SR.mpmSimulationStepFused(S.sim, S.mpm, S.xpbd)
    for k, S in enumerate(alone):
        for obj, name in FIELDS:
            want = getattr(getattr(S, obj), name).numpy()
            got = ens.scene_array(k, name, obj)
            assert np.array_equal(got, want), (k, name)
        for name in ("grid_m", "grid_v_in", "grid_v_out"):
            want = getattr(S.mpm, name).numpy()
            got = ens.scene_grid(k, name)
            assert np.array_equal(got, want.reshape(got.shape)), (k, name)
