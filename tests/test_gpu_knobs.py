"""The launch-shape knobs of the fused substep (mpmx_set_tuning) change where
and how the work runs, never the results: every combination gives the bits of
the default one over a dozen substeps with XPBD steps between them."""
import numpy as np
import pytest

from state_utils import host_state, make_sim

pytestmark = pytest.mark.gpu

COMBOS = [(1, 1, 1), (1, 2, 1), (1, 2, 0), (1, 0, 1), (1, 0, 0), (0, 1, 1), (0, 2, 1), (0, 0, 1)]  # (split_stress, g2p_tiles, xcd_tiles)


def test_knobs_leave_results_unchanged():
    from mpmwarpydw_amd import autotune
    from mpmwarpydw_amd import simulationRoutines as SR
    ref = None
    try:
        for split, g2p, xcd in COMBOS:
            autotune.set_split_stress(bool(split))
            autotune.set_g2p_tiles(g2p)
            autotune.set_xcd_tiles(bool(xcd))
            S = make_sim(seed=50)
            print(f"[knobs] split={split} g2p_tiles={g2p} xcd={xcd}", flush=True)
            for counter in range(1, 13):
                S.substep(counter)
            assert SR.misbinned_count(S.sim, S.mpm) == 0
            st = host_state(S)
            if ref is None:
                ref = st
                continue
            for k in ref:
                assert np.array_equal(st[k], ref[k]), (split, g2p, xcd, k)
    finally:
        autotune.set_split_stress(True)
        autotune.set_g2p_tiles(True)
        autotune.set_xcd_tiles(True)
