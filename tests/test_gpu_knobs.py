"""The launch-shape knobs of the fused substep (mpmx_set_tuning) change where
and how the work runs, never the results: every combination gives the bits of
the default one over a dozen substeps with XPBD steps between them."""
import numpy as np
import pytest

from state_utils import host_state, make_sim

pytestmark = pytest.mark.gpu

# (split_stress, g2p_tiles, xcd_tiles, p2g_chunk knob)
COMBOS = [(1, 1, 1, 0), (1, 2, 1, 0), (1, 2, 0, 0), (1, 0, 1, 0), (1, 0, 0, 0), (0, 1, 1, 0), (0, 2, 1, 0),
          (0, 0, 1, 0), (1, 1, 1, 1), (1, 2, 0, 2), (1, 1, 2, 0), (0, 2, 2, 0)]
# the slab at half size has more than 200 tiles (50 x 25 x 12 cells and more), so
# runs of 16 tiles in turn, which need at least 128, move blocks there
SLAB_COMBOS = [(1, 1, 1, 0), (1, 1, 2, 0), (1, 2, 2, 1), (0, 1, 2, 0), (1, 1, 0, 0)]


@pytest.mark.parametrize("scene,combos", [("block_collapse", COMBOS), ("undercut_slab", SLAB_COMBOS)])
def test_knobs_leave_results_unchanged(scene, combos):
    from mpmwarpydw_amd import autotune
    from mpmwarpydw_amd import simulationRoutines as SR
    ref = None
    try:
        for split, g2p, xcd, chunk in combos:
            autotune.set_split_stress(bool(split))
            autotune.set_g2p_tiles(g2p)
            autotune.set_xcd_tiles(xcd)
            autotune.set_p2g_chunk(chunk)
            S = make_sim(scene, seed=50)
            print(f"[knobs] {scene} split={split} g2p_tiles={g2p} xcd={xcd} p2g_chunk={chunk}", flush=True)
            for counter in range(1, 13):
                S.substep(counter)
            assert SR.misbinned_count(S.sim, S.mpm) == 0
            st = host_state(S)
            if ref is None:
                ref = st
                continue
            for k in ref:
                assert np.array_equal(st[k], ref[k]), (split, g2p, xcd, chunk, k)
    finally:
        autotune.set_split_stress(True)
        autotune.set_g2p_tiles(True)
        autotune.set_xcd_tiles(2)
        autotune.set_p2g_chunk(0)


@pytest.mark.parametrize("orders", [(False, True), (False,)])
def test_autotune_leaves_state_at_start_in_its_order(orders):
    """tune() times every candidate from the same starting state, in input
    order and then in tile order, and leaves the state as it started in the
    order of the best candidate with that candidate's knobs in force: the
    state then runs on to the bits of a state that never went through the
    tuner."""
    from mpmwarpydw_amd import autotune
    try:
        S = make_sim(seed=51)
        R = make_sim(seed=51)
        counter = [1]

        def step():
            S.substep(counter[0])
            counter[0] += 1

        log = autotune.tune(S, step, steps=2, warmup=1, trials=1, orders=orders)
        n_cand = len(log["candidates"])
        print(f"[autotune] orders {orders}: {n_cand} candidates, best {log['best']}", flush=True)
        # split: 3 chunk sizes x 3 G2P forms x 3 XCD mappings; fused: 3 G2P forms x 3 XCD mappings
        assert n_cand == len(orders) * (3 * 3 * 3 + 3 * 3)
        assert log["kept_reordered"] == log["best"]["reorder"]
        cur = autotune.current()
        for k in ("xcd_tiles", "split_stress", "g2p_tiles"):
            assert cur[k] == log["best"][k], k
        if log["kept_reordered"]:  # and the way back to input order keeps the state too
            S.restore_input_order()
            S.reorder()
        for c in range(1, 13):
            S.substep(c)
            R.substep(c)
        a, b = host_state(S), host_state(R)
        for k in a:
            assert np.array_equal(a[k], b[k]), k
    finally:
        autotune.apply({"xcd_tiles": 2, "split_stress": 1, "g2p_tiles": 1, "p2g_chunk": 512})
