"""Spatial reorder of the particle arrays (Simulation.reorder) changes where each
particle sits in HBM and nothing else: bins and hash tables order particles by
their input index (sim.particle_id), so a run that reorders gives the bits of
one that never does, and the host views (numpy) show the input order."""
import numpy as np
import pytest

from state_utils import F_FIELDS_MPM, F_FIELDS_SIM, F_FIELDS_XPBD, make_sim

pytestmark = pytest.mark.gpu


def _arrays(S):
    out = {}
    for obj, names in ((S.sim, F_FIELDS_SIM), (S.mpm, F_FIELDS_MPM), (S.xpbd, F_FIELDS_XPBD)):
        for k in names:
            out[k] = getattr(obj, k).numpy()
    return out


def test_reorder_is_a_pure_layout_change():
    A = make_sim(seed=21)
    B = make_sim(seed=21)
    n = A.sim.nPoints
    ratio = A.sim.mpmStepsPerXpbdStep
    print(f"[reorder] n={n} ratio={ratio}", flush=True)
    B.reorder()
    pid = B.sim.particle_id.t.cpu().numpy()  # the slots' input indices, as stored
    assert np.array_equal(np.sort(pid), np.arange(n))
    assert not np.array_equal(pid, np.arange(n))  # the random state is not in tile order
    x_slots = B.sim.particle_x.t.cpu().numpy().T
    assert np.array_equal(x_slots[np.argsort(pid)], A.sim.particle_x.numpy())
    for counter in range(1, ratio + 4):
        if counter == 7:
            B.reorder()
        A.substep(counter)
        B.substep(counter)
    a, b = _arrays(A), _arrays(B)
    for k in a:
        assert np.array_equal(b[k], a[k]), k


def test_reorder_with_many_non_mpm_particles():
    """Active MPM particles come first, then active XPBD particles (label 2) in
    tile order, then the rest (inactive particles) in input order; with most of
    a scene in the XPBD phase the reorder stays a pure layout change."""
    A = make_sim(seed=22)
    B = make_sim(seed=22)
    n = A.sim.nPoints
    rng = np.random.default_rng(22)
    lab = A.sim.materialLabel.numpy()
    lab[rng.random(n) < 0.8] = 2
    for S in (A, B):
        S.sim.materialLabel.assign(lab)
    print(f"[reorder-rest] n={n} non_mpm={(lab == 2).sum()}", flush=True)
    B.reorder()
    pid = B.sim.particle_id.t.cpu().numpy()
    assert np.array_equal(np.sort(pid), np.arange(n))
    act = A.sim.activeLabel.numpy()
    mpm = (act == 1) & (lab <= 1)
    xp = (act == 1) & (lab == 2)
    k, m = int(mpm.sum()), int(xp.sum())
    assert mpm[pid[:k]].all() and xp[pid[k:k + m]].all()  # the block collapse lies inside the grid
    assert np.array_equal(pid[k + m:], np.flatnonzero(~mpm & ~xp))  # the rest keep input order
    assert not np.array_equal(pid[k:k + m], np.flatnonzero(xp))  # XPBD particles in tile order
    for counter in range(1, A.sim.mpmStepsPerXpbdStep + 2):
        A.substep(counter)
        B.substep(counter)
    a, b = _arrays(A), _arrays(B)
    for key in a:
        assert np.array_equal(b[key], a[key]), key
