"""GPU parity of the XPBD contact step against the FP64 oracle.

The hash grid (point order and cell table) must match the oracle bit for bit;
contact deltas and the full step agree to 1e-5 relative to the array scale
(FP32; neighbour sums are taken in the same candidate order as the oracle)."""
import numpy as np
import pytest

from conftest import rel_err
from state_utils import host_state, make_sim

pytestmark = pytest.mark.gpu
TOL = 1e-5


def xpbd_scene(seed=0, scale=0.12, radius=None, cluster=0, **over):
    """Half-XPBD lattice block (contact_heavy at reduced size), jittered.  The
    default radius 0.25 m at 0.25 m spacing overlaps every neighbour by up to
    0.25 m (a stress case for single-kernel parity); radius=0.135 gives the
    few-centimetre overlaps of a settled pile.  cluster=k packs k XPBD particles
    into a 0.1 m cube inside one hash cell (far more near candidates per cell
    than a lattice has)."""
    S = make_sim("contact_heavy", scale=scale, seed=seed, perturb=False, **over)
    rng = np.random.default_rng(seed)
    n = S.sim.nPoints
    x = S.sim.particle_x.numpy().astype(np.float64) + 0.03 * rng.standard_normal((n, 3))
    if cluster:
        xp = np.flatnonzero(S.sim.materialLabel.numpy() == 2)
        pick = rng.choice(xp, cluster, replace=False)
        centre = (np.floor(x[pick[0]] / S.sim.dx) + 0.5) * S.sim.dx
        x[pick] = centre + rng.uniform(-0.05, 0.05, (cluster, 3))
    S.sim.particle_x.assign(x)
    S.sim.particle_v.assign(0.5 * rng.standard_normal((n, 3)))
    if radius is not None:
        S.sim.particle_radius.assign(np.full(n, radius))
        S.xpbd.max_radius = radius
    act = np.ones(n, np.int32)
    act[rng.random(n) < 0.02] = 0
    S.sim.activeLabel.assign(act)
    return S


def grid_tuple(g):
    cs = g.cell_start.numpy()
    return g.point_ids.numpy(), cs[:-1].copy(), cs[1:].copy(), g.cell_width, g.dims


def test_hashgrid_bit_exact(o64):
    S = xpbd_scene(1)
    sim, xpbd = S.sim, S.xpbd
    print(f"[hashgrid] n={sim.nPoints}", flush=True)
    xpbd.particle_grid.build(sim.particle_x, sim.dx)
    ids, cs, ce, _, dims = grid_tuple(xpbd.particle_grid)
    rid, rcs, rce = o64.hashgrid_build(sim.particle_x.numpy().astype(np.float32).astype(np.float64),
                                       np.float32(sim.dx), dims)
    assert np.array_equal(ids, rid)
    nonempty = rce > rcs
    assert np.array_equal(cs[nonempty], rcs[nonempty]) and np.array_equal(ce[nonempty], rce[nonempty])
    assert np.all(ce[~nonempty] == cs[~nonempty])


@pytest.mark.parametrize("dims", [(37, 5, 70), (64, 1, 65), (3, 130, 2)])
def test_hashgrid_odd_dims_bit_exact(o64, dims):
    """Dims that are not multiples of the 64-cell transpose tiles, points on
    both sides of 0 (wrapped cells), and cells shared by many points."""
    from mpmwarpydw_amd import runtime as rt
    rng = np.random.default_rng(sum(dims))
    n = 20000
    x = rng.uniform(-9.0, 11.0, (n, 3)).astype(np.float32)
    x[: n // 4] = np.round(x[: n // 4])  # a quarter on few cells
    pts = rt.array(x, rt.vec3)
    g = rt.HashGrid(*dims)
    print(f"[hashgrid] dims={dims} n={n}", flush=True)
    g.build(pts, 0.7)
    ids, cs, ce, _, _ = grid_tuple(g)
    rid, rcs, rce = o64.hashgrid_build(x.astype(np.float64), np.float32(0.7), dims)
    assert np.array_equal(ids, rid)
    nonempty = rce > rcs
    assert np.array_equal(cs[nonempty], rcs[nonempty]) and np.array_equal(ce[nonempty], rce[nonempty])
    assert np.all(ce[~nonempty] == cs[~nonempty]) and g.cell_start.numpy()[-1] == n


@pytest.mark.parametrize("ordered,lazy", [(False, False), (True, False), (True, True)])
def test_hashgrid_pair_matches_single_builds(ordered, lazy):
    """build_pair gives both tables (with different dims) bit for bit as two
    single builds (lazy or not, with points of the label present)."""
    from mpmwarpydw_amd import runtime as rt
    S = xpbd_scene(6)
    sim = S.sim
    n = sim.nPoints
    order = None
    if ordered:
        order = rt.array(np.random.default_rng(6).permutation(n).astype(np.int32), rt.int32)
    print(f"[hashgrid pair] n={n} ordered={ordered}", flush=True)
    a, b = rt.HashGrid(128, 128, 128), rt.HashGrid(64, 64, 64)
    a.build(sim.particle_x, sim.dx, order=order)
    b.build(sim.particle_x, sim.dx, labels=sim.materialLabel, label=2, order=order)
    c, d = rt.HashGrid(128, 128, 128), rt.HashGrid(64, 64, 64)
    c.build_pair(d, sim.particle_x, sim.dx, sim.materialLabel, 2, order, lazy=lazy)
    lab = sim.materialLabel.numpy()
    assert 0 < (lab == 2).sum() < n
    for single, pair in ((a, c), (b, d)):
        assert np.array_equal(single.point_ids.numpy(), pair.point_ids.numpy())
        assert np.array_equal(single.cell_start.numpy(), pair.cell_start.numpy())
    assert d.cell_start.numpy()[-1] == (lab == 2).sum()
    # point_pos inverts point_ids: every point's position in each table, -1 outside it
    for g, members in ((c, np.ones(n, bool)), (d, lab == 2)):
        pos, ids = g.point_pos.numpy(), g.point_ids.numpy()[:members.sum()]
        assert np.array_equal(pos[ids], np.arange(members.sum()))
        assert (pos[~members] == -1).all()


def test_lazy_pair_build_without_label_points():
    """With no point of the label, a lazy pair build leaves the first table as it
    was and empties the second; a build that is not lazy fills the first."""
    from mpmwarpydw_amd import runtime as rt
    S = xpbd_scene(8)
    sim = S.sim
    n = sim.nPoints
    sim.materialLabel.assign(np.ones(n, np.int32))
    print(f"[lazy pair] n={n}", flush=True)
    c, d = rt.HashGrid(128, 128, 128), rt.HashGrid(64, 64, 64)
    c.point_ids = rt.array(np.full(n, -7, np.int32), rt.int32)
    c.cell_start.assign(np.full(c.cell_start.shape[0], -5, np.int32))
    c.build_pair(d, sim.particle_x, sim.dx, sim.materialLabel, 2, None, lazy=True)
    assert (c.point_ids.numpy() == -7).all() and (c.cell_start.numpy() == -5).all()
    assert (d.cell_start.numpy() == 0).all()
    c.build_pair(d, sim.particle_x, sim.dx, sim.materialLabel, 2, None, lazy=False)
    a = rt.HashGrid(128, 128, 128)
    a.build(sim.particle_x, sim.dx)
    assert np.array_equal(a.point_ids.numpy(), c.point_ids.numpy())
    assert np.array_equal(a.cell_start.numpy(), c.cell_start.numpy())


@pytest.mark.parametrize("lock,cluster", [(0, 0), (1, 0), (0, 64)])
def test_contact_kernels_parity(o64, lock, cluster):
    from mpmwarpydw_amd import xpbdRoutines as X
    from mpmwarpydw_amd.runtime import launch
    S = xpbd_scene(2, cluster=cluster)
    print(f"[contacts] n={S.sim.nPoints} lock={lock} cluster={cluster}", flush=True)
    sim, xpbd = S.sim, S.xpbd
    xpbd.particle_grid.build(sim.particle_x, sim.dx)
    H = host_state(S)
    xpbd.particle_delta.zero_()
    launch(X.my_solve_particle_bound_contacts, sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, sim.particle_x, sim.particle_v, sim.particle_mass,
                   sim.particle_radius, xpbd.dynamicParticleFriction, xpbd.staticVelocityThreshold,
                   xpbd.staticParticleFriction, xpbd.minBoundsXPBD, xpbd.maxBoundsXPBD, sim.dtxpbd,
                   xpbd.xpbd_relaxation], outputs=[xpbd.particle_delta])
    launch(X.my_solve_particle_particle_contacts, sim.nPoints,
           inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_grid.id, sim.particle_x, sim.particle_v,
                   sim.particle_mass, sim.particle_radius, xpbd.dynamicParticleFriction,
                   xpbd.staticVelocityThreshold, xpbd.staticParticleFriction, xpbd.particle_cohesion,
                   xpbd.max_radius, sim.dtxpbd, xpbd.xpbd_relaxation, lock], outputs=[xpbd.particle_delta])
    d_gpu = xpbd.particle_delta.numpy()
    lab_gpu = sim.materialLabel.numpy()
    n = sim.nPoints
    d = np.zeros((n, 3))
    f32 = lambda v: float(np.float32(v))
    o64.my_solve_particle_bound_contacts(H["activeLabel"], H["materialLabel"], H["particle_x"], H["particle_v"],
                                         H["particle_mass"], H["particle_radius"], f32(xpbd.dynamicParticleFriction),
                                         f32(xpbd.staticVelocityThreshold), f32(xpbd.staticParticleFriction),
                                         np.float32(xpbd.minBoundsXPBD), np.float32(xpbd.maxBoundsXPBD),
                                         f32(sim.dtxpbd), f32(xpbd.xpbd_relaxation), d)
    ids, cs, ce = o64.hashgrid_build(H["particle_x"], np.float32(sim.dx), xpbd.particle_grid.dims)
    o64.my_solve_particle_particle_contacts(H["activeLabel"], H["materialLabel"],
                                            (ids, cs, ce, np.float32(sim.dx), xpbd.particle_grid.dims),
                                            H["particle_x"], H["particle_v"], H["particle_mass"],
                                            H["particle_radius"], f32(xpbd.dynamicParticleFriction),
                                            f32(xpbd.staticVelocityThreshold), f32(xpbd.staticParticleFriction),
                                            f32(xpbd.particle_cohesion), f32(xpbd.max_radius), f32(sim.dtxpbd),
                                            f32(xpbd.xpbd_relaxation), lock, d)
    assert np.isfinite(d).all()
    assert rel_err(d_gpu, d) < TOL
    assert np.array_equal(lab_gpu, H["materialLabel"])
    if lock:
        assert (lab_gpu == 0).sum() > 0


XPBD_STEP_CASES = [  # (fused, lock, xpbd_only, swelling)
    (False, 0, False, False), (True, 0, False, False), (False, 1, False, False), (True, 1, False, False),
    (False, 1, True, False), (True, 0, True, False), (False, 0, False, True), (True, 1, False, True)]


@pytest.mark.parametrize("fused,lock,xpbd_only,swell", XPBD_STEP_CASES)
def test_xpbd_step_parity(o64, fused, lock, xpbd_only, swell):
    """The whole xpbdSimulationStep (integrate, label reset under the transition
    lock, 4 iterations of bound + particle contacts and delta application,
    sleep, MPM reset, commit, swelling) from one state, launch for launch and
    fused, with the transition lock, in xpbd_only mode (MPM frozen: no reset,
    no lock, velocities not clamped) and with swelling, against the oracle."""
    from mpmwarpydw_amd import simulationRoutines as SR
    from state_utils import oracle_xpbd_step
    S = xpbd_scene(3, radius=0.135, **({"swellingRatio": 0.6} if swell else {}))
    sim, xpbd = S.sim, S.xpbd
    xpbd.mpm_contact_transition_lock = lock
    n = sim.nPoints
    rng = np.random.default_rng(33)
    if swell:  # XPBD birth positions 0-3 m away: below, inside and beyond the swelling ramp (1 m to 2 m)
        off = rng.standard_normal((n, 3))
        off *= (3.0 * rng.random(n) / np.linalg.norm(off, axis=1))[:, None]
        xpbd.particle_x_initial_xpbd.assign(sim.particle_x.numpy().astype(np.float64) - off)
    if lock:  # some MPM particles start locked (label 0); the reset must free those out of contact
        lab = sim.materialLabel.numpy()
        lab[(lab == 1) & (rng.random(n) < 0.5)] = 0
        sim.materialLabel.assign(lab)
    # velocities above particle_v_max on some MPM particles: the impulse kernel clamps them unless xpbd_only
    v = sim.particle_v.numpy()
    fast = (sim.materialLabel.numpy() <= 1) & (rng.random(n) < 0.05)
    v[fast] *= 4.0 * xpbd.particle_v_max / np.maximum(np.linalg.norm(v[fast], axis=1), 1e-6)[:, None]
    sim.particle_v.assign(v)
    H = host_state(S)
    print(f"[xpbd step] n={n} fused={fused} lock={lock} xpbd_only={xpbd_only} swell={swell}", flush=True)
    (SR.xpbdSimulationStepFused if fused else SR.xpbdSimulationStep)(sim, S.mpm, xpbd, xpbd_only=xpbd_only)
    G = host_state(S)
    total, sleeping = oracle_xpbd_step(o64, H, S, xpbd_only=xpbd_only)
    # Positions are FP32 coordinates of a several-metre domain: one rounding of x
    # is ulp(max|x|), and the step's velocities are (x_new - x_orig) / dt
    # (xpbdRoutines.py:246), so FP32 against FP64 differs by a few ulp(x) in x
    # and a few ulp(x) / dt in v after 4 iterations -- the resolution of the
    # reference's own FP32 state.  Bound: 16 ulp.
    ulp = float(np.spacing(np.float32(np.max(np.abs(H["particle_x"])))))
    ex = np.max(np.abs(G["particle_x"] - H["particle_x"]))
    ev = np.max(np.abs(G["particle_v"] - H["particle_v"]))
    print(f"  x err {ex:.2e} (ulp {ulp:.2e}), v err {ev:.2e} (ulp/dt {ulp / sim.dtxpbd:.2e})")
    assert ex <= 16 * ulp
    assert ev <= 16 * ulp / float(np.float32(sim.dtxpbd)) + TOL * np.max(np.abs(H["particle_v"]))
    assert np.max(np.abs(G["particle_cumDist_xpbd"] - H["particle_cumDist_xpbd"])) <= 16 * ulp
    assert np.array_equal(G["materialLabel"], H["materialLabel"])
    assert (int(sim.numTotalXPBD.numpy()[0]), int(sim.numSleepingXPBD.numpy()[0])) == (total, sleeping)
    lab = H["materialLabel"]
    if lock and not xpbd_only:
        assert ((lab == 0) & (H["activeLabel"] == 1)).sum() > 0  # contacts re-locked some
    if swell:
        grown = H["particle_radius"] > S.xpbd.particleBaseRadius.numpy() * (1 + 1e-6)
        assert grown.sum() > 0
        assert rel_err(G["particle_radius"], H["particle_radius"]) < TOL
        r = np.asarray(G["particle_radius"]); h = np.asarray(H["particle_radius"])
        assert np.all(np.abs(r - h) <= TOL * h)
    if not xpbd_only:  # MPM velocities above v_max were clamped
        mv = np.linalg.norm(H["particle_v"][(lab <= 1) & (H["activeLabel"] == 1)], axis=1)
        assert mv.max() <= xpbd.particle_v_max * (1 + 1e-6)


@pytest.mark.parametrize("lock,xdims", [(0, 64), (1, 64), (1, 8), (0, 4)])
def test_split_table_is_bit_identical(lock, xdims):
    """MPM particles walking the XPBD-only table give bit-identical deltas and
    labels to walking the full table (same candidates, same order).  The scene
    spans more than 8 cells per axis: with 8^3 XPBD cells, cells wrap onto each
    other and the far candidates must fail the contact test; with 4^3 the gap is
    shorter than the contact reach and the particles walk the full table."""
    from mpmwarpydw_amd import runtime as rt
    from mpmwarpydw_amd import xpbdRoutines as X
    from mpmwarpydw_amd.runtime import launch
    outs = []
    for split in (False, True):
        S = xpbd_scene(4)
        sim, xpbd = S.sim, S.xpbd
        xpbd.particle_grid_xpbd = rt.HashGrid(xdims, xdims, xdims)
        ext = np.ptp(sim.particle_x.numpy(), axis=0) / sim.dx
        assert ext.max() > 8  # wrapping happens in the 8^3 table
        print(f"[split table] lock={lock} xdims={xdims} extent cells={ext.round(1)}", flush=True)
        xpbd.particle_grid.build(sim.particle_x, sim.dx)
        xpbd.particle_delta.zero_()
        args = [sim.particle_x, sim.particle_v, sim.particle_mass, sim.particle_radius, xpbd.dynamicParticleFriction,
                xpbd.staticVelocityThreshold, xpbd.staticParticleFriction, xpbd.particle_cohesion, xpbd.max_radius,
                sim.dtxpbd, xpbd.xpbd_relaxation, lock]
        if split:
            xpbd.particle_grid_xpbd.build(sim.particle_x, sim.dx, labels=sim.materialLabel, label=2)
            launch(X.solve_particle_particle_contacts_split, sim.nPoints,
                   inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_grid.id, xpbd.particle_grid_xpbd.id,
                           *args], outputs=[xpbd.particle_delta])
        else:
            launch(X.my_solve_particle_particle_contacts, sim.nPoints,
                   inputs=[sim.activeLabel, sim.materialLabel, xpbd.particle_grid.id, *args],
                   outputs=[xpbd.particle_delta])
        outs.append((xpbd.particle_delta.numpy(), sim.materialLabel.numpy()))
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])


@pytest.mark.parametrize("lock,cluster,mpm_only,xpbd_only", [(0, 0, False, False), (1, 0, False, False),
                                                              (1, 64, False, False), (0, 0, True, False),
                                                              (1, 0, True, False), (0, 0, False, True),
                                                              (1, 64, False, True)])
def test_fused_xpbd_step_matches_launch_sequence(lock, cluster, mpm_only, xpbd_only):
    """The fused XPBD step (packed candidate records; with no XPBD particle the
    lazy tables and skipped delta applications) leaves x, v, labels, the
    cumulative distance and the sleep counters bit-identical to the
    launch-for-launch step, in the combined phase (with and without the
    contact lock) and in the XPBD-only phase, where MPM velocities come from
    the last delta application instead of the contact impulses."""
    from mpmwarpydw_amd import simulationRoutines as SR
    outs = []
    for fused in (False, True):
        S = xpbd_scene(5, cluster=cluster)
        if mpm_only:  # labels 0 and 1 only, some inactive particles, velocities above v_max
            n = S.sim.nPoints
            S.sim.materialLabel.assign(np.where(np.arange(n) % 3 == 0, 0, 1).astype(np.int32))
            S.sim.particle_v.assign(np.random.default_rng(5).normal(0, 30.0, (n, 3)))
        S.xpbd.mpm_contact_transition_lock = lock
        print(f"[fused xpbd] fused={fused} lock={lock} cluster={cluster} mpm_only={mpm_only} xpbd_only={xpbd_only}",
              flush=True)
        (SR.xpbdSimulationStepFused if fused else SR.xpbdSimulationStep)(S.sim, S.mpm, S.xpbd, xpbd_only=xpbd_only)
        h = host_state(S)
        outs.append((h, S.sim.numTotalXPBD.numpy()[0], S.sim.numSleepingXPBD.numpy()[0]))
    (a, ta, sa), (b, tb, sb) = outs
    for k in ("particle_x", "particle_v", "materialLabel", "particle_cumDist_xpbd"):
        assert np.array_equal(a[k], b[k]), k
    assert (ta, sa) == (tb, sb) and (ta == 0) == mpm_only


def test_contact_pack_records():
    """mpmx_xpbd_contact_pack writes, in table order, the flags of every
    candidate and the (x, radius), (v, 1/mass) records of the valid ones, the
    same from the particle arrays as from records in particle order."""
    import torch
    from mpmwarpydw_amd import runtime as rt
    from mpmwarpydw_amd._lib import lib
    S = xpbd_scene(6)
    sim, xpbd = S.sim, S.xpbd
    n = sim.nPoints
    grid, xgrid = xpbd.particle_grid, xpbd.particle_grid_xpbd
    grid.build_pair(xgrid, sim.particle_x, sim.dx, sim.materialLabel, 2, None)
    H = host_state(S)
    act, lab, m = H["activeLabel"], H["materialLabel"], np.float32(H["particle_mass"])
    valid = (act == 1) & (m > 0)
    pflags = (valid * 1 | (lab <= 1) * 2 | (lab == 2) * 4).astype(np.int32)
    pxr = np.concatenate([np.float32(H["particle_x"]), np.float32(H["particle_radius"])[:, None]], axis=1)
    with np.errstate(divide="ignore"):
        pvw = np.concatenate([np.float32(H["particle_v"]), (np.float32(1) / m)[:, None]], axis=1)
    ids = grid.point_ids.numpy()
    nx = int(xgrid.cell_start.numpy()[-1])
    order = np.concatenate([ids, xgrid.point_ids.numpy()[:nx]])
    assert nx > 0 and np.array_equal(np.sort(ids), np.arange(n))
    print(f"[contact pack] n={n} xpbd table={nx}", flush=True)
    dev = lambda a: torch.tensor(np.ascontiguousarray(a), device=rt.device())  # noqa: E731
    src = [dev(pxr), dev(pvw), dev(pflags)]
    outs = []
    for from_records in (False, True):
        xr = torch.zeros(2 * n, 4, device=rt.device())
        vw = torch.zeros(2 * n, 4, device=rt.device())
        fl = torch.full((2 * n,), -1, dtype=torch.int32, device=rt.device())
        soa = [sim.activeLabel.ptr, sim.materialLabel.ptr, sim.particle_mass.ptr, sim.particle_radius.ptr,
               sim.particle_x.ptr, sim.particle_v.ptr]
        rec = [t.data_ptr() for t in src] if from_records else [None] * 3
        lib("mpmx_xpbd_contact_pack", n, *([None] * 6 if from_records else soa), *rec, grid.point_ids.ptr,
            xgrid.point_ids.ptr, xgrid.cell_start.ptr, *xgrid.dims, xr.data_ptr(), vw.data_ptr(), fl.data_ptr(),
            rt.stream_handle())
        torch.cuda.synchronize()
        outs.append((xr.cpu().numpy(), vw.cpu().numpy(), fl.cpu().numpy()))
    k = n + nx
    for xr, vw, fl in outs:
        assert np.array_equal(fl[:k], pflags[order])
        assert (fl[k:] == -1).all()  # past the XPBD table's points nothing is written
        ok = valid[order]
        assert np.array_equal(xr[:k][ok], pxr[order][ok]) and np.array_equal(vw[:k][ok], pvw[order][ok])


@pytest.mark.parametrize("lock,few", [(0, False), (1, False), (1, True)])
def test_packed_contacts_from_gathered_records(lock, few):
    """The packed contact solve over records that mpmx_xpbd_contact_pack gathers
    from the particle arrays gives the deltas and labels of the split solve,
    bit for bit -- with the XPBD table's neighbourhood bitmap letting MPM
    particles far from every XPBD point skip their walk (few: only a handful of
    XPBD particles, as just after the first transitions of a run; with half the
    points in the XPBD phase no list is written and every position walks)."""
    import torch
    from mpmwarpydw_amd import runtime as rt
    from mpmwarpydw_amd._lib import lib
    outs = []
    for packed in (False, True):
        S = xpbd_scene(7, cluster=0 if few else 32)
        sim, xpbd = S.sim, S.xpbd
        n = sim.nPoints
        if few:
            lab = sim.materialLabel.numpy()
            keep = np.flatnonzero(lab == 2)[::97][:2]
            lab[lab == 2] = 1
            lab[keep] = 2
            sim.materialLabel.assign(lab)
        grid, xgrid = xpbd.particle_grid, xpbd.particle_grid_xpbd
        grid.build_pair(xgrid, sim.particle_x, sim.dx, sim.materialLabel, 2, None)
        xpbd.particle_delta.zero_()
        if packed:
            xr = torch.empty(2 * n, 4, device=rt.device())
            vw = torch.empty(2 * n, 4, device=rt.device())
            fl = torch.empty(2 * n, dtype=torch.int32, device=rt.device())
            lib("mpmx_xpbd_contact_pack", n, sim.activeLabel.ptr, sim.materialLabel.ptr, sim.particle_mass.ptr,
                sim.particle_radius.ptr, sim.particle_x.ptr, sim.particle_v.ptr, None, None, None, grid.point_ids.ptr,
                xgrid.point_ids.ptr, xgrid.cell_start.ptr, *xgrid.dims, xr.data_ptr(), vw.data_ptr(), fl.data_ptr(),
                rt.stream_handle())
            rec = [xr.data_ptr(), vw.data_ptr(), fl.data_ptr()]
            near = xgrid.near_cells().ptr  # MPM particles far from every XPBD point are not listed
            clist = torch.empty(n, dtype=torch.int32, device=rt.device())
            ccount = torch.empty(1, dtype=torch.int32, device=rt.device())
            lib("mpmx_xpbd_contact_list", n, sim.materialLabel.ptr, grid.point_pos.ptr, xgrid.cell_start.ptr,
                float(grid.cell_width), *xgrid.dims, sim.particle_x.ptr, sim.particle_radius.ptr,
                float(xpbd.particle_cohesion), float(xpbd.max_radius), near, clist.data_ptr(), ccount.data_ptr(),
                rt.stream_handle())
            listed = int(ccount.item())
            print(f"  contact list: {listed} of {n} table positions", flush=True)
            if few:
                assert 0 < listed <= n // 8  # the solve walks the listed positions only
            else:
                assert listed == n + 1  # half the points are XPBD: no list, every table position walks
        name = "mpmx_solve_particle_particle_contacts_" + ("packed" if packed else "split")
        print(f"[{name}] n={n} lock={lock}", flush=True)
        lib(name, n, sim.activeLabel.ptr, sim.materialLabel.ptr, grid.point_ids.ptr, grid.cell_start.ptr,
            xgrid.point_ids.ptr, xgrid.cell_start.ptr, float(grid.cell_width), *grid.dims, *xgrid.dims,
            sim.particle_x.ptr, sim.particle_v.ptr, sim.particle_mass.ptr, sim.particle_radius.ptr,
            *(rec if packed else []), float(xpbd.dynamicParticleFriction), float(xpbd.staticVelocityThreshold),
            float(xpbd.staticParticleFriction), float(xpbd.particle_cohesion), float(xpbd.max_radius),
            float(sim.dtxpbd), float(xpbd.xpbd_relaxation), lock, xpbd.particle_delta.ptr, None,
            *([near, clist.data_ptr(), ccount.data_ptr()] if packed else []), rt.stream_handle())
        outs.append((xpbd.particle_delta.numpy(), sim.materialLabel.numpy()))
    assert np.abs(outs[0][0]).max() > 0
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
