"""Pins the oracle's XPBD step (xpbdRoutines.py:3-376) and its wp.HashGrid
restatement: neighbour sets against brute force over the wrapped cell box,
two-particle and particle-wall contact responses in closed form, and the
bookkeeping kernels (NaN guard, velocity clamp, sleep, MPM reset, swelling).
The candidate ORDER of wp.HashGrid is taken from Warp's published query loop
(x fastest, then y, then z; ids ascending in a cell) and is parity unpinned:
no fixture in the reference records it.  CPU only."""
import numpy as np


def brute_candidates(x, i, qr, width, dims):
    """Points whose wrapped cell lies in the wrapped query box of point i."""
    inv = 1.0 / width
    lo = [int((x[i, a] - qr) * inv) for a in range(3)]
    hi = [min(int((x[i, a] + qr) * inv), lo[a] + dims[a] - 1) for a in range(3)]
    cells = set()
    for cz in range(lo[2], hi[2] + 1):
        for cy in range(lo[1], hi[1] + 1):
            for cx in range(lo[0], hi[0] + 1):
                cells.add(((cx + (1 << 20)) % dims[0], (cy + (1 << 20)) % dims[1], (cz + (1 << 20)) % dims[2]))
    out = []
    for j in range(len(x)):
        c = tuple((int(x[j, a] * inv) + (1 << 20)) % dims[a] for a in range(3))
        if c in cells:
            out.append(j)
    return sorted(out)


def test_hashgrid_cells_and_order(o64):
    rng = np.random.default_rng(0)
    n, width, dims = 400, 0.5, (5, 4, 6)  # small dims force wrapping
    x = rng.random((n, 3)) * np.array([4.0, 3.0, 5.0]) - 0.2
    ids, cs, ce = o64.hashgrid_build(x, width, dims)
    assert sorted(ids.tolist()) == list(range(n))
    inv = 1 / width
    key = lambda j: (((int(x[j, 2] * inv) + (1 << 20)) % dims[2]) * dims[0] * dims[1]
                     + ((int(x[j, 1] * inv) + (1 << 20)) % dims[1]) * dims[0]
                     + (int(x[j, 0] * inv) + (1 << 20)) % dims[0])
    keys = [key(j) for j in ids]
    assert keys == sorted(keys)  # sorted by cell
    for c in range(int(np.prod(dims))):
        run = ids[cs[c]:ce[c]].tolist()
        assert run == sorted(run) and all(key(j) == c for j in run)  # stable: ids ascending


def test_hashgrid_query_matches_brute_force(o64):
    """Every query yields exactly the points of the cells in its (wrapped) box,
    each once, in x-fastest cell order."""
    rng = np.random.default_rng(1)
    n, width = 300, 0.5
    for dims in ((5, 4, 6), (128, 128, 128)):
        x = rng.random((n, 3)) * 3.0 + 0.01
        ids, cs, ce = o64.hashgrid_build(x, width, dims)
        for i in range(0, n, 7):
            qr = 0.05 + 0.6 * rng.random()
            got = o64.hashgrid_query((ids, cs, ce, width, dims), x[i], qr)
            assert len(got) == len(set(got))
            assert sorted(got) == brute_candidates(x, i, qr, width, dims)


def contacts(o, x, v, mass, radius, labels, width=0.5, dims=(16, 16, 16), k_mu=0.3, thr=1e-5, smu=0.6, coh=0.0,
             dt=1e-3, relax=1.0, lock=0):
    n = len(x)
    ids, cs, ce = o.hashgrid_build(x, width, dims)
    d = np.zeros((n, 3))
    o.my_solve_particle_particle_contacts(np.ones(n, np.int32), labels, (ids, cs, ce, width, dims),
                                          np.ascontiguousarray(x), np.ascontiguousarray(v), mass, radius, k_mu, thr,
                                          smu, coh, float(radius.max()), dt, relax, lock, d)
    return d


def test_two_particle_contact_closed_form(o64):
    """Equal masses overlapping by e along x, no relative velocity: each moves
    e/2 apart; with tangential sliding the friction correction is
    -min(mu |err|, |vt| dt)... along -vt, shared by mass."""
    r = 0.13
    x = np.array([[1.0, 1.0, 1.0], [1.0 + 2 * r - 0.02, 1.0, 1.0]])
    d = contacts(o64, x, np.zeros((2, 3)), np.array([2.0, 2.0]), np.full(2, r), np.array([2, 2], np.int32))
    assert np.allclose(d, [[-0.01, 0, 0], [0.01, 0, 0]], atol=1e-14)
    # particle 0 slides along +y at 1 m/s relative to particle 1: dynamic friction (|vn| = 0 < thr r -> static!)
    v = np.array([[0.0, 1.0, 0.0], [0.0, 0.0, 0.0]])
    d = contacts(o64, x, v, np.array([1.0, 3.0]), np.full(2, r), np.array([2, 2], np.int32), k_mu=0.3, smu=0.6)
    err = -0.02
    w1, w2 = 1.0, 1 / 3
    lam_f = max(0.6 * err, -1.0 * 1e-3)  # static friction (vn = 0 is below the threshold)
    want0 = (np.array([0, lam_f, 0]) - np.array([-1, 0, 0]) * err) / (w1 + w2) * w1
    assert np.allclose(d[0], want0, atol=1e-14)


def test_mpm_pairs_skipped_and_lock(o64):
    """MPM-MPM pairs are skipped; with the lock on, an MPM particle touching an
    XPBD particle gets label 0 and the pair still responds."""
    r = 0.13
    x = np.array([[1.0, 1.0, 1.0], [1.2, 1.0, 1.0], [1.4, 1.0, 1.0]])
    labels = np.array([1, 1, 2], np.int32)
    d = contacts(o64, x, np.zeros((3, 3)), np.ones(3), np.full(3, r), labels, lock=1)
    assert np.all(d[0] == 0)
    assert d[1, 0] < 0 and d[2, 0] > 0
    assert labels.tolist() == [1, 0, 2]


def test_bound_contact_floor(o64):
    """A particle 0.02 m through the floor is pushed out by 0.02 and loses
    tangential displacement min(mu * 0.02, |vt| dt)."""
    r, dt = 0.1, 1e-2
    x = np.array([[0.5, 0.5, 0.08]])
    v = np.array([[1.0, 0.0, -0.1]])
    d = np.zeros((1, 3))
    o64.my_solve_particle_bound_contacts(np.ones(1, np.int32), np.array([2], np.int32), x, v, np.ones(1),
                                         np.full(1, r), 0.4, 1e-5, 0.9, np.zeros(3), np.full(3, 10.0), dt, 1.0, d)
    assert np.allclose(d[0], [max(0.4 * -0.02, -1.0 * dt), 0.0, 0.02], atol=1e-15)


def test_apply_deltas_nan_and_clamp(o64):
    x0 = np.zeros((3, 3))
    xp = np.array([[1.0, 0, 0], [0, 1.0, 0], [0, 0, 1e3]])
    d = np.array([[np.nan, 0, 0], [0.5, 0, 0], [0, 0, 0]])
    xo, vo = np.zeros((3, 3)), np.zeros((3, 3))
    o64.my_apply_particle_deltas(np.ones(3, np.int32), np.full(3, 2, np.int32), x0, xp, d, 0.1, 1000.0, xo, vo)
    assert np.array_equal(xo[0], xp[0]) and np.allclose(vo[0], [10.0, 0, 0])
    assert np.allclose(xo[1], [0.5, 1.0, 0]) and np.allclose(vo[1], [5.0, 10.0, 0])
    assert np.allclose(np.linalg.norm(vo[2]), 1000.0)


def test_sleep_impulse_swell(o64):
    act = np.ones(3, np.int32)
    lab = np.array([2, 2, 1], np.int32)
    xp = np.zeros((3, 3))
    xa = np.array([[1e-6, 0, 0], [0.1, 0, 0], [0.5, 0, 0]])
    dist = np.zeros(3)
    tot, slp = o64.sleepParticles(act, lab, 0.5, np.full(3, 0.1), xp, xa, dist, 1e-3)
    assert (tot, slp) == (2, 1)
    assert np.array_equal(xa[0], [0, 0, 0]) and np.allclose(dist, [1e-6, 0.1, 0.0])
    # MPM particles are put back and keep their (clamped) velocity; XPBD untouched
    xo, vo = np.zeros((3, 3)), np.array([[0, 0, 5.0], [0, 0, 5.0], [0, 0, 2000.0]])
    xf = np.ones((3, 3))
    vout = np.zeros((3, 3))
    o64.apply_mpm_contact_impulses(act, lab, xo, vo, xf, np.zeros((3, 3)), 1e-3, 1000.0, 0.25, 0, vout)
    assert np.array_equal(xf[2], [0, 0, 0]) and np.array_equal(xf[0], [1, 1, 1])
    assert np.allclose(vout[2], [0, 0, 1000.0]) and np.all(vout[0] == 0)
    # swelling: radius grows linearly between factor and factor_max (scaled by the base radius)
    base = np.full(3, 0.1)
    swell = np.full(3, 0.2)
    rad = base.copy()
    x0 = np.zeros((3, 3))
    x1 = np.array([[15.0, 0, 0], [25.0, 0, 0], [15.0, 0, 0]])
    o64.swellParticlesType2(act, lab, x0, x1, rad, 1.0, 2.0, swell, base)
    # dxl = |dx| * base = 1.5 -> (1.5 - 0.1) / (0.2 - 0.1) clipped to [r, swell]
    assert np.allclose(rad, [0.2, 0.2, 0.1])


def test_reductions(o64):
    rng = np.random.default_rng(1)
    n = 1000
    act = (rng.random(n) > 0.3).astype(np.int32)
    lab = rng.integers(0, 3, n).astype(np.int32)
    v = rng.standard_normal((n, 3))
    r = 0.1 + rng.random(n)
    s = o64.velocityConvergence(act, v, r)
    assert np.isclose(s, np.sum(act * np.linalg.norm(v, axis=1) / r), rtol=1e-13)
    assert o64.countActiveParticles(act) == act.sum()
    c = o64.countParticlesByType(act, lab)
    assert c == (int(((lab != 2) & (act == 1)).sum()), int(((lab == 2) & (act == 1)).sum()),
                 int(((lab != 2) & (act != 1)).sum()), int(((lab == 2) & (act != 1)).sum()))


def _contact_state(seed, vscale, n_side=10):
    """A contact_heavy-like block: 0.25 m lattice, 0.03 m jitter, radius
    0.25 m (every neighbour overlaps), XPBD (label 2) in one half and MPM
    (label 1) in the other, 2% inactive; velocities of scale vscale."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * 0.25 + 0.03 * rng.standard_normal(g.shape) + 1.0
    n = len(x)
    lab = np.where(g[:, 0] < n_side // 2, 2, 1).astype(np.int32)
    act = (rng.random(n) >= 0.02).astype(np.int32)
    v = vscale * rng.standard_normal((n, 3))
    return x, v, np.full(n, 39.0625), np.full(n, 0.25), lab, act


def _solve(o, st, order, static_mu, dyn_mu, thr):
    x, v, m, r, lab, act = st
    ids, cs, ce = o.hashgrid_build(x, 0.5, (128, 128, 128))
    d = np.zeros_like(x)
    o.my_solve_particle_particle_contacts(act, lab.copy(), (ids, cs, ce, 0.5, (128, 128, 128)), x, v, m, r, dyn_mu,
                                          thr, static_mu, 0.0, 0.25, 5e-4, 1.0, 0, d, order=order)
    return d


def _sliding_state(seed, radius, n_side=10):
    """Exact 0.25 m lattice, shallow overlaps (radius just above 0.125 m),
    alternate z-layers sliding at +-0.8 m/s along x, and half of the particles
    with a normal velocity: contacts with |v_n| below the static threshold and
    contacts above it meet in the same query."""
    rng = np.random.default_rng(seed)
    g = np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3)
    x = (g + 0.5) * 0.25 + 1.0
    n = len(x)
    lab = np.where(g[:, 0] < n_side // 2, 2, 1).astype(np.int32)
    v = np.zeros((n, 3))
    v[:, 0] = np.where(g[:, 2] % 2 == 0, 0.8, -0.8)
    half = rng.random(n) < 0.5
    v[half, 2] = 0.2 * rng.standard_normal(half.sum())
    return x, v, np.full(n, 39.0625), np.full(n, radius), lab, np.ones(n, np.int32)


def test_candidate_order_bound(o64):
    """What the unpinned wp.HashGrid candidate order can change.  The
    candidates are the same in every order (the neighbour set is pinned), so
    with equal static and dynamic friction only the summation order of the
    contact deltas differs (FP64 rounding).  With 0.6 static / 0.4 dynamic the
    sticky static switch (xpbdRoutines.py:211-212) makes k_mu static for every
    later candidate once one contact has |v_n| < 1e-5 r; but the friction is
    max(k_mu lambda_n, -|v_t| dt), so k_mu matters only where the tangential
    cap does not bind: |v_t| dt > k_mu |err|.  On contact_heavy-like states
    (every neighbour overlapping by centimetres, at rest or moving) the order
    changes nothing beyond rounding; it shows on sliding contacts whose
    overlap is under a millimetre.  Reported: the largest change of a
    particle's delta relative to the largest delta, for the reverse sweep
    (order 1) and for z-fastest (order 2)."""
    cases = {"c2_rest": _contact_state(5, 1e-6), "c2_moving": _contact_state(5, 0.5),
             "sliding_0.5mm": _sliding_state(7, 0.1255), "sliding_1mm": _sliding_state(7, 0.126)}
    report = {}
    for name, st in cases.items():
        for frictions in ((0.4, 0.4), (0.6, 0.4)):
            d0 = _solve(o64, st, 0, *frictions, 1e-5)
            assert np.abs(d0).max() > 0
            for order in (1, 2):
                d = _solve(o64, st, order, *frictions, 1e-5)
                rel = float(np.abs(d - d0).max() / np.abs(d0).max())
                report[(name, frictions, order)] = rel
                print(f"  {name} friction {frictions} order {order}: max |change| / max |delta| = {rel:.2e}")
                if frictions[0] == frictions[1] or not name.startswith("sliding_0.5"):
                    assert rel < 1e-12, (name, frictions, order, rel)  # summation order only
    # shallow sliding contacts: the order decides which contacts are static
    assert report[("sliding_0.5mm", (0.6, 0.4), 1)] > 1e-2
