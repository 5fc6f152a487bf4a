"""Pins the oracle's transfers and grid update (mpmRoutines.py:474-569 P2G,
:691-711 normalisation/damping, :715-802 G2P, :915-993 collideBounds) against
properties of quadratic B-spline MPM that are independent of the formulas:
partition of unity, linear-field reproduction, momentum conservation, the
first moment of the stress impulse, hand-computed single-particle weights and
per-node closed forms of the Coulomb boundary.  CPU only."""
import numpy as np

from conftest import rel_err


def make_grid(dims):
    return np.zeros(dims), np.zeros(dims + (3,))


def p2g(o, x, v, C, stress, vol, mass, dx, minB, dims, dt, rpic=0.0, active=None, mat=None):
    n = len(x)
    active = np.ones(n, np.int32) if active is None else active
    mat = np.ones(n, np.int32) if mat is None else mat
    gm, gv = make_grid(dims)
    bad = o.p2g_apic_with_stress(active, mat, np.ascontiguousarray(stress), np.ascontiguousarray(x),
                                 np.ascontiguousarray(v), np.ascontiguousarray(C), vol, mass, dx, 1.0 / dx, minB, rpic,
                                 gm, gv, dt)
    return gm, gv, bad


def test_single_particle_weights(o64):
    """A particle at grid coordinate base + 1 has per-axis weights (1/8, 3/4, 1/8)."""
    dx, minB, dims = 0.5, np.zeros(3), (6, 6, 6)
    x = np.array([[2 * dx, 2 * dx, 2 * dx]])  # gp = 2 -> base = int(1.5) = 1, fx = 1
    v = np.array([[1.0, -2.0, 3.0]])
    gm, gv, bad = p2g(o64, x, v, np.zeros((1, 3, 3)), np.zeros((1, 3, 3)), np.array([0.1]), np.array([4.0]), dx,
                      minB, dims, 1e-3)
    w = np.array([0.125, 0.75, 0.125])
    W = np.einsum("i,j,k->ijk", w, w, w)
    assert bad == 0
    assert np.allclose(gm[1:4, 1:4, 1:4], 4.0 * W, rtol=1e-15)
    assert gm.sum() == np.float64(4.0) or np.isclose(gm.sum(), 4.0, rtol=1e-15)
    assert np.allclose(gv[1:4, 1:4, 1:4], 4.0 * W[..., None] * v[0], rtol=1e-14)


def test_p2g_conservation(o64):
    """Sum of grid mass = sum of particle mass; sum of grid momentum = sum m v
    (the APIC term and the stress impulse both sum to zero over the stencil)."""
    rng = np.random.default_rng(0)
    n, dx, dims = 300, 0.5, (14, 12, 13)
    minB = np.array([-1.0, 0.3, 2.0])
    x = minB + dx * (2.0 + rng.random((n, 3)) * (np.array(dims) - 5))
    v = rng.standard_normal((n, 3))
    C = rng.standard_normal((n, 3, 3))
    S = 1e4 * rng.standard_normal((n, 3, 3))
    m = 1 + rng.random(n)
    vol = 0.01 + 0.01 * rng.random(n)
    gm, gv, bad = p2g(o64, x, v, C, S, vol, m, dx, minB, dims, 1e-3, rpic=0.3)
    assert bad == 0
    assert np.isclose(gm.sum(), m.sum(), rtol=1e-13)
    assert np.allclose(gv.reshape(-1, 3).sum(0), (m[:, None] * v).sum(0), rtol=1e-10, atol=1e-9)


def test_p2g_stress_first_moment(o64):
    """With v = C = 0 the grid impulse f_i = -dt vol tau grad w_i satisfies
    sum_i f_i (x_i - x_p)^T = -dt vol tau (gradient of a linear field)."""
    rng = np.random.default_rng(1)
    dx, dims, dt, vol = 0.5, (8, 8, 8), 2e-3, 0.03
    minB = np.zeros(3)
    for _ in range(10):
        x = (2.5 + 2 * rng.random((1, 3))) * dx
        tau = rng.standard_normal((1, 3, 3))
        gm, gv, _ = p2g(o64, x, np.zeros((1, 3)), np.zeros((1, 3, 3)), tau, np.array([vol]), np.array([1.0]), dx,
                        minB, dims, dt)
        idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1) * dx
        M = np.einsum("ijka,ijkb->ab", gv, idx - x[0])
        assert np.allclose(M, -dt * vol * tau[0], rtol=1e-12, atol=1e-15)


def test_pic_and_rpic(o64):
    """rpic < -0.001 drops C (PIC); rpic = 1 keeps only the skew part of C."""
    rng = np.random.default_rng(2)
    dx, dims, minB = 0.5, (8, 8, 8), np.zeros(3)
    x = (3 + rng.random((1, 3))) * dx
    C = rng.standard_normal((1, 3, 3))
    g0 = p2g(o64, x, np.zeros((1, 3)), np.zeros((1, 3, 3)), np.zeros((1, 3, 3)), np.ones(1), np.ones(1), dx, minB,
             dims, 1e-3)[1]
    gp = p2g(o64, x, np.zeros((1, 3)), C, np.zeros((1, 3, 3)), np.ones(1), np.ones(1), dx, minB, dims, 1e-3,
             rpic=-1.0)[1]
    assert np.array_equal(g0, gp)
    skew = 0.5 * (C - C.transpose(0, 2, 1))
    g1 = p2g(o64, x, np.zeros((1, 3)), C, np.zeros((1, 3, 3)), np.ones(1), np.ones(1), dx, minB, dims, 1e-3,
             rpic=1.0)[1]
    g2 = p2g(o64, x, np.zeros((1, 3)), skew, np.zeros((1, 3, 3)), np.ones(1), np.ones(1), dx, minB, dims, 1e-3)[1]
    assert np.allclose(g1, g2, rtol=1e-13, atol=1e-16)


def test_out_of_grid_particles_are_skipped(o64):
    """base = int(gp - 0.5) truncates toward zero, so gp = 0.2 still lands on base 0
    (the reference's quirk, weights still sum to one); stencils past the grid are skipped."""
    dx, dims, minB = 0.5, (6, 6, 6), np.zeros(3)
    x = np.array([[0.1, 1.5, 1.5], [2.9, 1.5, 1.5], [-0.3, 1.5, 1.5], [1.5, 1.5, 1.5]])
    gm, gv, bad = p2g(o64, x, np.zeros((4, 3)), np.zeros((4, 3, 3)), np.zeros((4, 3, 3)), np.ones(4), np.ones(4), dx,
                      minB, dims, 1e-3)
    assert bad == 2 and np.isclose(gm.sum(), 2.0, rtol=1e-14)
    assert np.isclose(gm[0].sum(), 0.845, rtol=1e-12)  # w0(0.2) = 0.5 * 1.3^2 on the x = 0 plane


def test_g2p_reproduces_affine_fields(o64):
    """Grid velocity a + B x gives v_p = a + B x_p, C_p = B, F_trial = (I + dt B) F."""
    rng = np.random.default_rng(3)
    n, dx, dims, dt = 50, 0.5, (10, 11, 12), 1e-3
    minB = np.array([0.2, -0.4, 1.0])
    a = rng.standard_normal(3)
    B = rng.standard_normal((3, 3))
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1) * dx + minB
    gv = a + idx @ B.T
    x = minB + dx * (2 + rng.random((n, 3)) * (np.array(dims) - 5))
    x0 = x.copy()
    v, C = np.zeros((n, 3)), np.zeros((n, 3, 3))
    F = np.array([np.eye(3) + 0.1 * rng.standard_normal((3, 3)) for _ in range(n)])
    Ft = np.zeros((n, 3, 3))
    cov = np.zeros(6 * n)
    bad = o64.g2p(dt, np.ones(n, np.int32), np.ones(n, np.int32), x, v, C, F, Ft, cov, 1.0 / dx,
                  np.ascontiguousarray(gv), True, minB)
    assert bad == 0
    assert np.allclose(v, a + x0 @ B.T, rtol=1e-12, atol=1e-12)
    assert np.allclose(C, np.broadcast_to(B, C.shape), rtol=1e-11, atol=1e-11)
    assert np.allclose(x, x0 + dt * v, rtol=1e-15)
    assert np.allclose(Ft, np.einsum("ij,pjk->pik", np.eye(3) + dt * B, F), rtol=1e-12, atol=1e-14)
    assert np.all(cov == 0)  # the homogeneous covariance update keeps zero at zero


def test_g2p_covariance_update(o64):
    dx, dims, dt = 0.5, (8, 8, 8), 1e-2
    B = np.array([[0.1, 0.2, 0.0], [-0.3, 0.0, 0.5], [0.0, 0.1, -0.2]])
    idx = np.stack(np.meshgrid(*[np.arange(d) for d in dims], indexing="ij"), -1) * dx
    gv = idx @ B.T
    c0 = np.array([1.0, 0.1, 0.2, 2.0, 0.3, 3.0])
    cov = c0.copy()
    x = np.array([[2.1, 1.9, 2.05]])
    o64.g2p(dt, np.ones(1, np.int32), np.ones(1, np.int32), x, np.zeros((1, 3)), np.zeros((1, 3, 3)),
            np.eye(3)[None].copy(), np.zeros((1, 3, 3)), cov, 1 / dx, np.ascontiguousarray(gv), True, np.zeros(3))
    S = np.array([[1.0, 0.1, 0.2], [0.1, 2.0, 0.3], [0.2, 0.3, 3.0]])
    S1 = S + dt * (B @ S + S @ B.T)
    assert np.allclose(cov, S1[np.triu_indices(3)], rtol=1e-12)


def test_grid_normalization(o64):
    rng = np.random.default_rng(4)
    dims = (3, 4, 5)
    m = rng.random(dims)
    m[0, 0, 0] = 1e-16
    vin = rng.standard_normal(dims + (3,))
    vout = np.zeros(dims + (3,))
    o64.grid_normalization_and_gravity(m, vin, vout, np.array([0.0, 0.0, -9.81]), 1e-3)
    want = vin / m[..., None] + 1e-3 * np.array([0, 0, -9.81])
    want[0, 0, 0] = 0.0
    assert np.allclose(vout, want, rtol=1e-14)
    o64.add_damping_via_grid(vout, 0.5)
    assert np.allclose(vout, 0.5 * want, rtol=1e-14)


def coul(v, mu, dt, g):
    if abs(v) < 1e-12:
        return 0.0
    return np.sign(v) * max(abs(v) - mu * g * dt, 0.0)


def test_collide_bounds_cases(o64):
    """Lower faces act for index <= pad with inward velocity, upper faces for
    index >= dim - pad; each match rewrites the vector from the original
    velocity, so at a corner the last matching face (z) wins."""
    dims, mu, dt, g, pad = (8, 8, 8), 0.5, 1e-2, 9.81, 2
    rng = np.random.default_rng(5)
    v = rng.standard_normal(dims + (3,))
    v[4, 4, 4] = [-5, -5, -5]
    v[0, 0, 0] = [-1.0, -0.02, -3.0]
    v[7, 6, 3] = [2.0, 0.5, 0.01]
    v[2, 5, 5] = [-1e-13, 0.3, -0.3]
    out = v.copy()
    o64.collideBounds(out, *dims, mu, dt, g, pad)
    assert np.array_equal(out[4, 4, 4], v[4, 4, 4])
    assert np.allclose(out[0, 0, 0], [coul(-1.0, mu, dt, g), coul(-0.02, mu, dt, g), 0.0])
    # x-upper and y-upper both match: the later (y) face wins
    assert np.allclose(out[7, 6, 3], [coul(2.0, mu, dt, g), 0.0, coul(0.01, mu, dt, g)])
    # brute force every node against the rule
    ref = v.copy()
    for ix in range(8):
        for iy in range(8):
            for iz in range(8):
                o = v[ix, iy, iz]
                r = o.copy()
                for ax, i in enumerate((ix, iy, iz)):
                    for lower in (True, False):
                        hit = (i <= pad and o[ax] < 0) if lower else (i >= dims[ax] - pad and o[ax] > 0)
                        if hit:
                            r = np.array([0.0 if a == ax else coul(o[a], mu, dt, g) for a in range(3)])
                ref[ix, iy, iz] = r
    assert np.allclose(out, ref, rtol=0, atol=0)


# ---- the other grid boundary conditions and the volumetric locking correction ----

def _grid(gx, gy, gz, v):
    g = np.zeros((gx, gy, gz, 3))
    g[...] = v
    return g


def test_bc_variants_hand_cases(o64):
    mu, dt, gm, pad = 0.5, 1e-3, 9.81, 2
    dv = mu * gm * dt
    # restitution: x face, approaching, tangential friction
    g = _grid(8, 8, 8, (-2.0, 0.5, 0.0))
    o64.collide_bounds_variant("restitution", g, 8, 8, 8, mu, dt, gm, pad, restitution=0.5)
    assert np.allclose(g[0, 4, 4], (1.0, 0.5 - dv, 0.0))
    assert np.allclose(g[4, 4, 4], (-2.0, 0.5, 0.0))  # interior untouched
    # gradual friction: fade = index / pad on the lower x face
    g = _grid(8, 8, 8, (-1.0, 0.5, 0.0))
    o64.collide_bounds_variant("friction_gradual", g, 8, 8, 8, mu, dt, gm, pad)
    assert np.allclose(g[1, 4, 4], (-0.5, 0.5 - mu * 0.5 * gm * dt, 0.0))
    assert np.allclose(g[0, 4, 4], (0.0, 0.5 - dv, 0.0))
    assert np.allclose(g[2, 4, 4], (-1.0, 0.5, 0.0))  # index 2 is not < pad
    # absorbing: a literal 3 outer layers; the last matching face scales the original v
    g = _grid(8, 8, 8, (2.0, 2.0, 2.0))
    o64.collide_bounds_variant("absorbing", g, 8, 8, 8, mu, dt, gm, pad)
    assert np.allclose(g[0, 4, 4], 0.0) and np.allclose(g[1, 4, 4], 1.0) and np.allclose(g[2, 4, 4], 2.0)
    assert np.allclose(g[0, 1, 4], 1.0)  # x says 0, y (tested later) says 0.5: y wins
    assert np.allclose(g[7, 4, 4], 0.0) and np.allclose(g[5, 4, 4], 2.0)
    # plane strain: v_y zeroed on the y faces whatever its sign; everywhere in the global form
    g = _grid(8, 8, 8, (1.0, 2.0, 3.0))
    o64.collide_bounds_variant("plane_strain_y", g, 8, 8, 8, mu, dt, gm, pad)
    assert np.allclose(g[4, 0, 4], (1.0, 0.0, 3.0)) and np.allclose(g[4, 4, 4], (1.0, 2.0, 3.0))
    g = _grid(8, 8, 8, (-1.0, 2.0, 3.0))
    o64.collide_bounds_variant("plane_strain_y_global", g, 8, 8, 8, mu, dt, gm, pad)
    assert np.allclose(g[4, 4, 4], (-1.0, 0.0, 3.0))
    assert np.allclose(g[0, 4, 4], (0.0, 0.0, 3.0 - dv))


def test_plane_strain_y_reread_feeds_the_z_faces(o64):
    """On a y face the node is read again after the x faces: the z faces then
    test the velocity the x faces produced (mpmRoutines.py:1035-1061)."""
    mu, dt, gm, pad = 0.5, 1e-3, 9.81, 2
    g = _grid(8, 8, 8, (-1.0, 2.0, -3.0))
    o64.collide_bounds_variant("plane_strain_y", g, 8, 8, 8, mu, dt, gm, pad)
    dv = mu * gm * dt
    # corner x=0, y=0, z=0: x face -> (0, f(2), f(-3)); the y face re-reads that node into v and
    # writes (0, 0, f(-3)); the z face then writes (f(0), f(v_y), 0) from the re-read v, so v_y
    # comes back as f(f(2)) (a quirk of the reference kept for parity)
    assert np.allclose(g[0, 0, 0], (0.0, 2.0 - 2 * dv, 0.0))
    # x=0, z=0, interior y: no re-read, z face on the original v
    assert np.allclose(g[0, 4, 0], (-1.0 + dv, 2.0 - dv, 0.0))


def test_volumetric_locking_correction_closed_form(o64):
    rng = np.random.default_rng(5)
    n = 400
    x = rng.uniform(0.1, 3.9, (n, 3))
    C = rng.normal(0, 1, (n, 3, 3))
    vol = rng.uniform(0.5, 2.0, n)
    act = np.ones(n, np.int32)
    lab = np.ones(n, np.int32)
    lab[:20] = 2  # XPBD particles neither contribute nor change
    C0 = C.copy()
    cd, cv = o64.volumetric_locking_correction(act, lab, x, C, vol, 2.0, np.zeros(3), 8, 8, 8)
    cell = (x * 2.0).astype(int)
    key = (cell[:, 0] * 8 + cell[:, 1]) * 8 + cell[:, 2]
    mpm = lab <= 1
    assert np.array_equal(C[~mpm], C0[~mpm])
    eye = np.eye(3)[None]
    dev = lambda M: M - np.trace(M, axis1=1, axis2=2)[:, None, None] / 3 * eye  # noqa: E731
    assert np.allclose(dev(C[mpm]), dev(C0[mpm]), atol=1e-12)  # shear part untouched
    tr, tr0 = np.trace(C, axis1=1, axis2=2), np.trace(C0, axis1=1, axis2=2)
    for k in np.unique(key[mpm]):
        sel = mpm & (key == k)
        want = np.sum(tr0[sel] * vol[sel]) / np.sum(vol[sel])
        assert np.allclose(tr[sel], want, atol=1e-12)  # every particle carries the cell's mean divergence
        assert np.isclose(cv.reshape(-1)[k], vol[sel].sum())


def test_volumetric_locking_cell_follows_float32(o64):
    """The cell is int((x - minBounds) * inv_dx) in float32 (mpmRoutines.py:840-844):
    a particle sitting on a cell face by construction (the lowest particle is
    grid_padding cells above minBounds) goes where float32 rounding puts it."""
    o = o64
    x0, mb = np.float32(0.07561769), np.float32(-3.9243823)
    c32 = int((x0 - mb) * np.float32(2.0))
    c64 = int((np.float64(x0) - np.float64(mb)) * 2.0)
    assert (c32, c64) == (8, 7)  # the two precisions disagree on this face
    x = np.array([[x0, 1.0, 1.0]], np.float64)
    C = np.zeros((1, 3, 3))
    cd, cv = o.volumetric_locking_correction(np.ones(1, np.int32), np.ones(1, np.int32), x, C, np.ones(1), 2.0,
                                             np.array([mb, 0.0, 0.0]), 16, 16, 16)
    assert cv[8, 2, 2] == 1.0 and cv.sum() == 1.0


def _kirchhoff_diag(F, mu, lam):
    """Closed-form Kirchhoff stress of a diagonal F: 2 mu ln F_ii + lam sum ln F_jj."""
    e = np.log(np.stack([F[:, 0], F[:, 4], F[:, 8]], axis=1))
    return 2 * mu[:, None] * e + lam[:, None] * e.sum(axis=1, keepdims=True)


def test_geostatic_F_solves_the_stress_inverse_problem(o64):
    """initialize_geostatic_F (mpmRoutines.py:52-124): the Kirchhoff stress of the
    F it returns is (K0 sv, K0 sv, sv) with sv = -rho g (z_top - z), at every
    depth (tension above z_top included), and F is diagonal."""
    rng = np.random.default_rng(11)
    n = 300
    E = np.exp(rng.uniform(np.log(1e7), np.log(1e9), n))
    nu = rng.uniform(0.05, 0.45, n)
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    rho = rng.uniform(1500, 3000, n)
    x = rng.uniform(-20, 20, (n, 3))
    F = np.zeros((n, 9))
    g, z_top, K0 = 9.81, 15.0, 0.6
    o64.initialize_geostatic_F(x, F, mu, lam, rho, g, z_top, K0)
    sv = -rho * g * (z_top - x[:, 2])
    tau = _kirchhoff_diag(F, mu, lam)
    want = np.stack([K0 * sv, K0 * sv, sv], axis=1)
    assert np.allclose(tau, want, rtol=1e-9, atol=1e-9 * np.abs(sv).max())
    assert np.array_equal(F[:, [1, 2, 3, 5, 6, 7]], np.zeros((n, 6)))


def test_geostatic_F_plane_strain(o64):
    """initialize_geostatic_F_planestrain (mpmRoutines.py:128-203): eps_yy = 0,
    tau_xx = K0 sv, tau_zz = sv, tau_yy = nu (tau_xx + tau_zz) below z_top, and
    F = I at or above it."""
    rng = np.random.default_rng(12)
    n = 300
    E = np.exp(rng.uniform(np.log(1e7), np.log(1e9), n))
    nu = rng.uniform(0.05, 0.45, n)
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    rho = rng.uniform(1500, 3000, n)
    x = rng.uniform(-20, 20, (n, 3))
    F = np.full((n, 9), 7.0)
    g, z_top, K0 = 9.81, 10.0, 0.45
    o64.initialize_geostatic_F(x, F, mu, lam, rho, g, z_top, K0, plane_strain=True)
    below = x[:, 2] < z_top
    assert np.array_equal(F[~below], np.tile(np.eye(3).ravel(), ((~below).sum(), 1)))
    assert np.array_equal(F[:, 4], np.ones(n))
    sv = -rho * g * (z_top - x[:, 2])
    tau = _kirchhoff_diag(F, mu, lam)[below]
    s = np.abs(sv).max()
    assert np.allclose(tau[:, 0], (K0 * sv)[below], rtol=1e-9, atol=1e-9 * s)
    assert np.allclose(tau[:, 2], sv[below], rtol=1e-9, atol=1e-9 * s)
    assert np.allclose(tau[:, 1], nu[below] * (tau[:, 0] + tau[:, 2]), rtol=1e-9, atol=1e-9 * s)


def test_deactivate_and_teleport_hand_case(o64):
    x = np.array([[0, 0, -1.0], [0, 0, -1.0], [0, 0, -1.0], [0, 0, 1.0], [0, 0, -1.0]])
    v = np.ones((5, 3))
    act = np.array([1, 1, 0, 1, 1], np.int32)
    lab = np.array([2, 1, 2, 2, 2], np.int32)
    o64.deactivate_and_teleport(x, v, act, lab, 0.0, 30.0)
    # only active XPBD particles below the datum move, stop and switch off
    assert act.tolist() == [0, 1, 0, 1, 0]
    assert x[:, 2].tolist() == [-31.0, -1.0, -1.0, 1.0, -31.0]
    assert v.tolist() == [[0, 0, 0], [1, 1, 1], [1, 1, 1], [1, 1, 1], [0, 0, 0]]


def test_openmp_build_matches_serial_fp32():
    """The all-cores CPU baseline (FP32 with OpenMP, P2G into a grid per
    thread) computes what the serial FP32 restatement computes: the particle
    kernels bit for bit, the grid sums up to their order (FP32 rounding)."""
    from oracle import Oracle
    s32, omp = Oracle(32), Oracle(32, omp=True)
    assert omp.threads() >= 1
    rng = np.random.default_rng(12)
    n, g = 3000, (12, 11, 10)
    x = (rng.uniform(1.2, 3.8, (n, 3)) * np.array([1.0, 0.9, 0.8])).astype(np.float32)
    v = rng.standard_normal((n, 3)).astype(np.float32)
    C_ = (0.3 * rng.standard_normal((n, 3, 3))).astype(np.float32)
    S = (1e4 * rng.standard_normal((n, 3, 3))).astype(np.float32)
    vol = np.full(n, 0.015625, np.float32)
    mass = np.full(n, 39.0625, np.float32)
    act = np.ones(n, np.int32)
    lab = np.ones(n, np.int32)
    out = []
    for o in (s32, omp):
        gm = np.zeros(g, np.float32)
        gv = np.zeros(g + (3,), np.float32)
        bad = o.p2g_apic_with_stress(act, lab, S, x, v, C_, vol, mass, 0.5, 2.0, np.zeros(3, np.float32), 0.0, gm,
                                     gv, 1e-4)
        go = np.zeros(g + (3,), np.float32)
        o.grid_normalization_and_gravity(gm, gv, go, np.array([0, 0, -9.81], np.float32), 1e-4)
        xo, vo, Co = x.copy(), v.copy(), C_.copy()
        Ft = np.zeros((n, 3, 3), np.float32)
        cov = np.zeros((n, 6), np.float32)
        F = np.broadcast_to(np.eye(3, dtype=np.float32), (n, 3, 3)).copy()
        o.g2p(1e-4, act, lab, xo, vo, Co, F, Ft, cov, 2.0, go, True, np.zeros(3, np.float32))
        out.append((bad, gm, gv, go, xo, vo, Co, Ft))
    assert out[0][0] == out[1][0] == 0
    for a, b in zip(out[0][1:4], out[1][1:4]):
        assert np.max(np.abs(a - b)) <= 1e-5 * np.max(np.abs(a))
    # G2P of the same grid is per particle: identical
    o = out[1]
    xo, vo, Co = x.copy(), v.copy(), C_.copy()
    Ft = np.zeros((n, 3, 3), np.float32)
    F = np.broadcast_to(np.eye(3, dtype=np.float32), (n, 3, 3)).copy()
    s32.g2p(1e-4, act, lab, xo, vo, Co, F, Ft, np.zeros((n, 6), np.float32), 2.0, o[3], True,
            np.zeros(3, np.float32))
    assert np.array_equal(xo, o[4]) and np.array_equal(vo, o[5]) and np.array_equal(Ft, o[7])
