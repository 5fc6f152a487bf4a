"""Seeded device states and their oracle copies for the GPU parity tests.

``make_sim`` builds a scene through the device driver and then perturbs it into
a state that exercises every branch of the substep (elastic, plastic,
transition, locked label 0, already-XPBD label 2, inactive particles), so that
one teacher-forced substep compares all code paths.  ``host_state`` downloads
every hot-path array in the reference's AoS layout as float64 (ints as int32)
for the FP64 oracle.
"""
from __future__ import annotations

import numpy as np

F_FIELDS_SIM = ["particle_x", "particle_v", "particle_mass", "particle_vol", "particle_radius", "particle_density",
                "materialLabel", "activeLabel"]
F_FIELDS_MPM = ["particle_F", "particle_F_trial", "particle_stress", "particle_accumulated_strain", "particle_damage",
                "particle_C", "particle_cov", "mu", "lam", "ys", "alpha", "hardening", "softening", "eta_shear",
                "eta_bulk", "strainCriteria", "grid_m", "grid_v_in", "grid_v_out"]
F_FIELDS_XPBD = ["particle_x_initial_xpbd", "particle_v_initial_xpbd", "particle_cumDist_xpbd"]


def make_sim(scene="block_collapse", scale=0.5, seed=0, perturb=True, **over):
    from mpmwarpydw_amd import driver, getArgs, scenes
    dom, ov = scenes.SCENES[scene](scale=scale)
    ov.update(over)
    args = getArgs.get_args([], ov)
    S = driver.setup(args, dom)
    if perturb:
        perturb_state(S, seed)
    return S


def perturb_state(S, seed):
    rng = np.random.default_rng(seed)
    sim, mpm = S.sim, S.mpm
    n = sim.nPoints
    v = 0.2 * rng.standard_normal((n, 3))
    sim.particle_v.assign(v)
    C = 0.3 * rng.standard_normal((n, 3, 3))
    mpm.particle_C.assign(C)
    Ft = np.eye(3) + 2e-3 * rng.standard_normal((n, 3, 3))
    mpm.particle_F_trial.assign(Ft)
    mpm.particle_F.assign(np.eye(3) + 2e-3 * rng.standard_normal((n, 3, 3)))
    ys = mpm.ys.numpy().astype(np.float64)
    weak = rng.random(n) < 0.3
    ys[weak] *= 0.05  # plastic
    mpm.ys.assign(ys)
    sc = mpm.strainCriteria.numpy().astype(np.float64)
    brittle = rng.random(n) < 0.1
    sc[brittle] = 1e-6  # transitions
    mpm.strainCriteria.assign(sc)
    dmg = 0.5 * rng.random(n)
    mpm.particle_damage.assign(dmg)
    lab = sim.materialLabel.numpy()
    r = rng.random(n)
    lab[r < 0.1] = 0
    lab[(r >= 0.1) & (r < 0.15)] = 2
    sim.materialLabel.assign(lab)
    act = np.ones(n, np.int32)
    act[rng.random(n) < 0.03] = 0
    sim.activeLabel.assign(act)
    mpm.eta_shear.assign(np.full(n, 50.0))
    mpm.eta_bulk.assign(np.full(n, 20.0))
    mpm.eff = 0.05
    stress = 1e4 * rng.standard_normal((n, 3, 3))
    mpm.particle_stress.assign(stress + stress.transpose(0, 2, 1))


def host_state(S):
    sim, mpm, xpbd = S.sim, S.mpm, S.xpbd
    out = {}
    for name in F_FIELDS_SIM:
        out[name] = getattr(sim, name).numpy()
    for name in F_FIELDS_MPM:
        out[name] = getattr(mpm, name).numpy()
    for name in F_FIELDS_XPBD:
        out[name] = getattr(xpbd, name).numpy()
    for k, a in out.items():
        out[k] = np.ascontiguousarray(a.astype(np.int32) if a.dtype == np.int32 else a.astype(np.float64))
    return out


def oracle_stress(o, H, eff):
    o.compute_stress_from_F_trial(H["activeLabel"], H["materialLabel"], H["particle_x"], H["particle_v"],
                                  H["particle_x_initial_xpbd"], H["particle_v_initial_xpbd"], H["particle_F"],
                                  H["particle_F_trial"], H["mu"], H["lam"], H["ys"], H["hardening"], H["softening"],
                                  H["particle_density"], H["strainCriteria"], eff, H["eta_shear"], H["eta_bulk"],
                                  H["particle_C"], H["particle_stress"], H["particle_accumulated_strain"],
                                  H["particle_damage"], 1, H["alpha"])


def oracle_p2g(o, H, S):
    sim, mpm = S.sim, S.mpm
    H["grid_m"][:] = 0
    H["grid_v_in"][:] = 0
    return o.p2g_apic_with_stress(H["activeLabel"], H["materialLabel"], H["particle_stress"], H["particle_x"],
                                  H["particle_v"], H["particle_C"], H["particle_vol"], H["particle_mass"],
                                  np.float32(sim.dx), np.float32(sim.invdx), np.float32(sim.minBounds),
                                  mpm.rpic_damping, H["grid_m"], H["grid_v_in"], np.float32(sim.dt))


def oracle_grid(o, H, S):
    sim, mpm = S.sim, S.mpm
    H["grid_v_out"][:] = 0
    o.grid_normalization_and_gravity(H["grid_m"], H["grid_v_in"], H["grid_v_out"],
                                     np.asarray(sim.gravity, np.float32), np.float32(sim.dt))
    if mpm.grid_v_damping_scale < 1.0:
        o.add_damping_via_grid(H["grid_v_out"], mpm.grid_v_damping_scale)
    args = (*mpm.gridDims, np.float32(mpm.boundFriction), np.float32(sim.dt), np.float32(abs(sim.gravity[2])),
            mpm.boundaryPadding)
    if mpm.boundaryCondition == "friction":
        o.collideBounds(H["grid_v_out"], *args)
    else:
        o.collide_bounds_variant(mpm.boundaryCondition, H["grid_v_out"], *args,
                                 restitution=np.float32(mpm.boundRestitution))


def oracle_vlc(o, H, S):
    """Step 8 of the substep (simulationRoutines.py:252-293) when it is on."""
    sim, mpm = S.sim, S.mpm
    if mpm.volumetric_locking_correction:
        o.volumetric_locking_correction(H["activeLabel"], H["materialLabel"], H["particle_x"], H["particle_C"],
                                        H["particle_vol"], np.float32(sim.invdx), np.float32(sim.minBounds),
                                        *mpm.gridDims)


def oracle_g2p(o, H, S):
    sim, mpm = S.sim, S.mpm
    return o.g2p(np.float32(sim.dt), H["activeLabel"], H["materialLabel"], H["particle_x"], H["particle_v"],
                 H["particle_C"], H["particle_F"], H["particle_F_trial"], H["particle_cov"], np.float32(sim.invdx),
                 H["grid_v_out"], bool(mpm.update_cov), np.float32(sim.minBounds))


def oracle_xpbd_step(o, H, S, xpbd_only=False):
    """xpbdSimulationStep (simulationRoutines.py:295-526) on the host copy H
    with the oracle's kernels, in the reference's launch order: hash grid of the
    pre-step positions, integration, the label reset when the transition lock
    is on (not in xpbd_only mode), the iterations of wall contacts, particle
    contacts and delta application, sleeping, the MPM reset, the commit and,
    when swellingRatio > 0, swelling.  Updates H in place; returns the sleep
    counters (total, sleeping)."""
    sim, xpbd = S.sim, S.xpbd
    f32 = lambda v: float(np.float32(v))  # noqa: E731
    n = sim.nPoints
    x0, v0 = H["particle_x"].copy(), H["particle_v"].copy()
    xi, vi = x0.copy(), v0.copy()
    dims = xpbd.particle_grid.dims
    ids, cs, ce = o.hashgrid_build(x0, np.float32(sim.dx), dims)
    o.integrateParticlesXPBD(H["activeLabel"], H["materialLabel"], xi, vi, np.float32(sim.gravity), f32(sim.dtxpbd),
                             xi, vi, f32(xpbd.particle_v_max))
    if xpbd.mpm_contact_transition_lock and not xpbd_only:
        o.reset_mpm_contact_labels(H["activeLabel"], H["materialLabel"])
    lock = 0 if xpbd_only else int(xpbd.mpm_contact_transition_lock)
    ft = H["particle_x"].dtype
    for _ in range(xpbd.xpbd_iterations):
        d = np.zeros((n, 3), ft)
        o.my_solve_particle_bound_contacts(H["activeLabel"], H["materialLabel"], xi, vi, H["particle_mass"],
                                           H["particle_radius"], f32(xpbd.dynamicParticleFriction),
                                           f32(xpbd.staticVelocityThreshold), f32(xpbd.staticParticleFriction),
                                           np.float32(xpbd.minBoundsXPBD), np.float32(xpbd.maxBoundsXPBD),
                                           f32(sim.dtxpbd), f32(xpbd.xpbd_relaxation), d)
        o.my_solve_particle_particle_contacts(H["activeLabel"], H["materialLabel"],
                                              (ids, cs, ce, np.float32(sim.dx), dims), xi, vi, H["particle_mass"],
                                              H["particle_radius"], f32(xpbd.dynamicParticleFriction),
                                              f32(xpbd.staticVelocityThreshold), f32(xpbd.staticParticleFriction),
                                              f32(xpbd.particle_cohesion), f32(xpbd.max_radius), f32(sim.dtxpbd),
                                              f32(xpbd.xpbd_relaxation), lock, d)
        xo, vo = xi.copy(), vi.copy()
        o.my_apply_particle_deltas(H["activeLabel"], H["materialLabel"], x0, xi, d, f32(sim.dtxpbd),
                                   f32(xpbd.particle_v_max), xo, vo)
        xi, vi = xo, vo
    counts = o.sleepParticles(H["activeLabel"], H["materialLabel"], f32(xpbd.sleepThreshold), H["particle_radius"],
                              x0, xi, H["particle_cumDist_xpbd"], f32(sim.dtxpbd))
    o.apply_mpm_contact_impulses(H["activeLabel"], H["materialLabel"], x0, v0, xi, vi, f32(sim.dtxpbd),
                                 f32(xpbd.particle_v_max), f32(xpbd.xpbd_mpm_coupling_strength), int(xpbd_only), vi)
    H["particle_x"], H["particle_v"] = xi, vi
    if xpbd.swellingRatio > 0:
        o.swellParticlesType2(H["activeLabel"], H["materialLabel"], H["particle_x_initial_xpbd"], H["particle_x"],
                              H["particle_radius"], f32(xpbd.swellingActivationFactor), f32(xpbd.swellingMaxFactor),
                              np.ascontiguousarray(xpbd.particleMaxRadius.numpy(), ft),
                              np.ascontiguousarray(xpbd.particleBaseRadius.numpy(), ft))
    return counts


def large_and_inverted_F(S, seed, large=0.15, inverted=0.1):
    """Overwrite F_trial of a fraction of the particles with states that take
    the SVD route of the constitutive update: rotated stretches of 20-60% (so
    |F F^T - I|_F > 0.25) and reflections (det F < 0, where the reference's
    clamps of ln s at 0.01 and 1e-6 act, mpmRoutines.py:341,460-462).  Returns
    the masks (large, inverted)."""
    rng = np.random.default_rng(seed)
    n = S.sim.nPoints
    Ft = S.mpm.particle_F_trial.numpy().astype(np.float64)
    r = rng.random(n)
    big = r < large
    inv = (r >= large) & (r < large + inverted)
    m = int(big.sum() + inv.sum())
    Q, _ = np.linalg.qr(rng.normal(size=(m, 3, 3)))
    Q *= np.sign(np.linalg.det(Q))[:, None, None]
    A = rng.normal(size=(m, 3, 3))
    A = A + A.transpose(0, 2, 1)
    A /= np.linalg.norm(A, axis=(1, 2))[:, None, None]
    nb = int(big.sum())
    mag = np.concatenate([rng.uniform(0.2, 0.6, nb), rng.uniform(0.0, 0.05, m - nb)])
    U = np.eye(3)[None] + A * mag[:, None, None]
    U[nb:, 2, :] *= -1.0  # reflections
    # a few inverted states squashed below the 0.01 clamp on one axis
    sq = np.flatnonzero(rng.random(m - nb) < 0.3) + nb
    U[sq, 2, :] *= 0.005
    sel = np.concatenate([np.flatnonzero(big), np.flatnonzero(inv)])
    Ft[sel] = Q @ U
    S.mpm.particle_F_trial.assign(Ft)
    return big, inv


# ---- error scales of the sums the substep takes ------------------------------
# A per-particle (per-node) error bound relative to the item's own value fails
# wherever the item is a sum whose terms cancel: the FP32 rounding of each term
# stays, the sum shrinks.  The scales below are the sums of the magnitudes of
# the terms (the condition of each sum), computed in FP64 from the same inputs
# with the reference's weights (mpmRoutines.py:474-569, :715-776).  They only
# scale errors; the expected values come from the oracle.

def _stencil(x, minb, inv_dx):
    gp = (x - np.asarray(minb, np.float64)) * inv_dx
    base = np.trunc(gp - 0.5).astype(np.int64)
    f = gp - base
    w = np.stack([0.5 * (1.5 - f) ** 2, 0.75 - (f - 1.0) ** 2, 0.5 * (f - 0.5) ** 2], -1)  # [n][axis][q]
    dw = np.stack([f - 1.5, -2.0 * (f - 1.0), f - 0.5], -1)
    return base, f, w, dw


def p2g_term_scale(H, S):
    """Per grid node: (sum of |mass terms|, largest over components of the sum
    of |momentum terms|) of P2G, over the particles that transfer (active,
    label <= 1 after the constitutive update, stencil inside the grid)."""
    sim, mpm = S.sim, S.mpm
    gx, gy, gz = mpm.gridDims
    dx, inv_dx, dt = float(np.float32(sim.dx)), float(np.float32(sim.invdx)), float(np.float32(sim.dt))
    x = H["particle_x"]
    base, f, w, dw = _stencil(x, sim.minBounds, inv_dx)
    sel = (H["activeLabel"] == 1) & (H["materialLabel"] <= 1) & np.all(base >= 0, 1) & \
        (base[:, 0] <= gx - 3) & (base[:, 1] <= gy - 3) & (base[:, 2] <= gz - 3)
    b, f, w, dw = base[sel], f[sel], w[sel], dw[sel]
    m, vol = H["particle_mass"][sel], H["particle_vol"][sel]
    v, C_, tau = H["particle_v"][sel], H["particle_C"][sel], H["particle_stress"][sel]
    r = float(mpm.rpic_damping)
    Cm = np.zeros_like(C_) if r < -0.001 else (1 - r) * C_ + 0.5 * r * (C_ - C_.transpose(0, 2, 1))
    sm = np.zeros(gx * gy * gz)
    sv = np.zeros((gx * gy * gz, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                wt = w[:, 0, i] * w[:, 1, j] * w[:, 2, k]
                g = np.stack([dw[:, 0, i] * w[:, 1, j] * w[:, 2, k], w[:, 0, i] * dw[:, 1, j] * w[:, 2, k],
                              w[:, 0, i] * w[:, 1, j] * dw[:, 2, k]], 1) * inv_dx
                dpos = (np.array([i, j, k]) - f) * dx
                t = (wt * m)[:, None] * (np.abs(v) + np.einsum("nij,nj->ni", np.abs(Cm), np.abs(dpos))) + \
                    (dt * vol)[:, None] * np.einsum("nij,nj->ni", np.abs(tau), np.abs(g))
                node = ((b[:, 0] + i) * gy + (b[:, 1] + j)) * gz + (b[:, 2] + k)
                np.add.at(sm, node, wt * m)
                np.add.at(sv, node, t)
    return sm, sv.max(axis=1)


def g2p_term_scale(H, S):
    """Per particle: the sums of the magnitudes of the G2P terms of v
    (sum w |v_i|), C (4 inv_dx sum w |v_i| |q - f|) and grad v
    (sum |v_i| |grad w|), each its largest over components, from the grid
    velocities H['grid_v_out']."""
    sim, mpm = S.sim, S.mpm
    gx, gy, gz = mpm.gridDims
    inv_dx = float(np.float32(sim.invdx))
    base, f, w, dw = _stencil(H["particle_x"], sim.minBounds, inv_dx)
    n = len(base)
    ok = np.all(base >= 0, 1) & (base[:, 0] <= gx - 3) & (base[:, 1] <= gy - 3) & (base[:, 2] <= gz - 3)
    bb = np.where(ok[:, None], base, 0)
    gv = H["grid_v_out"].reshape(-1, 3)
    sv, sc, sg = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    for i in range(3):
        for j in range(3):
            for k in range(3):
                wt = w[:, 0, i] * w[:, 1, j] * w[:, 2, k]
                g = np.stack([dw[:, 0, i] * w[:, 1, j] * w[:, 2, k], w[:, 0, i] * dw[:, 1, j] * w[:, 2, k],
                              w[:, 0, i] * w[:, 1, j] * dw[:, 2, k]], 1) * inv_dx
                node = ((bb[:, 0] + i) * gy + (bb[:, 1] + j)) * gz + (bb[:, 2] + k)
                av = np.abs(gv[node])
                sv += wt[:, None] * av
                sc += 4 * inv_dx * wt[:, None] * av * np.abs(np.array([i, j, k]) - f).max(1)[:, None]
                sg += av * np.abs(g).max(1)[:, None]
    return sv.max(1) * ok, sc.max(1) * ok, sg.max(1) * ok


def dev_strain_norm(F):
    """|dev ln V| of each F (principal stretches clamped at 0.01 like the return
    map, mpmRoutines.py:341): the scale of the plastic multiplier, which the
    return map forms as a difference of near-equal strains."""
    s = np.linalg.svd(F, compute_uv=False)
    e = np.log(np.maximum(s, 0.01))
    d = e - e.mean(axis=1, keepdims=True)
    return np.linalg.norm(d, axis=1)


def oracle_substep(o, H, S):
    """One mpmSimulationStep on the host copy H (stress, P2G, grid update, G2P,
    VLC when on); returns the error scales of its sums (p2g_term_scale before
    G2P moves the particles, g2p_term_scale of the new grid velocities)."""
    oracle_stress(o, H, S.mpm.eff)
    oracle_p2g(o, H, S)
    oracle_grid(o, H, S)
    sc = {"p2g": p2g_term_scale(H, S), "g2p": g2p_term_scale(H, S)}
    oracle_g2p(o, H, S)
    oracle_vlc(o, H, S)
    return sc


def oracle_transfers32(o32, Hin, G, S):
    """P2G, grid update, G2P and VLC of the FP32 oracle -- the reference's own
    FP32 arithmetic, whose grid coordinates (x - min) / dx and stencil offsets
    round exactly as the GPU's -- from the state before the substep, with the
    GPU's constitutive results (stress, F, labels, and the velocity of the
    particles that turned XPBD) in place of its own, so that only the transfers
    are compared.  Returns the result as float64 arrays and the error scales of
    its sums (oracle_substep)."""
    H = {k: (np.ascontiguousarray(v, np.float32) if v.dtype == np.float64 else v.copy()) for k, v in Hin.items()}
    H["particle_stress"] = np.ascontiguousarray(G["particle_stress"], np.float32)
    H["particle_F"] = np.ascontiguousarray(G["particle_F"], np.float32)
    H["materialLabel"] = G["materialLabel"].copy()
    turned = (Hin["materialLabel"] <= 1) & (G["materialLabel"] == 2)
    H["particle_v"][turned] = G["particle_v"][turned]
    oracle_p2g(o32, H, S)
    oracle_grid(o32, H, S)
    H64 = {k: v.astype(np.float64) if v.dtype == np.float32 else v for k, v in H.items()}
    sc = {"p2g": p2g_term_scale(H64, S), "g2p": g2p_term_scale(H64, S)}
    oracle_g2p(o32, H, S)
    oracle_vlc(o32, H, S)
    return {k: v.astype(np.float64) if v.dtype == np.float32 else v for k, v in H.items()}, sc


def c3_state(scale=1.61, seed=3, plastic_frac=0.5):
    """The learned-map workload: an undercut slab of about 4.0M particles whose
    materials follow the slab's distribution (E lognormal, median 5e8 Pa,
    sigma_ln 0.3; cohesion lognormal, median 2e5 Pa, sigma_ln 0.4; drawn from
    default_rng(seed)), with trial deformation gradients placed around the
    yield surface: compressive volumetric strain up to two cohesions of
    pressure, damage 0 (80%) or U(0, 0.9), and a deviator that is inside the
    cone (the elastic share) or past it by an overshoot log-uniform in
    [1e-5, 0.3] cohesions (``plastic_frac``), in random principal frames."""
    from mpmwarpydw_amd import driver, getArgs, scenes
    dom, ov = scenes.undercut_slab(scale=scale, seed=seed)
    S = driver.setup(getArgs.get_args([], ov), dom)
    n = S.sim.nPoints
    rng = np.random.default_rng(seed + 100)
    mu = S.mpm.mu.numpy().astype(np.float64)
    lam = S.mpm.lam.numpy().astype(np.float64)
    ys = S.mpm.ys.numpy().astype(np.float64)
    al = S.mpm.alpha.numpy().astype(np.float64)
    K = lam + 2 * mu / 3
    D = np.where(rng.random(n) < 0.8, 0.0, rng.uniform(0, 0.9, n))
    tr = -rng.uniform(0, 2, n) * ys / K
    Y = (1 - D) * (ys - al * K * tr)
    c = np.maximum(ys, 1.0)
    plastic = rng.random(n) < plastic_frac
    over = 10 ** rng.uniform(-5, np.log10(0.3), n)
    dmag = np.where(plastic, Y + c * over, Y * rng.uniform(0, 0.99, n)) / (2 * mu)
    w = rng.standard_normal((n, 3))
    w -= w.mean(1, keepdims=True)
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    e = tr[:, None] / 3 + dmag[:, None] * w
    Q1, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    Q2, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    Ft = np.einsum("nij,nj,nkj->nik", Q1, np.exp(e), Q2)
    S.mpm.particle_F_trial.assign(Ft)
    S.mpm.particle_damage.assign(D)
    return S
