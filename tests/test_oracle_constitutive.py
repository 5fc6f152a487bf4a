"""Pins the oracle's constitutive path (mpmRoutines.py:206-471: Lame constants,
svd3, Drucker-Prager / von Mises return map with damage and phase transition,
Kirchhoff stress, Kelvin-Voigt) against mathematics that does not repeat its
formulas: LAPACK's SVD, closed-form stresses of pure stretches, the
post-projection identity of the one-shot return map, isotropy, the geostatic
inverse problem and hand-computed transition cases.  CPU only."""
import numpy as np
import pytest

from conftest import rel_err


def rot(rng):
    q, r = np.linalg.qr(rng.standard_normal((3, 3)))
    q = q @ np.diag(np.sign(np.diag(r)))
    if np.linalg.det(q) < 0:
        q[:, 0] *= -1
    return q


def particles(n, mu=4e7, lam=4e7, ys=1e5, alpha=0.2, hard=0.0, soft=0.0, sc=0.05, damage=0.0, label=1):
    f = lambda v: np.full(n, v, dtype=np.float64)
    return dict(active=np.ones(n, np.int32), mat=np.full(n, label, np.int32), x=np.zeros((n, 3)),
                v=np.zeros((n, 3)), x_init=np.full((n, 3), 1e6), v_init=np.zeros((n, 3)), F=np.zeros((n, 3, 3)),
                mu=f(mu), lam=f(lam), ys=f(ys), hard=f(hard), soft=f(soft), density=f(2500.0), sc=f(sc),
                eta_s=f(0.0), eta_b=f(0.0), C=np.zeros((n, 3, 3)), stress=np.zeros((n, 3, 3)), acc=f(0.0),
                damage=f(damage), alpha=f(alpha))


def run_stress(o, P, F_trial, eff=0.0):
    o.compute_stress_from_F_trial(P["active"], P["mat"], P["x"], P["v"], P["x_init"], P["v_init"], P["F"],
                                  np.ascontiguousarray(F_trial, dtype=np.float64), P["mu"], P["lam"], P["ys"],
                                  P["hard"], P["soft"], P["density"], P["sc"], eff, P["eta_s"], P["eta_b"], P["C"],
                                  P["stress"], P["acc"], P["damage"], 1, P["alpha"])


def principal_log(F):
    s = np.linalg.svd(F, compute_uv=False)
    return np.log(s)


# ---------------------------------------------------------------- svd3
def test_svd3_matches_lapack_and_convention(o64):
    rng = np.random.default_rng(0)
    for t in range(200):
        F = np.eye(3) + 0.4 * rng.standard_normal((3, 3))
        if t % 3 == 0:
            F[:, rng.integers(3)] *= -1  # inverted elements: det F < 0
        U, s, V = o64.svd3(F)
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.abs(s), ref, rtol=1e-12, atol=1e-13)  # eigenvalue route: absolute accuracy
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-13)
        assert abs(np.linalg.det(U) - 1) < 1e-12 and abs(np.linalg.det(V) - 1) < 1e-12
        assert np.all(np.diff(np.abs(s)) <= 1e-14)
        assert np.sign(s[2]) == np.sign(np.linalg.det(F)) and s[0] > 0 and s[1] > 0


def test_svd3_keeps_small_strains(o64):
    """F = R diag(1+e) Q^T with e ~ 1e-8: the principal stretches minus one come
    back with ~1e-6 relative error (an SVD formed from F^T F loses them)."""
    rng = np.random.default_rng(1)
    for _ in range(50):
        e = 1e-8 * rng.standard_normal(3)
        R, Q = rot(rng), rot(rng)
        F = R @ np.diag(1 + e) @ Q.T
        _, s, _ = o64.svd3(F)
        assert np.allclose(np.sort(s - 1), np.sort(e), rtol=1e-5, atol=1e-15)


def test_svd3_repeated_values(o64):
    for F in (np.eye(3), 2 * np.eye(3), np.diag([1.5, 1.5, 0.5]), np.diag([-1.0, 1.0, 1.0])):
        U, s, V = o64.svd3(F)
        assert np.allclose(U @ np.diag(s) @ V.T, F, atol=1e-14)


# ---------------------------------------------------------------- A1
def test_lame_closed_form(o64):
    E = np.array([1e8, 5e8, 2.1e11])
    nu = np.array([0.25, 0.2, 0.3])
    mu, lam, bulk = np.zeros(3), np.zeros(3), np.zeros(3)
    o64.compute_mu_lam_bulk(E, nu, mu, lam, bulk)
    assert np.allclose(mu, [4e7, 5e8 / 2.4, 2.1e11 / 2.6], rtol=1e-14)
    assert np.allclose(lam, [4e7, 5e8 * 0.2 / (1.2 * 0.6), 2.1e11 * 0.3 / (1.3 * 0.4)], rtol=1e-14)
    # bulk modulus of linear elasticity E / (3 (1 - 2 nu))
    assert np.allclose(bulk, E / (3 * (1 - 2 * nu)), rtol=1e-13)


# ---------------------------------------------------------------- A4 + elastic branch
def test_elastic_stretch_closed_form(o64):
    """Elastic (huge ys): F unchanged and tau = R diag(2 mu ln s + lam sum ln s) R^T."""
    rng = np.random.default_rng(2)
    n = 20
    P = particles(n, ys=1e30)
    Ft = np.zeros((n, 3, 3))
    want = np.zeros((n, 3, 3))
    for p in range(n):
        s = 1 + 0.05 * rng.standard_normal(3)
        R, Q = rot(rng), rot(rng)
        Ft[p] = R @ np.diag(s) @ Q.T
        ls = np.log(s)
        want[p] = R @ np.diag(2 * 4e7 * ls + 4e7 * ls.sum()) @ R.T
    run_stress(o64, P, Ft)
    assert np.array_equal(P["F"], Ft)
    assert rel_err(P["stress"], want) < 1e-12
    assert np.all(P["acc"] == 0) and np.all(P["damage"] == 0)


def test_geostatic_inverse(o64):
    """The strain that initialize_geostatic_F solves for (mpmRoutines.py:51-124)
    gives back the target K0 stress state exactly."""
    mu, lam, rho, g, K0 = 4e7, 4e7, 2500.0, 9.81, 0.5
    depth = np.array([0.5, 3.0, 10.0])
    n = len(depth)
    sv = -rho * g * depth
    sh = K0 * sv
    A = np.array([[2 * lam + 2 * mu, lam], [2 * lam, lam + 2 * mu]])
    Ft = np.zeros((n, 3, 3))
    for p in range(n):
        eh, ev = np.linalg.solve(A, [sh[p], sv[p]])
        Ft[p] = np.diag(np.exp([eh, eh, ev]))
    P = particles(n, ys=1e30)
    run_stress(o64, P, Ft)
    for p in range(n):
        assert np.allclose(np.diag(P["stress"][p]), [sh[p], sh[p], sv[p]], rtol=1e-9)


def test_kelvin_voigt(o64):
    rng = np.random.default_rng(3)
    n = 5
    P = particles(n, ys=1e30)
    P["eta_s"][:] = 1e5
    P["eta_b"][:] = 3e4
    P["C"][:] = rng.standard_normal((n, 3, 3))
    run_stress(o64, P, np.tile(np.eye(3), (n, 1, 1)))
    for p in range(n):
        S = 0.5 * (P["C"][p] + P["C"][p].T)
        tr = np.trace(S)
        want = 2e5 * (S - tr / 3 * np.eye(3)) + 3e4 * tr * np.eye(3)
        assert np.allclose(P["stress"][p], want, rtol=1e-13, atol=1e-9)


# ---------------------------------------------------------------- A3 plastic branch
def trial_state(F, mu, lam, ys, alpha, D):
    eps = np.log(np.maximum(np.linalg.svd(F, compute_uv=False), 0.01))
    tau = 2 * mu * eps + lam * eps.sum()
    p = tau.mean()
    Y = (1 - D) * (ys - alpha * p)
    edev = eps - eps.mean()
    return eps, tau, p, Y, edev


@pytest.mark.parametrize("alpha,hard,soft", [(0.2309, 0.0, 0.0), (0.0, 0.0, 0.0), (0.3, 2.0, 0.5)])
def test_projection_identities(o64, alpha, hard, soft):
    """After the one-shot projection the deviatoric stress norm is
    Y |e_dev| / (|e_dev| + 1e-6) at the trial pressure and old damage, the
    volumetric strain grows by 3 (0.3 alpha) dgamma, the accumulated strain by
    sqrt(2/3) dgamma and ys by 2 mu (H - S) dgamma."""
    rng = np.random.default_rng(4)
    n = 40
    mu, lam, ys, sc, D0 = 4e7, 4e7, 1e5, 0.05, 0.1
    P = particles(n, mu=mu, lam=lam, ys=ys, alpha=alpha, hard=hard, soft=soft, sc=sc, damage=D0, label=0)
    Ft = np.zeros((n, 3, 3))
    for p in range(n):
        s = np.exp(np.array([3e-3, -1e-3, -2.5e-3]) * (1 + 0.5 * rng.random(3)) - 5e-4 * rng.random())
        Ft[p] = rot(rng) @ np.diag(s) @ rot(rng).T
    run_stress(o64, P, Ft)
    for p in range(n):
        eps, tau, pr, Y, edev = trial_state(Ft[p], mu, lam, ys, alpha, D0)
        sig_eq = np.linalg.norm(tau - pr)
        assert sig_eq > Y > 0  # the cases are plastic
        nrm = np.linalg.norm(edev) + 1e-6
        dgam = nrm - Y / (2 * mu)
        epsn = np.log(np.linalg.svd(P["F"][p], compute_uv=False))
        taun = 2 * mu * epsn + lam * epsn.sum()
        assert np.isclose(np.linalg.norm(taun - taun.mean()), Y * np.linalg.norm(edev) / nrm, rtol=1e-9)
        assert np.isclose(epsn.sum(), eps.sum() + 3 * 0.3 * alpha * dgam, rtol=1e-9, atol=1e-15)
        assert np.isclose(P["acc"][p], np.sqrt(2 / 3) * dgam, rtol=1e-12)
        assert np.isclose(P["damage"][p], min(1.0, D0 + np.sqrt(2 / 3) * dgam / sc), rtol=1e-12)
        assert np.isclose(P["ys"][p], max(0.0, ys + 2 * mu * (hard - soft) * dgam), rtol=1e-12)
        assert P["mat"][p] == 0  # label 0 particles never transition


def test_isotropy(o64):
    rng = np.random.default_rng(5)
    n = 30
    Ft = np.array([np.eye(3) + 4e-3 * rng.standard_normal((3, 3)) for _ in range(n)])
    R = rot(rng)
    P1 = particles(n, label=0)
    P2 = particles(n, label=0)
    run_stress(o64, P1, Ft)
    run_stress(o64, P2, np.einsum("ij,pjk,lk->pil", R, Ft, R))
    assert rel_err(P2["stress"], np.einsum("ij,pjk,lk->pil", R, P1["stress"], R)) < 1e-10
    assert np.allclose(P1["acc"], P2["acc"], rtol=1e-9, atol=1e-18)
    assert np.allclose(P1["damage"], P2["damage"], rtol=1e-9, atol=1e-18)


def test_phase_transition(o64):
    """Damage reaching 1 on a label-1 particle: label 2, birth position, energy
    kick along normalize(v + 1e-12), zeroed material, F_trial returned unprojected,
    stress = viscous part only (mpmRoutines.py:379-408)."""
    mu, lam, rho, eff = 4e7, 4e7, 2500.0, 0.05
    F = np.diag(np.exp([4e-3, -1e-3, -2e-3]))
    P = particles(1, mu=mu, lam=lam, ys=1e5, alpha=0.0, sc=1e-9)
    P["x"][0] = [1.0, 2.0, 3.0]
    P["v"][0] = [0.0, 0.0, -2.0]
    P["eta_s"][:] = 10.0
    P["C"][0] = np.diag([1.0, 2.0, 3.0])
    run_stress(o64, P, F[None], eff=eff)
    eps = np.array([4e-3, -1e-3, -2e-3])
    tau = 2 * mu * eps + lam * eps.sum()
    u = 0.5 * tau @ eps
    kick = np.sqrt(2 * eff * u / rho)
    d = np.array([1e-12, 1e-12, -2.0 + 1e-12])
    assert P["mat"][0] == 2 and P["damage"][0] == 1.0
    assert np.array_equal(P["x_init"][0], P["x"][0])
    assert np.allclose(P["v"][0], [0, 0, -2.0] + kick * d / np.linalg.norm(d), rtol=1e-12)
    assert np.array_equal(P["v_init"][0], P["v"][0])
    assert P["mu"][0] == 0 and P["lam"][0] == 0 and P["ys"][0] == 0 and P["alpha"][0] == 0
    assert np.array_equal(P["F"][0], F)
    assert np.allclose(P["stress"][0], 2 * 10.0 * (np.diag([1.0, 2, 3]) - 2 * np.eye(3)))


def test_forced_damage_and_lock(o64):
    """ys < 0 forces damage 1 (mpmRoutines.py:376-377); a label-0 (locked)
    particle is projected instead of transitioning; zero strainCriteria also
    means full damage."""
    F = np.diag(np.exp([1e-4, 0, -1e-4]))
    P = particles(3, ys=1e5, sc=0.05, label=1)
    P["ys"][0] = -1.0
    P["mat"][1] = 0
    P["ys"][1] = -1.0
    P["sc"][2] = 0.0
    P["ys"][2] = 1.0
    run_stress(o64, P, np.tile(F, (3, 1, 1)))
    assert list(P["damage"]) == [1.0, 1.0, 1.0]
    assert list(P["mat"]) == [2, 0, 2]


def test_fp32_restatement_close_to_fp64(o32, o64):
    """The FP32 restatement (reference arithmetic) stays within its expected
    FP32 distance of the FP64 oracle at moderate strain (1e-3)."""
    rng = np.random.default_rng(6)
    n = 200
    Ft = np.array([np.eye(3) + 1e-3 * rng.standard_normal((3, 3)) for _ in range(n)])
    P64 = particles(n, label=0)
    run_stress(o64, P64, Ft)
    P32 = {k: (v.astype(np.float32) if v.dtype == np.float64 else v.copy()) for k, v in particles(n, label=0).items()}
    o32.compute_stress_from_F_trial(P32["active"], P32["mat"], P32["x"], P32["v"], P32["x_init"], P32["v_init"],
                                    P32["F"], Ft.astype(np.float32), P32["mu"], P32["lam"], P32["ys"], P32["hard"],
                                    P32["soft"], P32["density"], P32["sc"], 0.0, P32["eta_s"], P32["eta_b"],
                                    P32["C"], P32["stress"], P32["acc"], P32["damage"], 1, P32["alpha"])
    assert rel_err(P32["stress"], P64["stress"]) < 5e-3
