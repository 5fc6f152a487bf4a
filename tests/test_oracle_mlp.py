"""CPU checks of the learned-return-map oracle (oracle/mlp_oracle.c).

The network's arithmetic is pinned against PyTorch's float64 autograd and its
library routines (bfloat16 conversion, tanh-GELU, Adam) and against central
finite differences; the features, exact targets and judge are pinned against
the full FP64 restatement of compute_stress_from_F_trial and closed forms.
"""
import numpy as np
import pytest
import torch

from oracle import MLPOracle, Oracle


@pytest.fixture(scope="module")
def mo():
    return MLPOracle()


def init_params(mo, rng, scale=0.3):
    p = rng.normal(0, scale, mo.nparams)
    return p


def torch_forward(mo, params, X, quant=False):
    W1, b1, W2, b2, W3, b3 = [torch.tensor(a, dtype=torch.float64) for a in mo.split(params)]
    q = (lambda t: t.to(torch.float32).to(torch.bfloat16).to(torch.float64)) if quant else (lambda t: t)
    # inputs and hidden activations enter the matrix units as hi + lo bfloat16 halves
    qs = (lambda t: q(t) + q(t - q(t))) if quant else (lambda t: t)
    x = qs(torch.tensor(X, dtype=torch.float64))
    h1 = qs(torch.nn.functional.gelu(x @ q(W1).T + b1, approximate="tanh"))
    h2 = qs(torch.nn.functional.gelu(h1 @ q(W2).T + b2, approximate="tanh"))
    return (h2 @ q(W3).T + b3).numpy()


def test_bf16_rounding_matches_torch(mo):
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-3, 2000), rng.normal(0, 1e4, 2000),
                        [1 + 2 ** -8, 1 + 3 * 2 ** -9, -(1 + 2 ** -8), 0.0, 1e-40]]).astype(np.float32)
    want = torch.tensor(v).to(torch.bfloat16).to(torch.float64).numpy()
    got = np.array([mo.bf16(float(a)) for a in v])
    assert np.array_equal(got, want)
    assert mo.bf16(1 + 2 ** -8) == 1.0  # a tie goes to the even mantissa
    # the split input keeps two 8-bit significands: 1 + 2^-12 survives (bf16 alone drops it), the
    # 2^-21 bit of 1 + 2^-12 + 2^-21 lies beyond the low half's significand and rounds away
    assert mo.bf16(1 + 2 ** -12) == 1.0 and mo.bf16_split(1 + 2 ** -12) == 1 + 2 ** -12
    assert mo.bf16_split(1 + 2 ** -12 + 2 ** -21) == 1 + 2 ** -12
    rel = np.abs(np.array([mo.bf16_split(float(a)) for a in v[:4000]]) / v[:4000].astype(np.float64) - 1)
    assert rel.max() <= 2.0 ** -16


def test_gelu_and_derivative(mo):
    x = torch.linspace(-8, 8, 1601, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(x, approximate="tanh")
    (g,) = torch.autograd.grad(y.sum(), x)
    got = np.array([mo.gelu(float(a)) for a in x.detach()])
    gg = np.array([mo.gelu_grad(float(a)) for a in x.detach()])
    assert np.max(np.abs(got - y.detach().numpy())) < 1e-14
    assert np.max(np.abs(gg - g.numpy())) < 1e-14


@pytest.mark.parametrize("quant", [False, True])
def test_forward_matches_torch(mo, quant):
    rng = np.random.default_rng(1)
    params = init_params(mo, rng)
    X = rng.normal(0, 1, (37, 10))
    got = mo.forward(params, X, quant=quant)
    want = torch_forward(mo, params, X, quant=quant)
    assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want)))


def test_gradient_matches_autograd_and_finite_differences(mo):
    rng = np.random.default_rng(2)
    params = init_params(mo, rng)
    X = rng.normal(0, 1, (23, 10))
    T = rng.normal(0, 1, (23, 3))
    loss, g = mo.loss_grad(params, X, T)
    pt = torch.tensor(params, requires_grad=True)
    W = torch.split(pt, [640, 64, 4096, 64, 192, 3])
    x = torch.tensor(X)
    h1 = torch.nn.functional.gelu(x @ W[0].view(64, 10).T + W[1], approximate="tanh")
    h2 = torch.nn.functional.gelu(h1 @ W[2].view(64, 64).T + W[3], approximate="tanh")
    y = h2 @ W[4].view(3, 64).T + W[5]
    L = ((y - torch.tensor(T)) ** 2).mean()
    (gt,) = torch.autograd.grad(L, pt)
    assert abs(loss - L.item()) < 1e-13 * L.item()
    assert np.max(np.abs(g - gt.numpy())) < 1e-12 * np.max(np.abs(gt.numpy()))
    # central differences on parameters of every layer
    idx = np.concatenate([rng.choice(640, 6, replace=False), 640 + rng.choice(64, 3, replace=False),
                          704 + rng.choice(4096, 6, replace=False), 4800 + rng.choice(64, 3, replace=False),
                          4864 + rng.choice(192, 4, replace=False), 5056 + np.arange(3)])
    h = 1e-6
    for i in idx:
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, X, T)[0] - mo.loss_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-7 * max(1.0, abs(g[i])), i


def test_quantized_gradient_is_close_to_exact(mo):
    """Rounding to bfloat16 at the matrix-unit inputs perturbs the gradient by
    a few bfloat16 ulps (2^-8 relative), not more: a dropped or misplaced term
    would move it by O(1)."""
    rng = np.random.default_rng(3)
    params = init_params(mo, rng)
    X = rng.normal(0, 1, (256, 10))
    T = rng.normal(0, 1, (256, 3))
    _, g = mo.loss_grad(params, X, T, quant=False)
    _, gq = mo.loss_grad(params, X, T, quant=True)
    assert np.max(np.abs(gq - g)) < 0.03 * np.max(np.abs(g))
    assert np.corrcoef(g, gq)[0, 1] > 0.999


def test_adam_matches_torch(mo):
    rng = np.random.default_rng(4)
    p0 = rng.normal(0, 1, 50)
    grads = [rng.normal(0, 1, 50) for _ in range(4)]
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, 1):
        mo.adam_step(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, t)
    pt = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([pt], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        opt.zero_grad()
        pt.grad = torch.tensor(g)
        opt.step()
    assert np.max(np.abs(p - pt.detach().numpy())) < 1e-14


# ---- features, exact targets and judge ---------------------------------------

def _material(n, rng):
    E = 10 ** rng.uniform(7, 9, n)
    nu = rng.uniform(0.15, 0.35, n)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    ys = 10 ** rng.uniform(4, 6, n)
    phi = np.deg2rad(rng.uniform(20, 40, n))
    s = np.sin(phi)
    alpha = 2 * s / (np.sqrt(3) * (3 - s))
    return E, nu, mu, lam, ys, alpha, phi


def _trial(e, mu, lam, ys, alpha, D):
    """Deviator, its norm and the yield value of principal trial strains."""
    d = e - e.mean(1, keepdims=True)
    dn = np.linalg.norm(d, axis=1)
    p = (2 * mu / 3 + lam) * e.sum(1)
    Y = (1 - D) * (ys - alpha * p)
    return d, dn, Y


def test_features_closed_form(mo):
    rng = np.random.default_rng(5)
    n = 64
    E, nu, mu, lam, ys, alpha, phi = _material(n, rng)
    e = rng.normal(0, 1e-3, (n, 3))
    acc = rng.uniform(0, 0.05, n)
    D = rng.uniform(0, 1, n)
    X = mo.features(e, mu, lam, ys, alpha, acc, D)
    d, dn, Y = _trial(e, mu, lam, ys, alpha, D)
    over, y0, kap = (2 * mu * dn - Y) / ys, Y / ys, 2 * mu * 1e-6 / ys
    assert np.allclose(X[:, 0], 0.1 * np.log(np.maximum(over + kap, 1e-30)), rtol=1e-10, atol=1e-12)
    assert np.allclose(X[:, 1], 0.1 * np.log(np.maximum(y0, 1e-12)), rtol=1e-12)
    assert np.allclose(X[:, 2], 0.1 * np.log(kap), rtol=1e-12)
    assert (over + kap > 0).sum() > 5 and (over + kap < 0).sum() > 5  # both branches of the clamp
    assert np.allclose(X[:, 3:6], d / dn[:, None], rtol=1e-10)
    assert np.allclose(X[:, 6], np.tan(phi), rtol=1e-10)
    assert np.allclose(X[:, 7], 4 * nu - 1, atol=1e-12)
    assert np.allclose(X[:, 8], np.log10(E) - 8, atol=1e-12)
    assert np.allclose(X[:, 9], D)


def test_targets_match_full_return_map():
    """The radial return plus the two predicted corrections,
    e - (1 - Y/(2 mu |d|)) d - eref kap s d/|d| + 0.3 alpha eref (phi + kap),
    with the exact targets (s, alpha) equals the principal log strains of F after
    the full FP64 constitutive update (so a dropped 1e-6 or dilatancy term, or a
    wrong sign, fails here); and s = y0 / (phi + y0 + kap) in terms of the
    features."""
    o = Oracle(64)
    mo = MLPOracle()
    rng = np.random.default_rng(6)
    n = 200
    E, nu, mu, lam, ys, alpha, _ = _material(n, rng)
    ys = ys * 10 ** rng.uniform(-1.5, 1.5, n)  # trial states on both sides of the yield surface
    F_trial = np.eye(3)[None] + rng.normal(0, 2e-3, (n, 3, 3))
    F = np.zeros((n, 3, 3))
    D = rng.uniform(0, 0.5, n)
    acc = np.zeros(n)
    lab = np.ones(n, np.int32)
    act = np.ones(n, np.int32)
    z3 = np.zeros((n, 3))
    ys0 = ys.copy()
    o.compute_stress_from_F_trial(act, lab, z3.copy(), z3.copy(), z3.copy(), z3.copy(), F, F_trial.copy(),
                                  mu.copy(), lam.copy(), ys, np.zeros(n), np.zeros(n), np.full(n, 2500.0),
                                  np.full(n, 1e9), 0.0, np.zeros(n), np.zeros(n), np.zeros((n, 3, 3)),
                                  np.zeros((n, 3, 3)), acc, D.copy(), 1, alpha.copy())
    e_trial = np.log(np.linalg.svd(F_trial, compute_uv=False))
    e_new = np.log(np.linalg.svd(F, compute_uv=False))
    T, plastic = mo.targets(e_trial, mu, lam, ys0, alpha, D)
    c = np.maximum(ys0, 1.0)
    eref = c / (2 * mu)
    assert plastic.sum() > 50 and (~plastic).sum() > 10
    assert np.array_equal(plastic, acc > 0)
    d, dn, Y = _trial(e_trial, mu, lam, ys0, alpha, D)
    phi, y0, kap = (2 * mu * dn - Y) / c, Y / c, 2 * mu * 1e-6 / c
    dh = d / dn[:, None]
    got = (e_trial - (1 - Y / (2 * mu * dn))[:, None] * d - (eref * kap * T[:, 0])[:, None] * dh
           + (0.3 * (0.3 + T[:, 1] / 20) * eref * (phi + kap))[:, None])
    stay = lab == 1  # a particle that reaches damage 1 turns XPBD and keeps F_trial instead
    assert stay.sum() > 100
    assert np.max(np.abs(got - e_new)[stay & plastic]) < 1e-9
    assert np.all(T[~plastic] == 0)
    assert np.allclose(T[plastic, 1], 20 * (alpha[plastic] - 0.3), rtol=0, atol=1e-14) and np.all(T[:, 2] == 0)
    assert np.allclose(mo.apply(e_trial, T, mu, lam, ys0, alpha, D)[plastic], got[plastic], atol=1e-15)
    X = mo.features(e_trial, mu, lam, ys0, alpha, acc, D)
    pk, y0f = np.exp(10 * X[:, 0]), np.exp(10 * X[:, 1])
    kapf = np.exp(10 * X[:, 2])
    closed = y0f / (pk + y0f)
    assert np.max(np.abs(closed - T[:, 0])[plastic & (Y > 0)]) < 1e-9
    assert np.allclose(kapf, kap, rtol=1e-12)


def test_judge_accepts_exact_and_rejects_each_kind_of_error(mo):
    rng = np.random.default_rng(7)
    n = 300
    E, nu, mu, lam, ys, alpha, _ = _material(n, rng)
    ys = ys * 0.05
    e = rng.normal(0, 2e-3, (n, 3))
    D = rng.uniform(0, 0.5, n)
    T, plastic = mo.targets(e, mu, lam, ys, alpha, D)
    e, T, mu, lam, ys, alpha, D = (a[plastic] for a in (e, T, mu, lam, ys, alpha, D))
    exact = mo.apply(e, T, mu, lam, ys, alpha, D)
    assert mo.judge(e, exact, mu, lam, ys, alpha, D).all()
    K = lam + 2 * mu / 3
    m = exact.mean(1, keepdims=True)
    dev = exact - m
    dh = dev / np.linalg.norm(dev, axis=1, keepdims=True)
    # (1) deviatoric norm off by 3 tolerances of stress
    bad1 = exact + dh * (3e-4 * ys / (2 * mu))[:, None]
    # (2) a deviatoric component orthogonal to the flow direction
    w = np.cross(dh, np.ones(3))  # orthogonal to dh and to the trace direction
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    bad2 = exact + w * (3e-4 * ys / (2 * mu))[:, None]
    # (3) the volume change off
    bad3 = exact + (3e-4 * ys / (3 * K))[:, None]
    for bad in (bad1, bad2, bad3):
        assert not mo.judge(e, bad, mu, lam, ys, alpha, D).any()
    # within a third of the tolerance it still accepts
    ok = exact + dh * (0.3e-4 * ys / (2 * mu))[:, None]
    assert mo.judge(e, ok, mu, lam, ys, alpha, D).all()
