"""The device math of libmpmx compiled for the host (tests/native/host_math.cpp)
against closed forms, SciPy and the FP64 oracle, on the CPU.

The constitutive update evaluates the Hencky strain and the plastic correction
with matrix series when F_trial is near a rotation and with the SVD otherwise;
both routes are checked here against the oracle's literal two-SVD restatement
of compute_stress_from_F_trial (mpmRoutines.py:229-471).
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest
import scipy.linalg

from oracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "native", "host_math.cpp")
OUT = os.path.join(HERE, "native", "_build", "libhostmath.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.fixture(scope="module")
def hm():
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    subprocess.check_call([HIPCC, "-O2", "-fPIC", "-shared", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
                           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", SRC, "-o", OUT],
                          stderr=subprocess.DEVNULL)
    return C.CDLL(OUT)


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _sym6(M):
    return np.stack([M[:, 0, 0], M[:, 1, 1], M[:, 2, 2], M[:, 0, 1], M[:, 0, 2], M[:, 1, 2]], 1)


def _full(s):
    return np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])


def _random_sym(rng, n, norms):
    A = rng.normal(size=(n, 3, 3))
    A = A + A.transpose(0, 2, 1)
    A /= np.linalg.norm(A, axis=(1, 2))[:, None, None]
    return A * norms[:, None, None]


def test_series_log_and_exp_match_scipy(hm):
    rng = np.random.default_rng(0)
    norms = np.concatenate([10 ** rng.uniform(-7, -1, 300), rng.uniform(0.1, 0.25, 100)])
    B = _random_sym(rng, len(norms), norms)
    b6 = np.ascontiguousarray(_sym6(B), np.float32)
    L = np.zeros_like(b6)
    hm.host_sym_log1p(C.c_long(len(b6)), _p(b6), _p(L))
    for k in range(len(b6)):
        Bk = _full(b6[k].astype(np.float64))
        want = scipy.linalg.logm(np.eye(3) + Bk).real
        assert np.max(np.abs(_full(L[k]) - want)) <= 3e-7 * np.max(np.abs(want)), k
    norms = np.concatenate([10 ** rng.uniform(-7, -1, 300), rng.uniform(0.1, 0.9, 100)])
    S = np.ascontiguousarray(_sym6(_random_sym(rng, len(norms), norms)), np.float32)
    X = np.zeros_like(S)
    hm.host_sym_exp(C.c_long(len(S)), _p(S), _p(X))
    for k in range(len(S)):
        want = scipy.linalg.expm(_full(S[k].astype(np.float64)))
        assert np.max(np.abs(_full(X[k]) - want)) <= 3e-7, k


def _state(rng, n, kind):
    """Random particle state in SoA float32 for the host update and AoS float64 for the oracle."""
    Q, _ = np.linalg.qr(rng.normal(size=(n, 3, 3)))
    Q *= np.sign(np.linalg.det(Q))[:, None, None]
    if kind == "tiny":  # strains far below the rotation: F F^T - I cancels to 1e-8 of its terms
        U = np.eye(3)[None] + _random_sym(rng, n, 10 ** rng.uniform(-8, -6, n))
    elif kind == "small":
        U = np.eye(3)[None] + _random_sym(rng, n, 10 ** rng.uniform(-6, -2.5, n))
    elif kind == "moderate":
        U = np.eye(3)[None] + _random_sym(rng, n, rng.uniform(0.003, 0.1, n))
    elif kind == "large":
        U = np.eye(3)[None] + _random_sym(rng, n, rng.uniform(0.2, 0.6, n))
    else:  # inverted
        U = np.eye(3)[None] + _random_sym(rng, n, rng.uniform(0.0, 0.05, n))
        U[:, 2, :] *= -1
    F_trial = Q @ U
    E = 10 ** rng.uniform(7, 9, n)
    nu = rng.uniform(0.15, 0.35, n)
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    ys = 10 ** rng.uniform(3.5, 6, n)
    ys[rng.random(n) < 0.02] = -1.0  # the reference's ys < 0 failure branch
    alpha = rng.uniform(0.1, 0.35, n)
    D = rng.uniform(0, 0.9, n)
    D[rng.random(n) < 0.1] = 0.999  # close to the phase transition
    sc = rng.uniform(0.01, 0.1, n)
    sc[rng.random(n) < 0.05] = 0.0
    st = dict(F_trial=F_trial, mu=mu, lam=lam, ys=ys, alpha=alpha, damage=D, strainCriteria=sc,
              hardening=rng.uniform(0, 0.1, n), softening=rng.uniform(0, 0.1, n), eta_shear=rng.uniform(0, 10, n),
              eta_bulk=rng.uniform(0, 10, n), C=rng.normal(0, 0.5, (n, 3, 3)), x=rng.normal(0, 1, (n, 3)),
              v=rng.normal(0, 1, (n, 3)), density=np.full(n, 2500.0), acc=rng.uniform(0, 0.01, n))
    # round every input to FP32 so both sides see the same numbers
    return {k: v.astype(np.float32).astype(np.float64) for k, v in st.items()}


def _run_host(hm, st, eff):
    n = len(st["mu"])
    f = lambda a: np.ascontiguousarray(a, np.float32)  # noqa: E731
    soa = lambda a: f(np.moveaxis(a.reshape(n, -1), 0, 1))  # noqa: E731
    A = {"activeLabel": np.ones(n, np.int32), "materialLabel": np.ones(n, np.int32),
         "x": soa(st["x"]), "v": soa(st["v"]), "x_init_xpbd": np.zeros((3, n), np.float32),
         "v_init_xpbd": np.zeros((3, n), np.float32), "F": np.zeros((9, n), np.float32),
         "F_trial": soa(st["F_trial"]), "C": soa(st["C"]), "stress": np.zeros((9, n), np.float32),
         "cov": np.zeros((6, n), np.float32), "acc_strain": f(st["acc"]), "vol": np.ones(n, np.float32),
         "mass": np.ones(n, np.float32)}
    for k in ("mu", "lam", "ys", "alpha", "hardening", "softening", "density", "strainCriteria", "eta_shear",
              "eta_bulk", "damage"):
        A[k] = f(st[k])
    from mpmwarpydw_amd._lib import Particles
    P = Particles(n, *[_p(A[name]) if name in A else None for name, _ in Particles._fields_[1:]])
    hm.host_constitutive(C.byref(P), C.c_float(eff))
    aos = lambda a: np.moveaxis(a, 0, 1).astype(np.float64)  # noqa: E731
    return {"F": aos(A["F"]).reshape(n, 3, 3), "stress": aos(A["stress"]).reshape(n, 3, 3),
            "acc": A["acc_strain"].astype(np.float64), "damage": A["damage"].astype(np.float64),
            "ys": A["ys"].astype(np.float64), "label": A["materialLabel"].copy(), "v": aos(A["v"])}


def _run_oracle(st, eff):
    o = Oracle(64)
    n = len(st["mu"])
    H = {k: np.ascontiguousarray(v) for k, v in st.items()}
    lab = np.ones(n, np.int32)
    F = np.zeros((n, 3, 3))
    stress = np.zeros((n, 3, 3))
    v = H["v"].copy()
    ys = H["ys"].copy()
    acc = H["acc"].copy()
    D = H["damage"].copy()
    o.compute_stress_from_F_trial(np.ones(n, np.int32), lab, H["x"], v, np.zeros((n, 3)), np.zeros((n, 3)), F,
                                  H["F_trial"], H["mu"].copy(), H["lam"].copy(), ys, H["hardening"],
                                  H["softening"], H["density"], H["strainCriteria"], eff, H["eta_shear"],
                                  H["eta_bulk"], H["C"], stress, acc, D, 1, H["alpha"].copy())
    return {"F": F, "stress": stress, "acc": acc, "damage": D, "ys": ys, "label": lab, "v": v}


@pytest.mark.parametrize("kind", ["tiny", "small", "moderate", "large", "inverted"])
def test_constitutive_update_matches_oracle(hm, kind):
    rng = np.random.default_rng({"tiny": 5, "small": 1, "moderate": 2, "large": 3, "inverted": 4}[kind])
    n = 3000
    st = _state(rng, n, kind)
    route = np.zeros(n, np.int32)
    Ft = np.ascontiguousarray(np.moveaxis(st["F_trial"].reshape(n, 9), 0, 1), np.float32)
    hm.host_route(C.c_long(n), _p(Ft), _p(route))
    want_fast = {"tiny": 1.0, "small": 1.0, "moderate": 1.0, "large": 0.0, "inverted": 0.0}[kind]
    assert abs(route.mean() - want_fast) < 0.02, route.mean()
    eff = 0.3
    h = _run_host(hm, st, eff)
    o = _run_oracle(st, eff)
    assert np.array_equal(h["label"], o["label"])
    moved = o["acc"] != st["acc"]
    assert np.array_equal(h["acc"] != st["acc"].astype(np.float32), moved)
    # FP32 against FP64, particle by particle: 1e-5 of the larger of the
    # particle's largest stress component and its yield stress (1 Pa floor);
    # F to a few FP32 ulps of 1
    scale_tau = np.maximum(np.max(np.abs(o["stress"]), axis=(1, 2)), np.maximum(np.abs(o["ys"]), 1.0))
    err_tau = np.max(np.abs(h["stress"] - o["stress"]), axis=(1, 2)) / scale_tau
    err_F = np.max(np.abs(h["F"] - o["F"]), axis=(1, 2))
    assert err_tau.max() < 1e-5, err_tau.max()
    assert err_F.max() < 2e-6, err_F.max()
    # plastic increments, relative to the increment
    dacc = np.abs((h["acc"] - st["acc"]) - (o["acc"] - st["acc"]))
    assert np.all(dacc <= 1e-4 * np.abs(o["acc"] - st["acc"]) + 2e-8)
    assert np.max(np.abs(h["damage"] - o["damage"])) < 1e-5
    assert np.max(np.abs(h["ys"] - o["ys"]) / np.maximum(np.abs(o["ys"]), 1.0)) < 1e-5
    # the phase transition adds sqrt(2 eff u / rho) along v: relative FP32 precision
    assert np.max(np.abs(h["v"] - o["v"]) / np.maximum(np.abs(o["v"]), 1.0)) < 1e-5


def test_g2p_gather_matches_the_27_node_sums(hm):
    """The axis-factored G2P sums (mpmx_device.hpp g2p_gather) equal the
    reference's 27 separate node terms (mpmRoutines.py:738-770) evaluated in FP64."""
    rng = np.random.default_rng(9)
    n = 2000
    fx = rng.uniform(0.5, 1.5, (n, 3)).astype(np.float32)
    v = rng.standard_normal((n, 27, 3)).astype(np.float32)
    v[: n // 2] += 20.0  # a large uniform part on half of them
    inv_dx = 2.0
    nv = np.zeros((n, 3), np.float32)
    nB = np.zeros((n, 9), np.float32)
    nG = np.zeros((n, 9), np.float32)
    hm.host_g2p_gather(C.c_long(n), _p(fx), _p(v), C.c_float(inv_dx), _p(nv), _p(nB), _p(nG))
    f = fx.astype(np.float64)
    w = np.stack([0.5 * (1.5 - f) ** 2, 0.75 - (f - 1.0) ** 2, 0.5 * (f - 0.5) ** 2], axis=2)  # [n][axis][q]
    dw = np.stack([f - 1.5, -2.0 * (f - 1.0), f - 0.5], axis=2)
    ev = np.zeros((n, 3)); eB = np.zeros((n, 3, 3)); eG = np.zeros((n, 3, 3))
    vd = v.astype(np.float64)
    for i in range(3):
        for j in range(3):
            for k in range(3):
                wt = w[:, 0, i] * w[:, 1, j] * w[:, 2, k]
                g = np.stack([dw[:, 0, i] * w[:, 1, j] * w[:, 2, k], w[:, 0, i] * dw[:, 1, j] * w[:, 2, k],
                              w[:, 0, i] * w[:, 1, j] * dw[:, 2, k]], axis=1) * inv_dx
                d = np.stack([i - f[:, 0], j - f[:, 1], k - f[:, 2]], axis=1)
                vn = vd[:, 9 * i + 3 * j + k]
                ev += wt[:, None] * vn
                eB += wt[:, None, None] * vn[:, :, None] * d[:, None, :]
                eG += vn[:, :, None] * g[:, None, :]
    # the uniform part cancels in nB; 27 separate FP32 terms in the reference's
    # order miss the FP64 value by 1.8e-6 of the array's scale on these inputs
    for got, want in ((nv, ev), (nB, eB.reshape(n, 9)), (nG, eG.reshape(n, 9))):
        assert np.abs(got - want).max() <= 4e-6 * np.abs(want).max()
