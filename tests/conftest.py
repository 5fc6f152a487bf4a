"""Shared fixtures.  Tests marked ``gpu`` need an MI355X and the built
libmpmx.so; the rest run on the CPU (oracle pinning, host logic, ABI)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built libmpmx.so")


@pytest.fixture(scope="session")
def o64():
    from oracle import Oracle
    return Oracle(64)


@pytest.fixture(scope="session")
def o32():
    from oracle import Oracle
    return Oracle(32)


def rel_err(a, b):
    """max |a - b| / max |b| over the whole array (error relative to the array's scale)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = np.max(np.abs(b))
    if scale == 0:
        return float(np.max(np.abs(a - b)))
    return float(np.max(np.abs(a - b)) / scale)


def _rows(a):
    a = np.asarray(a, dtype=np.float64)
    return a.reshape(a.shape[0], -1) if a.ndim > 1 else a.reshape(-1, 1)


def pp_err(a, b, scale=None, floor=1e-3):
    """Largest error of one particle (or grid node) relative to its own size:
    max over items p of max_c |a_pc - b_pc| / s_p, with s_p the item's largest
    |b_pc| (or the given per-item ``scale``, whichever is larger), floored at
    ``floor`` times the array's largest |b|.  Grid arrays are taken per node."""
    a, b = _rows(a), _rows(b)
    if a.shape[0] != b.shape[0] and a.size == b.size:
        a = a.reshape(b.shape)
    s = np.max(np.abs(b), axis=1)
    if scale is not None:
        s = np.maximum(s, np.abs(np.asarray(scale, dtype=np.float64).reshape(-1)))
    top = np.max(np.abs(b)) if b.size else 0.0
    s = np.maximum(s, floor * top)
    d = np.max(np.abs(a - b), axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(s > 0, d / s, d)
    return float(r.max()) if r.size else 0.0


def incr_err(a, b, base, tol=1e-5, ulps=2.0, extra=None):
    """Worst ratio, over entries, of |a - b| to
        tol * (|b - base| + extra_p) + ulps * u_p:
    the update b - base of an FP32 quantity (x += dt v, F_trial = (I + dt G) F)
    agrees to ``tol`` of its own size, or of the per-item scale ``extra`` of
    the terms it is formed from, up to ``ulps`` roundings at the item's scale
    (u_p = the FP32 spacing at the item's largest |b|).  Items are the rows of
    a [n][...] array.  A value <= 1 passes."""
    a, b, base = (_rows(t) for t in (a, b, base))
    u = np.spacing(np.max(np.abs(b), axis=1).astype(np.float32)).astype(np.float64)[:, None]
    ex = 0.0 if extra is None else np.abs(np.asarray(extra, dtype=np.float64)).reshape(-1, 1)
    lim = tol * (np.abs(b - base) + ex) + ulps * u
    with np.errstate(invalid="ignore", divide="ignore"):
        r = np.where(lim > 0, np.abs(a - b) / lim, np.where(a == b, 0.0, np.inf))
    return float(r.max()) if r.size else 0.0
