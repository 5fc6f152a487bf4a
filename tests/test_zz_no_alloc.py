"""mpmx_velocity_convergence_noalloc, the allocation-free form of the
per-substep velocity residual (velocityConvergence + countActiveParticles,
simulationRoutines.py:538-558), against the workspace form through the C ABI:
two calls in a row, the second on fewer particles, bit for bit (the same
kernels in the same order), and against an FP64 sum on the host."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_noalloc_residual_matches_workspace_form():
    import torch
    from mpmwarpydw_amd import runtime as rt
    from mpmwarpydw_amd._lib import lib
    dev = rt.device()
    rng = np.random.default_rng(41)
    n0 = 300_000
    act = torch.tensor((rng.random(n0) > 0.1).astype(np.int32), device=dev)
    v = torch.tensor(rng.standard_normal((3, n0)).astype(np.float32), device=dev)
    r = torch.tensor(rng.uniform(0.05, 0.3, n0).astype(np.float32), device=dev)
    ws_bytes = int(lib.raw("mpmx_reduce_workspace_bytes")(n0))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    s = rt.stream_handle()
    p = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731
    for n in (n0, 12_345):  # the second call on a smaller input reuses the scratch
        # views of the first n particles: the SoA planes of v are n0 apart, so copy
        vn = v[:, :n].contiguous()
        a, rr = act[:n].contiguous(), r[:n].contiguous()
        out_ws = (torch.zeros(1, dtype=torch.float64, device=dev), torch.zeros(1, dtype=torch.int32, device=dev))
        out_na = (torch.full((1,), -1.0, dtype=torch.float64, device=dev),
                  torch.full((1,), -1, dtype=torch.int32, device=dev))
        print(f"[noalloc residual] n={n}", flush=True)
        lib("mpmx_velocity_convergence", n, p(a), p(vn), p(rr), p(out_ws[0]), p(out_ws[1]), p(ws), ws_bytes, s)
        lib("mpmx_velocity_convergence_noalloc", n, p(a), p(vn), p(rr), p(out_na[0]), p(out_na[1]), s)
        torch.cuda.synchronize()
        assert out_na[0].item() == out_ws[0].item() and out_na[1].item() == out_ws[1].item()
        hv = vn.cpu().numpy().astype(np.float64)
        ha = a.cpu().numpy() == 1
        want = np.sum((np.sqrt((hv ** 2).sum(0)) / rr.cpu().numpy())[ha])
        assert abs(out_na[0].item() - want) <= 1e-9 * want and out_na[1].item() == ha.sum()


def test_noalloc_residual_rejects_bad_arguments():
    from mpmwarpydw_amd._lib import MpmxError, lib
    with pytest.raises(MpmxError, match="n < 0"):
        lib("mpmx_velocity_convergence_noalloc", -1, None, None, None, None, None, None)
    with pytest.raises(MpmxError, match="null output"):
        lib("mpmx_velocity_convergence_noalloc", 0, None, None, None, None, None, None)
