"""The context / buffer / graph entry points of the C ABI, used the way a
caller without torch would: library-owned buffers, uploads and downloads, and
a captured sequence of fused substeps replayed as a graph.  Results equal the
Python layer's bit for bit (same kernels, same inputs)."""
import ctypes as C

import numpy as np
import pytest

from state_utils import make_sim

pytestmark = pytest.mark.gpu

FIELDS = [("activeLabel", "sim", "activeLabel"), ("materialLabel", "sim", "materialLabel"), ("x", "sim", "particle_x"),
          ("v", "sim", "particle_v"), ("x_init_xpbd", "xpbd", "particle_x_initial_xpbd"),
          ("v_init_xpbd", "xpbd", "particle_v_initial_xpbd"), ("F", "mpm", "particle_F"),
          ("F_trial", "mpm", "particle_F_trial"), ("C", "mpm", "particle_C"), ("stress", "mpm", "particle_stress"),
          ("cov", "mpm", "particle_cov"), ("mu", "mpm", "mu"), ("lam", "mpm", "lam"), ("ys", "mpm", "ys"),
          ("alpha", "mpm", "alpha"), ("hardening", "mpm", "hardening"), ("softening", "mpm", "softening"),
          ("density", "sim", "particle_density"), ("strainCriteria", "mpm", "strainCriteria"),
          ("eta_shear", "mpm", "eta_shear"), ("eta_bulk", "mpm", "eta_bulk"),
          ("acc_strain", "mpm", "particle_accumulated_strain"), ("damage", "mpm", "particle_damage"),
          ("vol", "sim", "particle_vol"), ("mass", "sim", "particle_mass"), ("order", "sim", "particle_id")]


def test_context_buffers_and_graph_replay():
    from mpmwarpydw_amd import simulationRoutines as SR
    from mpmwarpydw_amd._lib import Grid, Particles, lib
    A = make_sim(seed=50)
    n = A.sim.nPoints
    objs = {"sim": A.sim, "mpm": A.mpm, "xpbd": A.xpbd}
    ctx = C.c_void_p()
    lib("mpmx_ctx_create", 0, C.byref(ctx))
    stream = lib.raw("mpmx_ctx_stream")(ctx)
    bufs = {}
    try:
        def alloc(nbytes):
            p = C.c_void_p()
            lib("mpmx_buf_alloc", ctx, nbytes, C.byref(p))
            return p

        host = {}
        for field, o, name in FIELDS:
            h = getattr(objs[o], name).t.cpu().numpy()  # raw device layout (SoA planes)
            host[field] = np.ascontiguousarray(h)
            bufs[field] = alloc(h.nbytes)
            lib("mpmx_buf_upload", ctx, bufs[field], host[field].ctypes.data_as(C.c_void_p), h.nbytes)
        gx, gy, gz = A.mpm.gridDims
        G = gx * gy * gz
        gm, gvi, gvo = alloc(4 * G), alloc(12 * G), alloc(12 * G)
        bufs.update({"_gm": gm, "_gvi": gvi, "_gvo": gvo})
        P = Particles(n=n, **{f: bufs[f].value for f, _, _ in FIELDS})
        Gr = Grid(gx=gx, gy=gy, gz=gz, m=gm.value, v_in=gvi.value, v_out=gvo.value)
        _, _, prm = SR.step_structs(A.sim, A.mpm, A.xpbd)
        wsb = lib.raw("mpmx_substep_workspace_bytes")(n, gx, gy, gz)
        ws = alloc(wsb)
        bufs["_ws"] = ws
        # eager: bin + 2 substeps on the library's stream; then a captured 3-substep graph
        lib("mpmx_substep_bin", C.byref(P), C.byref(Gr), C.byref(prm), ws, wsb, stream)
        for _ in range(2):
            lib("mpmx_mpm_substep", C.byref(P), C.byref(Gr), C.byref(prm), ws, wsb, stream)
        graph = C.c_void_p()
        lib("mpmx_graph_begin", ctx)
        for _ in range(3):
            lib("mpmx_mpm_substep", C.byref(P), C.byref(Gr), C.byref(prm), ws, wsb, stream)
        lib("mpmx_graph_end", ctx, C.byref(graph))
        for _ in range(2):
            lib("mpmx_graph_launch", graph, stream)
        lib("mpmx_stream_sync", stream)
        lib("mpmx_graph_destroy", graph)
        # the same 8 substeps through the Python layer
        for _ in range(8):
            SR.mpmSimulationStepFused(A.sim, A.mpm, A.xpbd)
        for field, o, name in FIELDS:
            got = np.empty_like(host[field])  # noqa: F841
            lib("mpmx_buf_download", ctx, got.ctypes.data_as(C.c_void_p), bufs[field], got.nbytes)
            want = getattr(objs[o], name).t.cpu().numpy()
            assert np.array_equal(got, want), field
        gv = np.empty(3 * G, np.float32)
        lib("mpmx_buf_download", ctx, gv.ctypes.data_as(C.c_void_p), gvo, gv.nbytes)
        assert np.array_equal(gv, A.mpm.grid_v_out.t.cpu().numpy().reshape(-1))
    finally:
        for p in bufs.values():
            lib("mpmx_buf_free", ctx, p)
        lib("mpmx_ctx_destroy", ctx)
