// Standalone P2G entry points (p2g_apic_with_stress, mpmRoutines.py:485-569):
// bin the particles, reduce each tile in LDS, combine the tile partials per node
// (see p2g_tile.hpp for the deterministic scheme).
#include "mpmx_internal.hpp"
#include "p2g_tile.hpp"

namespace mpmx {

__global__ void fill_i32_one(int32_t* __restrict__ p) { *p = 0; }

__global__ void p2g_key_kernel(int n, const int32_t* __restrict__ activeLabel,
                               const int32_t* __restrict__ materialLabel, const float* __restrict__ x,
                               float inv_dx, Geo geo, int32_t* __restrict__ keys, int32_t* __restrict__ oob) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int key = -1;
    if (activeLabel[p] == 1 && materialLabel[p] <= 1) {
        key = tile_key(ld3(x, n, p), inv_dx, geo.one);
        if (key < 0) atomicAdd(oob, 1);
    }
    keys[p] = key;
}

struct P2GWork {
    int32_t* keys;
    int32_t* sorted;
    float* partial;
    int32_t* tile_flag;
    int32_t* oob;
    BinWork bin;
};

static size_t p2g_ws_bytes(long n, const Geo& tg) {
    long nkeys = tg.ntiles() * kTileCells;
    long nn = n > 0 ? n : 1;
    return Carver::bytes_for(nn, 4) * 2 + Carver::bytes_for(tg.ntiles() * kTileNodes * 4, 4) +
           Carver::bytes_for(tg.ntiles(), 4) + Carver::bytes_for(1, 4) + bin_workspace_bytes(nn, nkeys);
}

static bool carve_p2g(void* ws, size_t bytes, long n, const Geo& tg, P2GWork& w) {
    Carver cv(ws, bytes);
    long nn = n > 0 ? n : 1;
    w.keys = cv.take<int32_t>(nn);
    w.sorted = cv.take<int32_t>(nn);
    w.partial = cv.take<float>(tg.ntiles() * kTileNodes * 4);
    w.tile_flag = cv.take<int32_t>(tg.ntiles());
    w.oob = cv.take<int32_t>(1);
    carve_bin(cv, nn, tg.ntiles() * kTileCells, w.bin);
    return !cv.overflow;
}

static int p2g_run(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* stress,
                   const float* x, const float* v, const float* C, const float* vol, const float* mass, float dx,
                   float inv_dx, float minx, float miny, float minz, float rpic, int gx, int gy, int gz,
                   float* grid_m, float* grid_v_in, float* grid_v_out, int do_update, float gxv, float gyv,
                   float gzv, float dt, float damping, int bc, float fric, float gmag, int pad, float restitution,
                   void* ws, size_t ws_bytes, int* host_oob, hipStream_t s) {
    MPMX_REQUIRE(n >= 0 && gx >= 3 && gy >= 3 && gz >= 3, "p2g: grid must have at least 3 nodes per axis");
    MPMX_REQUIRE((long)gx * gy * gz < (1L << 31), "p2g: a grid needs fewer than 2^31 nodes");
    MPMX_REQUIRE(grid_m && grid_v_in && (!do_update || grid_v_out), "p2g: null grid pointer");
    MPMX_REQUIRE(n == 0 || (activeLabel && materialLabel && stress && x && v && C && vol && mass),
                 "p2g: null particle pointer");
    MPMX_REQUIRE(bc >= MPMX_BC_NONE && bc <= MPMX_BC_PLANE_STRAIN_Y_GLOBAL, "p2g: unsupported boundary condition %d",
                 bc);
    Geo tg = make_geom(gx, gy, gz, minx, miny, minz);
    MPMX_REQUIRE(tg.ntiles() * kTileCells < (1L << 31), "p2g: grid too large for int32 keys");
    size_t need = p2g_ws_bytes(n, tg);
    MPMX_REQUIRE(ws && ws_bytes >= need, "p2g: workspace of %zu bytes given, %zu needed", ws_bytes, need);
    P2GWork w;
    MPMX_REQUIRE(carve_p2g(ws, ws_bytes, n, tg, w), "p2g: workspace carve failed");
    int rc;
    MPMX_LAUNCH(fill_i32_one, dim3(1), dim3(1), 0, s, w.oob);
    if ((rc = check_launch("p2g_oob_zero", s))) return rc;
    if (n > 0) {
        MPMX_LAUNCH(p2g_key_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, activeLabel,
                           materialLabel, x, inv_dx, tg, w.keys, w.oob);
        if ((rc = check_launch("p2g_key", s))) return rc;
    }
    if ((rc = counting_sort_ids(w.keys, n, tg.ntiles() * kTileCells, w.sorted, w.bin, s))) return rc;
    TileArgs ta{};
    ta.P.n = n;
    ta.P.activeLabel = const_cast<int32_t*>(activeLabel);
    ta.P.materialLabel = const_cast<int32_t*>(materialLabel);  // the tile kernel tests the label
    ta.P.x = const_cast<float*>(x);
    ta.P.v = const_cast<float*>(v);
    ta.P.C = const_cast<float*>(C);
    ta.P.vol = const_cast<float*>(vol);
    ta.P.mass = const_cast<float*>(mass);
    ta.stress = stress;
    ta.sorted = w.sorted;
    ta.cell_start = w.bin.cell_start;
    ta.dx = dx;
    ta.inv_dx = inv_dx;
    ta.rpic = rpic;
    ta.dt = dt;
    ta.partial = w.partial;
    ta.tile_flag = w.tile_flag;
    ta.ntiles = (int)tg.ntiles();
    ta.geo = tg;
    ta.xcd = tuning(MPMX_TUNE_XCD_TILES);
    MPMX_LAUNCH(p2g_tile_kernel<false>, dim3((unsigned)tg.ntiles()), dim3(kBlock), 0, s, ta);
    if ((rc = check_launch("p2g_tile", s))) return rc;
    long G = (long)gx * gy * gz;
    CombineArgs ca{};
    ca.geo = tg;
    ca.partial = w.partial;
    ca.tile_flag = w.tile_flag;
    ca.grid_m = grid_m;
    ca.grid_v_in = grid_v_in;
    ca.grid_v_out = grid_v_out;
    ca.do_update = do_update;
    ca.gxv = gxv;
    ca.gyv = gyv;
    ca.gzv = gzv;
    ca.dt = dt;
    ca.damping = damping;
    ca.fric = fric;
    ca.gmag = gmag;
    ca.bc = bc;
    ca.pad = pad;
    ca.restitution = restitution;
    MPMX_LAUNCH(p2g_combine_kernel<>, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, s, ca);
    if ((rc = check_launch("p2g_combine", s))) return rc;
    if (host_oob) {
        MPMX_HIP(hipMemcpyAsync(host_oob, w.oob, sizeof(int), hipMemcpyDeviceToHost, s));
        MPMX_HIP(hipStreamSynchronize(s));
    }
    return 0;
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

size_t mpmx_p2g_workspace_bytes(int n, int gx, int gy, int gz) {
    if (n < 0 || gx < 3 || gy < 3 || gz < 3) return 0;
    return p2g_ws_bytes(n, make_geom(gx, gy, gz));
}

int mpmx_p2g_apic_with_stress(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* stress,
                              const float* x, const float* v, const float* C, const float* vol, const float* mass,
                              float dx, float inv_dx, float minx, float miny, float minz, float rpic_damping, int gx,
                              int gy, int gz, float* grid_m, float* grid_v_in, float dt, void* workspace,
                              size_t ws_bytes, int* host_out_of_grid, void* stream) {
    return p2g_run(n, activeLabel, materialLabel, stress, x, v, C, vol, mass, dx, inv_dx, minx, miny, minz,
                   rpic_damping, gx, gy, gz, grid_m, grid_v_in, nullptr, 0, 0.f, 0.f, 0.f, dt, 1.0f, MPMX_BC_NONE,
                   0.f, 0.f, 0, 0.f, workspace, ws_bytes, host_out_of_grid, as_stream(stream));
}

int mpmx_p2g_grid_update(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* stress,
                         const float* x, const float* v, const float* C, const float* vol, const float* mass,
                         float dx, float inv_dx, float minx, float miny, float minz, float rpic_damping, int gx,
                         int gy, int gz, float* grid_m, float* grid_v_in, float* grid_v_out, float gxv, float gyv,
                         float gzv, float dt, float damping_scale, int bc, float friction, float gravity_magnitude,
                         int boundary_padding, float restitution, void* workspace, size_t ws_bytes, void* stream) {
    return p2g_run(n, activeLabel, materialLabel, stress, x, v, C, vol, mass, dx, inv_dx, minx, miny, minz,
                   rpic_damping, gx, gy, gz, grid_m, grid_v_in, grid_v_out, 1, gxv, gyv, gzv, dt, damping_scale, bc,
                   friction, gravity_magnitude, boundary_padding, restitution, workspace, ws_bytes, nullptr,
                   as_stream(stream));
}

}  // extern "C"
