// Deterministic P2G (p2g_apic_with_stress, mpmRoutines.py:485-569) for gfx950.
//
// The reference scatters 27 x 4 float atomics per particle into the grid, so its
// sums depend on thread timing.  Here:
//   1. bin:   every active MPM particle gets the key of its base cell,
//             key = tile * 64 + local cell, tiles being 4x4x4 base cells;
//             a counting sort (ids ascending inside a cell) groups them.
//   2. tile:  one 256-thread workgroup per tile stages a 25-float record per
//             particle in LDS (fractional offset, mass, affine momentum terms,
//             stress impulse), then each (cell, i, j) task sums the cell's
//             particles in sorted order into its 3 nodes (k = 0..2) and adds
//             the result to an LDS array indexed by (cell, i, j, k).  Every
//             (cell, offset) slot has one owner, so no atomics are needed.
//             Finally each of the 6x6x6 tile nodes sums its <= 27 slots in a
//             fixed offset order and the tile writes its partial node block.
//   3. combine: each grid node adds the partial blocks of the <= 8 tiles that
//             cover it in a fixed tile order, and optionally applies the grid
//             update (normalisation, gravity, damping, boundary condition).
// The result does not depend on scheduling: identical inputs give identical bits.
#include "mpmx_device.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

constexpr int kTB = 4;                    // base cells per tile edge
constexpr int kTN = kTB + 2;              // nodes per tile edge
constexpr int kTileCells = kTB * kTB * kTB;  // 64
constexpr int kTileNodes = kTN * kTN * kTN;  // 216
constexpr int kRec = 25;                  // floats per staged particle record
constexpr int kChunk = 256;               // particles staged per LDS round
constexpr int kTasks = kTileCells * 9;    // (cell, i, j)

struct TileGeom {
    int gx, gy, gz;     // grid nodes
    int ntx, nty, ntz;  // tiles
    long ntiles() const { return (long)ntx * nty * ntz; }
};

static TileGeom make_geom(int gx, int gy, int gz) {
    TileGeom g;
    g.gx = gx; g.gy = gy; g.gz = gz;
    g.ntx = (gx - 2 + kTB - 1) / kTB;
    g.nty = (gy - 2 + kTB - 1) / kTB;
    g.ntz = (gz - 2 + kTB - 1) / kTB;
    return g;
}

__global__ void fill_i32_one(int32_t* __restrict__ p) { *p = 0; }

// ---- 1. keys -------------------------------------------------------------
__global__ void p2g_key_kernel(int n, const int32_t* __restrict__ activeLabel,
                               const int32_t* __restrict__ materialLabel, const float* __restrict__ x,
                               float minx, float miny, float minz, float inv_dx, TileGeom tg,
                               int32_t* __restrict__ keys, int32_t* __restrict__ oob) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int key = -1;
    if (activeLabel[p] == 1 && materialLabel[p] <= 1) {
        float f;
        int bx = base_axis(x[p], minx, inv_dx, f);
        int by = base_axis(x[n + p], miny, inv_dx, f);
        int bz = base_axis(x[2 * (long)n + p], minz, inv_dx, f);
        // compare against g - 3 (not b + 2 < g) so a saturated huge index cannot wrap
        if (bx >= 0 && by >= 0 && bz >= 0 && bx <= tg.gx - 3 && by <= tg.gy - 3 && bz <= tg.gz - 3) {
            int tx = bx / kTB, ty = by / kTB, tz = bz / kTB;
            int lx = bx - tx * kTB, ly = by - ty * kTB, lz = bz - tz * kTB;
            key = ((tx * tg.nty + ty) * tg.ntz + tz) * kTileCells + (lx * kTB + ly) * kTB + lz;
        } else {
            atomicAdd(oob, 1);
        }
    }
    keys[p] = key;
}

// B-spline weight and derivative for a compile-time-unknown stencil offset i
__device__ __forceinline__ void wsel(float f, int i, float& w, float& dw) {
    if (i == 0) {
        float a = 1.5f - f;
        w = a * a * 0.5f;
        dw = f - 1.5f;
    } else if (i == 1) {
        float b = f - 1.0f;
        w = (0.0f - b * b) + 0.75f;
        dw = -2.0f * b;
    } else {
        float c = f - 0.5f;
        w = c * c * 0.5f;
        dw = c;
    }
}

// ---- 2. tile reduction -----------------------------------------------------
__global__ void __launch_bounds__(kBlock) p2g_tile_kernel(
    int n, const int32_t* __restrict__ sorted, const int32_t* __restrict__ cell_start,
    const float* __restrict__ stress, const float* __restrict__ x, const float* __restrict__ v,
    const float* __restrict__ Cmat, const float* __restrict__ vol, const float* __restrict__ mass, float dx,
    float inv_dx, float minx, float miny, float minz, float rpic, float dt, float* __restrict__ partial,
    int32_t* __restrict__ tile_flag) {
    __shared__ float rec[kRec][kChunk];
    __shared__ float4 part[kTileCells * 27];
    __shared__ int cs[kTileCells + 1];

    const int t = blockIdx.x;
    const int tid = threadIdx.x;
    if (tid <= kTileCells) cs[tid] = cell_start[(long)t * kTileCells + tid];
    __syncthreads();
    const int p0 = cs[0];
    const int np = cs[kTileCells] - p0;
    if (np == 0) {
        if (tid == 0) tile_flag[t] = 0;
        return;
    }
    for (int s = tid; s < kTileCells * 27; s += kBlock) part[s] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);

    const float sfac = -dt * inv_dx;
    for (int c0 = 0; c0 < np; c0 += kChunk) {
        const int cn = min(kChunk, np - c0);
        __syncthreads();  // previous round's readers are done with rec
        for (int s = tid; s < cn; s += kBlock) {
            const int id = sorted[p0 + c0 + s];
            f3 xp = ld3(x, n, id);
            f3 vp = ld3(v, n, id);
            float f0, f1, f2;
            base_axis(xp.x, minx, inv_dx, f0);
            base_axis(xp.y, miny, inv_dx, f1);
            base_axis(xp.z, minz, inv_dx, f2);
            float Cp[9], Cm[9];
            ld9(Cmat, n, id, Cp);
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int b = 0; b < 3; b++)
                    Cm[a * 3 + b] = (1.0f - rpic) * Cp[a * 3 + b] + rpic / 2.0f * (Cp[a * 3 + b] - Cp[b * 3 + a]);
            if (rpic < -0.001f) {
#pragma unroll
                for (int k = 0; k < 9; k++) Cm[k] = 0.0f;
            }
            const float m = mass[id];
            const float md = m * dx;
            const float sv = sfac * vol[id];
            rec[0][s] = f0;
            rec[1][s] = f1;
            rec[2][s] = f2;
            rec[3][s] = m;
            rec[4][s] = m * vp.x - md * (Cm[0] * f0 + Cm[1] * f1 + Cm[2] * f2);
            rec[5][s] = m * vp.y - md * (Cm[3] * f0 + Cm[4] * f1 + Cm[5] * f2);
            rec[6][s] = m * vp.z - md * (Cm[6] * f0 + Cm[7] * f1 + Cm[8] * f2);
#pragma unroll
            for (int k = 0; k < 9; k++) rec[7 + k][s] = md * Cm[k];
#pragma unroll
            for (int k = 0; k < 9; k++) rec[16 + k][s] = sv * stress[(long)k * n + id];
        }
        __syncthreads();
        for (int task = tid; task < kTasks; task += kBlock) {
            const int c = task / 9;
            const int i = (task % 9) / 3;
            const int j = task % 3;
            const int lo = max(cs[c] - p0, c0) - c0;
            const int hi = min(cs[c + 1] - p0, c0 + cn) - c0;
            if (lo >= hi) continue;
            float acc[3][4];
#pragma unroll
            for (int k = 0; k < 3; k++) acc[k][0] = acc[k][1] = acc[k][2] = acc[k][3] = 0.0f;
            const float fi = (float)i, fj = (float)j;
            for (int s = lo; s < hi; s++) {
                float wx, dwx, wy, dwy;
                wsel(rec[0][s], i, wx, dwx);
                wsel(rec[1][s], j, wy, dwy);
                float wz[3], dwz[3];
                bspline(rec[2][s], wz, dwz);
                const float m = rec[3][s];
                const float wxy = wx * wy, gxw = dwx * wy, gyw = wx * dwy;
                // affine momentum at k = 0: A + Q (i, j, 0)
                float bx = rec[4][s] + rec[7][s] * fi + rec[8][s] * fj;
                float by = rec[5][s] + rec[10][s] * fi + rec[11][s] * fj;
                float bz = rec[6][s] + rec[13][s] * fi + rec[14][s] * fj;
                const float qx = rec[9][s], qy = rec[12][s], qz = rec[15][s];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const float w = wxy * wz[k];
                    const float g0 = gxw * wz[k], g1 = gyw * wz[k], g2 = wxy * dwz[k];
                    const float ax = bx + qx * (float)k, ay = by + qy * (float)k, az = bz + qz * (float)k;
                    acc[k][0] += w * m;
                    acc[k][1] += w * ax + (rec[16][s] * g0 + rec[17][s] * g1 + rec[18][s] * g2);
                    acc[k][2] += w * ay + (rec[19][s] * g0 + rec[20][s] * g1 + rec[21][s] * g2);
                    acc[k][3] += w * az + (rec[22][s] * g0 + rec[23][s] * g1 + rec[24][s] * g2);
                }
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                float4& dst = part[c * 27 + i * 9 + j * 3 + k];
                dst.x += acc[k][0];
                dst.y += acc[k][1];
                dst.z += acc[k][2];
                dst.w += acc[k][3];
            }
        }
    }
    __syncthreads();
    float* out = partial + (long)t * kTileNodes * 4;
    for (int node = tid; node < kTileNodes; node += kBlock) {
        const int nx = node / (kTN * kTN), ny = (node / kTN) % kTN, nz = node % kTN;
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int cx = nx - i;
            if (cx < 0 || cx >= kTB) continue;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int cy = ny - j;
                if (cy < 0 || cy >= kTB) continue;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const int cz = nz - k;
                    if (cz < 0 || cz >= kTB) continue;
                    const float4 q = part[((cx * kTB + cy) * kTB + cz) * 27 + i * 9 + j * 3 + k];
                    s.x += q.x;
                    s.y += q.y;
                    s.z += q.z;
                    s.w += q.w;
                }
            }
        }
        out[node] = s.x;
        out[kTileNodes + node] = s.y;
        out[2 * kTileNodes + node] = s.z;
        out[3 * kTileNodes + node] = s.w;
    }
    if (tid == 0) tile_flag[t] = 1;
}

// ---- 3. combine ----------------------------------------------------------
__global__ void p2g_combine_kernel(TileGeom tg, const float* __restrict__ partial,
                                   const int32_t* __restrict__ tile_flag, float* __restrict__ grid_m,
                                   float* __restrict__ grid_v_in, float* __restrict__ grid_v_out, int do_update,
                                   float gxv, float gyv, float gzv, float dt, float damping, int bc, float fric,
                                   float gmag, int pad) {
    const long G = (long)tg.gx * tg.gy * tg.gz;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const int iz = (int)(i % tg.gz);
    const int iy = (int)((i / tg.gz) % tg.gy);
    const int ix = (int)(i / ((long)tg.gy * tg.gz));
    float m = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
    const int tx1 = ix / kTB, ty1 = iy / kTB, tz1 = iz / kTB;
    for (int tx = tx1 - 1; tx <= tx1; tx++) {
        const int lx = ix - tx * kTB;
        if (tx < 0 || tx >= tg.ntx || lx >= kTN) continue;
        for (int ty = ty1 - 1; ty <= ty1; ty++) {
            const int ly = iy - ty * kTB;
            if (ty < 0 || ty >= tg.nty || ly >= kTN) continue;
            for (int tz = tz1 - 1; tz <= tz1; tz++) {
                const int lz = iz - tz * kTB;
                if (tz < 0 || tz >= tg.ntz || lz >= kTN) continue;
                const long t = ((long)tx * tg.nty + ty) * tg.ntz + tz;
                if (!tile_flag[t]) continue;
                const float* pb = partial + t * kTileNodes * 4;
                const int node = (lx * kTN + ly) * kTN + lz;
                m += pb[node];
                mx += pb[kTileNodes + node];
                my += pb[2 * kTileNodes + node];
                mz += pb[3 * kTileNodes + node];
            }
        }
    }
    grid_m[i] = m;
    grid_v_in[i] = mx;
    grid_v_in[G + i] = my;
    grid_v_in[2 * G + i] = mz;
    if (do_update) {
        f3 vo = grid_update_node(m, f3{mx, my, mz}, ix, iy, iz, tg.gx, tg.gy, tg.gz, gxv, gyv, gzv, dt, damping, bc,
                                 fric, gmag, pad);
        st3(grid_v_out, G, i, vo);
    }
}

struct P2GWork {
    int32_t* keys;
    int32_t* sorted;
    float* partial;
    int32_t* tile_flag;
    int32_t* oob;
    BinWork bin;
};

static size_t p2g_ws_bytes(long n, const TileGeom& tg) {
    long nkeys = tg.ntiles() * kTileCells;
    long nn = n > 0 ? n : 1;
    return Carver::bytes_for(nn, 4) * 2 + Carver::bytes_for(tg.ntiles() * kTileNodes * 4, 4) +
           Carver::bytes_for(tg.ntiles(), 4) + Carver::bytes_for(1, 4) + bin_workspace_bytes(nn, nkeys);
}

static bool carve_p2g(void* ws, size_t bytes, long n, const TileGeom& tg, P2GWork& w) {
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
                   float gzv, float dt, float damping, int bc, float fric, float gmag, int pad, void* ws,
                   size_t ws_bytes, int* host_oob, hipStream_t s) {
    MPMX_REQUIRE(n >= 0 && gx >= 3 && gy >= 3 && gz >= 3, "p2g: grid must have at least 3 nodes per axis");
    MPMX_REQUIRE(grid_m && grid_v_in && (!do_update || grid_v_out), "p2g: null grid pointer");
    MPMX_REQUIRE(n == 0 || (activeLabel && materialLabel && stress && x && v && C && vol && mass),
                 "p2g: null particle pointer");
    MPMX_REQUIRE(bc == MPMX_BC_NONE || bc == MPMX_BC_FRICTION, "p2g: unsupported boundary condition %d", bc);
    TileGeom tg = make_geom(gx, gy, gz);
    MPMX_REQUIRE(tg.ntiles() * kTileCells < (1L << 31), "p2g: grid too large for int32 keys");
    size_t need = p2g_ws_bytes(n, tg);
    MPMX_REQUIRE(ws && ws_bytes >= need, "p2g: workspace of %zu bytes given, %zu needed", ws_bytes, need);
    P2GWork w;
    MPMX_REQUIRE(carve_p2g(ws, ws_bytes, n, tg, w), "p2g: workspace carve failed");
    int rc;
    hipLaunchKernelGGL(fill_i32_one, dim3(1), dim3(1), 0, s, w.oob);
    if ((rc = check_launch("p2g_oob_zero", s))) return rc;
    if (n > 0) {
        hipLaunchKernelGGL(p2g_key_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, activeLabel,
                           materialLabel, x, minx, miny, minz, inv_dx, tg, w.keys, w.oob);
        if ((rc = check_launch("p2g_key", s))) return rc;
    }
    if ((rc = counting_sort_ids(w.keys, n, tg.ntiles() * kTileCells, w.sorted, w.bin, s))) return rc;
    hipLaunchKernelGGL(p2g_tile_kernel, dim3((unsigned)tg.ntiles()), dim3(kBlock), 0, s, n, w.sorted,
                       w.bin.cell_start, stress, x, v, C, vol, mass, dx, inv_dx, minx, miny, minz, rpic, dt,
                       w.partial, w.tile_flag);
    if ((rc = check_launch("p2g_tile", s))) return rc;
    long G = (long)gx * gy * gz;
    hipLaunchKernelGGL(p2g_combine_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, s, tg, w.partial,
                       w.tile_flag, grid_m, grid_v_in, grid_v_out, do_update, gxv, gyv, gzv, dt, damping, bc, fric,
                       gmag, pad);
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
                   0.f, 0.f, 0, workspace, ws_bytes, host_out_of_grid, as_stream(stream));
}

int mpmx_p2g_grid_update(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* stress,
                         const float* x, const float* v, const float* C, const float* vol, const float* mass,
                         float dx, float inv_dx, float minx, float miny, float minz, float rpic_damping, int gx,
                         int gy, int gz, float* grid_m, float* grid_v_in, float* grid_v_out, float gxv, float gyv,
                         float gzv, float dt, float damping_scale, int bc, float friction, float gravity_magnitude,
                         int boundary_padding, void* workspace, size_t ws_bytes, void* stream) {
    return p2g_run(n, activeLabel, materialLabel, stress, x, v, C, vol, mass, dx, inv_dx, minx, miny, minz,
                   rpic_damping, gx, gy, gz, grid_m, grid_v_in, grid_v_out, 1, gxv, gyv, gzv, dt, damping_scale, bc,
                   friction, gravity_magnitude, boundary_padding, workspace, ws_bytes, nullptr, as_stream(stream));
}

}  // extern "C"
