// Deterministic tile reduction for P2G (p2g_apic_with_stress,
// mpmRoutines.py:485-569), shared by the standalone P2G entry points and the
// fused substep.
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
//
// In the fused substep the record stage first runs the constitutive update
// (return map + Kirchhoff stress) of the particle, so the stress goes from
// registers into the record; particles whose stencil leaves the grid are
// collected in one extra "orphan" bin after the last tile and get the
// constitutive update only.
#pragma once
#include "mpmx_constitutive.hpp"
#include "mpmx_device.hpp"

namespace mpmx {

constexpr int kTB = 4;                       // base cells per tile edge
constexpr int kTN = kTB + 2;                 // nodes per tile edge
constexpr int kTileCells = kTB * kTB * kTB;  // 64
constexpr int kTileNodes = kTN * kTN * kTN;  // 216
constexpr int kRec = 25;                     // floats per staged particle record
constexpr int kTasks = kTileCells * 3;       // (cell, i): each task owns the 9 nodes (i, j, k) of its cell
// Particles staged per LDS round: 512 (two per thread; a whole tile of 8 per
// cell in one round, 51 KB, 3 workgroups per CU), 384 (38 KB) or 256 (27.6 KB,
// the size of the final node-sum slots).  The smaller chunks are compiled for 4
// and 5 waves per SIMD, so that their LDS saving turns into occupancy.  The
// chunk changes no bit of the result: every task sums its cell's particles in
// slot order, in registers, across rounds.  Chosen at launch (MPMX_TUNE_P2G_CHUNK).
constexpr int kChunkMax = 512;
// Record slot s lives at column s + s/8: the particles of neighbouring cells
// (about 8 apart in the list) then fall in different LDS banks when the tasks
// of a wavefront read their cells' records in step.
__host__ __device__ constexpr int rec_cols(int chunk) { return chunk + chunk / 8; }
__host__ __device__ constexpr int rcol(int s) { return s + (s >> 3); }

// Grid geometry of one scene: its nodes, its 4x4x4-cell tiles, where its tiles
// and nodes start in the global numbering, and its grid origin.  Several
// scenes (an ensemble) share one launch of every substep kernel: tiles are
// numbered scene after scene, grid nodes likewise (scene s owns the global
// nodes node0 .. node0 + gx*gy*gz), and each particle belongs to one scene.
struct SceneDesc {
    int gx, gy, gz;     // grid nodes
    int ntx, nty, ntz;  // tiles
    int tile0;          // first global tile
    int unused;
    long node0;         // first global node
    float minx, miny, minz, unused2;
};

inline SceneDesc make_scene(int gx, int gy, int gz, float minx, float miny, float minz, int tile0, long node0) {
    SceneDesc d{};
    d.gx = gx;
    d.gy = gy;
    d.gz = gz;
    d.ntx = (gx - 2 + kTB - 1) / kTB;
    d.nty = (gy - 2 + kTB - 1) / kTB;
    d.ntz = (gz - 2 + kTB - 1) / kTB;
    d.tile0 = tile0;
    d.node0 = node0;
    d.minx = minx;
    d.miny = miny;
    d.minz = minz;
    return d;
}

// All scenes of a launch.  With one scene its description travels in the
// kernel arguments; with more, in a device table of nscenes + 1 entries whose
// last entry holds the totals (tile0 = tiles, node0 = nodes).
struct Geo {
    SceneDesc one;
    const SceneDesc* table;
    int nscenes;
    int ntiles_;
    long nnodes;
    __host__ __device__ long ntiles() const { return ntiles_; }
    __device__ __forceinline__ SceneDesc at(int s) const { return nscenes == 1 ? one : table[s]; }
    // the scene holding global tile t (node i): the last entry with tile0 <= t
    __device__ __forceinline__ int scene_of_tile(int t) const {
        if (nscenes == 1) return 0;
        int lo = 0, hi = nscenes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].tile0 <= t) lo = mid; else hi = mid - 1;
        }
        return lo;
    }
    __device__ __forceinline__ int scene_of_node(long i) const {
        if (nscenes == 1) return 0;
        int lo = 0, hi = nscenes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (table[mid].node0 <= i) lo = mid; else hi = mid - 1;
        }
        return lo;
    }
};

inline Geo make_geom(int gx, int gy, int gz, float minx = 0.0f, float miny = 0.0f, float minz = 0.0f) {
    Geo g{};
    g.one = make_scene(gx, gy, gz, minx, miny, minz, 0, 0);
    g.table = nullptr;
    g.nscenes = 1;
    g.ntiles_ = g.one.ntx * g.one.nty * g.one.ntz;
    g.nnodes = (long)gx * gy * gz;
    return g;
}

// Key of the base cell of a particle of scene d at x, or -1 when its 3x3x3
// stencil leaves the scene's grid.  Compares against g - 3 (not b + 2 < g) so
// that a saturated huge index cannot wrap around.
__device__ __forceinline__ int tile_key(f3 x, float inv_dx, const SceneDesc& d) {
    float f;
    int bx = base_axis(x.x, d.minx, inv_dx, f);
    int by = base_axis(x.y, d.miny, inv_dx, f);
    int bz = base_axis(x.z, d.minz, inv_dx, f);
    if (bx < 0 || by < 0 || bz < 0 || bx > d.gx - 3 || by > d.gy - 3 || bz > d.gz - 3) return -1;
    int tx = bx / kTB, ty = by / kTB, tz = bz / kTB;
    int lx = bx - tx * kTB, ly = by - ty * kTB, lz = bz - tz * kTB;
    return (d.tile0 + (tx * d.nty + ty) * d.ntz + tz) * kTileCells + (lx * kTB + ly) * kTB + lz;
}

struct TileArgs {
    mpmx_particles P;             // particle state (read; written by the constitutive update when fused)
    const float* stress;          // stress to scatter when not fused
    const int32_t* sorted;        // particle ids in key order
    const int32_t* cell_start;    // [nkeys + 1]
    float dx, inv_dx, rpic, dt, eff;
    float* partial;               // [ntiles][4][216]
    int32_t* tile_flag;           // [ntiles]
    int ntiles;
    Geo geo;
    int xcd;                      // deal tiles to XCDs in contiguous ranges (xcd_tile) or in block order
    int32_t* misbinned;           // optional: counts particles whose base cell is not their bin's cell
};

// Constitutive update of the particles in the orphan bin (fused substep only).
__device__ __forceinline__ void orphan_bin(const TileArgs& a) {
    const long b = a.cell_start[(long)a.ntiles * kTileCells];
    const long e = a.cell_start[(long)a.ntiles * kTileCells + 1];
    for (long s = b + threadIdx.x; s < e; s += kBlock) {
        float T[9];
        constitutive_update(a.P, a.sorted[s], a.eff, T);
    }
}

// LDS of the tile kernel: the staged records of one round, and afterwards (the
// same bytes) the 27 offset slots of every cell for the final node sums.
constexpr int kPartFloats = kTileCells * 27 * 4;

// FUSED: the constitutive update runs here (stress from registers); else the
// stress is read.  SYM: the stress is symmetric by construction (the substep's
// own constitutive update writes T[i][j] = T[j][i]), so 6 components are read
// and staged; the standalone P2G reads all 9 of whatever it is given.
// P2G inputs of one particle
struct P2GIn {
    float T[9];  // stress
    f3 x, v;
    float C[9];
    float m, vol;
    bool mpm;
};

template <bool FUSED, bool SYM>
__device__ __forceinline__ void p2g_load(const TileArgs& a, long n, int id, P2GIn& q) {
    if (FUSED) {
        q.mpm = constitutive_update(a.P, id, a.eff, q.T);
    } else if (SYM) {
        // the constitutive update ran before this kernel and may have turned it XPBD
        q.mpm = a.P.materialLabel[id] <= 1;
        q.T[0] = a.stress[id];
        q.T[1] = a.stress[1L * n + id];
        q.T[2] = a.stress[2L * n + id];
        q.T[4] = a.stress[4L * n + id];
        q.T[5] = a.stress[5L * n + id];
        q.T[8] = a.stress[8L * n + id];
        q.T[3] = q.T[1];
        q.T[6] = q.T[2];
        q.T[7] = q.T[5];
    } else {
        q.mpm = a.P.materialLabel[id] <= 1;
#pragma unroll
        for (int k = 0; k < 9; k++) q.T[k] = a.stress[(long)k * n + id];
    }
    q.x = ld3(a.P.x, n, id);
    q.v = ld3(a.P.v, n, id);
    ld9(a.P.C, n, id, q.C);
    q.m = a.P.mass[id];
    q.vol = a.P.vol[id];
}

// the P2G record of one particle in LDS column r (slot: its place in the bins)
template <bool FUSED, bool SYM, int COLS>
__device__ __forceinline__ void p2g_record(const TileArgs& a, const P2GIn& q, float (*rec)[COLS], int r, int slot,
                                           int t, const int* cs, const SceneDesc& sd, float sfac) {
    constexpr int kRows = (SYM || FUSED) ? kRec - 3 : kRec;
    if (!q.mpm) {
        // became XPBD in this substep: it no longer transfers to the grid
        // (p2g tests the label after the stress pass, mpmRoutines.py:507)
#pragma unroll
        for (int k = 0; k < kRows; k++) rec[k][r] = 0.0f;
        rec[0][r] = 1.0f;
        rec[1][r] = 1.0f;
        rec[2][r] = 1.0f;
        return;
    }
    const f3 xp = q.x, vp = q.v;
    if (a.misbinned) {
        // bins come from the previous substep's G2P: check that the particle is still in its cell
        const int key = tile_key(xp, a.inv_dx, sd);
        const int c = key - t * kTileCells;
        if (c < 0 || c >= kTileCells || slot < cs[c] || slot >= cs[c + 1]) atomicAdd(a.misbinned, 1);
    }
    float f0, f1, f2;
    base_axis(xp.x, sd.minx, a.inv_dx, f0);
    base_axis(xp.y, sd.miny, a.inv_dx, f1);
    base_axis(xp.z, sd.minz, a.inv_dx, f2);
    const float* Cp = q.C;
    float Cm[9];
    const float rpic = a.rpic;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            Cm[i * 3 + j] = (1.0f - rpic) * Cp[i * 3 + j] + rpic / 2.0f * (Cp[i * 3 + j] - Cp[j * 3 + i]);
    if (rpic < -0.001f) {
#pragma unroll
        for (int k = 0; k < 9; k++) Cm[k] = 0.0f;
    }
    const float m = q.m;
    const float md = m * a.dx;
    const float sv = sfac * q.vol;
    const float* T = q.T;
    rec[0][r] = f0;
    rec[1][r] = f1;
    rec[2][r] = f2;
    rec[3][r] = m;
    rec[4][r] = m * vp.x - md * (Cm[0] * f0 + Cm[1] * f1 + Cm[2] * f2);
    rec[5][r] = m * vp.y - md * (Cm[3] * f0 + Cm[4] * f1 + Cm[5] * f2);
    rec[6][r] = m * vp.z - md * (Cm[6] * f0 + Cm[7] * f1 + Cm[8] * f2);
#pragma unroll
    for (int k = 0; k < 9; k++) rec[7 + k][r] = md * Cm[k];
    if (SYM || FUSED) {  // symmetric: xx yy zz xy xz yz
        rec[16][r] = sv * T[0];
        rec[17][r] = sv * T[4];
        rec[18][r] = sv * T[8];
        rec[19][r] = sv * T[1];
        rec[20][r] = sv * T[2];
        rec[21][r] = sv * T[5];
    } else {
#pragma unroll
        for (int k = 0; k < 9; k++) rec[16 + k][r] = sv * T[k];
    }
}

template <bool FUSED, bool SYM, int CHUNK = kChunkMax>
__device__ __forceinline__ void p2g_tile_body(const TileArgs& a) {
    // a symmetric stress takes 6 record rows instead of 9: 22 rows x 576 columns
    // (51 KB) stage a whole tile of up to 512 particles in one round
    constexpr int kRows = (SYM || FUSED) ? kRec - 3 : kRec;
    constexpr int kRecCols = rec_cols(CHUNK);
    constexpr int kLds = kRows * kRecCols > kPartFloats ? kRows * kRecCols : kPartFloats;
    __shared__ float4 smem4[(kLds + 3) / 4];
    __shared__ int cs[kTileCells + 1];
    float (*rec)[kRecCols] = reinterpret_cast<float (*)[kRecCols]>(smem4);
    float4* part = smem4;

    const int t = tile_of_block(a.xcd, (int)blockIdx.x, (int)gridDim.x);
    const int tid = threadIdx.x;
    if (FUSED && t == a.ntiles) {
        orphan_bin(a);
        return;
    }
    const long n = a.P.n;
    const SceneDesc sd = a.geo.at(a.geo.scene_of_tile(t));
    // the tile's range straight from memory, so that the first round's loads need
    // not wait for the cell table in LDS (the barrier after staging publishes it)
    if (tid <= kTileCells) cs[tid] = a.cell_start[(long)t * kTileCells + tid];
    const int p0 = a.cell_start[(long)t * kTileCells];
    const int np = a.cell_start[(long)t * kTileCells + kTileCells] - p0;
    if (np == 0) {
        if (tid == 0) a.tile_flag[t] = 0;
        return;
    }

    // task (cell, i) of this thread: it owns the nodes (i, j, k) of the cell and
    // keeps their sums in registers over all rounds
    const int tc = tid / 3, ti = tid % 3;
    const bool has_task = tid < kTasks;
    float acc[9][4];
#pragma unroll
    for (int q = 0; q < 9; q++) acc[q][0] = acc[q][1] = acc[q][2] = acc[q][3] = 0.0f;

    const float sfac = -a.dt * a.inv_dx;
    for (int c0 = 0; c0 < np; c0 += CHUNK) {
        const int cn = min(CHUNK, np - c0);
        // the previous round's readers are done with rec; the misbinned check reads cs
        if (c0 > 0 || a.misbinned) __syncthreads();
        if constexpr (FUSED) {
            for (int s = tid; s < cn; s += kBlock) {
                P2GIn q;
                p2g_load<FUSED, SYM>(a, n, a.sorted[p0 + c0 + s], q);
                p2g_record<FUSED, SYM, kRecCols>(a, q, rec, rcol(s), p0 + c0 + s, t, cs, sd, sfac);
            }
        } else {
            // two particles per thread and pass: both load before either is used
            for (int s = tid; s < cn; s += 2 * kBlock) {
                const int s2 = s + kBlock;
                const bool two = s2 < cn;
                const int id0 = a.sorted[p0 + c0 + s];
                const int id1 = a.sorted[p0 + c0 + (two ? s2 : s)];
                P2GIn q0, q1;
                p2g_load<FUSED, SYM>(a, n, id0, q0);
                p2g_load<FUSED, SYM>(a, n, id1, q1);
                p2g_record<FUSED, SYM, kRecCols>(a, q0, rec, rcol(s), p0 + c0 + s, t, cs, sd, sfac);
                if (two) p2g_record<FUSED, SYM, kRecCols>(a, q1, rec, rcol(s2), p0 + c0 + s2, t, cs, sd, sfac);
            }
        }
        __syncthreads();
        if (has_task) {
            const int lo = max(cs[tc] - p0, c0) - c0;
            const int hi = min(cs[tc + 1] - p0, c0 + cn) - c0;
            const float fi = (float)ti;
            // the x weight and derivative of offset ti without branches (the lanes of a
            // wavefront hold all three offsets): with u = f - (1.5, 1, 0.5)[ti],
            // w = u^2 / 2, 0.75 - u^2, u^2 / 2 and dw = u, -2u, u (bspline's products:
            // 1.5 - f is -u exactly)
            const float wshift = ti == 0 ? 1.5f : (ti == 1 ? 1.0f : 0.5f);
            const bool wmid = ti == 1;
            for (int s = lo; s < hi; s++) {
                const int r = rcol(s);
                const float u = rec[0][r] - wshift, uu = u * u;
                const float wx = wmid ? (0.0f - uu) + 0.75f : uu * 0.5f;
                const float dwx = wmid ? -2.0f * u : u;
                float wy[3], dwy[3], wz[3], dwz[3];
                bspline(rec[1][r], wy, dwy);
                bspline(rec[2][r], wz, dwz);
                const float m = rec[3][r];
                // Momentum of node (i, j, k): w (A + Q (i, j, k)) + S grad w, with w the
                // product of the axis weights, A the staged momentum, Q = m dx Cm and S
                // the stress impulse (columns S_d).  Factored along the axes, every FMA
                // landing on an accumulator:
                //   node (j, k) += wz_k E(j) + a_j H(k),   mass += wz_k wy_j wx m
                //   E(j) = wy_j (P + j Qy) + dwy_j R,      a_j = wx wy_j
                //   P = wx (A + i Q_0) + dwx S_0,  Qy = wx Q_1,  R = wx S_1
                //   H(k) = k wz_k Q_2 + dwz_k S_2
                float s00, s01, s02, s10, s11, s12, s20, s21, s22;
                if (SYM || FUSED) {
                    s00 = rec[16][r];
                    s11 = rec[17][r];
                    s22 = rec[18][r];
                    s01 = s10 = rec[19][r];
                    s02 = s20 = rec[20][r];
                    s12 = s21 = rec[21][r];
                } else {
                    s00 = rec[16][r], s01 = rec[17][r], s02 = rec[18][r];
                    s10 = rec[19][r], s11 = rec[20][r], s12 = rec[21][r];
                    s20 = rec[22][r], s21 = rec[23][r], s22 = rec[24][r];
                }
                const float P[3] = {fmaf(wx, fmaf(rec[7][r], fi, rec[4][r]), dwx * s00),
                                    fmaf(wx, fmaf(rec[10][r], fi, rec[5][r]), dwx * s10),
                                    fmaf(wx, fmaf(rec[13][r], fi, rec[6][r]), dwx * s20)};
                const float Qy[3] = {wx * rec[8][r], wx * rec[11][r], wx * rec[14][r]};
                const float R[3] = {wx * s01, wx * s11, wx * s21};
                const float q2[3] = {rec[9][r], rec[12][r], rec[15][r]};
                const float sz[3] = {s02, s12, s22};
                float H[3][3];
#pragma unroll
                for (int k = 0; k < 3; k++)
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        H[k][c] = k == 0 ? dwz[0] * sz[c] : fmaf((float)k * wz[k], q2[c], dwz[k] * sz[c]);
                const float wxm = wx * m;
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float aj = wx * wy[j], Mj = wxm * wy[j];
                    float E[3];
#pragma unroll
                    for (int c = 0; c < 3; c++)
                        E[c] = fmaf(wy[j], j == 0 ? P[c] : fmaf((float)j, Qy[c], P[c]), dwy[j] * R[c]);
#pragma unroll
                    for (int k = 0; k < 3; k++) {
                        float* A = acc[j * 3 + k];
                        A[0] = fmaf(wz[k], Mj, A[0]);
#pragma unroll
                        for (int c = 0; c < 3; c++) A[c + 1] = fmaf(wz[k], E[c], fmaf(aj, H[k][c], A[c + 1]));
                    }
                }
            }
        }
    }
    __syncthreads();  // every record has been read: the LDS now holds the offset slots
    if (has_task) {
#pragma unroll
        for (int q = 0; q < 9; q++) part[tc * 27 + ti * 9 + q] = make_float4(acc[q][0], acc[q][1], acc[q][2], acc[q][3]);
    }
    __syncthreads();
    float* out = a.partial + (long)t * kTileNodes * 4;
    for (int node = tid; node < kTileNodes; node += kBlock) {
        const int nx = node / (kTN * kTN), ny = (node / kTN) % kTN, nz = node % kTN;
        float4 s = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        // offsets (i, j, k) whose cell (n - offset) lies in the tile, i ascending, then j, then k
        const int i0 = max(0, nx - (kTB - 1)), i1 = min(2, nx);
        const int j0 = max(0, ny - (kTB - 1)), j1 = min(2, ny);
        const int k0 = max(0, nz - (kTB - 1)), k1 = min(2, nz);
        for (int i = i0; i <= i1; i++) {
            for (int j = j0; j <= j1; j++) {
                for (int k = k0; k <= k1; k++) {
                    const float4 q = part[(((nx - i) * kTB + (ny - j)) * kTB + (nz - k)) * 27 + i * 9 + j * 3 + k];
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
    if (tid == 0) a.tile_flag[t] = 1;
}

// P2G only (the stress is read); SYM as for p2g_tile_body.  The waves-per-SIMD
// hint asks the register allocator for the occupancy the chunk's LDS allows.
template <bool SYM, int CHUNK = kChunkMax>
__global__ void __launch_bounds__(kBlock)
    __attribute__((amdgpu_waves_per_eu(CHUNK >= 512 ? 1 : (CHUNK >= 384 ? 4 : 5))))
    p2g_tile_kernel(TileArgs a) {
    p2g_tile_body<false, SYM, CHUNK>(a);
}

// the constitutive update of every particle inside the tile pass
template <int UNUSED = 0>
__global__ void __launch_bounds__(kBlock) p2g_tile_fused_kernel(TileArgs a) {
    p2g_tile_body<true, true>(a);
}

// ---- combine ----------------------------------------------------------------
struct CombineArgs {
    Geo geo;
    const float* partial;
    const int32_t* tile_flag;
    float *grid_m, *grid_v_in, *grid_v_out;
    int do_update;
    float gxv, gyv, gzv, dt, damping, fric, gmag, restitution;
    int bc, pad;
    int32_t* zero_ints;  // optional: the next substep's bin counts are cleared here
    long n_zero;
    unsigned long long* zero_u64;  // optional: cell sums cleared before G2P
    long n_zero_u64;
    int32_t* zero_chunks;          // optional: the next substep's per-chunk bin totals
    long n_chunks;
};

// a template so that every translation unit including this header can instantiate it
template <int UNUSED = 0>
__global__ void p2g_combine_kernel(CombineArgs a) {
    const long G = a.geo.nnodes;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (a.zero_ints) {
        for (long z = i; z < a.n_zero; z += (long)gridDim.x * blockDim.x) a.zero_ints[z] = 0;
    }
    if (a.zero_u64) {
        for (long z = i; z < a.n_zero_u64; z += (long)gridDim.x * blockDim.x) a.zero_u64[z] = 0ull;
    }
    if (a.zero_chunks) {
        for (long z = i; z < a.n_chunks; z += (long)gridDim.x * blockDim.x) a.zero_chunks[z] = 0;
    }
    if (i >= G) return;
    const SceneDesc d = a.geo.at(a.geo.scene_of_node(i));
    const int li = (int)(i - d.node0);  // a scene has fewer than 2^31 nodes: 32-bit division
    const int iz = li % d.gz;
    const int iy = (li / d.gz) % d.gy;
    const int ix = li / (d.gy * d.gz);
    // the (up to 8) tiles whose node blocks hold this node, tx ascending, then ty,
    // then tz; every flag and partial is loaded before the first sum, and a tile
    // whose flag is clear adds nothing
    const int tx1 = ix / kTB, ty1 = iy / kTB, tz1 = iz / kTB;
    float pv[8][4];
    int on[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int tx = tx1 - 1 + (q >> 2), ty = ty1 - 1 + ((q >> 1) & 1), tz = tz1 - 1 + (q & 1);
        const int lx = ix - tx * kTB, ly = iy - ty * kTB, lz = iz - tz * kTB;
        on[q] = 0;
        pv[q][0] = pv[q][1] = pv[q][2] = pv[q][3] = 0.0f;
        if (tx >= 0 && tx < d.ntx && lx < kTN && ty >= 0 && ty < d.nty && ly < kTN && tz >= 0 && tz < d.ntz &&
            lz < kTN) {
            const long t = d.tile0 + ((long)tx * d.nty + ty) * d.ntz + tz;
            const float* pb = a.partial + t * kTileNodes * 4;
            const int node = (lx * kTN + ly) * kTN + lz;
            on[q] = a.tile_flag[t];
            pv[q][0] = pb[node];
            pv[q][1] = pb[kTileNodes + node];
            pv[q][2] = pb[2 * kTileNodes + node];
            pv[q][3] = pb[3 * kTileNodes + node];
        }
    }
    float m = 0.0f, mx = 0.0f, my = 0.0f, mz = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        if (on[q]) {
            m += pv[q][0];
            mx += pv[q][1];
            my += pv[q][2];
            mz += pv[q][3];
        }
    }
    a.grid_m[i] = m;
    a.grid_v_in[i] = mx;
    a.grid_v_in[G + i] = my;
    a.grid_v_in[2 * G + i] = mz;
    if (a.do_update) {
        f3 vo = grid_update_node(m, f3{mx, my, mz}, ix, iy, iz, d.gx, d.gy, d.gz, a.gxv, a.gyv, a.gzv, a.dt,
                                 a.damping, a.bc, a.fric, a.gmag, a.pad, a.restitution);
        st3(a.grid_v_out, G, i, vo);
    }
}

}  // namespace mpmx
