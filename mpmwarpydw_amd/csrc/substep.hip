// The fused MPM substep (mpmSimulationStep, simulationRoutines.py:6-250) for
// gfx950: three passes over HBM instead of the reference's launch chain.
//
//   K1 tile pass   per 4x4x4-cell tile: the constitutive update of every binned
//                  particle (return map + Kirchhoff stress, written back) feeds
//                  the P2G record straight from registers; the tile is reduced
//                  in LDS (p2g_tile.hpp).  Out-of-grid particles sit in an
//                  extra orphan bin and get the constitutive update only.
//   K2 node pass   tile partials -> grid_m, grid_v_in, and the grid update
//                  (normalisation, gravity, damping, walls) -> grid_v_out; it
//                  also clears the bin counts of the next substep.
//   K3 tile pass   G2P per tile with the tile's 6x6x6 grid velocities staged in
//                  LDS; each particle then computes its new base cell and
//                  counts itself into the next substep's bins.
//   bin            scan of the counts, scatter of ids, ascending ids per cell.
//
// Particles keep their bins between substeps: K3 bins them at their new
// positions, and the XPBD step between substeps restores MPM positions
// (apply_mpm_contact_impulses, xpbdRoutines.py:293) and turns no particle into
// MPM.  After any other change of x, labels or activity, call
// mpmx_substep_bin() again.
#include "mpmx_internal.hpp"
#include "p2g_tile.hpp"

namespace mpmx {

struct SubstepWork {
    int32_t* misbinned;  // [1]
    int32_t* keys;
    int32_t* sorted;
    float* partial;
    int32_t* tile_flag;
    SceneDesc* table;    // [nscenes + 1] for an ensemble
    unsigned long long* vlc;  // [2 * nodes] cell sums of the volumetric locking correction
    uint8_t* lrm_flag;        // [n] candidates of the learned projection (mpmx_*_substep_learned)
    BinWork bin;
};

// keys: tile * 64 + cell, then the orphan bin (out-of-grid MPM particles), then
// a bin used only by the spatial reorder for every particle that is not MPM
static long substep_nkeys(const Geo& g) { return g.ntiles() * kTileCells + 2; }

// The spatial reorder's keys: active MPM particles as in the substep (tile
// keys, then the orphan bin), then active XPBD particles by the same tile keys
// (their contact walk gathers the particles of neighbouring cells), then every
// other particle in bins of kRestRun consecutive input indices, so that no bin of
// them holds more than kRestRun (a bin is sorted by one thread).
constexpr long kRestRun = 16;
static long reorder_nkeys(const Geo& g, long n) {
    return 2 * g.ntiles() * kTileCells + 1 + (n + kRestRun - 1) / kRestRun;
}
static long bin_nkeys(const Geo& g, long n) {
    const long a = substep_nkeys(g), b = reorder_nkeys(g, n);
    return a > b ? a : b;
}

static size_t substep_ws_bytes(long n, const Geo& g) {
    long nn = n > 0 ? n : 1;
    return Carver::bytes_for(1, 4) + Carver::bytes_for(nn, 4) * 2 + Carver::bytes_for(g.ntiles() * kTileNodes * 4, 4) +
           Carver::bytes_for(g.ntiles(), 4) + Carver::bytes_for(g.nscenes + 1, sizeof(SceneDesc)) +
           Carver::bytes_for(2 * g.nnodes, sizeof(unsigned long long)) + Carver::bytes_for(nn, 1) +
           bin_workspace_bytes(nn, bin_nkeys(g, nn));
}

static bool carve_substep(void* ws, size_t bytes, long n, const Geo& g, SubstepWork& w) {
    Carver cv(ws, bytes);
    long nn = n > 0 ? n : 1;
    w.misbinned = cv.take<int32_t>(1);
    w.keys = cv.take<int32_t>(nn);
    w.sorted = cv.take<int32_t>(nn);
    w.partial = cv.take<float>(g.ntiles() * kTileNodes * 4);
    w.tile_flag = cv.take<int32_t>(g.ntiles());
    w.table = cv.take<SceneDesc>(g.nscenes + 1);
    w.vlc = cv.take<unsigned long long>(2 * g.nnodes);
    w.lrm_flag = cv.take<uint8_t>(nn);
    carve_bin(cv, nn, bin_nkeys(g, nn), w.bin);
    return !cv.overflow;
}

__global__ void zero_one_kernel(int32_t* __restrict__ p) { *p = 0; }

// the scene of particle p (its scene array, or scene 0 without one)
__device__ __forceinline__ int scene_of(const mpmx_particles& P, long p) { return P.scene ? P.scene[p] : 0; }

// key of an active MPM particle: its base-cell key, or the orphan bin
__global__ void substep_key_kernel(mpmx_particles P, float inv_dx, Geo geo, int32_t* __restrict__ keys) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    int key = -1;
    if (P.activeLabel[p] == 1 && P.materialLabel[p] <= 1) {
        const int sc = scene_of(P, p);
        key = (sc >= 0 && sc < geo.nscenes) ? tile_key(ld3(P.x, P.n, p), inv_dx, geo.at(sc)) : -1;
        if (key < 0) key = (int)(geo.ntiles() * kTileCells);
    }
    keys[p] = key;
}

// key of every particle for the spatial reorder (the key ranges of reorder_nkeys)
__global__ void reorder_key_kernel(mpmx_particles P, float inv_dx, Geo geo, int32_t* __restrict__ keys) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    const long tiles = geo.ntiles() * kTileCells;
    const long input_index = P.order ? P.order[p] : p;
    long key = 2 * tiles + 1 + input_index / kRestRun;
    const int label = P.materialLabel[p];
    if (P.activeLabel[p] == 1 && (label <= 1 || label == 2)) {
        const int sc = scene_of(P, p);
        const int t = (sc >= 0 && sc < geo.nscenes) ? tile_key(ld3(P.x, P.n, p), inv_dx, geo.at(sc)) : -1;
        if (label <= 1)
            key = t >= 0 ? t : tiles;
        else if (t >= 0)
            key = tiles + 1 + t;
    }
    keys[p] = (int)key;
}

// the scene table of an ensemble, written into the workspace from kernel
// arguments (no host memory is read by a replayed graph)
constexpr int kSceneChunk = 16;
struct SceneChunk {
    SceneDesc d[kSceneChunk];
};
__global__ void write_scenes_kernel(SceneChunk c, int count, SceneDesc* __restrict__ dst) {
    const int k = threadIdx.x;
    if (k < count) dst[k] = c.d[k];
}

struct G2PArgs {
    mpmx_particles P;
    const int32_t* sorted;
    const int32_t* cell_start;
    const float* v_out;
    Geo geo;
    float dt, inv_dx;
    int update_cov;
    int32_t* keys;   // next substep: key per particle
    int32_t* count;  // next substep: particles per key (cleared by K2)
    int32_t* rank;   // next substep: arrival rank inside the key
    int32_t* chunks; // next substep: particles per chunk of kScanChunk keys (cleared by K2)
    int32_t* misbinned;
    int xcd;
    int segs;  // tile kernel: blocks per tile
};

__device__ __forceinline__ void bin_next(const G2PArgs& a, long id, int key) {
    a.keys[id] = key;
    if (key >= 0) {
        a.rank[id] = atomicAdd(&a.count[key], 1);
        atomicAdd(&a.chunks[key / kScanChunk], 1);
    }
}

// Grid velocity of node (b0+i, b1+j, b2+k) relative to the tile origin: from the
// tile's LDS block, or from HBM for a particle outside its tile's node block.
template <bool LDS>
__device__ __forceinline__ f3 g2p_node(const G2PArgs& a, const float (*gvs)[kTileNodes], int row, long grow, long G) {
    if (LDS) return f3{gvs[0][row], gvs[1][row], gvs[2][row]};
    return f3{a.v_out[grow], a.v_out[G + grow], a.v_out[2 * G + grow]};
}

// G2P of one particle (mpmRoutines.py:715-802): new v, x, C, F_trial (and the
// covariance when it is tracked); returns the particle's key for the next substep.
template <bool LDS>
__device__ __forceinline__ int g2p_one(const G2PArgs& a, const SceneDesc& sd, const float (*gvs)[kTileNodes], int id,
                                       f3 xp, const float Fp[9], int b0, int b1, int b2, int ox, int oy, int oz,
                                       const float fx[3]) {
    const mpmx_particles& P = a.P;
    const long n = P.n;
    const long G = a.geo.nnodes;
    const float inv_dx = a.inv_dx, dt = a.dt;
    float nv[3], nB[9], nG[9];
    g2p_gather(fx, inv_dx,
               [&](int i, int j, int k) {
                   const int row = ((b0 + i) * kTN + (b1 + j)) * kTN + b2 + k;
                   const long grow = sd.node0 + ((long)(b0 + ox + i) * sd.gy + (b1 + oy + j)) * sd.gz + (b2 + oz + k);
                   return g2p_node<LDS>(a, gvs, row, grow, G);
               },
               nv, nB, nG);
    const f3 xn = f3{xp.x + dt * nv[0], xp.y + dt * nv[1], xp.z + dt * nv[2]};
    st3(P.v, n, id, f3{nv[0], nv[1], nv[2]});
    st3(P.x, n, id, xn);
    const float sC = inv_dx * 4.0f;
    float Cn[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Cn[k] = nB[k] * sC;
    st9(P.C, n, id, Cn);
    float Fn[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            Fn[i * 3 + j] = Fp[i * 3 + j] + dt * (nG[i * 3 + 0] * Fp[0 * 3 + j] + nG[i * 3 + 1] * Fp[1 * 3 + j] +
                                                  nG[i * 3 + 2] * Fp[2 * 3 + j]);
    st9(P.F_trial, n, id, Fn);
    if (a.update_cov) {
        float* c6 = P.cov + 6L * id;
        float cn[9] = {c6[0], c6[1], c6[2], c6[1], c6[3], c6[4], c6[2], c6[4], c6[5]};
        float c1[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                float gc = nG[i * 3 + 0] * cn[0 * 3 + j] + nG[i * 3 + 1] * cn[1 * 3 + j] + nG[i * 3 + 2] * cn[2 * 3 + j];
                float cg = cn[i * 3 + 0] * nG[j * 3 + 0] + cn[i * 3 + 1] * nG[j * 3 + 1] + cn[i * 3 + 2] * nG[j * 3 + 2];
                c1[i * 3 + j] = cn[i * 3 + j] + dt * (gc + cg);
            }
        c6[0] = c1[0]; c6[1] = c1[1]; c6[2] = c1[2]; c6[3] = c1[4]; c6[4] = c1[5]; c6[5] = c1[8];
    }
    const int key = tile_key(xn, inv_dx, sd);
    return key < 0 ? (int)(a.geo.ntiles() * kTileCells) : key;
}

constexpr int kDeferRounds = 4;  // rounds of kBlock particles whose ranks wait for one count flush

// One wavefront adds the tile's LDS cell counts to the global counts (the
// returned bases go to lbase), clears them, and adds their total to the
// tile's scan chunk (the tile's keys all lie in one chunk: kScanChunk is a
// multiple of 64).
__device__ __forceinline__ void tile_counts_to_global(const G2PArgs& a, int* lcnt, int* lbase, int key0) {
    if (threadIdx.x < kTileCells) {
        const int c = lcnt[threadIdx.x];
        lbase[threadIdx.x] = c ? atomicAdd(&a.count[key0 + threadIdx.x], c) : 0;
        lcnt[threadIdx.x] = 0;
        int tot = c;
#pragma unroll
        for (int off = kTileCells / 2; off > 0; off >>= 1) tot += __shfl_xor(tot, off, kTileCells);
        if (threadIdx.x == 0 && tot) atomicAdd(&a.chunks[key0 / kScanChunk], tot);
    }
}

// G2P per tile with the tile's grid velocities staged in LDS; F_trial = (I + dt G) F
// is evaluated as F + dt (G F).  The particles then count themselves into the
// next substep's bins: a particle that stays in this tile takes its rank from an
// LDS counter, and each cell of the tile adds its LDS count to the global count
// with one atomic, whose return value is the base of those ranks: once after the
// block's last round when it has at most kDeferRounds rounds (the ranks wait in
// LDS), else after every round.  Particles that leave the tile count globally.
// The ranks only place ids inside their bin; the bin is sorted by id afterwards.
// The first round's id and position are loaded before the grid staging.
__global__ void __launch_bounds__(kBlock) g2p_tile_kernel(G2PArgs a) {
    __shared__ float gvs[3][kTileNodes];
    __shared__ int lcnt[kTileCells];
    __shared__ int lbase[kTileCells];
    __shared__ int lslot[kDeferRounds][kBlock];  // (local cell << 16) | rank in the LDS count, or -1
    __shared__ int lid[kDeferRounds][kBlock];
    const mpmx_particles& P = a.P;
    const long n = P.n;
    // a.segs blocks share a tile, block h of them taking rounds h, h + segs, ...
    const int b = tile_of_block(a.xcd, (int)blockIdx.x, (int)gridDim.x);
    const int t = b / a.segs, h = b - t * a.segs;
    const long ntiles = a.geo.ntiles();
    const int orphan_key = (int)(ntiles * kTileCells);
    if (t == ntiles) {
        // out-of-grid particles: no transfer (the reference would read out of bounds); they stay orphans
        const long b = a.cell_start[orphan_key], e = a.cell_start[orphan_key + 1];
        for (long s = b + threadIdx.x; s < e; s += kBlock) {
            const int id = a.sorted[s];
            bin_next(a, id, P.materialLabel[id] <= 1 ? orphan_key : -1);
        }
        return;
    }
    const long p0 = a.cell_start[(long)t * kTileCells] + (long)h * kBlock;
    const long p1 = a.cell_start[(long)t * kTileCells + kTileCells];
    if (p1 <= p0) return;
    const SceneDesc sd = a.geo.at(a.geo.scene_of_tile(t));
    const int lt = t - sd.tile0;
    const int tz = lt % sd.ntz, ty = (lt / sd.ntz) % sd.nty, tx = lt / (sd.ntz * sd.nty);
    const int ox = tx * kTB, oy = ty * kTB, oz = tz * kTB;
    const long G = a.geo.nnodes;
    const long step = (long)a.segs * kBlock;
    const int rounds = (int)((p1 - p0 + step - 1) / step);
    // the first round's particle: its loads overlap the grid staging
    int id = -1, label = 2;
    f3 xp{};
    float Fp[9];
    {
        const long s = p0 + threadIdx.x;
        if (s < p1) {
            id = a.sorted[s];
            label = P.materialLabel[id];
            xp = ld3(P.x, n, id);
            ld9(P.F, n, id, Fp);
        }
    }
    for (int node = threadIdx.x; node < kTileNodes; node += kBlock) {
        const int nx = ox + node / (kTN * kTN), ny = oy + (node / kTN) % kTN, nz = oz + node % kTN;
        float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f;
        if (nx < sd.gx && ny < sd.gy && nz < sd.gz) {
            const long gi = sd.node0 + ((long)nx * sd.gy + ny) * sd.gz + nz;
            v0 = a.v_out[gi];
            v1 = a.v_out[G + gi];
            v2 = a.v_out[2 * G + gi];
        }
        gvs[0][node] = v0;
        gvs[1][node] = v1;
        gvs[2][node] = v2;
    }
    if (threadIdx.x < kTileCells) lcnt[threadIdx.x] = 0;
    __syncthreads();
    const int key0 = t * kTileCells;
    // up to kDeferRounds rounds, the tile's cells add their LDS counts to the
    // global counts once, after the last round; longer tiles do it every round
    const bool defer = rounds <= kDeferRounds;
    for (int r = 0; r < rounds; r++) {  // uniform trip count: the barriers below are block-wide
        const long s = p0 + r * step + threadIdx.x;
        if (r > 0) {
            id = -1;
            if (s < p1) {
                id = a.sorted[s];
                label = P.materialLabel[id];
                xp = ld3(P.x, n, id);
                ld9(P.F, n, id, Fp);
            }
        }
        int key = -1, lrank = -1;
        if (id >= 0) {
            if (label <= 1) {  // else it became XPBD in this substep's constitutive update
                float fx[3];
                const int b0 = base_axis(xp.x, sd.minx, a.inv_dx, fx[0]) - ox;
                const int b1 = base_axis(xp.y, sd.miny, a.inv_dx, fx[1]) - oy;
                const int b2 = base_axis(xp.z, sd.minz, a.inv_dx, fx[2]) - oz;
                if (b0 >= 0 && b1 >= 0 && b2 >= 0 && b0 < kTB && b1 < kTB && b2 < kTB) {
                    key = g2p_one<true>(a, sd, gvs, id, xp, Fp, b0, b1, b2, ox, oy, oz, fx);
                } else {
                    // stale bin: the particle is outside its tile's node block
                    if (a.misbinned) atomicAdd(a.misbinned, 1);
                    if (b0 + ox < 0 || b1 + oy < 0 || b2 + oz < 0 || b0 + ox > sd.gx - 3 || b1 + oy > sd.gy - 3 ||
                        b2 + oz > sd.gz - 3)
                        key = orphan_key;  // its stencil leaves the grid: no transfer
                    else
                        key = g2p_one<false>(a, sd, gvs, id, xp, Fp, b0, b1, b2, ox, oy, oz, fx);
                }
            }
            a.keys[id] = key;
            if (key >= key0 && key < key0 + kTileCells) {
                lrank = atomicAdd(&lcnt[key - key0], 1);
            } else if (key >= 0) {
                a.rank[id] = atomicAdd(&a.count[key], 1);
                atomicAdd(&a.chunks[key / kScanChunk], 1);
            }
        }
        if (defer) {
            lslot[r][threadIdx.x] = lrank >= 0 ? ((key - key0) << 16) | lrank : -1;
            lid[r][threadIdx.x] = id;
        } else {
            __syncthreads();
            tile_counts_to_global(a, lcnt, lbase, key0);
            __syncthreads();
            if (lrank >= 0) a.rank[id] = lbase[key - key0] + lrank;
        }
    }
    if (defer) {
        __syncthreads();
        tile_counts_to_global(a, lcnt, lbase, key0);
        __syncthreads();
        for (int r = 0; r < rounds; r++) {
            const int v = lslot[r][threadIdx.x];
            if (v >= 0) a.rank[lid[r][threadIdx.x]] = lbase[v >> 16] + (v & 0xffff);
        }
    }
}

// volumetric locking correction after G2P (simulationRoutines.py:252-293), per
// scene: a scene's cells are its grid's nodes, numbered from its node0
__global__ void vlc_accumulate_substep(mpmx_particles P, Geo geo, float inv_dx, unsigned long long* __restrict__ acc) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n || P.activeLabel[p] != 1 || P.materialLabel[p] > 1) return;
    const SceneDesc d = geo.at(scene_of(P, p));
    long k;
    if (!vlc_cell(ld3(P.x, P.n, p), d.minx, d.miny, d.minz, inv_dx, d.gx, d.gy, d.gz, k)) return;
    k += d.node0;
    const long n = P.n;
    const float div = (P.C[p] + P.C[4 * n + p]) + P.C[8 * n + p];
    atomicAdd(&acc[k], vlc_fixed(div, P.vol[p]));
    atomicAdd(&acc[geo.nnodes + k], vlc_fixed(P.vol[p], 1.0f));
}

__global__ void vlc_apply_substep(mpmx_particles P, Geo geo, float inv_dx, const unsigned long long* __restrict__ acc) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n || P.activeLabel[p] != 1 || P.materialLabel[p] > 1) return;
    const SceneDesc d = geo.at(scene_of(P, p));
    long k;
    if (!vlc_cell(ld3(P.x, P.n, p), d.minx, d.miny, d.minz, inv_dx, d.gx, d.gy, d.gz, k)) return;
    k += d.node0;
    const long n = P.n;
    float* C = P.C;
    const float c00 = C[p], c11 = C[4 * n + p], c22 = C[8 * n + p];
    const float corr = vlc_shift(c00, c11, c22, vlc_unfixed(acc[k]), vlc_unfixed(acc[geo.nnodes + k]));
    C[p] = c00 + corr;
    C[4 * n + p] = c11 + corr;
    C[8 * n + p] = c22 + corr;
}

// G2P over the sorted particle list, one thread per particle, the grid
// velocities read through the caches.  The list is in bin order and most
// particles keep their cell, so consecutive lanes mostly share their next key:
// each run of equal keys takes one atomic for its ranks, and each run of
// equal key chunks one atomic for the chunk totals.
__global__ void __launch_bounds__(kBlock) g2p_particles_kernel(G2PArgs a) {
    const mpmx_particles& P = a.P;
    const long n = P.n;
    const int orphan_key = (int)(a.geo.ntiles() * kTileCells);
    const long total = a.cell_start[orphan_key + 1];  // particles in the bins (tiles and orphans)
    const long s = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int key = -1;
    int id = -1;
    if (s < total) {
        id = a.sorted[s];
        if (P.materialLabel[id] <= 1) {  // else it became XPBD in this substep's constitutive update
            const SceneDesc sd = a.geo.at(scene_of(P, id));
            const f3 xp = ld3(P.x, n, id);
            float fx[3];
            const int b0 = base_axis(xp.x, sd.minx, a.inv_dx, fx[0]);
            const int b1 = base_axis(xp.y, sd.miny, a.inv_dx, fx[1]);
            const int b2 = base_axis(xp.z, sd.minz, a.inv_dx, fx[2]);
            const bool out = b0 < 0 || b1 < 0 || b2 < 0 || b0 > sd.gx - 3 || b1 > sd.gy - 3 || b2 > sd.gz - 3;
            if (a.misbinned) {
                // the bins come from the key written last time: is the particle still there?
                const int now = out ? orphan_key : tile_key(xp, a.inv_dx, sd);
                if (now != a.keys[id]) atomicAdd(a.misbinned, 1);
            }
            if (out) {
                key = orphan_key;
            } else {
                float Fp[9];
                ld9(P.F, n, id, Fp);
                key = g2p_one<false>(a, sd, nullptr, id, xp, Fp, b0, b1, b2, 0, 0, 0, fx);
            }
        }
        a.keys[id] = key;
    }
    // every lane stays for the cross-lane operations
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
    {
        const int prev = __shfl_up(key, 1);
        const bool head = lane == 0 || prev != key;
        const unsigned long long heads = __ballot(head);
        const int my_head = 63 - __clzll((long long)(heads & upto));
        const unsigned long long above = heads & ~upto;
        const int next_head = above ? __ffsll((long long)above) - 1 : 64;
        int base = 0;
        if (head && key >= 0) base = atomicAdd(&a.count[key], next_head - lane);
        base = __shfl(base, my_head);
        if (key >= 0) a.rank[id] = base + (lane - my_head);
    }
    {
        const int ch = key >= 0 ? key / kScanChunk : -1;
        const int prev = __shfl_up(ch, 1);
        const bool head = lane == 0 || prev != ch;
        const unsigned long long heads = __ballot(head);
        const unsigned long long above = heads & ~upto;
        const int next_head = above ? __ffsll((long long)above) - 1 : 64;
        if (head && ch >= 0) atomicAdd(&a.chunks[ch], next_head - lane);
    }
}

// the constitutive update of every active MPM particle (compute_stress_from_F_trial,
// mpmRoutines.py:229-312) when the substep runs it as its own kernel
__global__ void __launch_bounds__(kBlock) substep_stress_kernel(mpmx_particles P, float eff) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n || P.materialLabel[p] > 1 || P.activeLabel[p] != 1) return;
    float T[9];
    constitutive_update(P, p, eff, T);
}

static int check_particles(const mpmx_particles* P, const mpmx_step_params* prm) {
    MPMX_REQUIRE(P && prm, "substep: null argument");
    MPMX_REQUIRE(P->n >= 0, "substep: bad sizes");
    MPMX_REQUIRE(P->n == 0 || (P->activeLabel && P->materialLabel && P->x && P->v && P->x_init_xpbd &&
                               P->v_init_xpbd && P->F && P->F_trial && P->C && P->stress && P->mu && P->lam &&
                               P->ys && P->alpha && P->hardening && P->softening && P->density &&
                               P->strainCriteria && P->eta_shear && P->eta_bulk && P->acc_strain && P->damage &&
                               P->vol && P->mass && (!prm->update_cov || P->cov)),
                 "substep: null particle array");
    MPMX_REQUIRE(prm->bc >= MPMX_BC_NONE && prm->bc <= MPMX_BC_PLANE_STRAIN_Y_GLOBAL,
                 "substep: unsupported boundary condition %d", prm->bc);
    return 0;
}

static int check_args(const mpmx_particles* P, const mpmx_grid* G, const mpmx_step_params* prm) {
    MPMX_REQUIRE(P && G && prm, "substep: null argument");
    MPMX_REQUIRE(G->gx >= 3 && G->gy >= 3 && G->gz >= 3, "substep: bad sizes");
    MPMX_REQUIRE((long)G->gx * G->gy * G->gz < (1L << 31), "substep: a grid needs fewer than 2^31 nodes");
    MPMX_REQUIRE(G->m && G->v_in && G->v_out, "substep: null grid array");
    return check_particles(P, prm);
}

// Geometry of an ensemble from its scene list: tiles and nodes numbered scene
// after scene; the device table goes into the workspace (w.table).
static int ensemble_geo(const mpmx_scene* scenes, int nscenes, Geo& g) {
    MPMX_REQUIRE(scenes && nscenes >= 1, "ensemble: no scenes");
    long tiles = 0, nodes = 0;
    for (int k = 0; k < nscenes; k++) {
        const mpmx_scene& c = scenes[k];
        MPMX_REQUIRE(c.gx >= 3 && c.gy >= 3 && c.gz >= 3, "ensemble: scene %d has fewer than 3 nodes on an axis", k);
        MPMX_REQUIRE((long)c.gx * c.gy * c.gz < (1L << 31), "ensemble: scene %d has 2^31 nodes or more", k);
        const SceneDesc d = make_scene(c.gx, c.gy, c.gz, c.minx, c.miny, c.minz, (int)tiles, nodes);
        tiles += (long)d.ntx * d.nty * d.ntz;
        nodes += (long)c.gx * c.gy * c.gz;
        if (k == 0) g.one = d;
    }
    MPMX_REQUIRE(tiles * kTileCells + 2 < (1L << 31), "ensemble: too many tiles for int32 keys");
    g.table = nullptr;
    g.nscenes = nscenes;
    g.ntiles_ = (int)tiles;
    g.nnodes = nodes;
    return 0;
}

// write the scene table (nscenes + 1 entries, the last with the totals) from kernel arguments
static int write_scene_table(const mpmx_scene* scenes, int nscenes, const Geo& g, SceneDesc* dst, hipStream_t s) {
    SceneChunk c;
    int tile0 = 0;
    long node0 = 0;
    for (int base = 0; base <= nscenes; base += kSceneChunk) {
        const int cnt = nscenes + 1 - base < kSceneChunk ? nscenes + 1 - base : kSceneChunk;
        for (int k = 0; k < cnt; k++) {
            const int sidx = base + k;
            if (sidx < nscenes) {
                const mpmx_scene& sc = scenes[sidx];
                c.d[k] = make_scene(sc.gx, sc.gy, sc.gz, sc.minx, sc.miny, sc.minz, tile0, node0);
                tile0 += c.d[k].ntx * c.d[k].nty * c.d[k].ntz;
                node0 += (long)sc.gx * sc.gy * sc.gz;
            } else {
                c.d[k] = SceneDesc{};
                c.d[k].tile0 = (int)g.ntiles();
                c.d[k].node0 = g.nnodes;
            }
        }
        MPMX_LAUNCH(write_scenes_kernel, dim3(1), dim3(kSceneChunk), 0, s, c, cnt, dst + base);
        int rc = check_launch("write_scenes", s);
        if (rc) return rc;
    }
    return 0;
}

static int substep_bin_geo(const mpmx_particles* P, const Geo& geo, float inv_dx, SubstepWork& w, hipStream_t s) {
    int rc;
    if (P->n > 0) {
        MPMX_LAUNCH(substep_key_kernel, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, inv_dx, geo, w.keys);
        if ((rc = check_launch("substep_key", s))) return rc;
    }
    MPMX_LAUNCH(zero_one_kernel, dim3(1), dim3(1), 0, s, w.misbinned);
    if ((rc = check_launch("substep_clear", s))) return rc;
    return counting_sort_ids(w.keys, P->n, substep_nkeys(geo), w.sorted, w.bin, s, P->order);
}

static int reorder_geo(const mpmx_particles* P, const Geo& geo, float inv_dx, SubstepWork& w, int32_t* perm,
                       hipStream_t s) {
    MPMX_LAUNCH(reorder_key_kernel, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, inv_dx, geo, w.keys);
    int rc = check_launch("reorder_key", s);
    if (rc) return rc;
    return counting_sort_ids(w.keys, P->n, reorder_nkeys(geo, P->n), perm, w.bin, s, P->order);
}

// the learned projection of the constitutive update (mpmx_*_substep_learned)
struct LearnedStress {
    const float* params;
    float tol;
    int64_t* counters;
};

static int substep_geo(const mpmx_particles* P, float* grid_m, float* grid_v_in, float* grid_v_out,
                       const mpmx_step_params* prm, const Geo& geo, SubstepWork& w, hipStream_t s,
                       const LearnedStress* ls = nullptr) {
    int rc;
    const long nkeys = substep_nkeys(geo);
    const unsigned ntile_blocks = (unsigned)geo.ntiles() + 1;  // + orphan bin

    TileArgs ta{};
    ta.P = *P;
    ta.sorted = w.sorted;
    ta.cell_start = w.bin.cell_start;
    ta.dx = prm->dx;
    ta.inv_dx = prm->inv_dx;
    ta.rpic = prm->rpic_damping;
    ta.dt = prm->dt;
    ta.eff = prm->eff;
    ta.partial = w.partial;
    ta.tile_flag = w.tile_flag;
    ta.ntiles = (int)geo.ntiles();
    ta.geo = geo;
    ta.xcd = tuning(MPMX_TUNE_XCD_TILES);
    ta.misbinned = w.misbinned;
    if (ls || tuning(MPMX_TUNE_SPLIT_STRESS)) {  // the learned update runs as its own kernel
        if (ls) {
            if ((rc = launch_stress_learned(*P, prm->eff, ls->params, ls->tol, ls->counters, w.lrm_flag, s)))
                return rc;
        } else if (P->n > 0) {
            MPMX_LAUNCH(substep_stress_kernel, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, prm->eff);
            if ((rc = check_launch("substep_stress", s))) return rc;
        }
        ta.stress = P->stress;
        const dim3 tgrid(ntile_blocks - 1);  // no orphan block
        switch (tuning(MPMX_TUNE_P2G_CHUNK)) {
            case 1: MPMX_LAUNCH((p2g_tile_kernel<true, 384>), tgrid, dim3(kBlock), 0, s, ta); break;
            case 2: MPMX_LAUNCH((p2g_tile_kernel<true, 256>), tgrid, dim3(kBlock), 0, s, ta); break;
            default: MPMX_LAUNCH((p2g_tile_kernel<true, 512>), tgrid, dim3(kBlock), 0, s, ta); break;
        }
        if ((rc = check_launch("substep_tile(p2g)", s))) return rc;
    } else {
        MPMX_LAUNCH(p2g_tile_fused_kernel<>, dim3(ntile_blocks), dim3(kBlock), 0, s, ta);
        if ((rc = check_launch("substep_tile(stress+p2g)", s))) return rc;
    }

    CombineArgs ca{};
    ca.geo = geo;
    ca.partial = w.partial;
    ca.tile_flag = w.tile_flag;
    ca.grid_m = grid_m;
    ca.grid_v_in = grid_v_in;
    ca.grid_v_out = grid_v_out;
    ca.do_update = 1;
    ca.gxv = prm->gravity_x;
    ca.gyv = prm->gravity_y;
    ca.gzv = prm->gravity_z;
    ca.dt = prm->dt;
    ca.damping = prm->damping_scale;
    ca.fric = prm->friction;
    ca.gmag = prm->gravity_magnitude;
    ca.bc = prm->bc;
    ca.pad = prm->boundary_padding;
    ca.restitution = prm->restitution;
    ca.zero_ints = w.bin.count;
    ca.n_zero = nkeys;
    ca.zero_chunks = w.bin.partials;
    ca.n_chunks = (nkeys + kScanChunk - 1) / kScanChunk;
    if (prm->vlc) {
        ca.zero_u64 = w.vlc;
        ca.n_zero_u64 = 2 * geo.nnodes;
    }
    const long Gn = geo.nnodes;
    MPMX_LAUNCH(p2g_combine_kernel<>, dim3(grid_for(Gn > nkeys ? Gn : nkeys, kBlock)), dim3(kBlock), 0, s, ca);
    if ((rc = check_launch("substep_combine(grid)", s))) return rc;

    G2PArgs ga{};
    ga.P = *P;
    ga.sorted = w.sorted;
    ga.cell_start = w.bin.cell_start;
    ga.v_out = grid_v_out;
    ga.geo = geo;
    ga.dt = prm->dt;
    ga.inv_dx = prm->inv_dx;
    ga.update_cov = prm->update_cov;
    ga.keys = w.keys;
    ga.count = w.bin.count;
    ga.rank = w.bin.rank;
    ga.chunks = w.bin.partials;
    ga.misbinned = w.misbinned;
    ga.xcd = tuning(MPMX_TUNE_XCD_TILES);
    ga.segs = tuning(MPMX_TUNE_G2P_TILES);
    if (ga.segs) {
        // segs blocks per tile, then the orphan block
        MPMX_LAUNCH(g2p_tile_kernel, dim3(ga.segs * geo.ntiles() + 1), dim3(kBlock), 0, s, ga);
    } else if (P->n > 0) {
        MPMX_LAUNCH(g2p_particles_kernel, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, ga);
    }
    if ((rc = check_launch("substep_g2p", s))) return rc;

    if (prm->vlc && P->n > 0) {
        MPMX_LAUNCH(vlc_accumulate_substep, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, geo, prm->inv_dx,
                    w.vlc);
        if ((rc = check_launch("substep_vlc_accumulate", s))) return rc;
        MPMX_LAUNCH(vlc_apply_substep, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, geo, prm->inv_dx, w.vlc);
        if ((rc = check_launch("substep_vlc_apply", s))) return rc;
    }

    return counting_sort_from_chunk_counts(w.keys, P->n, nkeys, w.sorted, w.bin, s, P->order);
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

size_t mpmx_substep_workspace_bytes(int n, int gx, int gy, int gz) {
    if (n < 0 || gx < 3 || gy < 3 || gz < 3) return 0;
    return substep_ws_bytes(n, make_geom(gx, gy, gz));
}

int mpmx_substep_bin(const mpmx_particles* P, const mpmx_grid* G, const mpmx_step_params* prm, void* workspace,
                     size_t ws_bytes, void* stream) {
    int rc = check_args(P, G, prm);
    if (rc) return rc;
    Geo geo = make_geom(G->gx, G->gy, G->gz, prm->minx, prm->miny, prm->minz);
    MPMX_REQUIRE(substep_nkeys(geo) < (1L << 31), "substep: grid too large for int32 keys");
    MPMX_REQUIRE(workspace && ws_bytes >= substep_ws_bytes(P->n, geo), "substep: workspace too small");
    SubstepWork w;
    MPMX_REQUIRE(carve_substep(workspace, ws_bytes, P->n, geo, w), "substep: workspace carve failed");
    return substep_bin_geo(P, geo, prm->inv_dx, w, as_stream(stream));
}

int mpmx_reorder_permutation(const mpmx_particles* P, const mpmx_grid* G, const mpmx_step_params* prm, int32_t* perm,
                             void* workspace, size_t ws_bytes, void* stream) {
    int rc = check_args(P, G, prm);
    if (rc) return rc;
    MPMX_REQUIRE(P->n == 0 || perm, "mpmx_reorder_permutation: null perm");
    if (P->n == 0) return 0;
    Geo geo = make_geom(G->gx, G->gy, G->gz, prm->minx, prm->miny, prm->minz);
    MPMX_REQUIRE(reorder_nkeys(geo, P->n) < (1L << 31), "substep: grid too large for int32 keys");
    MPMX_REQUIRE(workspace && ws_bytes >= substep_ws_bytes(P->n, geo), "mpmx_reorder_permutation: workspace too small");
    SubstepWork w;
    MPMX_REQUIRE(carve_substep(workspace, ws_bytes, P->n, geo, w), "mpmx_reorder_permutation: workspace carve failed");
    return reorder_geo(P, geo, prm->inv_dx, w, perm, as_stream(stream));
}

int mpmx_substep_misbinned(void* workspace, size_t ws_bytes, int gx, int gy, int gz, int n, int* host_count,
                           void* stream) {
    MPMX_REQUIRE(host_count && gx >= 3 && gy >= 3 && gz >= 3, "mpmx_substep_misbinned: bad arguments");
    Geo geo = make_geom(gx, gy, gz);
    SubstepWork w;
    MPMX_REQUIRE(workspace && carve_substep(workspace, ws_bytes, n, geo, w), "mpmx_substep_misbinned: workspace");
    MPMX_HIP(hipMemcpyAsync(host_count, w.misbinned, sizeof(int), hipMemcpyDeviceToHost, as_stream(stream)));
    MPMX_HIP(hipStreamSynchronize(as_stream(stream)));
    return 0;
}

int mpmx_mpm_substep(const mpmx_particles* P, const mpmx_grid* G, const mpmx_step_params* prm, void* workspace,
                     size_t ws_bytes, void* stream) {
    int rc = check_args(P, G, prm);
    if (rc) return rc;
    Geo geo = make_geom(G->gx, G->gy, G->gz, prm->minx, prm->miny, prm->minz);
    MPMX_REQUIRE(workspace && ws_bytes >= substep_ws_bytes(P->n, geo), "substep: workspace too small");
    SubstepWork w;
    MPMX_REQUIRE(carve_substep(workspace, ws_bytes, P->n, geo, w), "substep: workspace carve failed");
    return substep_geo(P, G->m, G->v_in, G->v_out, prm, geo, w, as_stream(stream));
}

int mpmx_mpm_substep_learned(const mpmx_particles* P, const mpmx_grid* G, const mpmx_step_params* prm,
                             const float* mlp_params, float tol, int64_t* counters, void* workspace, size_t ws_bytes,
                             void* stream) {
    int rc = check_args(P, G, prm);
    if (rc) return rc;
    MPMX_REQUIRE(tol >= 0.0f, "mpmx_mpm_substep_learned: negative tolerance");
    Geo geo = make_geom(G->gx, G->gy, G->gz, prm->minx, prm->miny, prm->minz);
    MPMX_REQUIRE(workspace && ws_bytes >= substep_ws_bytes(P->n, geo), "substep: workspace too small");
    SubstepWork w;
    MPMX_REQUIRE(carve_substep(workspace, ws_bytes, P->n, geo, w), "substep: workspace carve failed");
    const LearnedStress ls{mlp_params, tol, counters};
    return substep_geo(P, G->m, G->v_in, G->v_out, prm, geo, w, as_stream(stream), &ls);
}

size_t mpmx_ensemble_workspace_bytes(int n, const mpmx_scene* scenes, int nscenes) {
    Geo geo;
    if (n < 0 || ensemble_geo(scenes, nscenes, geo)) return 0;
    return substep_ws_bytes(n, geo);
}

// shared set-up of the ensemble entry points: geometry, workspace, scene table
static int ensemble_setup(const mpmx_particles* P, const mpmx_scene* scenes, int nscenes,
                          const mpmx_step_params* prm, void* workspace, size_t ws_bytes, hipStream_t s, Geo& geo,
                          SubstepWork& w) {
    int rc = check_particles(P, prm);
    if (rc) return rc;
    MPMX_REQUIRE(P->n == 0 || P->scene, "ensemble: particles need a scene array");
    if ((rc = ensemble_geo(scenes, nscenes, geo))) return rc;
    MPMX_REQUIRE(workspace && ws_bytes >= substep_ws_bytes(P->n, geo), "ensemble: workspace too small");
    MPMX_REQUIRE(carve_substep(workspace, ws_bytes, P->n, geo, w), "ensemble: workspace carve failed");
    geo.table = w.table;
    return write_scene_table(scenes, nscenes, geo, w.table, s);
}

int mpmx_ensemble_bin(const mpmx_particles* P, const mpmx_scene* scenes, int nscenes, const mpmx_step_params* prm,
                      void* workspace, size_t ws_bytes, void* stream) {
    Geo geo;
    SubstepWork w;
    hipStream_t s = as_stream(stream);
    int rc = ensemble_setup(P, scenes, nscenes, prm, workspace, ws_bytes, s, geo, w);
    if (rc) return rc;
    return substep_bin_geo(P, geo, prm->inv_dx, w, s);
}

int mpmx_ensemble_reorder_permutation(const mpmx_particles* P, const mpmx_scene* scenes, int nscenes,
                                      const mpmx_step_params* prm, int32_t* perm, void* workspace, size_t ws_bytes,
                                      void* stream) {
    MPMX_REQUIRE(P && (P->n == 0 || perm), "mpmx_ensemble_reorder_permutation: null argument");
    Geo geo;
    SubstepWork w;
    hipStream_t s = as_stream(stream);
    int rc = ensemble_setup(P, scenes, nscenes, prm, workspace, ws_bytes, s, geo, w);
    if (rc) return rc;
    if (P->n == 0) return 0;
    return reorder_geo(P, geo, prm->inv_dx, w, perm, s);
}

int mpmx_ensemble_substep(const mpmx_particles* P, const mpmx_scene* scenes, int nscenes, float* grid_m,
                          float* grid_v_in, float* grid_v_out, const mpmx_step_params* prm, void* workspace,
                          size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(grid_m && grid_v_in && grid_v_out, "ensemble: null grid array");
    Geo geo;
    SubstepWork w;
    hipStream_t s = as_stream(stream);
    int rc = ensemble_setup(P, scenes, nscenes, prm, workspace, ws_bytes, s, geo, w);
    if (rc) return rc;
    return substep_geo(P, grid_m, grid_v_in, grid_v_out, prm, geo, w, s);
}

int mpmx_ensemble_substep_learned(const mpmx_particles* P, const mpmx_scene* scenes, int nscenes, float* grid_m,
                                  float* grid_v_in, float* grid_v_out, const mpmx_step_params* prm,
                                  const float* mlp_params, float tol, int64_t* counters, void* workspace,
                                  size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(grid_m && grid_v_in && grid_v_out, "ensemble: null grid array");
    MPMX_REQUIRE(tol >= 0.0f, "mpmx_ensemble_substep_learned: negative tolerance");
    Geo geo;
    SubstepWork w;
    hipStream_t s = as_stream(stream);
    int rc = ensemble_setup(P, scenes, nscenes, prm, workspace, ws_bytes, s, geo, w);
    if (rc) return rc;
    const LearnedStress ls{mlp_params, tol, counters};
    return substep_geo(P, grid_m, grid_v_in, grid_v_out, prm, geo, w, s, &ls);
}

}  // extern "C"
