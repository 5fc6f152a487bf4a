// XPBD contact step kernels (utils/xpbdRoutines.py) and the spatial hash grid
// that replaces wp.HashGrid (simStates.py:174, simulationRoutines.py:311).
//
// Hash grid: cell coordinates int(p / width) (truncation toward zero), offset by
// 2^20 and wrapped modulo the hash dims per axis, key = cz*dx*dy + cy*dx + cx.
// Points are counting-sorted by key with ids ascending inside a cell (the order
// of a stable radix sort of (key, id) pairs), and thread tid of the contact
// kernel works on point_ids[tid].  A query box visits cells x fastest, then y,
// then z, never visiting a wrapped cell twice.
#include "mpmx_device.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

// wrapped cell coordinates of cell (x, y, z)
__device__ __forceinline__ void hash_cell(int dimx, int dimy, int dimz, int x, int y, int z, int& cx, int& cy,
                                          int& cz) {
    const int origin = 1 << 20;
    x += origin;
    y += origin;
    z += origin;
    // wp.HashGrid wraps x + 2^20 modulo the dims; the extra "+ dim) % dim" only
    // matters for coordinates below -2^20 cells, where it keeps the index in range
    cx = ((x % dimx) + dimx) % dimx;
    cy = ((y % dimy) + dimy) % dimy;
    cz = ((z % dimz) + dimz) % dimz;
}

__device__ __forceinline__ int hash_index(int dimx, int dimy, int dimz, int x, int y, int z) {
    int cx, cy, cz;
    hash_cell(dimx, dimy, dimz, x, y, z, cx, cy, cz);
    return cz * (dimx * dimy) + cy * dimx + cx;
}

// One or two hash tables over the same points, built in one set of launches.
// Counting: points stored in tile order step through z fastest, so the keys of
// neighbouring points lie dimx*dimy apart and their count atomics land on
// scattered cache lines (about 3x slower, measured).  The counts are taken in
// the transposed numbering t = (cx*dimy + cy)*dimz + cz, where those keys are
// adjacent, and hash_counts_transpose_kernel brings them back to key order
// together with the per-chunk totals of the scan.
struct HashTables {
    int n;
    const float* x;
    float inv;
    int ntab;
    // lazy pair build: label_any_kernel sets this word to 1 when the second table
    // gets a point, and the first table is left as it was while it reads 0 (null:
    // not lazy)
    int32_t* label_any;
    int dimx[kMaxSortSegs], dimy[kMaxSortSegs], dimz[kMaxSortSegs];  // per table
    long nkeys[kMaxSortSegs], nchunks[kMaxSortSegs];
    long nblk0;  // transpose blocks of table 0 (table 1's follow)
    const int32_t* labels[kMaxSortSegs];  // table filter (null: every point)
    int label[kMaxSortSegs];
    int32_t* keys[kMaxSortSegs];
    int32_t* rank[kMaxSortSegs];
    int32_t* tcount[kMaxSortSegs];  // [nkeys], transposed numbering
    int32_t* count[kMaxSortSegs];   // [nkeys], key order
    int32_t* totals[kMaxSortSegs];  // [nchunks]
};

// key of each point per table (-1 when the table leaves it out) and its rank
// among the points of its cell: runs of equal keys in a wavefront take one
// atomic, whose return value is the base of the run's ranks
__global__ void __launch_bounds__(kBlock) hash_count_kernel(HashTables h) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int ux = 0, uy = 0, uz = 0;  // the point's cell before wrapping
    if (i < h.n) {
        const long n = h.n;
        ux = (int)(h.x[i] * h.inv);
        uy = (int)(h.x[n + i] * h.inv);
        uz = (int)(h.x[2 * n + i] * h.inv);
    }
    const int lane = (int)(threadIdx.x & 63);
    const unsigned long long upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);  // lanes 0..lane
#pragma unroll
    for (int g = 0; g < kMaxSortSegs; g++) {  // every lane stays for the cross-lane operations
        if (g >= h.ntab) break;
        if (g == 0 && h.label_any && *h.label_any == 0) continue;  // lazy: the same word for every lane
        const int dx = h.dimx[g], dy = h.dimy[g], dz = h.dimz[g];
        int cx, cy, cz;
        hash_cell(dx, dy, dz, ux, uy, uz, cx, cy, cz);
        const bool in = i < h.n && (h.labels[g] == nullptr || h.labels[g][i] == h.label[g]);
        const int k = in ? (cx * dy + cy) * dz + cz : -1;
        if (i < h.n) h.keys[g][i] = in ? cz * (dx * dy) + cy * dx + cx : -1;
        const int prev = __shfl_up(k, 1);
        const bool head = lane == 0 || prev != k;
        const unsigned long long heads = __ballot(head);
        const int my_head = 63 - __clzll((long long)(heads & upto));
        const unsigned long long above = heads & ~upto;
        const int next_head = above ? __ffsll((long long)above) - 1 : 64;
        int base = 0;
        if (head && k >= 0) base = atomicAdd(&h.tcount[g][k], next_head - lane);
        base = __shfl(base, my_head);
        if (in) h.rank[g][i] = base + (lane - my_head);
    }
}

// *flag = 1 when some labels[i] == label (the word is zeroed before; kernels after
// this one read it): one check per block that finds one, and a store unless the
// check already reads 1.  Every writer writes 1, so their order is moot, and the
// accesses need no coherence beyond the kernel's end: workgroup-scope relaxed
// atomics are plain cached loads and stores (agent scope bypassed the caches,
// and thousands of blocks then queued on the one word in memory).
__global__ void __launch_bounds__(kBlock) label_any_kernel(int n, const int32_t* __restrict__ labels, int label,
                                                           int32_t* flag) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int any = __syncthreads_or(i < n && labels[i] == label);
    if (threadIdx.x == 0 && any && __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
        __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// dst[0..n) = 0 unless *gate reads 0
__global__ void fill_zero_gated_kernel(int32_t* __restrict__ dst, long n, const int32_t* __restrict__ gate) {
    if (*gate == 0) return;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) dst[i] = 0;
}

// tcount -> count, per table and y plane, through a 64 x 64 (x, z) LDS tile;
// each wavefront then adds a 64-key row of the output to the totals of the
// (at most two) scan chunks the row touches
__global__ void __launch_bounds__(kBlock) hash_counts_transpose_kernel(HashTables h) {
    __shared__ int tile[64][65];
    long b = blockIdx.x;
    const int g = b < h.nblk0 ? 0 : 1;
    if (g == 1) b -= h.nblk0;
    if (g == 0 && h.label_any && *h.label_any == 0) return;  // lazy: the first table stays as it was
    const int dimx = h.dimx[g], dimy = h.dimy[g], dimz = h.dimz[g];
    const int nbx = (dimx + 63) / 64, nbz = (dimz + 63) / 64;
    const int bx = (int)(b % nbx);
    b /= nbx;
    const int bz = (int)(b % nbz);
    const int cy = (int)(b / nbz);
    const int cx0 = bx * 64, cz0 = bz * 64;
    const int lane = (int)(threadIdx.x & 63), w = (int)(threadIdx.x >> 6);
    const int32_t* tc = h.tcount[g];
    for (int r = w; r < 64; r += kBlock / 64) {
        const int cx = cx0 + r, cz = cz0 + lane;
        tile[r][lane] = (cx < dimx && cz < dimz) ? tc[((long)cx * dimy + cy) * dimz + cz] : 0;
    }
    __syncthreads();
    int32_t* out = h.count[g];
    int32_t* tot = h.totals[g];
    for (int r = w; r < 64; r += kBlock / 64) {
        const int cz = cz0 + r;
        if (cz >= dimz) break;  // uniform over the wavefront
        const int cx = cx0 + lane;
        const bool valid = cx < dimx;
        const int v = valid ? tile[lane][r] : 0;
        const long row = ((long)cz * dimy + cy) * dimx;
        if (valid) out[row + cx] = v;
        const long c0 = (row + cx0) / kScanChunk;
        const bool second = valid && (row + cx) / kScanChunk != c0;
        int a = second ? 0 : v, s2 = second ? v : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            a += __shfl_xor(a, off);
            s2 += __shfl_xor(s2, off);
        }
        if (lane == 0) {
            if (a) atomicAdd(&tot[c0], a);
            if (s2) atomicAdd(&tot[c0 + 1], s2);
        }
    }
}

// reset_mpm_contact_labels (xpbdRoutines.py:3-21)
__global__ void reset_labels_kernel(int n, const int32_t* __restrict__ active, int32_t* __restrict__ mat) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (active[p] == 1 && mat[p] == 0) mat[p] = 1;
}

// integrateParticlesXPBD (xpbdRoutines.py:23-48)
__global__ void integrate_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                                 const float* x, const float* v, float gx, float gy, float gz, float dt, float* x_new,
                                 float* v_new, float v_max) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (active[p] != 1 || mat[p] != 2) return;
    f3 x0 = ld3(x, n, p), v0 = ld3(v, n, p);
    f3 v1 = f3{v0.x + gx * dt, v0.y + gy * dt, v0.z + gz * dt};
    float m = len3(v1);
    if (m > v_max) {
        float s = v_max / m;
        v1 = f3{v1.x * s, v1.y * s, v1.z * s};
    }
    f3 x1 = f3{x0.x + v1.x * dt, x0.y + v1.y * dt, x0.z + v1.z * dt};
    st3(x_new, n, p, x1);
    st3(v_new, n, p, v1);
}

// my_solve_particle_bound_contacts (xpbdRoutines.py:49-136)
__global__ void bound_contacts_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                                      const float* __restrict__ x, const float* __restrict__ v,
                                      const float* __restrict__ mass, const float* __restrict__ radius, float mu,
                                      float thr, float smu, f3 bmin, f3 bmax, float dt, float relax,
                                      float* __restrict__ delta) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (active[t] != 1 || mat[t] != 2 || mass[t] == 0.0f) return;
    float wi = 1.0f / mass[t];
    f3 xp = ld3(x, n, t), vp = ld3(v, n, t);
    float r = radius[t];
    f3 d = ld3(delta, n, t);
    const float xa[3] = {xp.x, xp.y, xp.z}, va[3] = {vp.x, vp.y, vp.z};
    const float lo[3] = {bmin.x, bmin.y, bmin.z}, hi[3] = {bmax.x, bmax.y, bmax.z};
    float da[3] = {d.x, d.y, d.z};
#pragma unroll
    for (int ax = 0; ax < 3; ax++) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
            float sg = side == 0 ? 1.0f : -1.0f;
            float b = side == 0 ? lo[ax] : hi[ax];
            float c = fminf(sg * (xa[ax] - b) - r, 0.0f);
            if (c < 0.0f) {
                float lambda_n = c;
                float vn = sg * va[ax];
                float fmu = (fabsf(vn) < thr * r) ? smu : mu;
                float vt[3] = {va[0], va[1], va[2]};
                vt[ax] = va[ax] - sg * vn;
                float vtl = len3(f3{vt[0], vt[1], vt[2]});
                float a_ = fmu * lambda_n, b_ = -vtl * dt;
                float lf = a_ > b_ ? a_ : b_;  // wp.max
                float df[3] = {0.0f, 0.0f, 0.0f};
                if (vtl > 1.0e-8f) {
                    f3 nt = normalize3(f3{vt[0], vt[1], vt[2]});
                    df[0] = nt.x * lf; df[1] = nt.y * lf; df[2] = nt.z * lf;
                }
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    float dn = (a == ax) ? sg * lambda_n : 0.0f;
                    da[a] += (df[a] - dn) / wi * relax * wi;
                }
            }
        }
    }
    st3(delta, n, t, f3{da[0], da[1], da[2]});
}

// Candidate records of the contact walk in hash-table order: record s of a
// table describes the particle at position s of that table, so the candidates
// of a cell are contiguous records, loaded without loading their ids first and
// whatever order the particle arrays are in.  Records [0, n) follow the full
// table, [n, n + m) the m points of the XPBD-only table.  The flags hold what
// the walk tests of a candidate's labels: lock writes turn label 1 into 0 only,
// so "label <= 1" and "label == 2" cannot change while a contact solve runs,
// and records gathered before it give the tests of the live labels.
struct ContactPack {
    const float4* xr;       // x, y, z, radius
    const float4* vw;       // vx, vy, vz, 1 / mass
    const int32_t* flags;   // kCand* bits
    const uint64_t* near;   // optional: bit c = a cell of the XPBD-only table within one cell of c holds a point
    const int32_t* list;        // optional: the table positions to solve (mpmx_xpbd_contact_list)
    const int32_t* list_count;  // their number
};

// Neighbourhood bitmap of a hash table, one bit per cell: set when the cell or
// one of its 26 neighbours (the table's wrapped cell coordinates) holds a point.
// occ first gets the occupancy bits; both hold cells / 64 rounded up words.
__global__ void __launch_bounds__(kBlock) table_occupancy_kernel(const int32_t* __restrict__ cell_start, long ncell,
                                                                 uint64_t* __restrict__ occ) {
    if (cell_start[ncell] == 0) return;  // empty table: the contact solve reads no bitmap
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const bool full = c < ncell && cell_start[c + 1] > cell_start[c];
    const uint64_t m = __ballot(full);
    if ((threadIdx.x & 63) == 0 && c < ncell) occ[c >> 6] = m;
}

__global__ void __launch_bounds__(kBlock) table_near_kernel(const int32_t* __restrict__ cell_start, int dimx, int dimy,
                                                            int dimz, const uint64_t* __restrict__ occ,
                                                            uint64_t* __restrict__ near) {
    const long ncell = (long)dimx * dimy * dimz;
    if (cell_start[ncell] == 0) return;
    const long c = (long)blockIdx.x * blockDim.x + threadIdx.x;
    bool any = false;
    if (c < ncell) {
        const int cx = (int)(c % dimx), cy = (int)((c / dimx) % dimy), cz = (int)(c / ((long)dimx * dimy));
        // the wrapped neighbour coordinates, as hash_index wraps cell coordinates
        const int xs[3] = {cx == 0 ? dimx - 1 : cx - 1, cx, cx + 1 == dimx ? 0 : cx + 1};
        const int ys[3] = {cy == 0 ? dimy - 1 : cy - 1, cy, cy + 1 == dimy ? 0 : cy + 1};
        const int zs[3] = {cz == 0 ? dimz - 1 : cz - 1, cz, cz + 1 == dimz ? 0 : cz + 1};
        for (int a = 0; a < 3 && !any; a++)
            for (int b = 0; b < 3 && !any; b++)
                for (int d = 0; d < 3 && !any; d++) {
                    const long q = ((long)zs[a] * dimy + ys[b]) * dimx + xs[d];
                    any = (occ[q >> 6] >> (q & 63)) & 1ull;
                }
    }
    const uint64_t m = __ballot(any);
    if ((threadIdx.x & 63) == 0 && c < ncell) near[c >> 6] = m;
}
constexpr int kCandValid = 1;  // active == 1 and mass > 0
constexpr int kCandMpm = 2;    // label <= 1
constexpr int kCandXpbd = 4;   // label == 2

// thread t < n: record t of the full table; n <= t < 2n: record t of the
// XPBD-only table; nothing while that table is empty (the walk then stops at
// once).  From particle-order records (pk, when given: three loads per record)
// or from the particle arrays.
__global__ void contact_pack_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                                    const float* __restrict__ mass, const float* __restrict__ radius,
                                    const float* __restrict__ x, const float* __restrict__ v, ContactPack pk,
                                    const int32_t* __restrict__ point_ids, const int32_t* __restrict__ xp_ids,
                                    const int32_t* __restrict__ xp_count, float4* __restrict__ xr,
                                    float4* __restrict__ vw, int32_t* __restrict__ flags) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int m = *xp_count;
    if (m == 0 || t >= (long)n + m) return;
    const int j = t < n ? point_ids[t] : xp_ids[t - n];
    if (j < 0 || j >= n) {
        flags[t] = 0;
        return;
    }
    if (pk.xr) {
        const int f = pk.flags[j];
        flags[t] = f;
        if (!(f & kCandValid)) return;
        xr[t] = pk.xr[j];
        vw[t] = pk.vw[j];
        return;
    }
    const float pm = mass[j];
    const int lab = mat[j];
    const bool valid = active[j] == 1 && pm > 0.0f;
    flags[t] = (valid ? kCandValid : 0) | (lab <= 1 ? kCandMpm : 0) | (lab == 2 ? kCandXpbd : 0);
    if (!valid) return;  // the walk reads no more of an invalid candidate
    const f3 xp = ld3(x, n, j), vp = ld3(v, n, j);
    xr[t] = make_float4(xp.x, xp.y, xp.z, radius[j]);
    vw[t] = make_float4(vp.x, vp.y, vp.z, 1.0f / pm);
}

// The contact walk runs in two passes over each lane's candidates.  The scan
// keeps the candidates within a margin of contact reach, queued per lane in LDS
// (the unpacked walk applies the reference's validity and label filters there
// too, the packed one leaves them to the drain); the drain takes them in the
// same order, makes the filters and the exact contact test and solves the
// contacts.  The lanes of a wave
// scan in step, and the drain's divisions run for near candidates only (solved
// inline, they ran for the whole wave whenever one of its 64 lanes had a contact).
constexpr int kContactQueue = 24;  // queued candidates per lane
constexpr int kScanBatch = 4;  // candidates whose records one lane loads together in the scan (8: slower)
constexpr int kContactFlush = 12;  // a wave drains once one of its lanes holds this many

// False only where the exact test d - ri - rj <= k_coh fails: the squared distance
// d2 exceeds (1.001 R)^2, R = |ri| + |rj| + |k_coh|, so d exceeds R by far more
// than the rounding of that test (a few float ulps of R); the 1e-30 floor keeps
// the bound where the square of a tiny R would underflow.  NaN keeps a candidate.
__device__ __forceinline__ bool maybe_in_reach(float d2, float base, float rj) {
    const float r = 1.001f * (base + fabsf(rj));
    return !(d2 > fmaxf(r * r, 1.0e-30f));
}

// The table positions whose contact walk can find a contact in this XPBD step:
// every position the contact solve does not return from at once, except MPM
// particles that walk the XPBD-only table from a cell the neighbourhood bitmap
// shows to be beside no XPBD point.  Those find no candidate in any iteration:
// the table is fixed during the step, and such a particle gets no delta, so its
// position (and query box) stays where it is.  Each wave appends its kept
// positions with one atomic; the list order is free, because every position
// writes only its own particle's delta and the lock writes only turn label 1
// into 0, which no test of the solve distinguishes.
__global__ void __launch_bounds__(kBlock) contact_list_kernel(
    int n, const int32_t* __restrict__ mat, const int32_t* __restrict__ point_pos, const int32_t* __restrict__ xp_start,
    float inv, int xdimx, int xdimy, int xdimz, const float* __restrict__ x, const float* __restrict__ radius,
    float k_coh, float max_radius, const uint64_t* __restrict__ near, int32_t* __restrict__ list,
    int32_t* __restrict__ count) {
    // one thread per particle in array order (coalesced); the list holds table positions
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxpbd = xp_start[(long)xdimx * xdimy * xdimz];
    if (nxpbd > (n >> 4)) {
        // every XPBD point walks, so the list would be long and the solve walks every
        // table position instead: no list is written, and a count above n says so
        if (i == 0) *count = n + 1;
        return;
    }
    bool keep = false;
    if (i < n && nxpbd > 0) {
        keep = true;
        if (mat[i] <= 1) {  // an MPM particle: does its walk meet an XPBD point?
            const f3 xi = ld3(x, n, i);
            const float ri = radius[i];
            const float qr = ri + max_radius + k_coh * ri;
            const int xs = (int)((xi.x - qr) * inv), ys = (int)((xi.y - qr) * inv), zs = (int)((xi.z - qr) * inv);
            const int xe = min((int)((xi.x + qr) * inv), xs + xdimx - 1);
            const int ye = min((int)((xi.y + qr) * inv), ys + xdimy - 1);
            const int ze = min((int)((xi.z + qr) * inv), zs + xdimz - 1);
            // (the solve's own test for walking the XPBD-only table, on the same box)
            const float reach = (ri + max_radius + k_coh) * inv;
            const int gapx = xdimx - (xe - xs + 1) - 1, gapy = xdimy - (ye - ys + 1) - 1;
            const int gapz = xdimz - (ze - zs + 1) - 1;
            const int cx = (int)(xi.x * inv), cy = (int)(xi.y * inv), cz = (int)(xi.z * inv);
            if ((float)gapx > reach && (float)gapy > reach && (float)gapz > reach && xs >= cx - 1 && xe <= cx + 1 &&
                ys >= cy - 1 && ye <= cy + 1 && zs >= cz - 1 && ze <= cz + 1) {
                const int c = hash_index(xdimx, xdimy, xdimz, cx, cy, cz);
                keep = (near[c >> 6] >> (c & 63)) & 1ull;
            }
        }
    }
    const uint64_t m = __ballot(keep);
    if (m == 0ull) return;
    const int lane = threadIdx.x & 63;
    const int first = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, first);
    if (keep) list[base + __popcll(m & ((1ull << lane) - 1ull))] = point_pos[i];
}

// my_solve_particle_particle_contacts (xpbdRoutines.py:138-219).
// With xp_ids/xp_start (a table of the label-2 points only, same hashing and
// stable order) an MPM particle i walks that table instead of the full one: it
// skips every MPM candidate anyway (:185-186), and the XPBD candidates come in
// the same relative order, so deltas and lock writes are unchanged.  PACKED
// reads the candidates from ContactPack records instead of the particle arrays.
// the contact walk of table position tid (one per thread); queue is the
// workgroup's LDS queue of candidate positions, [kContactQueue][kBlock]
template <bool PACKED>
__device__ __forceinline__ void contact_one(
    int tid, int32_t (*queue)[kBlock], int n, const int32_t* __restrict__ active, int32_t* mat,
    const int32_t* __restrict__ point_ids, const int32_t* __restrict__ cell_start, const int32_t* __restrict__ xp_ids,
    const int32_t* __restrict__ xp_start, float inv, int dimx, int dimy, int dimz, int xdimx, int xdimy, int xdimz,
    const float* __restrict__ x, const float* __restrict__ v, const float* __restrict__ mass,
    const float* __restrict__ radius, float k_mu_in, float thr, float smu, float k_coh, float max_radius, float dt,
    float relax, int lock, float* __restrict__ deltas, const int32_t* __restrict__ slot_of_id, const ContactPack& pk) {
    // with no XPBD particle at all there is no pair to solve (MPM-MPM pairs are skipped)
    if (xp_ids && xp_start[(long)xdimx * xdimy * xdimz] == 0) return;
    // The reference tests the labels and the mass of array index tid, the thread's
    // position in the hash order, not of the particle i it solves
    // (xpbdRoutines.py:151-175).  After a spatial permutation of the arrays that
    // index means the particle whose input index is tid.
    const int q = slot_of_id ? slot_of_id[tid] : tid;
    if (active[q] != 1 || mat[q] < 0) return;
    int i = point_ids[tid];
    if (i == -1) return;
    f3 xi = ld3(x, n, i), vi = ld3(v, n, i);
    float ri = radius[i];
    if (mass[q] == 0.0f) return;
    float w1 = 1.0f / mass[i];
    float qr = ri + max_radius + k_coh * ri;
    int xs = (int)((xi.x - qr) * inv), ys = (int)((xi.y - qr) * inv), zs = (int)((xi.z - qr) * inv);
    int xe = min((int)((xi.x + qr) * inv), xs + dimx - 1);
    int ye = min((int)((xi.y + qr) * inv), ys + dimy - 1);
    int ze = min((int)((xi.z + qr) * inv), zs + dimz - 1);
    const int32_t* ids = point_ids;
    const int32_t* starts = cell_start;
    int hx = dimx, hy = dimy, hz = dimz;
    long rec = 0;    // the walked table's first record
    int self = tid;  // i's position in the walked table (i is not in the XPBD-only one)
    // labels <= 1 and == 2 do not change while this kernel runs (lock turns 1 into 0)
    const int lab_i = mat[i];
    const bool i_mpm = lab_i <= 1, i_xpbd = lab_i == 2;
    // The XPBD table may have fewer cells: a cell there also holds the XPBD points of
    // the cells that wrap onto it.  Those lie more than `gap` cells beyond the query
    // box in some axis, so while gap cells are longer than the contact reach
    // (ri + max_radius + k_coh) they fail the contact test below, and the candidates
    // in reach, and their order, are those of the full table.  Otherwise the
    // particle walks the full table.
    const float reach = (ri + max_radius + k_coh) * inv;  // in cells
    const int gapx = xdimx - (xe - xs + 1) - 1, gapy = xdimy - (ye - ys + 1) - 1, gapz = xdimz - (ze - zs + 1) - 1;
    if (xp_ids && i_mpm && (float)gapx > reach && (float)gapy > reach && (float)gapz > reach) {
        ids = xp_ids;
        starts = xp_start;
        hx = xdimx;
        hy = xdimy;
        hz = xdimz;
        rec = n;
        self = -1;
        if constexpr (PACKED) {
            // A query box within one cell of the particle's own cell, whose cell is not
            // beside any XPBD point, walks only empty cells: no contact, no lock write.
            const int cx = (int)(xi.x * inv), cy = (int)(xi.y * inv), cz = (int)(xi.z * inv);
            if (pk.near && xs >= cx - 1 && xe <= cx + 1 && ys >= cy - 1 && ye <= cy + 1 && zs >= cz - 1 &&
                ze <= cz + 1) {
                const int c = hash_index(hx, hy, hz, cx, cy, cz);
                if (!((pk.near[c >> 6] >> (c & 63)) & 1ull)) return;
            }
        }
    }
    const float base = fabsf(ri) + fabsf(k_coh);
    float k_mu = k_mu_in;
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f;
    int queued = 0;
    // the exact contact test and the contact of every queued candidate, in order
    auto drain = [&]() {
        // PACKED: the three records of a queued candidate load together, one
        // candidate ahead of the one being solved (table positions, all in bounds)
        int next_sq = 0, next_f = 0;
        float4 next_p = {}, next_q = {};
        auto fetch = [&](int k) {
            next_sq = queue[k][threadIdx.x];
            next_f = pk.flags[rec + next_sq];
            next_p = pk.xr[rec + next_sq];
            next_q = pk.vw[rec + next_sq];
        };
        if constexpr (PACKED) {
            if (queued > 0) fetch(0);
        }
        for (int k = 0; k < queued; k++) {
            int sq;
            f3 xj, vj;
            float rj, w2;
            bool j_mpm, j_xpbd;
            int j = -1;
            int fj = 0;
            float4 qj;  // PACKED: (v, 1 / mass)
            if constexpr (PACKED) {
                sq = next_sq;
                fj = next_f;
                const float4 pj = next_p;
                qj = next_q;
                if (k + 1 < queued) fetch(k + 1);
                // the scan's candidate filters: valid, not i itself, no MPM-MPM pair
                if (!(fj & kCandValid) || sq == self || (i_mpm && (fj & kCandMpm))) continue;
                xj = f3{pj.x, pj.y, pj.z};
                rj = pj.w;
            } else {
                sq = queue[k][threadIdx.x];
                j = ids[sq];
                xj = ld3(x, n, j);
                rj = radius[j];
            }
            const float nx = xi.x - xj.x, ny = xi.y - xj.y, nz = xi.z - xj.z;
            const float d = len3(f3{nx, ny, nz});
            const float err = d - ri - rj;
            if (!(err <= k_coh)) continue;
            if constexpr (PACKED) {
                vj = f3{qj.x, qj.y, qj.z};
                w2 = qj.w;
            } else {
                vj = ld3(v, n, j);
                w2 = 1.0f / mass[j];
            }
            const float denom = w1 + w2;
            if (!(denom > 0.0f)) continue;
            if (lock == 1) {
                if constexpr (PACKED) {
                    j_mpm = (fj & kCandMpm) != 0;
                    j_xpbd = (fj & kCandXpbd) != 0;
                } else {
                    const int lab_j = mat[j];
                    j_mpm = lab_j <= 1;
                    j_xpbd = lab_j == 2;
                }
                if (i_mpm && j_xpbd) mat[i] = 0;
                if (i_xpbd && j_mpm) mat[PACKED ? ids[sq] : j] = 0;
            }
            // Divisions by the same value are one reciprocal and products (within an
            // ulp or two of the reference's quotients): the contact's roughly 200
            // instructions held 11 IEEE divisions of about 10 instructions each.
            const float inv_d = 1.0f / d;
            const float ux = nx * inv_d, uy = ny * inv_d, uz = nz * inv_d;
            const float rx = vi.x - vj.x, ry = vi.y - vj.y, rz = vi.z - vj.z;
            // err / (1 + alpha dt^2) with the reference's compliance alpha = 0 is err exactly
            const float lambda_n = err;
            const float vn = ux * rx + uy * ry + uz * rz;
            if (fabsf(vn) < thr * ri) k_mu = smu;
            const float tx = rx - ux * vn, ty = ry - uy * vn, tz = rz - uz * vn;
            const float tl = len3(f3{tx, ty, tz});
            const float a_ = k_mu * lambda_n, b_ = -tl * dt;
            const float lf = a_ > b_ ? a_ : b_;  // wp.max
            const float inv_tl = tl > 0.0f ? 1.0f / tl : 0.0f;  // Warp's normalize: zero stays zero
            const float share = w1 / denom;
            d0 += (tx * inv_tl * lf - ux * lambda_n) * share;
            d1 += (ty * inv_tl * lf - uy * lambda_n) * share;
            d2 += (tz * inv_tl * lf - uz * lambda_n) * share;
        }
        queued = 0;
    };
    if constexpr (PACKED) {
        // The cells of one x-row of the query box are consecutive keys, so their
        // candidates are one contiguous range of the table, or two where the row
        // wraps round the hash dims (the wrapped part follows, as the walk visits
        // it).  A row costs two cell-start loads instead of two per cell, and its
        // candidates are scanned kScanBatch at a time with their records loaded
        // together; the order of the queued candidates is the walk's.  The scan
        // reads only the (x, radius) records: the flag tests and the self test
        // are filters like the reach test, so the drain makes them on the few
        // queued candidates (an invalid candidate's record is stale, and any
        // value there only queues it).
        const float4* __restrict__ xr = pk.xr + rec;
        int wx0, wy0, wz0;
        hash_cell(hx, hy, hz, xs, ys, zs, wx0, wy0, wz0);
        const int wx1 = wx0 + (xe - xs);  // past hx-1 where the row wraps
        const int npiece = wx1 < hx ? 1 : 2;
        for (int cz = zs, wz = wz0; cz <= ze; cz++, wz = wz + 1 == hz ? 0 : wz + 1)
            for (int cy = ys, wy = wy0; cy <= ye; cy++, wy = wy + 1 == hy ? 0 : wy + 1) {
                const int row = (wz * hy + wy) * hx;
                for (int piece = 0; piece < npiece; piece++) {
                    const int lo = piece == 0 ? wx0 : 0;
                    const int hi = piece == 0 ? (npiece == 1 ? wx1 : hx - 1) : wx1 - hx;
                    const int b = starts[row + lo], e = starts[row + hi + 1];
                    for (int s0 = b; s0 < e; s0 += kScanBatch) {
                        float4 pv[kScanBatch];
#pragma unroll
                        for (int u = 0; u < kScanBatch; u++) pv[u] = xr[s0 + u < e ? s0 + u : s0];
                        // every test evaluated for the whole batch (no branch per candidate)
                        unsigned keep = 0;
#pragma unroll
                        for (int u = 0; u < kScanBatch; u++) {
                            const float nx = xi.x - pv[u].x, ny = xi.y - pv[u].y, nz = xi.z - pv[u].z;
                            const bool k = (s0 + u < e) &
                                           maybe_in_reach(fmaf(nz, nz, fmaf(ny, ny, nx * nx)), base, pv[u].w);
                            keep |= (unsigned)k << u;
                        }
                        if (keep) {
                            // room for the whole batch (a lane with many near candidates)
                            if (queued > kContactQueue - kScanBatch) drain();
#pragma unroll
                            for (int u = 0; u < kScanBatch; u++)
                                if (keep & (1u << u)) queue[queued++][threadIdx.x] = s0 + u;
                        }
                    }
                }
                if (__any(queued >= kContactFlush)) drain();
            }
    } else {
        for (int cz = zs; cz <= ze; cz++)
            for (int cy = ys; cy <= ye; cy++)
                for (int cx = xs; cx <= xe; cx++) {
                    const int cell = hash_index(hx, hy, hz, cx, cy, cz);
                    const int b = starts[cell], e = starts[cell + 1];
                    // the next candidate's id loads while this one is tested
                    int j_next = b < e ? ids[b] : 0;
                    for (int s = b; s < e; s++) {
                        const int j = j_next;
                        if (s + 1 < e) j_next = ids[s + 1];
                        if (!(active[j] == 1 && j != i && mass[j] > 0.0f)) continue;
                        if (i_mpm && mat[j] <= 1) continue;
                        const float rj = radius[j];
                        const f3 xj = ld3(x, n, j);
                        const float nx = xi.x - xj.x, ny = xi.y - xj.y, nz = xi.z - xj.z;
                        if (!maybe_in_reach(fmaf(nz, nz, fmaf(ny, ny, nx * nx)), base, rj)) continue;
                        if (queued == kContactQueue) drain();  // a lane with many near candidates
                        queue[queued++][threadIdx.x] = s;
                    }
                    if (__any(queued >= kContactFlush)) drain();
                }
    }
    drain();
    // point_ids is a permutation, so particle i has exactly one writer here
    deltas[i] += d0 * relax;
    deltas[n + i] += d1 * relax;
    deltas[2L * n + i] += d2 * relax;
}

template <bool PACKED>
__global__ void __launch_bounds__(kBlock) particle_contacts_kernel(
    int n, const int32_t* __restrict__ active, int32_t* mat, const int32_t* __restrict__ point_ids,
    const int32_t* __restrict__ cell_start, const int32_t* __restrict__ xp_ids, const int32_t* __restrict__ xp_start,
    float inv, int dimx, int dimy, int dimz, int xdimx, int xdimy, int xdimz, const float* __restrict__ x,
    const float* __restrict__ v, const float* __restrict__ mass, const float* __restrict__ radius, float k_mu_in,
    float thr, float smu, float k_coh, float max_radius, float dt, float relax, int lock, float* __restrict__ deltas,
    const int32_t* __restrict__ slot_of_id, ContactPack pk) {
    __shared__ int32_t queue[kContactQueue][kBlock];
    // one thread per table position, or per listed position while the list is
    // short (it holds the walks that can meet a contact); the positions go to
    // the XCDs in runs of 16 blocks in turn, so the cells a block walks are
    // mostly those of the blocks beside it in the table and share one L2, and a
    // crowded part of the table spreads over every XCD.  A list holding more than
    // an eighth of the particles is not used: its order is that of the particle
    // arrays, and table order walks neighbouring cells from neighbouring lanes
    // (2.0x faster on a 2M-particle scene with 1M XPBD points).  The results are
    // the same either way: each walk writes only its own particle's delta.
    const int tid = xcd_run_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    int pos = tid;
    if (PACKED && pk.list) {
        const int cnt = *pk.list_count;
        if (cnt <= (n >> 3)) {
            if (tid >= cnt) return;
            pos = pk.list[tid];
        }
    }
    if (pos >= n) return;
    contact_one<PACKED>(pos, queue, n, active, mat, point_ids, cell_start, xp_ids, xp_start, inv, dimx, dimy, dimz,
                        xdimx, xdimy, xdimz, x, v, mass, radius, k_mu_in, thr, smu, k_coh, max_radius, dt, relax, lock,
                        deltas, slot_of_id, pk);
}

// my_apply_particle_deltas (xpbdRoutines.py:221-257)
__global__ void apply_deltas_kernel(int n, const int32_t* __restrict__ active, const float* __restrict__ x_orig,
                                    const float* x_pred, const float* __restrict__ delta, float dt, float v_max,
                                    float* x_out, float* v_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (active[t] != 1) return;
    f3 x0 = ld3(x_orig, n, t), xp = ld3(x_pred, n, t), d = ld3(delta, n, t);
    if (isnan(d.x) || isnan(d.y) || isnan(d.z)) d = f3{0.0f, 0.0f, 0.0f};
    f3 xn = f3{xp.x + d.x, xp.y + d.y, xp.z + d.z};
    f3 vn = f3{(xn.x - x0.x) / dt, (xn.y - x0.y) / dt, (xn.z - x0.z) / dt};
    float m = len3(vn);
    if (m > v_max) {
        float s = v_max / m;
        vn = f3{vn.x * s, vn.y * s, vn.z * s};
    }
    st3(x_out, n, t, xn);
    st3(v_out, n, t, vn);
}

// counter += the number of lanes of this wave with pred set, in one atomic (the
// counts are integers, so the total does not depend on the order); every lane
// of the wave that has not returned must reach it
__device__ __forceinline__ void wave_count_add(int32_t* counter, bool pred) {
    const unsigned long long m = __ballot(pred);
    if (m && (int)(threadIdx.x & 63) == __ffsll((long long)m) - 1) atomicAdd(counter, (int)__popcll(m));
}

// sleepParticles (xpbdRoutines.py:313-334)
__global__ void sleep_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat, float thr,
                             const float* __restrict__ radius, const float* __restrict__ x_prev,
                             float* __restrict__ x_after, float* __restrict__ dist, float dt,
                             int32_t* __restrict__ counters) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const bool xp = active[t] == 1 && mat[t] == 2;
    bool slept = false;
    if (xp) {
        f3 a = ld3(x_after, n, t), p = ld3(x_prev, n, t);
        float d = len3(f3{a.x - p.x, a.y - p.y, a.z - p.z});
        if (d / dt < thr * radius[t]) {
            st3(x_after, n, t, p);
            slept = true;
        }
        dist[t] = dist[t] + d;
    }
    wave_count_add(&counters[0], xp);
    wave_count_add(&counters[1], slept);
}

// apply_mpm_contact_impulses (xpbdRoutines.py:259-310)
__global__ void impulse_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                               const float* x_orig, const float* v_orig, float* x_final, float dt, float v_max,
                               float k, int xpbd_only, float* v_out) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (!(active[t] == 1 && mat[t] <= 1)) return;
    f3 xo = ld3(x_orig, n, t);
    st3(x_final, n, t, xo);
    if (xpbd_only == 0) {
        f3 vo = ld3(v_orig, n, t);
        // (x_final - x_orig) is identically zero after the reset above
        f3 vc = f3{vo.x + (xo.x - xo.x) / dt * k, vo.y + (xo.y - xo.y) / dt * k, vo.z + (xo.z - xo.z) / dt * k};
        float m = len3(vc);
        if (m > v_max) {
            float s = v_max / m;
            vc = f3{vc.x * s, vc.y * s, vc.z * s};
        }
        st3(v_out, n, t, vc);
    }
}

// swellParticlesType2 (xpbdRoutines.py:354-376)
__global__ void swell_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                             const float* __restrict__ x_init, const float* __restrict__ x, float* __restrict__ radius,
                             float factor_in, float factor_max_in, const float* __restrict__ swell_r,
                             const float* __restrict__ base_r) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (active[t] != 1 || mat[t] != 2) return;
    f3 a = ld3(x, n, t), b = ld3(x_init, n, t);
    float br = base_r[t];
    float dxl = len3(f3{a.x - b.x, a.y - b.y, a.z - b.z}) * br;
    float factor = factor_in * br;
    float factor_max = factor_max_in * br;
    if (dxl > factor) {
        float nr = br + (swell_r[t] - br) * (dxl - factor) / (factor_max - factor);
        float r0 = radius[t];
        nr = nr > r0 ? nr : r0;
        float sr = swell_r[t];
        nr = nr < sr ? nr : sr;
        radius[t] = nr;
    }
}

}  // namespace mpmx

using namespace mpmx;

#define LAUNCH1(kernel, n, ...)                                                                         \
    MPMX_LAUNCH(kernel, dim3(grid_for((n), kBlock)), dim3(kBlock), 0, as_stream(stream), __VA_ARGS__); \
    return check_launch(#kernel, as_stream(stream));

extern "C" {

static long table_keys(const int* d) { return (long)d[0] * d[1] * d[2]; }
static long table_chunks(const int* d) { return (table_keys(d) + kScanChunk - 1) / kScanChunk; }

// dims[g] = {dimx, dimy, dimz} of table g
static size_t hash_tables_bytes(int ntab, int n, const int (*dims)[3]) {
    if (n < 0) return 0;
    const long nn = n > 0 ? n : 1;
    size_t bytes = 2 * ntab * Carver::bytes_for(nn, 4);
    for (int g = 0; g < ntab; g++) {
        if (dims[g][0] <= 0 || dims[g][1] <= 0 || dims[g][2] <= 0) return 0;
        bytes += Carver::bytes_for(table_keys(dims[g]) + table_chunks(dims[g]) + 1, 4) +
                 Carver::bytes_for(table_keys(dims[g]), 4);
    }
    return bytes;
}

size_t mpmx_hashgrid_workspace_bytes(int n, int dimx, int dimy, int dimz) {
    const int dims[1][3] = {{dimx, dimy, dimz}};
    return hash_tables_bytes(1, n, dims);
}

size_t mpmx_hashgrid_pair_workspace_bytes(int n, int dimx, int dimy, int dimz, int label_dimx, int label_dimy,
                                          int label_dimz) {
    const int dims[2][3] = {{dimx, dimy, dimz}, {label_dimx, label_dimy, label_dimz}};
    return hash_tables_bytes(2, n, dims);
}

// ntab tables: table g has dims[g] and holds the points whose labels[g] entry
// equals label[g] (every point when labels[g] is null), with ids ascending in
// order[id] (in id when order is null) inside each cell
static int hash_tables_build(int ntab, int n, const float* x, float cell_width, const int (*dims)[3],
                             const int32_t* const* labels, const int* label, int32_t* const* point_ids,
                             int32_t* const* cell_start, void* workspace, size_t ws_bytes, void* stream,
                             const int32_t* order, int32_t* const* point_pos = nullptr,
                             const int32_t* pos_gate = nullptr, bool lazy = false) {
    MPMX_REQUIRE(n >= 0 && cell_width > 0.0f, "mpmx_hashgrid_build: bad arguments");
    for (int g = 0; g < ntab; g++) {
        MPMX_REQUIRE(dims[g][0] > 0 && dims[g][1] > 0 && dims[g][2] > 0, "mpmx_hashgrid_build: bad dims");
        MPMX_REQUIRE(table_keys(dims[g]) < (1L << 31), "mpmx_hashgrid_build: too many cells");
        MPMX_REQUIRE(cell_start[g] && (n == 0 || (x && point_ids[g])), "mpmx_hashgrid_build: null pointer");
    }
    MPMX_REQUIRE(workspace && ws_bytes >= hash_tables_bytes(ntab, n, dims), "mpmx_hashgrid_build: workspace too small");
    hipStream_t s = as_stream(stream);
    const long nn = n > 0 ? n : 1;
    HashTables h{};
    h.n = n;
    h.x = x;
    h.inv = 1.0f / cell_width;
    h.ntab = ntab;
    Carver cv(workspace, ws_bytes);
    long nblk = 0, nkeys_max = 0;
    for (int g = 0; g < ntab; g++) {
        h.dimx[g] = dims[g][0];
        h.dimy[g] = dims[g][1];
        h.dimz[g] = dims[g][2];
        h.nkeys[g] = table_keys(dims[g]);
        h.nchunks[g] = table_chunks(dims[g]);
        h.labels[g] = labels[g];
        h.label[g] = label[g];
        h.keys[g] = cv.take<int32_t>(nn);
        h.rank[g] = cv.take<int32_t>(nn);
        const long blk = (long)dims[g][1] * ((dims[g][2] + 63) / 64) * ((dims[g][0] + 63) / 64);
        if (g == 0) h.nblk0 = blk;
        nblk += blk;
        if (h.nkeys[g] > nkeys_max) nkeys_max = h.nkeys[g];
    }
    for (int g = 0; g < ntab; g++) {
        h.tcount[g] = cv.take<int32_t>(h.nkeys[g] + h.nchunks[g] + 1);  // the chunk totals follow the counts
        h.totals[g] = h.tcount[g] + h.nkeys[g];
        h.count[g] = cv.take<int32_t>(h.nkeys[g]);
    }
    MPMX_REQUIRE(!cv.overflow, "mpmx_hashgrid_build: workspace carve failed");
    MPMX_REQUIRE(!lazy || ntab == 2, "mpmx_hashgrid_build: a lazy build needs two tables");
    if (lazy) h.label_any = h.tcount[1] + h.nkeys[1] + h.nchunks[1];  // zeroed with the counts
    int rc;
    for (int g = ntab - 1; g >= 0; g--) {  // the second table (and the flag word) first
        if (g == 0 && lazy) {
            // lazy: first whether any point has the label; the first table's
            // counts are then zeroed and counted only if one has
            if (n > 0) {
                MPMX_LAUNCH(label_any_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, labels[1], label[1],
                            h.label_any);
                if ((rc = check_launch("label_any", s))) return rc;
            }
            const long cnt = h.nkeys[0] + h.nchunks[0] + 1;
            const long blocks = (cnt + kBlock - 1) / kBlock;
            MPMX_LAUNCH(fill_zero_gated_kernel, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(kBlock), 0, s,
                        h.tcount[0], cnt, (const int32_t*)h.label_any);
            if ((rc = check_launch("fill_i32", s))) return rc;
            continue;
        }
        if ((rc = fill_i32(h.tcount[g], h.nkeys[g] + h.nchunks[g] + 1, 0, s))) return rc;
    }
    if (n > 0) {
        MPMX_LAUNCH(hash_count_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, h);
        if ((rc = check_launch("hash_count", s))) return rc;
    }
    MPMX_LAUNCH(hash_counts_transpose_kernel, dim3((unsigned)nblk), dim3(kBlock), 0, s, h);
    if ((rc = check_launch("hash_counts_transpose", s))) return rc;
    SortSegs t{};
    for (int g = 0; g < ntab; g++) {
        t.seg[g].keys = h.keys[g];
        t.seg[g].rank = h.rank[g];
        t.seg[g].count = h.count[g];
        t.seg[g].totals = h.totals[g];
        t.seg[g].cell_start = cell_start[g];
        t.seg[g].sorted = point_ids[g];
        t.seg[g].nkeys = h.nkeys[g];
        t.seg[g].pos = point_pos ? point_pos[g] : nullptr;
        t.seg[g].pos_gate = pos_gate;
        t.seg[g].skip_gate = (g == 0) ? h.label_any : nullptr;
    }
    return counting_sort_segments(t, ntab, n, nkeys_max, s, order);
}

static int hashgrid_build(int n, const float* x, float cell_width, int dimx, int dimy, int dimz,
                          const int32_t* labels, int label, int32_t* point_ids, int32_t* cell_start, void* workspace,
                          size_t ws_bytes, void* stream, const int32_t* order = nullptr) {
    const int dims[1][3] = {{dimx, dimy, dimz}};
    return hash_tables_build(1, n, x, cell_width, dims, &labels, &label, &point_ids, &cell_start, workspace, ws_bytes,
                             stream, order);
}

int mpmx_hashgrid_build_pair(int n, const float* x, float cell_width, int dimx, int dimy, int dimz,
                             const int32_t* materialLabel, int label, int label_dimx, int label_dimy, int label_dimz,
                             const int32_t* order, int32_t* point_ids, int32_t* cell_start, int32_t* label_point_ids,
                             int32_t* label_cell_start, int32_t* point_pos, int32_t* label_point_pos,
                             int lazy, void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n == 0 || materialLabel, "mpmx_hashgrid_build_pair: null label array");
    MPMX_REQUIRE((point_pos == nullptr) == (label_point_pos == nullptr),
                 "mpmx_hashgrid_build_pair: give both position arrays or neither");
    const int dims[2][3] = {{dimx, dimy, dimz}, {label_dimx, label_dimy, label_dimz}};
    const int32_t* labels[2] = {nullptr, materialLabel};
    const int lab[2] = {0, label};
    int32_t* ids[2] = {point_ids, label_point_ids};
    int32_t* starts[2] = {cell_start, label_cell_start};
    int32_t* pos[2] = {point_pos, label_point_pos};
    // the positions serve the contact records, which nothing reads while the second table is empty
    return hash_tables_build(2, n, x, cell_width, dims, labels, lab, ids, starts, workspace, ws_bytes, stream, order,
                             point_pos ? pos : nullptr, label_cell_start + table_keys(dims[1]), lazy != 0);
}

size_t mpmx_hashgrid_near_words(int dimx, int dimy, int dimz) {
    const long ncell = (long)dimx * dimy * dimz;
    return (size_t)((ncell + 63) / 64);
}

int mpmx_hashgrid_near_cells(const int32_t* cell_start, int dimx, int dimy, int dimz, uint64_t* occ, uint64_t* near,
                             void* stream) {
    MPMX_REQUIRE(dimx > 0 && dimy > 0 && dimz > 0 && (long)dimx * dimy * dimz < (1L << 31),
                 "mpmx_hashgrid_near_cells: bad dims");
    MPMX_REQUIRE(cell_start && occ && near, "mpmx_hashgrid_near_cells: null pointer");
    const long ncell = (long)dimx * dimy * dimz;
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(table_occupancy_kernel, dim3(grid_for(ncell, kBlock)), dim3(kBlock), 0, s, cell_start, ncell, occ);
    int rc = check_launch("table_occupancy", s);
    if (rc) return rc;
    MPMX_LAUNCH(table_near_kernel, dim3(grid_for(ncell, kBlock)), dim3(kBlock), 0, s, cell_start, dimx, dimy, dimz,
                (const uint64_t*)occ, near);
    return check_launch("table_near", s);
}

int mpmx_xpbd_contact_list(int n, const int32_t* materialLabel, const int32_t* point_pos,
                           const int32_t* xpbd_cell_start, float cell_width, int xpbd_dimx, int xpbd_dimy,
                           int xpbd_dimz, const float* x, const float* radius, float k_cohesion, float max_radius,
                           const uint64_t* xpbd_near, int32_t* list, int32_t* count, void* stream) {
    MPMX_REQUIRE(n >= 0 && xpbd_dimx > 0 && xpbd_dimy > 0 && xpbd_dimz > 0 && cell_width > 0.0f,
                 "mpmx_xpbd_contact_list: bad sizes");
    MPMX_REQUIRE(count && xpbd_cell_start, "mpmx_xpbd_contact_list: null pointer");
    hipStream_t s = as_stream(stream);
    int rc = fill_i32(count, 1, 0, s);
    if (rc || n == 0) return rc;
    MPMX_REQUIRE(materialLabel && point_pos && x && radius && xpbd_near && list, "mpmx_xpbd_contact_list: null pointer");
    MPMX_LAUNCH(contact_list_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, materialLabel, point_pos,
                xpbd_cell_start, 1.0f / cell_width, xpbd_dimx, xpbd_dimy, xpbd_dimz, x, radius, k_cohesion, max_radius,
                xpbd_near, list, count);
    return check_launch("contact_list", s);
}

int mpmx_hashgrid_build(int n, const float* x, float cell_width, int dimx, int dimy, int dimz, int32_t* point_ids,
                        int32_t* cell_start, void* workspace, size_t ws_bytes, void* stream) {
    return hashgrid_build(n, x, cell_width, dimx, dimy, dimz, nullptr, 0, point_ids, cell_start, workspace, ws_bytes,
                          stream);
}

int mpmx_hashgrid_build_label(int n, const float* x, float cell_width, int dimx, int dimy, int dimz,
                              const int32_t* materialLabel, int label, int32_t* point_ids, int32_t* cell_start,
                              void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n == 0 || materialLabel, "mpmx_hashgrid_build_label: null label array");
    return hashgrid_build(n, x, cell_width, dimx, dimy, dimz, materialLabel, label, point_ids, cell_start, workspace,
                          ws_bytes, stream);
}

int mpmx_hashgrid_build_ordered(int n, const float* x, float cell_width, int dimx, int dimy, int dimz,
                                const int32_t* materialLabel, int label, const int32_t* order, int32_t* point_ids,
                                int32_t* cell_start, void* workspace, size_t ws_bytes, void* stream) {
    return hashgrid_build(n, x, cell_width, dimx, dimy, dimz, materialLabel, label, point_ids, cell_start, workspace,
                          ws_bytes, stream, order);
}

int mpmx_reset_mpm_contact_labels(int n, const int32_t* activeLabel, int32_t* materialLabel, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel)), "mpmx_reset_mpm_contact_labels: bad args");
    if (n == 0) return 0;
    LAUNCH1(reset_labels_kernel, n, n, activeLabel, materialLabel)
}

int mpmx_integrate_particles_xpbd(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* x,
                                  const float* v, float gxv, float gyv, float gzv, float dt, float* x_new,
                                  float* v_new, float v_max, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x && v && x_new && v_new)),
                 "mpmx_integrate_particles_xpbd: bad args");
    if (n == 0) return 0;
    LAUNCH1(integrate_kernel, n, n, activeLabel, materialLabel, x, v, gxv, gyv, gzv, dt, x_new, v_new, v_max)
}

int mpmx_solve_particle_bound_contacts(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                                       const float* x, const float* v, const float* mass, const float* radius,
                                       float mu, float static_threshold, float static_friction, float minx,
                                       float miny, float minz, float maxx, float maxy, float maxz, float dt,
                                       float relaxation, float* delta, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x && v && mass && radius && delta)),
                 "mpmx_solve_particle_bound_contacts: bad args");
    if (n == 0) return 0;
    LAUNCH1(bound_contacts_kernel, n, n, activeLabel, materialLabel, x, v, mass, radius, mu, static_threshold,
            static_friction, f3{minx, miny, minz}, f3{maxx, maxy, maxz}, dt, relaxation, delta)
}

int mpmx_solve_particle_particle_contacts(int n, const int32_t* activeLabel, int32_t* materialLabel,
                                          const int32_t* point_ids, const int32_t* cell_start, float cell_width,
                                          int dimx, int dimy, int dimz, const float* x, const float* v,
                                          const float* mass, const float* radius, float k_mu, float static_threshold,
                                          float static_friction, float k_cohesion, float max_radius, float dt,
                                          float relaxation, int contact_lock, float* deltas, void* stream) {
    MPMX_REQUIRE(n >= 0 && dimx > 0 && dimy > 0 && dimz > 0 && cell_width > 0.0f,
                 "mpmx_solve_particle_particle_contacts: bad sizes");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && point_ids && cell_start && x && v && mass && radius && deltas,
                 "mpmx_solve_particle_particle_contacts: null pointer");
    LAUNCH1(particle_contacts_kernel<false>, n, n, activeLabel, materialLabel, point_ids, cell_start, nullptr,
            nullptr, 1.0f / cell_width, dimx, dimy, dimz, dimx, dimy, dimz, x, v, mass, radius, k_mu,
            static_threshold, static_friction, k_cohesion, max_radius, dt, relaxation, contact_lock, deltas, nullptr,
            ContactPack{})
}

int mpmx_solve_particle_particle_contacts_split(int n, const int32_t* activeLabel, int32_t* materialLabel,
                                                const int32_t* point_ids, const int32_t* cell_start,
                                                const int32_t* xpbd_point_ids, const int32_t* xpbd_cell_start,
                                                float cell_width, int dimx, int dimy, int dimz, int xpbd_dimx,
                                                int xpbd_dimy, int xpbd_dimz, const float* x,
                                                const float* v, const float* mass, const float* radius, float k_mu,
                                                float static_threshold, float static_friction, float k_cohesion,
                                                float max_radius, float dt, float relaxation, int contact_lock,
                                                float* deltas, const int32_t* slot_of_id, void* stream) {
    MPMX_REQUIRE(n >= 0 && dimx > 0 && dimy > 0 && dimz > 0 && xpbd_dimx > 0 && xpbd_dimy > 0 && xpbd_dimz > 0 &&
                     cell_width > 0.0f,
                 "mpmx_solve_particle_particle_contacts_split: bad sizes");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && point_ids && cell_start && xpbd_point_ids && xpbd_cell_start && x &&
                     v && mass && radius && deltas,
                 "mpmx_solve_particle_particle_contacts_split: null pointer");
    LAUNCH1(particle_contacts_kernel<false>, n, n, activeLabel, materialLabel, point_ids, cell_start,
            xpbd_point_ids, xpbd_cell_start, 1.0f / cell_width, dimx, dimy, dimz, xpbd_dimx, xpbd_dimy, xpbd_dimz, x,
            v, mass, radius, k_mu, static_threshold, static_friction, k_cohesion, max_radius, dt, relaxation,
            contact_lock, deltas, slot_of_id, ContactPack{})
}

int mpmx_xpbd_contact_pack(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* mass,
                           const float* radius, const float* x, const float* v, const float* particle_xr,
                           const float* particle_vw, const int32_t* particle_flags, const int32_t* point_ids,
                           const int32_t* xpbd_point_ids, const int32_t* xpbd_cell_start, int xpbd_dimx,
                           int xpbd_dimy, int xpbd_dimz, float* xr, float* vw, int32_t* flags, void* stream) {
    MPMX_REQUIRE(n >= 0 && xpbd_dimx > 0 && xpbd_dimy > 0 && xpbd_dimz > 0, "mpmx_xpbd_contact_pack: bad sizes");
    if (n == 0) return 0;
    const bool from_records = particle_xr != nullptr;
    MPMX_REQUIRE(point_ids && xpbd_point_ids && xpbd_cell_start && xr && vw && flags &&
                     (from_records ? (particle_vw && particle_flags)
                                   : (activeLabel && materialLabel && mass && radius && x && v)),
                 "mpmx_xpbd_contact_pack: null pointer");
    MPMX_REQUIRE(((uintptr_t)xr | (uintptr_t)vw | (uintptr_t)particle_xr | (uintptr_t)particle_vw) % 16 == 0,
                 "mpmx_xpbd_contact_pack: records need 16-byte alignment");
    const long cells = (long)xpbd_dimx * xpbd_dimy * xpbd_dimz;
    const ContactPack src{(const float4*)particle_xr, (const float4*)particle_vw, particle_flags};
    LAUNCH1(contact_pack_kernel, 2L * n, n, activeLabel, materialLabel, mass, radius, x, v, src, point_ids,
            xpbd_point_ids, xpbd_cell_start + cells, (float4*)xr, (float4*)vw, flags)
}

int mpmx_solve_particle_particle_contacts_packed(
    int n, const int32_t* activeLabel, int32_t* materialLabel, const int32_t* point_ids, const int32_t* cell_start,
    const int32_t* xpbd_point_ids, const int32_t* xpbd_cell_start, float cell_width, int dimx, int dimy, int dimz,
    int xpbd_dimx, int xpbd_dimy, int xpbd_dimz, const float* x, const float* v, const float* mass,
    const float* radius, const float* xr, const float* vw, const int32_t* flags, float k_mu, float static_threshold,
    float static_friction, float k_cohesion, float max_radius, float dt, float relaxation, int contact_lock,
    float* deltas, const int32_t* slot_of_id, const uint64_t* xpbd_near, const int32_t* contact_list,
    const int32_t* contact_count, void* stream) {
    MPMX_REQUIRE(n >= 0 && dimx > 0 && dimy > 0 && dimz > 0 && xpbd_dimx > 0 && xpbd_dimy > 0 && xpbd_dimz > 0 &&
                     cell_width > 0.0f,
                 "mpmx_solve_particle_particle_contacts_packed: bad sizes");
    MPMX_REQUIRE((contact_list == nullptr) == (contact_count == nullptr),
                 "mpmx_solve_particle_particle_contacts_packed: give the contact list with its count");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && point_ids && cell_start && xpbd_point_ids && xpbd_cell_start && x &&
                     v && mass && radius && xr && vw && flags && deltas,
                 "mpmx_solve_particle_particle_contacts_packed: null pointer");
    MPMX_REQUIRE(((uintptr_t)xr | (uintptr_t)vw) % 16 == 0,
                 "mpmx_solve_particle_particle_contacts_packed: records need 16-byte alignment");
    // one thread per table position (a short contact list is read by the first threads)
    const unsigned blocks = grid_for(n, kBlock);
    MPMX_LAUNCH(particle_contacts_kernel<true>, dim3(blocks), dim3(kBlock), 0, as_stream(stream), n, activeLabel,
                materialLabel, point_ids, cell_start, xpbd_point_ids, xpbd_cell_start, 1.0f / cell_width, dimx, dimy,
                dimz, xpbd_dimx, xpbd_dimy, xpbd_dimz, x, v, mass, radius, k_mu, static_threshold, static_friction,
                k_cohesion, max_radius, dt, relaxation, contact_lock, deltas, slot_of_id,
                (ContactPack{(const float4*)xr, (const float4*)vw, flags, xpbd_near, contact_list, contact_count}));
    return check_launch("particle_contacts_kernel<true>", as_stream(stream));
}

int mpmx_apply_particle_deltas(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* x_orig,
                               const float* x_pred, const float* delta, float dt, float v_max, float* x_out,
                               float* v_out, void* stream) {
    (void)materialLabel;
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && x_orig && x_pred && delta && x_out && v_out)),
                 "mpmx_apply_particle_deltas: bad args");
    if (n == 0) return 0;
    LAUNCH1(apply_deltas_kernel, n, n, activeLabel, x_orig, x_pred, delta, dt, v_max, x_out, v_out)
}

int mpmx_sleep_particles(int n, const int32_t* activeLabel, const int32_t* materialLabel, float threshold,
                         const float* radius, const float* x_prev, float* x_after, float* dist_total, float dt,
                         int32_t* counters, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && radius && x_prev && x_after && dist_total &&
                                       counters)),
                 "mpmx_sleep_particles: bad args");
    if (n == 0) return 0;
    LAUNCH1(sleep_kernel, n, n, activeLabel, materialLabel, threshold, radius, x_prev, x_after, dist_total, dt,
            counters)
}

int mpmx_apply_mpm_contact_impulses(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                                    const float* x_orig, const float* v_orig, float* x_final, const float* v_final,
                                    float dt, float v_max, float coupling, int xpbd_only, float* v_out,
                                    void* stream) {
    (void)v_final;
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x_orig && v_orig && x_final && v_out)),
                 "mpmx_apply_mpm_contact_impulses: bad args");
    if (n == 0) return 0;
    LAUNCH1(impulse_kernel, n, n, activeLabel, materialLabel, x_orig, v_orig, x_final, dt, v_max, coupling, xpbd_only,
            v_out)
}

int mpmx_swell_particles_type2(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* x_init,
                               const float* x, float* radius, float factor, float factor_max,
                               const float* swell_radius, const float* base_radius, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x_init && x && radius && swell_radius &&
                                       base_radius)),
                 "mpmx_swell_particles_type2: bad args");
    if (n == 0) return 0;
    LAUNCH1(swell_kernel, n, n, activeLabel, materialLabel, x_init, x, radius, factor, factor_max, swell_radius,
            base_radius)
}

}  // extern "C"

// ===========================================================================
// Fused XPBD step: the same persistent results as the launch sequence of
// xpbdSimulationStep (simulationRoutines.py:295-526) with the copies folded in.
// x_int/v_int are updated in place; the scratch arrays particle_x_deltaInt,
// particle_v_deltaInt and particle_v_before_contact are not written.
// ===========================================================================
namespace mpmx {

// the particle-wall delta of my_solve_particle_bound_contacts (xpbdRoutines.py:49-136)
// for particle t at (xp, vp); zero unless it is an active XPBD particle with mass
__device__ __forceinline__ f3 bound_delta(const mpmx_xpbd_bounds& B, int n, int t, int a, int m, f3 xp, f3 vp) {
    float da[3] = {0.0f, 0.0f, 0.0f};
    if (a == 1 && m == 2 && B.mass[t] != 0.0f) {
        const float wi = 1.0f / B.mass[t];
        const float r = B.radius[t];
        const float xa[3] = {xp.x, xp.y, xp.z}, va[3] = {vp.x, vp.y, vp.z};
        const float lo[3] = {B.minx, B.miny, B.minz}, hi[3] = {B.maxx, B.maxy, B.maxz};
        const float thr = B.static_threshold, dt = B.dt, relax = B.relaxation;
#pragma unroll
        for (int ax = 0; ax < 3; ax++) {
#pragma unroll
            for (int side = 0; side < 2; side++) {
                float sg = side == 0 ? 1.0f : -1.0f;
                float b = side == 0 ? lo[ax] : hi[ax];
                float c = fminf(sg * (xa[ax] - b) - r, 0.0f);
                if (c < 0.0f) {
                    float lambda_n = c;
                    float vn = sg * va[ax];
                    float fmu = (fabsf(vn) < thr * r) ? B.static_friction : B.mu;
                    float vt[3] = {va[0], va[1], va[2]};
                    vt[ax] = va[ax] - sg * vn;
                    float vtl = len3(f3{vt[0], vt[1], vt[2]});
                    float a_ = fmu * lambda_n, b_ = -vtl * dt;
                    float lf = a_ > b_ ? a_ : b_;  // wp.max
                    float df[3] = {0.0f, 0.0f, 0.0f};
                    if (vtl > 1.0e-8f) {
                        f3 nt = normalize3(f3{vt[0], vt[1], vt[2]});
                        df[0] = nt.x * lf; df[1] = nt.y * lf; df[2] = nt.z * lf;
                    }
#pragma unroll
                    for (int q = 0; q < 3; q++) {
                        float dn = (q == ax) ? sg * lambda_n : 0.0f;
                        da[q] += (df[q] - dn) / wi * relax * wi;
                    }
                }
            }
        }
    }
    return f3{da[0], da[1], da[2]};
}

// v as a kernel that loads it from memory sees it: the empty asm hides how it was
// computed, so no multiply before it is contracted with an add after it (a fused
// multiply-add would round once where the stored value rounded twice)
__device__ __forceinline__ f3 as_stored(f3 v) {
    asm volatile("" : "+v"(v.x), "+v"(v.y), "+v"(v.z));
    return v;
}

// my_apply_particle_deltas (xpbdRoutines.py:221-257) of one particle
__device__ __forceinline__ void apply_delta(f3 x0, f3 xp, f3 d, float dt, float v_max, f3& xn, f3& vn) {
    if (isnan(d.x) || isnan(d.y) || isnan(d.z)) d = f3{0.0f, 0.0f, 0.0f};
    xn = f3{xp.x + d.x, xp.y + d.y, xp.z + d.z};
    vn = f3{(xn.x - x0.x) / dt, (xn.y - x0.y) / dt, (xn.z - x0.z) / dt};
    float m = len3(vn);
    if (m > v_max) {
        float s = v_max / m;
        vn = f3{vn.x * s, vn.y * s, vn.z * s};
    }
}

// whether the kernel taking B writes the contact records (mpmx_xpbd_bounds)
__device__ __forceinline__ bool records_wanted(const mpmx_xpbd_bounds& B) {
    return B.xr && !(B.xpbd_count && *B.xpbd_count == 0);
}

// particle p's records, in the full table and (when it is in it) the XPBD-only one
__device__ __forceinline__ void write_records(const mpmx_xpbd_bounds& B, int n, int p, float4 xr, float4 vw) {
    const long q0 = B.pos[p], q1 = B.xpbd_pos[p];
    ((float4*)B.xr)[q0] = xr;
    ((float4*)B.vw)[q0] = vw;
    if (q1 >= 0) {
        ((float4*)B.xr)[n + q1] = xr;
        ((float4*)B.vw)[n + q1] = vw;
    }
}

// x_int, v_int <- x, v; integrateParticlesXPBD on active XPBD particles
// (xpbdRoutines.py:23-48); reset_mpm_contact_labels when lock (:3-21)
__global__ void xpbd_begin_kernel(int n, const int32_t* __restrict__ active, int32_t* __restrict__ mat,
                                  const float* __restrict__ x, const float* __restrict__ v, float gx, float gy,
                                  float gz, float dt, float v_max, int reset_labels, float* __restrict__ x_int,
                                  float* __restrict__ v_int, int32_t* __restrict__ counters, int bound,
                                  mpmx_xpbd_bounds B) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (counters && p < 2) counters[p] = 0;  // for xpbd_finish's sleep counts
    if (p >= n) return;
    f3 x0 = ld3(x, n, p), v0 = ld3(v, n, p);
    const int a = active[p], m = mat[p];
    if (a == 1 && m == 2) {
        f3 v1 = f3{v0.x + gx * dt, v0.y + gy * dt, v0.z + gz * dt};
        float mag = len3(v1);
        if (mag > v_max) {
            float s = v_max / mag;
            v1 = f3{v1.x * s, v1.y * s, v1.z * s};
        }
        x0 = f3{x0.x + v1.x * dt, x0.y + v1.y * dt, x0.z + v1.z * dt};
        v0 = v1;
    }
    st3(x_int, n, p, x0);
    st3(v_int, n, p, v0);
    const bool reset = reset_labels && a == 1 && m == 0;
    if (reset) mat[p] = 1;
    // the first iteration's particle-wall deltas (the label reset never makes an XPBD particle)
    if (bound) {
        st3(B.delta, n, p,
            (a == 1 && m == 2) ? bound_delta(B, n, p, a, m, as_stored(x0), as_stored(v0)) : f3{0.0f, 0.0f, 0.0f});
        if (records_wanted(B)) {
            const int lab = reset ? 1 : m;
            const float pm = B.mass[p];
            const bool valid = a == 1 && pm > 0.0f;
            const int f = (valid ? kCandValid : 0) | (lab <= 1 ? kCandMpm : 0) | (lab == 2 ? kCandXpbd : 0);
            const int q1 = B.xpbd_pos[p];
            B.flags[B.pos[p]] = f;
            if (q1 >= 0) B.flags[n + q1] = f;
            if (valid)  // no record of an invalid candidate is read
                write_records(B, n, p, make_float4(x0.x, x0.y, x0.z, B.radius[p]),
                              make_float4(v0.x, v0.y, v0.z, 1.0f / pm));
        }
    }
}

// my_solve_particle_bound_contacts writing (not adding) the delta of every
// particle: 0 where the reference leaves the zeroed delta untouched
__global__ void bound_contacts_write_kernel(int n, const int32_t* __restrict__ active,
                                            const int32_t* __restrict__ mat, const float* __restrict__ x,
                                            const float* __restrict__ v, mpmx_xpbd_bounds B) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int a = active[t], m = mat[t];
    f3 d{0.0f, 0.0f, 0.0f};
    if (a == 1 && m == 2) d = bound_delta(B, n, t, a, m, ld3(x, n, t), ld3(v, n, t));
    st3(B.delta, n, t, d);
}

// my_apply_particle_deltas (xpbdRoutines.py:221-257) in place on x_int / v_int; with
// bound, then the next iteration's particle-wall deltas from the new x_int / v_int
// (each particle's bound delta depends on that particle alone)
__global__ void apply_deltas_inplace_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat,
                                            const float* __restrict__ x_orig, const float* delta, float dt,
                                            float v_max, float* __restrict__ x_int, float* __restrict__ v_int,
                                            int bound, mpmx_xpbd_bounds B) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    // With no XPBD point every delta is 0: wall deltas go to label-2 particles only,
    // and there is no MPM-XPBD pair.  The application would then leave x_int and set
    // v_int to 0, which the finish kernel recomputes or overwrites for every
    // particle, so it is skipped.
    if (bound && B.xpbd_count && *B.xpbd_count == 0) return;
    const int a = active[t];
    f3 nd{0.0f, 0.0f, 0.0f};
    if (a == 1) {
        f3 xn, vn;
        apply_delta(ld3(x_orig, n, t), ld3(x_int, n, t), ld3(delta, n, t), dt, v_max, xn, vn);
        st3(x_int, n, t, xn);
        st3(v_int, n, t, vn);
        if (bound) {
            const int m = mat[t];
            if (m == 2) nd = bound_delta(B, n, t, a, m, as_stored(xn), as_stored(vn));
            // labels, mass and radius are those the records were written from (lock
            // writes change no flag): only the positions and velocities are new
            if (records_wanted(B)) {
                const long q[2] = {B.pos[t], B.xpbd_pos[t] >= 0 ? n + (long)B.xpbd_pos[t] : -1L};
                for (int k = 0; k < 2; k++) {
                    if (q[k] < 0) continue;
                    float* r = B.xr + 4 * q[k];
                    float* w = B.vw + 4 * q[k];
                    r[0] = xn.x, r[1] = xn.y, r[2] = xn.z;
                    w[0] = vn.x, w[1] = vn.y, w[2] = vn.z;
                }
            }
        }
    }
    if (bound) st3(B.delta, n, t, nd);
}

// sleepParticles (xpbdRoutines.py:313-334), apply_mpm_contact_impulses (:259-310)
// and the commit x <- x_int, v <- v_int (simulationRoutines.py:469-470) of particle t;
// returns whether it is an active XPBD particle and whether it went to sleep
__device__ __forceinline__ void finish_one(int n, int t, const int32_t* __restrict__ active,
                                           const int32_t* __restrict__ mat, float thr,
                                           const float* __restrict__ radius, float* __restrict__ dist, float dt,
                                           float v_max, float k, int xpbd_only, float* __restrict__ x,
                                           float* __restrict__ v, float* __restrict__ x_int,
                                           float* __restrict__ v_int, const float* __restrict__ delta, bool& xp,
                                           bool& slept) {
    f3 xi = ld3(x_int, n, t), vi;
    const f3 x0 = ld3(x, n, t);
    const int a = active[t], m = mat[t];
    if (delta && a == 1) {
        // the last iteration's my_apply_particle_deltas (x_orig is x)
        apply_delta(x0, xi, ld3(delta, n, t), dt, v_max, xi, vi);
        xi = as_stored(xi);
        vi = as_stored(vi);
    } else {
        vi = ld3(v_int, n, t);
    }
    xp = a == 1 && m == 2;
    slept = false;
    if (xp) {
        float d = len3(f3{xi.x - x0.x, xi.y - x0.y, xi.z - x0.z});
        if (d / dt < thr * radius[t]) {
            xi = x0;
            slept = true;
        }
        dist[t] = dist[t] + d;
    }
    if (a == 1 && m <= 1) {
        xi = x0;
        if (xpbd_only == 0) {
            f3 vo = ld3(v, n, t);
            f3 vc = f3{vo.x + (xi.x - x0.x) / dt * k, vo.y + (xi.y - x0.y) / dt * k, vo.z + (xi.z - x0.z) / dt * k};
            float mag = len3(vc);
            if (mag > v_max) {
                float s = v_max / mag;
                vc = f3{vc.x * s, vc.y * s, vc.z * s};
            }
            vi = vc;
        }
    }
    st3(x_int, n, t, xi);
    st3(v_int, n, t, vi);
    st3(x, n, t, xi);
    st3(v, n, t, vi);
}

// finish_one over every particle, a bounded grid striding over them, and the two
// sleep counts summed per block into one atomic each: atomics on one address
// serialise in L2, so one per wave cost ~0.2 ms at 2M particles
constexpr int kFinishBlocks = 2048;
__global__ void __launch_bounds__(kBlock) xpbd_finish_kernel(
    int n, const int32_t* __restrict__ active, const int32_t* __restrict__ mat, float thr,
    const float* __restrict__ radius, float* __restrict__ dist, float dt, float v_max, float k, int xpbd_only,
    float* __restrict__ x, float* __restrict__ v, float* __restrict__ x_int, float* __restrict__ v_int,
    int32_t* __restrict__ counters, const float* __restrict__ delta) {
    __shared__ int32_t block_counts[2];
    if (threadIdx.x < 2) block_counts[threadIdx.x] = 0;
    __syncthreads();
    int nxp = 0, nslept = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        bool xp, slept;
        finish_one(n, t, active, mat, thr, radius, dist, dt, v_max, k, xpbd_only, x, v, x_int, v_int, delta, xp,
                   slept);
        nxp += xp;
        nslept += slept;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        nxp += __shfl_xor(nxp, off);
        nslept += __shfl_xor(nslept, off);
    }
    if ((threadIdx.x & 63) == 0) {
        if (nxp) atomicAdd(&block_counts[0], nxp);
        if (nslept) atomicAdd(&block_counts[1], nslept);
    }
    __syncthreads();
    if (threadIdx.x < 2 && block_counts[threadIdx.x]) atomicAdd(&counters[threadIdx.x], block_counts[threadIdx.x]);
}

}  // namespace mpmx

extern "C" {

static int check_bounds(int n, const mpmx_xpbd_bounds* B, const char* who) {
    MPMX_REQUIRE(n == 0 || (B->mass && B->radius && B->delta), "%s: null array in the bounds", who);
    MPMX_REQUIRE(!B->xr || (B->vw && B->flags && B->pos && B->xpbd_pos), "%s: contact records without positions",
                 who);
    MPMX_REQUIRE(((uintptr_t)B->xr | (uintptr_t)B->vw) % 16 == 0, "%s: records need 16-byte alignment", who);
    return 0;
}

int mpmx_xpbd_begin(int n, const int32_t* activeLabel, int32_t* materialLabel, const float* x, const float* v,
                    float gxv, float gyv, float gzv, float dt, float v_max, int reset_labels, float* x_int,
                    float* v_int, int32_t* counters, const mpmx_xpbd_bounds* bounds, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x && v && x_int && v_int)),
                 "mpmx_xpbd_begin: bad args");
    mpmx_xpbd_bounds B{};
    if (bounds) {
        int rc = check_bounds(n, bounds, "mpmx_xpbd_begin");
        if (rc) return rc;
        B = *bounds;
    }
    const int threads = n > 2 ? n : 2;  // the counters are cleared even with no particle
    LAUNCH1(xpbd_begin_kernel, threads, n, activeLabel, materialLabel, x, v, gxv, gyv, gzv, dt, v_max, reset_labels,
            x_int, v_int, counters, bounds ? 1 : 0, B)
}

int mpmx_xpbd_bound_contacts_write(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* x,
                                   const float* v, const float* mass, const float* radius, float mu,
                                   float static_threshold, float static_friction, float minx, float miny, float minz,
                                   float maxx, float maxy, float maxz, float dt, float relaxation, float* delta,
                                   void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && x && v && mass && radius && delta)),
                 "mpmx_xpbd_bound_contacts_write: bad args");
    if (n == 0) return 0;
    const mpmx_xpbd_bounds B{mass, radius, mu, static_threshold, static_friction, minx, miny, minz,
                             maxx, maxy, maxz, dt, relaxation, delta};
    LAUNCH1(bound_contacts_write_kernel, n, n, activeLabel, materialLabel, x, v, B)
}

int mpmx_xpbd_apply_deltas_inplace(int n, const int32_t* activeLabel, const float* x_orig, const float* delta,
                                   float dt, float v_max, float* x_int, float* v_int, const int32_t* materialLabel,
                                   const mpmx_xpbd_bounds* next_bounds, void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && x_orig && delta && x_int && v_int)),
                 "mpmx_xpbd_apply_deltas_inplace: bad args");
    if (n == 0) return 0;
    mpmx_xpbd_bounds B{};
    if (next_bounds) {
        MPMX_REQUIRE(materialLabel, "mpmx_xpbd_apply_deltas_inplace: the bounds need materialLabel");
        int rc = check_bounds(n, next_bounds, "mpmx_xpbd_apply_deltas_inplace");
        if (rc) return rc;
        B = *next_bounds;
    }
    LAUNCH1(apply_deltas_inplace_kernel, n, n, activeLabel, materialLabel, x_orig, delta, dt, v_max, x_int, v_int,
            next_bounds ? 1 : 0, B)
}

int mpmx_xpbd_finish(int n, const int32_t* activeLabel, const int32_t* materialLabel, float sleep_threshold,
                     const float* radius, float* dist_total, float dt, float v_max, float coupling, int xpbd_only,
                     float* x, float* v, float* x_int, float* v_int, int32_t* counters, const float* delta,
                     void* stream) {
    MPMX_REQUIRE(n >= 0 && (n == 0 || (activeLabel && materialLabel && radius && dist_total && x && v && x_int &&
                                       v_int && counters)),
                 "mpmx_xpbd_finish: bad args");
    if (n == 0) return 0;
    const long blocks = grid_for(n, kBlock);
    MPMX_LAUNCH(xpbd_finish_kernel, dim3((unsigned)(blocks < kFinishBlocks ? blocks : kFinishBlocks)), dim3(kBlock), 0,
                as_stream(stream), n, activeLabel, materialLabel, sleep_threshold, radius, dist_total, dt, v_max,
                coupling, xpbd_only, x, v, x_int, v_int, counters, delta);
    return check_launch("xpbd_finish_kernel", as_stream(stream));
}

}  // extern "C"
