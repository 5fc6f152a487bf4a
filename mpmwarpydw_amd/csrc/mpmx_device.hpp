// Device-side building blocks shared by the gfx950 kernels of libmpmx.
//
// Layout in HBM (structure of arrays): a per-particle vec3 array of N items is
// three float planes [3][N]; a mat33 array is nine planes [9][N] in row-major
// component order (r*3+c); a grid vec3 array of G = gx*gy*gz nodes is [3][G]
// with node index (ix*gy + iy)*gz + iz.  Every particle kernel therefore reads
// and writes each component with unit stride across a wavefront.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpmx.h"

// Math helpers are host+device so that a CPU build can exercise them.
#define MPMX_HD __host__ __device__ __forceinline__

namespace mpmx {

// Workgroups are dealt to the 8 XCDs of the chip round-robin (block b runs on
// XCD b % 8), and every XCD has its own L2.  Neighbouring tiles (or ranges of
// a table) share cache lines of the particle and grid arrays, so they are
// handed out in 8 contiguous ranges, one per XCD: block b works on tile
// xcd_tile(b, nb).  A bijection of [0, nb) whatever the dispatch order.
MPMX_HD int xcd_tile(int b, int nb) {
    const int q = nb / 8, r = nb % 8, x = b % 8, j = b / 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + j;
}

// Runs of kXcdRun consecutive tiles dealt to the XCDs in turn: neighbouring
// tiles still share an L2 within a run, and a scene that fills only part of
// the tile range (a block beside an empty region) spreads over every XCD,
// where 8 contiguous ranges left some XCDs with empty tiles only.  A bijection
// of [0, nb): blocks past the last whole round of 8 runs keep their own tile.
constexpr int kXcdRun = 16;
MPMX_HD int xcd_run_tile(int b, int nb) {
    const int round = 8 * kXcdRun, full = nb / round * round;
    if (b >= full) return b;
    const int x = b % 8, j = b / 8;  // the XCD and the block's turn on it
    return (j / kXcdRun) * round + x * kXcdRun + j % kXcdRun;
}

// the tile of block b under tuning mode xcd (0: block order, 1: 8 contiguous
// ranges, 2: runs of kXcdRun tiles in turn)
MPMX_HD int tile_of_block(int xcd, int b, int nb) {
    return xcd == 1 ? xcd_tile(b, nb) : xcd == 2 ? xcd_run_tile(b, nb) : b;
}

constexpr int kBlock = 256;

struct f3 {
    float x, y, z;
};

MPMX_HD f3 ld3(const float* __restrict__ a, long n, long i) {
    return f3{a[i], a[n + i], a[2 * n + i]};
}
MPMX_HD void st3(float* __restrict__ a, long n, long i, f3 v) {
    a[i] = v.x;
    a[n + i] = v.y;
    a[2 * n + i] = v.z;
}
MPMX_HD void ld9(const float* __restrict__ a, long n, long i, float m[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) m[k] = a[k * n + i];
}
MPMX_HD void st9(float* __restrict__ a, long n, long i, const float m[9]) {
#pragma unroll
    for (int k = 0; k < 9; k++) a[k * n + i] = m[k];
}
// |v| with the squares summed as fma(z, z, fma(y, y, x x)) wherever it is inlined:
// left to the compiler, whether the sum is contracted depends on the code around
// it, and kernels that split the same work differently would not agree bit for bit
MPMX_HD float len3(f3 v) { return sqrtf(fmaf(v.z, v.z, fmaf(v.y, v.y, v.x * v.x))); }

// Warp's normalize(): the zero vector stays zero, otherwise divide by length.
MPMX_HD f3 normalize3(f3 v) {
    float l = len3(v);
    if (l > 0.0f) return f3{v.x / l, v.y / l, v.z / l};
    return f3{0.0f, 0.0f, 0.0f};
}

// a0 b0 + a1 b1 + a2 b2 + c in FP32 arithmetic with about twice FP32's
// precision, rounded once at the end: the compensated dot product "Dot2" of
// Ogita, Rump & Oishi (SIAM J. Sci. Comput. 26(6), 2005).  Each product is
// split exactly into p + e with an FMA (TwoProduct), the running sum is kept
// exact as hi + err with Knuth's TwoSum, and the error terms are added at the
// end.  The error is at most one rounding of the result plus about 9 u^2
// (u = 2^-24) of the sum of the magnitudes of the terms, so F F^T - I keeps the
// relative precision of a small strain when F holds a rotation of order 1 (the
// terms are of order 1 and cancel to the strain).  Contraction is off so that
// the compiler cannot fuse a product into a TwoSum.
MPMX_HD float dot3_plus(float a0, float b0, float a1, float b1, float a2, float b2, float c) {
#pragma clang fp contract(off)
    float p = a0 * b0;
    float s = fmaf(a0, b0, -p);
    float h = a1 * b1, r = fmaf(a1, b1, -h);
    float t = p + h, z = t - p;
    s += ((p - (t - z)) + (h - z)) + r;
    p = t;
    h = a2 * b2;
    r = fmaf(a2, b2, -h);
    t = p + h;
    z = t - p;
    s += ((p - (t - z)) + (h - z)) + r;
    p = t;
    t = p + c;
    z = t - p;
    s += (p - (t - z)) + (c - z);
    return t + s;
}

// ---------------------------------------------------------------------------
// 3x3 SVD  F = U diag(s) V^T,  U, V in SO(3), |s| descending, sign(det F) on
// s[2] (the McAdams et al. 2011 convention of the reference's wp.svd3,
// mpmRoutines.py:288,339).
//
// Method (written for near-identity deformation gradients):
//  1. A = F^T F - I formed by the compensated dot product above and rounded
//     once, so A carries the strain with relative rather than absolute FP32
//     error even when F holds a large rotation or a reflection.
//  2. Cyclic Jacobi on A with exact rotations (Golub & Van Loan, sym.schur2),
//     sweeping until the off-diagonal mass is below FP32 resolution.
//  3. Sort descending with a sign flip per column swap (V stays in SO(3)).
//  4. U from a Givens QR of B = F V; R00, R11 >= 0 and R22 carries det F.
// lam[k] = s_k^2 - 1 is returned so that ln s_k = 0.5*log1p(lam[k]) keeps the
// precision of small principal strains.
// ---------------------------------------------------------------------------
struct Svd3 {
    float U[9], V[9], s[3], lam[3];
};

MPMX_HD void jacobi_rot(float A[9], float V[9], int p, int q) {
    float apq = A[p * 3 + q];
    if (apq == 0.0f) return;
    float theta = (A[q * 3 + q] - A[p * 3 + p]) / (2.0f * apq);
    float at = fabsf(theta);
    float t = (at > 1.0e10f) ? 0.5f / theta : copysignf(1.0f / (at + sqrtf(1.0f + theta * theta)), theta);
    float c = 1.0f / sqrtf(1.0f + t * t);
    float s = t * c;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float akp = A[k * 3 + p], akq = A[k * 3 + q];
        A[k * 3 + p] = c * akp - s * akq;
        A[k * 3 + q] = s * akp + c * akq;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float apk = A[p * 3 + k], aqk = A[q * 3 + k];
        A[p * 3 + k] = c * apk - s * aqk;
        A[q * 3 + k] = s * apk + c * aqk;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float vkp = V[k * 3 + p], vkq = V[k * 3 + q];
        V[k * 3 + p] = c * vkp - s * vkq;
        V[k * 3 + q] = s * vkp + c * vkq;
    }
}

MPMX_HD void givens_rows(float B[9], float Q[9], int p, int q, int col) {
    float a = B[p * 3 + col], b = B[q * 3 + col];
    float r = sqrtf(a * a + b * b);
    float c = 1.0f, s = 0.0f;
    if (r > 0.0f) {
        c = a / r;
        s = b / r;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float bp = B[p * 3 + k], bq = B[q * 3 + k];
        B[p * 3 + k] = c * bp + s * bq;
        B[q * 3 + k] = -s * bp + c * bq;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        float qp = Q[k * 3 + p], qq = Q[k * 3 + q];
        Q[k * 3 + p] = c * qp + s * qq;
        Q[k * 3 + q] = -s * qp + c * qq;
    }
}

MPMX_HD void svd3(const float F[9], Svd3& r) {
    float A[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            const float v = dot3_plus(F[0 * 3 + i], F[0 * 3 + j], F[1 * 3 + i], F[1 * 3 + j], F[2 * 3 + i],
                                      F[2 * 3 + j], i == j ? -1.0f : 0.0f);
            A[i * 3 + j] = v;
            A[j * 3 + i] = v;
        }
    float* V = r.V;
    V[0] = 1.0f; V[1] = 0.0f; V[2] = 0.0f;
    V[3] = 0.0f; V[4] = 1.0f; V[5] = 0.0f;
    V[6] = 0.0f; V[7] = 0.0f; V[8] = 1.0f;
    for (int sweep = 0; sweep < 8; sweep++) {
        float off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        float dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if (!(off > 1.0e-15f * dia)) break;
        jacobi_rot(A, V, 0, 1);
        jacobi_rot(A, V, 0, 2);
        jacobi_rot(A, V, 1, 2);
    }
    float l0 = A[0], l1 = A[4], l2 = A[8];
    // sort descending; each swap negates one column so that det V stays +1
    if (l0 < l1) {
        float t = l0; l0 = l1; l1 = t;
#pragma unroll
        for (int k = 0; k < 3; k++) { float tv = V[k * 3 + 0]; V[k * 3 + 0] = V[k * 3 + 1]; V[k * 3 + 1] = -tv; }
    }
    if (l1 < l2) {
        float t = l1; l1 = l2; l2 = t;
#pragma unroll
        for (int k = 0; k < 3; k++) { float tv = V[k * 3 + 1]; V[k * 3 + 1] = V[k * 3 + 2]; V[k * 3 + 2] = -tv; }
    }
    if (l0 < l1) {
        float t = l0; l0 = l1; l1 = t;
#pragma unroll
        for (int k = 0; k < 3; k++) { float tv = V[k * 3 + 0]; V[k * 3 + 0] = V[k * 3 + 1]; V[k * 3 + 1] = -tv; }
    }
    float B[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) B[i * 3 + j] = F[i * 3 + 0] * V[0 * 3 + j] + F[i * 3 + 1] * V[1 * 3 + j] + F[i * 3 + 2] * V[2 * 3 + j];
    float* U = r.U;
    U[0] = 1.0f; U[1] = 0.0f; U[2] = 0.0f;
    U[3] = 0.0f; U[4] = 1.0f; U[5] = 0.0f;
    U[6] = 0.0f; U[7] = 0.0f; U[8] = 1.0f;
    givens_rows(B, U, 0, 1, 0);
    givens_rows(B, U, 0, 2, 0);
    givens_rows(B, U, 1, 2, 1);
    r.lam[0] = l0; r.lam[1] = l1; r.lam[2] = l2;
    r.s[0] = sqrtf(fmaxf(1.0f + l0, 0.0f));
    r.s[1] = sqrtf(fmaxf(1.0f + l1, 0.0f));
    float m2 = sqrtf(fmaxf(1.0f + l2, 0.0f));
    r.s[2] = (B[8] < 0.0f) ? -m2 : m2;
}

// log(1 + x) to a few ulps from logf alone (Goldberg, "What every computer
// scientist should know about floating-point arithmetic", 1991, Thm. 4): with
// u = fl(1 + x), log(u) x / (u - 1) cancels the rounding of 1 + x.  The device
// library's log1pf extracts exponents through FP64 instructions; this keeps the
// per-particle kernels free of them.
MPMX_HD float log1p_f32(float x) {
#pragma clang fp contract(off)
    const float u = 1.0f + x;
    if (u == 1.0f) return x;
    const float d = u - 1.0f;
    if (!(fabsf(x) < 1.0e30f)) return logf(u);  // inf, NaN and huge x: u - 1 carries no rounding to undo
    return logf(u) * (x / d);
}

// ln(max(s, floor)) with the precision of log1p for s >= floor.
MPMX_HD float log_clamped(float s, float lam, float floorv) {
    return (s >= floorv) ? 0.5f * log1p_f32(lam) : logf(floorv);
}

// ---------------------------------------------------------------------------
// Quadratic B-spline stencil (mpmRoutines.py:514-530, :733-749): base index by
// truncation toward zero (wp.int), fx in [0.5, 1.5) for interior particles.
// ---------------------------------------------------------------------------
MPMX_HD void bspline(float f, float w[3], float dw[3]) {
    float wa = 1.5f - f, wb = f - 1.0f, wc = f - 0.5f;
    w[0] = wa * wa * 0.5f;
    w[1] = (0.0f - wb * wb) + 0.75f;
    w[2] = wc * wc * 0.5f;
    dw[0] = f - 1.5f;
    dw[1] = -2.0f * (f - 1.0f);
    dw[2] = f - 0.5f;
}

// Grid coordinate, base index and fractional offset of a particle on one axis.
MPMX_HD int base_axis(float x, float minb, float inv_dx, float& fx) {
    float gp = (x - minb) * inv_dx;
    int b = (int)(gp - 0.5f);
    fx = gp - (float)b;
    return b;
}

// Coulomb friction on one tangential component (mpmRoutines.py:915-940).
// G2P sums of one particle (mpmRoutines.py:730-776) over its 27 stencil nodes:
//   nv = sum w v,  nB = sum w v (x) (node - fx),  nG = sum v (x) grad w,
// with w = wx_i wy_j wz_k and grad w = (dwx_i wy_j wz_k, wx_i dwy_j wz_k,
// wx_i wy_j dwz_k) inv_dx.  The weights are products along the axes, so the
// sums are taken over k first, then j, then i: about half the arithmetic of
// the 27 separate node terms.  fetch(i, j, k) returns the node velocity.
template <class Fetch>
MPMX_HD void g2p_gather(const float fx[3], float inv_dx, Fetch fetch, float nv[3], float nB[9], float nG[9]) {
    float w[3][3], dw[3][3], wd[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        bspline(fx[a], w[a], dw[a]);
#pragma unroll
        for (int q = 0; q < 3; q++) wd[a][q] = w[a][q] * ((float)q - fx[a]);  // weight times node offset
    }
    float sv[3] = {0.0f, 0.0f, 0.0f}, sx[3] = {0.0f, 0.0f, 0.0f}, sy[3] = {0.0f, 0.0f, 0.0f};
    float sz[3] = {0.0f, 0.0f, 0.0f}, gx[3] = {0.0f, 0.0f, 0.0f}, gy[3] = {0.0f, 0.0f, 0.0f};
    float gz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        // sums over j and k at fixed i
        float a_[3] = {0.0f, 0.0f, 0.0f}, ay[3] = {0.0f, 0.0f, 0.0f}, az[3] = {0.0f, 0.0f, 0.0f};
        float ady[3] = {0.0f, 0.0f, 0.0f}, adz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 3; j++) {
            float b[3] = {0.0f, 0.0f, 0.0f}, bz[3] = {0.0f, 0.0f, 0.0f}, bdz[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const f3 v = fetch(i, j, k);
                const float vv[3] = {v.x, v.y, v.z};
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    b[r] += w[2][k] * vv[r];
                    bz[r] += wd[2][k] * vv[r];
                    bdz[r] += dw[2][k] * vv[r];
                }
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
                a_[r] += w[1][j] * b[r];
                ay[r] += wd[1][j] * b[r];
                az[r] += w[1][j] * bz[r];
                ady[r] += dw[1][j] * b[r];
                adz[r] += w[1][j] * bdz[r];
            }
        }
#pragma unroll
        for (int r = 0; r < 3; r++) {
            sv[r] += w[0][i] * a_[r];
            sx[r] += wd[0][i] * a_[r];
            sy[r] += w[0][i] * ay[r];
            sz[r] += w[0][i] * az[r];
            gx[r] += dw[0][i] * a_[r];
            gy[r] += w[0][i] * ady[r];
            gz[r] += w[0][i] * adz[r];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; r++) {
        nv[r] = sv[r];
        nB[r * 3 + 0] = sx[r];
        nB[r * 3 + 1] = sy[r];
        nB[r * 3 + 2] = sz[r];
        nG[r * 3 + 0] = gx[r] * inv_dx;
        nG[r * 3 + 1] = gy[r] * inv_dx;
        nG[r * 3 + 2] = gz[r] * inv_dx;
    }
}

MPMX_HD float coulomb(float v, float mu, float dt, float g) {
    if (fabsf(v) < 1.0e-12f) return 0.0f;
    float dv = mu * g * dt;
    float s = (v < 0.0f) ? -1.0f : 1.0f;
    return s * fmaxf(fabsf(v) - dv, 0.0f);
}


// collideBounds on one node (mpmRoutines.py:942-993): v is read once, every
// matching face overwrites the whole vector, the last match wins; lower faces
// act for index <= pad, upper faces for index >= dim - pad.
MPMX_HD f3 collide_node(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, float mu, float dt,
                                           float g, int pad) {
    f3 o = v;
    if (ix <= pad && v.x < 0.0f) o = f3{0.0f, coulomb(v.y, mu, dt, g), coulomb(v.z, mu, dt, g)};
    if (ix >= gx - pad && v.x > 0.0f) o = f3{0.0f, coulomb(v.y, mu, dt, g), coulomb(v.z, mu, dt, g)};
    if (iy <= pad && v.y < 0.0f) o = f3{coulomb(v.x, mu, dt, g), 0.0f, coulomb(v.z, mu, dt, g)};
    if (iy >= gy - pad && v.y > 0.0f) o = f3{coulomb(v.x, mu, dt, g), 0.0f, coulomb(v.z, mu, dt, g)};
    if (iz <= pad && v.z < 0.0f) o = f3{coulomb(v.x, mu, dt, g), coulomb(v.y, mu, dt, g), 0.0f};
    if (iz >= gz - pad && v.z > 0.0f) o = f3{coulomb(v.x, mu, dt, g), coulomb(v.y, mu, dt, g), 0.0f};
    return o;
}


MPMX_HD float& comp(f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }
MPMX_HD float compc(const f3& v, int a) { return a == 0 ? v.x : (a == 1 ? v.y : v.z); }

// collideBoundsPlaneStrainY (mpmRoutines.py:996-1062): x and z faces as
// collideBounds; on a y face the node is read again (it becomes the v that the
// z faces test) and its v_y is zeroed whatever its sign.
MPMX_HD f3 collide_plane_strain_y(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, float mu, float dt, float g,
                                  int pad) {
    f3 o = v;
    if (ix <= pad && v.x < 0.0f) o = f3{0.0f, coulomb(v.y, mu, dt, g), coulomb(v.z, mu, dt, g)};
    if (ix >= gx - pad && v.x > 0.0f) o = f3{0.0f, coulomb(v.y, mu, dt, g), coulomb(v.z, mu, dt, g)};
    if (iy <= pad) {
        v = o;
        o = f3{v.x, 0.0f, v.z};
    }
    if (iy >= gy - pad) {
        v = o;
        o = f3{v.x, 0.0f, v.z};
    }
    if (iz <= pad && v.z < 0.0f) o = f3{coulomb(v.x, mu, dt, g), coulomb(v.y, mu, dt, g), 0.0f};
    if (iz >= gz - pad && v.z > 0.0f) o = f3{coulomb(v.x, mu, dt, g), coulomb(v.y, mu, dt, g), 0.0f};
    return o;
}

// collideBoundsPlaneStrainYGlobal (mpmRoutines.py:1065-1123): v_y = 0 on every
// node, then x and z faces as collideBounds on that velocity.
MPMX_HD f3 collide_plane_strain_y_global(f3 v, int ix, int iz, int gx, int gz, float mu, float dt, float g, int pad) {
    v.y = 0.0f;
    f3 o = v;
    if (ix <= pad && v.x < 0.0f) o = f3{0.0f, 0.0f, coulomb(v.z, mu, dt, g)};
    if (ix >= gx - pad && v.x > 0.0f) o = f3{0.0f, 0.0f, coulomb(v.z, mu, dt, g)};
    if (iz <= pad && v.z < 0.0f) o = f3{coulomb(v.x, mu, dt, g), 0.0f, 0.0f};
    if (iz >= gz - pad && v.z > 0.0f) o = f3{coulomb(v.x, mu, dt, g), 0.0f, 0.0f};
    return o;
}

// collideBoundsAbsorbing (mpmRoutines.py:1126-1176): the three outer node
// layers of every face (a literal 3) scale the node's original v by
// 0 * lf + 1 * (1 - lf), lf = (pad - distance) / pad; the last matching face wins.
MPMX_HD f3 collide_absorbing(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, int pad) {
    const float P = (float)pad;
    const int idx[3] = {ix, iy, iz}, dim[3] = {gx, gy, gz};
    f3 o = v;
    for (int a = 0; a < 3; a++) {
        for (int side = 0; side < 2; side++) {
            const bool hit = side == 0 ? idx[a] < 3 : idx[a] >= dim[a] - 3;
            if (!hit) continue;
            const int dist = side == 0 ? idx[a] : dim[a] - 1 - idx[a];
            const float lf = (P - (float)dist) / P;
            const float d = 0.0f * lf + 1.0f * (1.0f - lf);
            o = f3{v.x * d, v.y * d, v.z * d};
        }
    }
    return o;
}

// collideBoundsRestitution (mpmRoutines.py:1179-1247): the normal component of
// an approaching node is reversed and scaled by e, the tangential ones get
// Coulomb friction; all faces from the original v, the last match wins.
MPMX_HD f3 collide_restitution(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, float e, float mu, float dt,
                               float g, int pad) {
    f3 o = v;
    const int idx[3] = {ix, iy, iz}, dim[3] = {gx, gy, gz};
    for (int a = 0; a < 3; a++) {
        const float vn = compc(v, a);
        for (int side = 0; side < 2; side++) {
            const bool hit = side == 0 ? (idx[a] <= pad && vn < 0.0f) : (idx[a] >= dim[a] - pad && vn > 0.0f);
            if (!hit) continue;
            for (int b = 0; b < 3; b++) comp(o, b) = b == a ? -e * vn : coulomb(compc(v, b), mu, dt, g);
        }
    }
    return o;
}

// collideBoundsGradualFriction (mpmRoutines.py:1249-1340): inside the padding an
// approaching node keeps fade = distance / pad of its normal velocity and gets
// friction mu (1 - fade) on the others; per axis the upper face is tested only
// when the lower one does not apply (elif); axes in turn, all from the original v.
MPMX_HD f3 collide_gradual(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, float mu, float dt, float g,
                           int pad) {
    const float P = (float)pad;
    const int idx[3] = {ix, iy, iz}, dim[3] = {gx, gy, gz};
    f3 o = v;
    for (int a = 0; a < 3; a++) {
        const float vn = compc(v, a);
        float fade;
        if (idx[a] < pad && vn < 0.0f)
            fade = (float)idx[a] / P;
        else if (idx[a] >= dim[a] - pad && vn > 0.0f)
            fade = (float)(dim[a] - 1 - idx[a]) / P;
        else
            continue;
        const float m = mu * (1.0f - fade);
        for (int b = 0; b < 3; b++)
            comp(o, b) = b == a ? 0.0f * (1.0f - fade) + vn * fade : coulomb(compc(v, b), m, dt, g);
    }
    return o;
}

MPMX_HD f3 apply_bc(f3 v, int ix, int iy, int iz, int gx, int gy, int gz, int bc, float fric, float dt, float gmag,
                    int pad, float restitution) {
    switch (bc) {
        case MPMX_BC_FRICTION: return collide_node(v, ix, iy, iz, gx, gy, gz, fric, dt, gmag, pad);
        case MPMX_BC_FRICTION_GRADUAL: return collide_gradual(v, ix, iy, iz, gx, gy, gz, fric, dt, gmag, pad);
        case MPMX_BC_RESTITUTION:
            return collide_restitution(v, ix, iy, iz, gx, gy, gz, restitution, fric, dt, gmag, pad);
        case MPMX_BC_ABSORBING: return collide_absorbing(v, ix, iy, iz, gx, gy, gz, pad);
        case MPMX_BC_PLANE_STRAIN_Y: return collide_plane_strain_y(v, ix, iy, iz, gx, gy, gz, fric, dt, gmag, pad);
        case MPMX_BC_PLANE_STRAIN_Y_GLOBAL: return collide_plane_strain_y_global(v, ix, iz, gx, gz, fric, dt, gmag, pad);
        default: return v;
    }
}

// Steps 4-6 of mpmSimulationStep on one node (simulationRoutines.py:94-140):
// normalisation + gravity (v = 0 where m <= 1e-15), damping when the scale is
// below 1, then the boundary condition.
MPMX_HD f3 grid_update_node(float m, f3 vin, int ix, int iy, int iz, int gx, int gy, int gz, float gxv, float gyv,
                               float gzv, float dt, float damping, int bc, float fric, float gmag, int pad,
                               float restitution) {
    f3 v = f3{0.0f, 0.0f, 0.0f};
    if (m > 1.0e-15f) {
        float inv = 1.0f / m;
        v = f3{vin.x * inv + dt * gxv, vin.y * inv + dt * gyv, vin.z * inv + dt * gzv};
    }
    if (damping < 1.0f) v = f3{v.x * damping, v.y * damping, v.z * damping};
    return apply_bc(v, ix, iy, iz, gx, gy, gz, bc, fric, dt, gmag, pad, restitution);
}

// ---------------------------------------------------------------------------
// volumetric locking correction (mpmRoutines.py:818-911).  Per cell the sums
// of tr(C) V and of V are kept in 64-bit fixed point with 2^-40 units, so that
// integer atomics make them independent of the order of the additions.
// ---------------------------------------------------------------------------
constexpr float kVlcScale = 1099511627776.0f;     // 2^40
constexpr float kVlcUnscale = 9.094947017729282e-13f;  // 2^-40

// a * b in 2^-40 units, rounded to the nearest integer, in FP32 arithmetic:
// the product is split exactly into p + e (TwoProduct), both parts are scaled
// by 2^40 (exact), and the integer part of p and the rounded remainder
// (p - rint p) + e are summed as 64-bit integers.  Clamped to +-4e18.
MPMX_HD unsigned long long vlc_fixed(float a, float b) {
#pragma clang fp contract(off)
    const float p = a * b, e = fmaf(a, b, -p);
    const float P = p * kVlcScale;
    const float lim = 4.0e18f;
    if (!(fabsf(P) <= lim)) return (unsigned long long)(P > 0.0f ? (long long)4000000000000000000LL
                                                                   : (long long)-4000000000000000000LL);
    const float qh = rintf(P);
    const float ql = rintf((P - qh) + e * kVlcScale);
    return (unsigned long long)((long long)qh + (long long)ql);
}

MPMX_HD float vlc_unfixed(unsigned long long q) { return (float)(long long)q * kVlcUnscale; }

// the particle's cell: wp.int truncation of (x - min) / dx, as :841-844
MPMX_HD bool vlc_cell(f3 x, float minx, float miny, float minz, float inv_dx, int gx, int gy, int gz, long& k) {
    const int cx = (int)((x.x - minx) * inv_dx), cy = (int)((x.y - miny) * inv_dx), cz = (int)((x.z - minz) * inv_dx);
    if (cx < 0 || cy < 0 || cz < 0 || cx >= gx || cy >= gy || cz >= gz) return false;
    k = ((long)cx * gy + cy) * gz + cz;
    return true;
}

// the diagonal shift of C toward the cell's volume-weighted divergence (:888-908)
MPMX_HD float vlc_shift(float c00, float c11, float c22, float cell_div, float cell_vol) {
    const float dc = cell_vol > 1e-20f ? cell_div / cell_vol : 0.0f;
    return (dc - ((c00 + c11) + c22)) / 3.0f;
}

}  // namespace mpmx
