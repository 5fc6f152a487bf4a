// Per-particle and per-node MPM kernels of the hot path (gfx950):
//   compute_mu_lam_bulk_from_E_nu, compute_stress_from_F_trial (return map +
//   Kirchhoff stress + Kelvin-Voigt), grid normalisation / damping /
//   collideBounds (separate and fused), g2p.
// The deterministic P2G lives in p2g.hip, the reductions in reduce.hip.
#include "mpmx_constitutive.hpp"
#include "mpmx_device.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

// --------------------------------------------------------------------------
// compute_mu_lam_bulk_from_E_nu (mpmRoutines.py:206-226)
// --------------------------------------------------------------------------
__global__ void mu_lam_bulk_kernel(int n, const float* __restrict__ E, const float* __restrict__ nu,
                                   float* __restrict__ mu, float* __restrict__ lam, float* __restrict__ bulk) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float e = E[p], v = nu[p];
    float m = e / (2.0f * (1.0f + v));
    float l = e * v / ((1.0f + v) * (1.0f - 2.0f * v));
    mu[p] = m;
    lam[p] = l;
    bulk[p] = l + (2.0f / 3.0f) * m;
}

// --------------------------------------------------------------------------
// initialize_geostatic_F (mpmRoutines.py:52-124) and, with plane_strain,
// initialize_geostatic_F_planestrain (:128-203): F = diag(exp eps) of the
// linear-elastic strain of the geostatic stress at the particle's depth
// --------------------------------------------------------------------------
__global__ void geostatic_F_kernel(int n, const float* __restrict__ x, float* __restrict__ F,
                                   const float* __restrict__ mu, const float* __restrict__ lam,
                                   const float* __restrict__ density, float g, float z_top, float K0,
                                   int plane_strain) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float depth = z_top - x[2L * n + p];
    float exx = 0.0f, eyy = 0.0f, ezz = 0.0f;
    if (!plane_strain) {
        const float sv = -density[p] * g * depth;
        const float sh = K0 * sv;
        const float m = mu[p], l = lam[p];
        const float a11 = 2.0f * l + 2.0f * m, a12 = l, a21 = 2.0f * l, a22 = l + 2.0f * m;
        const float det = a11 * a22 - a12 * a21;
        if (fabsf(det) > 1e-12f) {
            exx = eyy = (sh * a22 - sv * a12) / det;  // Cramer's rule on the (eps_h, eps_v) system
            ezz = (a11 * sv - a21 * sh) / det;
        }
    } else if (depth > 0.0f) {
        const float m = mu[p], l = lam[p];
        const float nu = l / (2.0f * (l + m));
        const float E = m * (3.0f * l + 2.0f * m) / (l + m);
        const float sv = -density[p] * g * depth;
        const float sh = K0 * sv;
        const float syy = nu * (sh + sv);  // eps_yy = 0
        exx = (sh - nu * (syy + sv)) / E;
        ezz = (sv - nu * (sh + syy)) / E;
    }
    const float d[3] = {expf(exx), expf(eyy), expf(ezz)};
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) F[(long)(3 * r + c) * n + p] = r == c ? d[r] : 0.0f;
}

// --------------------------------------------------------------------------
// deactivate_and_teleport_xpbd_below_datum (mpmRoutines.py:1446-1475)
// --------------------------------------------------------------------------
__global__ void deactivate_teleport_kernel(int n, float* __restrict__ x, float* __restrict__ v,
                                           int32_t* __restrict__ activeLabel,
                                           const int32_t* __restrict__ materialLabel, float z_datum,
                                           float teleport) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || activeLabel[p] != 1 || materialLabel[p] != 2) return;
    const float z = x[2L * n + p];
    if (z < z_datum) {
        activeLabel[p] = 0;
        x[2L * n + p] = z - teleport;
        v[p] = 0.0f;
        v[(long)n + p] = 0.0f;
        v[2L * n + p] = 0.0f;
    }
}

// --------------------------------------------------------------------------
// compute_stress_from_F_trial (mpmRoutines.py:229-312), one thread per particle
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) stress_kernel(mpmx_particles P, float eff) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    if (P.materialLabel[p] > 1 || P.activeLabel[p] != 1) return;
    float T[9];
    constitutive_update(P, p, eff, T);
}

// --------------------------------------------------------------------------
// grid kernels (mpmRoutines.py:691-711, 915-993)
// --------------------------------------------------------------------------
__global__ void grid_norm_kernel(long G, const float* __restrict__ m, const float* __restrict__ vin,
                                 float* __restrict__ vout, float gx, float gy, float gz, float dt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    float mm = m[i];
    if (mm > 1.0e-15f) {
        float inv = 1.0f / mm;
        vout[i] = vin[i] * inv + dt * gx;
        vout[G + i] = vin[G + i] * inv + dt * gy;
        vout[2 * G + i] = vin[2 * G + i] * inv + dt * gz;
    }
}

__global__ void grid_damp_kernel(long G, float* __restrict__ vout, float scale) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3 * G) return;
    vout[i] = vout[i] * scale;
}

__global__ void collide_kernel(float* __restrict__ vout, int gx, int gy, int gz, float mu, float dt, float g,
                               int pad) {
    long G = (long)gx * gy * gz;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    int iz = (int)(i % gz);
    int iy = (int)((i / gz) % gy);
    int ix = (int)(i / ((long)gy * gz));
    // interior nodes are never touched: skip the load
    if (ix > pad && ix < gx - pad && iy > pad && iy < gy - pad && iz > pad && iz < gz - pad) return;
    f3 v = ld3(vout, G, i);
    f3 o = collide_node(v, ix, iy, iz, gx, gy, gz, mu, dt, g, pad);
    st3(vout, G, i, o);
}

}  // namespace mpmx

namespace mpmx {
__global__ void grid_update_kernel(int gx, int gy, int gz, const float* __restrict__ m, const float* __restrict__ vin,
                                   float* __restrict__ vout, float gxv, float gyv, float gzv, float dt, float damping,
                                   int bc, float fric, float gmag, int pad, float restitution) {
    long G = (long)gx * gy * gz;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    int iz = (int)(i % gz);
    int iy = (int)((i / gz) % gy);
    int ix = (int)(i / ((long)gy * gz));
    f3 v = grid_update_node(m[i], ld3(vin, G, i), ix, iy, iz, gx, gy, gz, gxv, gyv, gzv, dt, damping, bc, fric,
                            gmag, pad, restitution);
    st3(vout, G, i, v);
}

// the boundary-condition kernels other than collideBounds, on every node
__global__ void collide_variant_kernel(float* __restrict__ vout, int gx, int gy, int gz, int bc, float mu, float dt,
                                       float g, int pad, float restitution) {
    long G = (long)gx * gy * gz;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    int iz = (int)(i % gz);
    int iy = (int)((i / gz) % gy);
    int ix = (int)(i / ((long)gy * gz));
    st3(vout, G, i, apply_bc(ld3(vout, G, i), ix, iy, iz, gx, gy, gz, bc, mu, dt, g, pad, restitution));
}

// --------------------------------------------------------------------------
// volumetric locking correction (mpmRoutines.py:805-911)
// --------------------------------------------------------------------------
__global__ void vlc_accumulate_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ label,
                                      const float* __restrict__ x, const float* __restrict__ C,
                                      const float* __restrict__ vol, float inv_dx, float minx, float miny, float minz,
                                      int gx, int gy, int gz, unsigned long long* __restrict__ acc) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || active[p] != 1 || label[p] > 1) return;
    long k;
    if (!vlc_cell(ld3(x, n, p), minx, miny, minz, inv_dx, gx, gy, gz, k)) return;
    const float div = (C[p] + C[4L * n + p]) + C[8L * n + p];
    const long G = (long)gx * gy * gz;
    atomicAdd(&acc[k], vlc_fixed(div, vol[p]));
    atomicAdd(&acc[G + k], vlc_fixed(vol[p], 1.0f));
}

__global__ void vlc_finalize_kernel(long G, unsigned long long* __restrict__ acc, float* __restrict__ cell_div,
                                    float* __restrict__ cell_vol) {
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G) return;
    cell_div[k] = vlc_unfixed(acc[k]);
    cell_vol[k] = vlc_unfixed(acc[G + k]);
}

__global__ void vlc_apply_kernel(int n, const int32_t* __restrict__ active, const int32_t* __restrict__ label,
                                 const float* __restrict__ x, float* __restrict__ C, float inv_dx, float minx,
                                 float miny, float minz, int gx, int gy, int gz, const float* __restrict__ cell_div,
                                 const float* __restrict__ cell_vol) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || active[p] != 1 || label[p] > 1) return;
    long k;
    if (!vlc_cell(ld3(x, n, p), minx, miny, minz, inv_dx, gx, gy, gz, k)) return;
    const float c00 = C[p], c11 = C[4L * n + p], c22 = C[8L * n + p];
    const float corr = vlc_shift(c00, c11, c22, cell_div[k], cell_vol[k]);
    C[p] = c00 + corr;
    C[4L * n + p] = c11 + corr;
    C[8L * n + p] = c22 + corr;
}

__global__ void zero_u64_kernel(unsigned long long* __restrict__ p, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

// --------------------------------------------------------------------------
// g2p + update_cov (mpmRoutines.py:715-802)
// --------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) g2p_kernel(
    float dt, int n, const int32_t* __restrict__ activeLabel, const int32_t* __restrict__ materialLabel,
    float* __restrict__ x, float* __restrict__ v, float* __restrict__ C, const float* __restrict__ F,
    float* __restrict__ F_trial, float* __restrict__ cov, float inv_dx, const float* __restrict__ gv, int gx, int gy,
    int gz, int update_cov, float minx, float miny, float minz) {
    int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (activeLabel[p] != 1 || materialLabel[p] > 1) return;
    f3 xp = ld3(x, n, p);
    float fx[3];
    int b0 = base_axis(xp.x, minx, inv_dx, fx[0]);
    int b1 = base_axis(xp.y, miny, inv_dx, fx[1]);
    int b2 = base_axis(xp.z, minz, inv_dx, fx[2]);
    if (b0 < 0 || b1 < 0 || b2 < 0 || b0 > gx - 3 || b1 > gy - 3 || b2 > gz - 3) return;
    const long G = (long)gx * gy * gz;
    float nv[3], nB[9], nG[9];
    g2p_gather(fx, inv_dx,
               [&](int i, int j, int k) {
                   const long idx = ((long)(b0 + i) * gy + (b1 + j)) * gz + b2 + k;
                   return f3{gv[idx], gv[G + idx], gv[2 * G + idx]};
               },
               nv, nB, nG);
    st3(v, n, p, f3{nv[0], nv[1], nv[2]});
    st3(x, n, p, f3{xp.x + dt * nv[0], xp.y + dt * nv[1], xp.z + dt * nv[2]});
    float sC = inv_dx * 4.0f;
    float Cn[9];
#pragma unroll
    for (int k = 0; k < 9; k++) Cn[k] = nB[k] * sC;
    st9(C, n, p, Cn);
    // F_trial = (I + dt G) F, evaluated as F + dt (G F)
    float Fp[9], Fn[9];
    ld9(F, n, p, Fp);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            Fn[i * 3 + j] = Fp[i * 3 + j] +
                            dt * (nG[i * 3 + 0] * Fp[0 * 3 + j] + nG[i * 3 + 1] * Fp[1 * 3 + j] + nG[i * 3 + 2] * Fp[2 * 3 + j]);
    st9(F_trial, n, p, Fn);
    if (update_cov) {
        float* c6 = cov + 6L * p;
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
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

int mpmx_compute_mu_lam_bulk(int n, const float* E, const float* nu, float* mu, float* lam, float* bulk,
                             void* stream) {
    MPMX_REQUIRE(n >= 0, "mpmx_compute_mu_lam_bulk: n < 0");
    if (n == 0) return 0;
    MPMX_REQUIRE(E && nu && mu && lam && bulk, "mpmx_compute_mu_lam_bulk: null pointer");
    MPMX_LAUNCH(mu_lam_bulk_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), n, E, nu,
                       mu, lam, bulk);
    return check_launch("compute_mu_lam_bulk", as_stream(stream));
}

int mpmx_initialize_geostatic_F(int n, const float* x, float* F, const float* mu, const float* lam,
                                const float* density, float gravity, float z_top, float K0, int plane_strain,
                                void* stream) {
    MPMX_REQUIRE(n >= 0, "mpmx_initialize_geostatic_F: n < 0");
    if (n == 0) return 0;
    MPMX_REQUIRE(x && F && mu && lam && density, "mpmx_initialize_geostatic_F: null pointer");
    MPMX_LAUNCH(geostatic_F_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), n, x, F, mu, lam,
                density, gravity, z_top, K0, plane_strain ? 1 : 0);
    return check_launch("initialize_geostatic_F", as_stream(stream));
}

int mpmx_deactivate_and_teleport_xpbd_below_datum(int n, float* x, float* v, int32_t* activeLabel,
                                                  const int32_t* materialLabel, float z_datum,
                                                  float teleport_distance, void* stream) {
    MPMX_REQUIRE(n >= 0, "mpmx_deactivate_and_teleport_xpbd_below_datum: n < 0");
    if (n == 0) return 0;
    MPMX_REQUIRE(x && v && activeLabel && materialLabel, "mpmx_deactivate_and_teleport_xpbd_below_datum: null pointer");
    MPMX_LAUNCH(deactivate_teleport_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), n, x, v,
                activeLabel, materialLabel, z_datum, teleport_distance);
    return check_launch("deactivate_and_teleport_xpbd_below_datum", as_stream(stream));
}

int mpmx_compute_stress_from_F_trial(int n, const int32_t* activeLabel, int32_t* materialLabel, const float* x,
                                     float* v, float* x_init, float* v_init, float* F, const float* F_trial,
                                     float* mu, float* lam, float* ys, const float* hardening,
                                     const float* softening, const float* density, const float* strainCriteria,
                                     float eff, const float* eta_shear, const float* eta_bulk, const float* C,
                                     float* stress, float* acc_strain, float* damage, int constitutive_model,
                                     float* alpha, void* stream) {
    (void)constitutive_model;  // passed but never read by the reference kernel (mpmRoutines.py:253)
    MPMX_REQUIRE(n >= 0, "mpmx_compute_stress_from_F_trial: n < 0");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && x && v && x_init && v_init && F && F_trial && mu && lam && ys &&
                     hardening && softening && density && strainCriteria && eta_shear && eta_bulk && C && stress &&
                     acc_strain && damage && alpha,
                 "mpmx_compute_stress_from_F_trial: null pointer");
    mpmx_particles P{};
    P.n = n;
    P.activeLabel = const_cast<int32_t*>(activeLabel);
    P.materialLabel = materialLabel;
    P.x = const_cast<float*>(x);
    P.v = v;
    P.x_init_xpbd = x_init;
    P.v_init_xpbd = v_init;
    P.F = F;
    P.F_trial = const_cast<float*>(F_trial);
    P.C = const_cast<float*>(C);
    P.stress = stress;
    P.mu = mu;
    P.lam = lam;
    P.ys = ys;
    P.alpha = alpha;
    P.hardening = const_cast<float*>(hardening);
    P.softening = const_cast<float*>(softening);
    P.density = const_cast<float*>(density);
    P.strainCriteria = const_cast<float*>(strainCriteria);
    P.eta_shear = const_cast<float*>(eta_shear);
    P.eta_bulk = const_cast<float*>(eta_bulk);
    P.acc_strain = acc_strain;
    P.damage = damage;
    MPMX_LAUNCH(stress_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), P, eff);
    return check_launch("compute_stress_from_F_trial", as_stream(stream));
}

int mpmx_grid_normalization_and_gravity(int gx, int gy, int gz, const float* grid_m, const float* grid_v_in,
                                        float* grid_v_out, float gxv, float gyv, float gzv, float dt, void* stream) {
    MPMX_REQUIRE(gx > 0 && gy > 0 && gz > 0 && grid_m && grid_v_in && grid_v_out,
                 "mpmx_grid_normalization_and_gravity: bad arguments");
    long G = (long)gx * gy * gz;
    MPMX_LAUNCH(grid_norm_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, as_stream(stream), G, grid_m,
                       grid_v_in, grid_v_out, gxv, gyv, gzv, dt);
    return check_launch("grid_normalization_and_gravity", as_stream(stream));
}

int mpmx_add_damping_via_grid(int gx, int gy, int gz, float* grid_v_out, float scale, void* stream) {
    MPMX_REQUIRE(gx > 0 && gy > 0 && gz > 0 && grid_v_out, "mpmx_add_damping_via_grid: bad arguments");
    long G = (long)gx * gy * gz;
    MPMX_LAUNCH(grid_damp_kernel, dim3(grid_for(3 * G, kBlock)), dim3(kBlock), 0, as_stream(stream), G,
                       grid_v_out, scale);
    return check_launch("add_damping_via_grid", as_stream(stream));
}

int mpmx_collide_bounds(float* grid_v_out, int gx, int gy, int gz, float friction, float dt, float gmag,
                        int boundary_padding, void* stream) {
    MPMX_REQUIRE(gx > 0 && gy > 0 && gz > 0 && grid_v_out, "mpmx_collide_bounds: bad arguments");
    long G = (long)gx * gy * gz;
    MPMX_LAUNCH(collide_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, as_stream(stream), grid_v_out, gx,
                       gy, gz, friction, dt, gmag, boundary_padding);
    return check_launch("collideBounds", as_stream(stream));
}

int mpmx_grid_update(int gx, int gy, int gz, const float* grid_m, const float* grid_v_in, float* grid_v_out,
                     float gxv, float gyv, float gzv, float dt, float damping_scale, int bc, float friction,
                     float gmag, int boundary_padding, float restitution, void* stream) {
    MPMX_REQUIRE(gx > 0 && gy > 0 && gz > 0 && grid_m && grid_v_in && grid_v_out, "mpmx_grid_update: bad arguments");
    MPMX_REQUIRE(bc >= MPMX_BC_NONE && bc <= MPMX_BC_PLANE_STRAIN_Y_GLOBAL,
                 "mpmx_grid_update: unsupported boundary condition %d", bc);
    long G = (long)gx * gy * gz;
    MPMX_LAUNCH(grid_update_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, as_stream(stream), gx, gy, gz,
                       grid_m, grid_v_in, grid_v_out, gxv, gyv, gzv, dt, damping_scale, bc, friction, gmag,
                       boundary_padding, restitution);
    return check_launch("grid_update", as_stream(stream));
}

int mpmx_collide_bounds_variant(float* grid_v_out, int gx, int gy, int gz, int bc, float friction, float dt,
                                float gmag, int boundary_padding, float restitution, void* stream) {
    MPMX_REQUIRE(gx > 0 && gy > 0 && gz > 0 && grid_v_out, "mpmx_collide_bounds_variant: bad arguments");
    MPMX_REQUIRE(bc >= MPMX_BC_FRICTION && bc <= MPMX_BC_PLANE_STRAIN_Y_GLOBAL,
                 "mpmx_collide_bounds_variant: unsupported boundary condition %d", bc);
    long G = (long)gx * gy * gz;
    MPMX_LAUNCH(collide_variant_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, as_stream(stream), grid_v_out, gx,
                gy, gz, bc, friction, dt, gmag, boundary_padding, restitution);
    return check_launch("collide_bounds_variant", as_stream(stream));
}

size_t mpmx_vlc_workspace_bytes(int gx, int gy, int gz) {
    if (gx <= 0 || gy <= 0 || gz <= 0) return 0;
    return Carver::bytes_for(2L * gx * gy * gz, sizeof(unsigned long long));
}

int mpmx_accumulate_cell_divergence(int n, const int32_t* activeLabel, const int32_t* materialLabel, const float* x,
                                    const float* C, const float* vol, float inv_dx, float minx, float miny,
                                    float minz, int gx, int gy, int gz, float* cell_div, float* cell_vol,
                                    void* workspace, size_t ws_bytes, void* stream) {
    MPMX_REQUIRE(n >= 0 && gx > 0 && gy > 0 && gz > 0 && cell_div && cell_vol,
                 "mpmx_accumulate_cell_divergence: bad arguments");
    MPMX_REQUIRE(n == 0 || (activeLabel && materialLabel && x && C && vol),
                 "mpmx_accumulate_cell_divergence: null particle array");
    MPMX_REQUIRE(workspace && ws_bytes >= mpmx_vlc_workspace_bytes(gx, gy, gz),
                 "mpmx_accumulate_cell_divergence: workspace too small");
    hipStream_t s = as_stream(stream);
    const long G = (long)gx * gy * gz;
    unsigned long long* acc = static_cast<unsigned long long*>(workspace);
    MPMX_LAUNCH(zero_u64_kernel, dim3(grid_for(2 * G, kBlock)), dim3(kBlock), 0, s, acc, 2 * G);
    int rc = check_launch("vlc_zero", s);
    if (rc) return rc;
    if (n > 0) {
        MPMX_LAUNCH(vlc_accumulate_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, activeLabel,
                    materialLabel, x, C, vol, inv_dx, minx, miny, minz, gx, gy, gz, acc);
        if ((rc = check_launch("vlc_accumulate", s))) return rc;
    }
    MPMX_LAUNCH(vlc_finalize_kernel, dim3(grid_for(G, kBlock)), dim3(kBlock), 0, s, G, acc, cell_div, cell_vol);
    return check_launch("vlc_finalize", s);
}

int mpmx_apply_volumetric_locking_correction(int n, const int32_t* activeLabel, const int32_t* materialLabel,
                                             const float* x, float* C, float inv_dx, float minx, float miny,
                                             float minz, int gx, int gy, int gz, const float* cell_div,
                                             const float* cell_vol, void* stream) {
    MPMX_REQUIRE(n >= 0 && gx > 0 && gy > 0 && gz > 0 && cell_div && cell_vol,
                 "mpmx_apply_volumetric_locking_correction: bad arguments");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && x && C, "mpmx_apply_volumetric_locking_correction: null pointer");
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(vlc_apply_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, s, n, activeLabel, materialLabel, x, C,
                inv_dx, minx, miny, minz, gx, gy, gz, cell_div, cell_vol);
    return check_launch("vlc_apply", s);
}

int mpmx_g2p(float dt, int n, const int32_t* activeLabel, const int32_t* materialLabel, float* x, float* v,
             float* C, const float* F, float* F_trial, float* cov, float inv_dx, const float* grid_v_out, int gx,
             int gy, int gz, int update_cov, float minx, float miny, float minz, void* stream) {
    MPMX_REQUIRE(n >= 0 && gx > 0 && gy > 0 && gz > 0, "mpmx_g2p: bad sizes");
    if (n == 0) return 0;
    MPMX_REQUIRE(activeLabel && materialLabel && x && v && C && F && F_trial && grid_v_out && (!update_cov || cov),
                 "mpmx_g2p: null pointer");
    MPMX_LAUNCH(g2p_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, as_stream(stream), dt, n, activeLabel,
                       materialLabel, x, v, C, F, F_trial, cov, inv_dx, grid_v_out, gx, gy, gz, update_cov, minx,
                       miny, minz);
    return check_launch("g2p", as_stream(stream));
}

}  // extern "C"
