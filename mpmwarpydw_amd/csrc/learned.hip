// Learned return map in the constitutive update (compute_stress_from_F_trial,
// mpmRoutines.py:229-471, with the projection of drucker_prager_return_mapping
// :360-413 predicted by the network of mlp.hpp).
//
// The projection acts on the principal trial log strains e (descending, from
// the SVD of F_trial).  With d = e - mean(e), edn = |d| + 1e-6, the trial
// pressure p = mean(2 mu e + lam sum e) and Y = (1 - D)(ys - alpha p), a state
// is plastic when 2 mu |d| > Y or ys < 0 or Y < 0, and the exact projection is
//     e' = e - (dgam / edn) d + 0.3 alpha dgam,   dgam = edn - Y / (2 mu).
// With c = max(ys, 1), eref = c / (2 mu), the overshoot phi = (2 mu |d| - Y)/c,
// y0 = Y / c and kap = 2 mu 1e-6 / c this is the classical radial return
// e0' = e - (1 - Y / (2 mu |d|)) d plus two corrections,
//     e' = e0' - eref kap s d/|d| + 0.3 alpha eref (phi + kap) (1, 1, 1),
// where s = y0 / (phi + y0 + kap) is the part of the deviator that the 1e-6
// keeps and alpha the cone coefficient that sets the dilatancy.  The network
// predicts y = (s, 20 (alpha - 0.3), 0) (alpha centred and scaled, so that the
// loss weighs its error as the dilatancy needs) from 10 features:
// 0.1 ln(phi + kap), 0.1 ln(max(y0, 1e-12)), 0.1 ln kap, d / |d|, tan(friction
// angle), 4 nu - 1, log10 E - 8 and D.  The exact map is the judge: a
// prediction is used only when both residuals
//     2 mu |dev e' - Y/(2 mu edn) d|              (yield surface and flow direction)
//     K |tr e' - tr e - 0.9 alpha dgam|           (dilatancy)
// are within tol * max(cohesion, 1); otherwise, and for Y < 0 (tension past
// the cone's apex), the plain update's projection stands.  Everything else
// (which particles are plastic, plastic strain, damage, hardening, phase
// transition, Kirchhoff stress, viscosity) is the plain update of
// mpmx_constitutive.hpp.
#include "mlp.hpp"
#include "mpmx_constitutive.hpp"
#include "mpmx_internal.hpp"

namespace mpmx {

MPMX_HD float lrm_tan_of_alpha(float a) {
    const float r3 = 1.7320508075688772f;
    float s = 3.0f * r3 * a / (2.0f + r3 * a);
    if (s > 0.999999f) s = 0.999999f;
    return s / sqrtf(1.0f - s * s);
}

// principal-space quantities of one trial state
struct Principal {
    float e[3], d[3], esum, dn, pr, Y;
    bool plastic;
};

MPMX_HD Principal principal_state(const float eps[3], float mu, float lam, float ys, float al, float D) {
    Principal s;
    s.esum = eps[0] + eps[1] + eps[2];
    const float mean = s.esum / 3.0f;
    for (int k = 0; k < 3; k++) {
        s.e[k] = eps[k];
        s.d[k] = eps[k] - mean;
    }
    s.dn = sqrtf(s.d[0] * s.d[0] + s.d[1] * s.d[1] + s.d[2] * s.d[2]);
    s.pr = (2.0f * mu / 3.0f + lam) * s.esum;
    s.Y = (1.0f - D) * (ys - al * s.pr);
    s.plastic = (2.0f * mu * s.dn > s.Y) || ys < 0.0f || s.Y < 0.0f;
    return s;
}

MPMX_HD void lrm_features(const Principal& s, float mu, float lam, float ys, float al, float acc, float D, float x[10]) {
    (void)acc;
    const float E = mu * (3.0f * lam + 2.0f * mu) / (lam + mu);
    const float nu = lam / (2.0f * (lam + mu));
    const float c = ys > 1.0f ? ys : 1.0f;
    const float kap = 2.0f * mu * 1.0e-6f / c, y0 = s.Y / c;
    const float pk = (2.0f * mu * s.dn - s.Y) / c + kap;
    x[0] = 0.1f * logf(pk > 1.0e-30f ? pk : 1.0e-30f);
    x[1] = 0.1f * logf(y0 > 1.0e-12f ? y0 : 1.0e-12f);
    x[2] = 0.1f * logf(kap);
    for (int k = 0; k < 3; k++) x[3 + k] = s.dn > 0.0f ? s.d[k] / s.dn : 0.0f;
    x[6] = lrm_tan_of_alpha(al);
    x[7] = 4.0f * nu - 1.0f;
    x[8] = log10f(E) - 8.0f;
    x[9] = D;
}

// the radial return e0' = e - (1 - Y / (2 mu |d|)) d (e for |d| = 0)
MPMX_HD void radial_return(const Principal& s, float mu, float e0[3]) {
    const float r0 = s.dn > 0.0f ? 1.0f - s.Y / (2.0f * mu * s.dn) : 0.0f;
    for (int k = 0; k < 3; k++) e0[k] = s.e[k] - r0 * s.d[k];
}

MPMX_HD bool lrm_judge(const Principal& s, const float q[3], float mu, float lam, float ys, float al, float tol) {
    const float edn = s.dn + 1.0e-6f;
    const float c = s.Y / (2.0f * mu * edn);
    const float qm = (q[0] + q[1] + q[2]) / 3.0f;
    const float r0 = q[0] - qm - c * s.d[0], r1 = q[1] - qm - c * s.d[1], r2 = q[2] - qm - c * s.d[2];
    const float res1 = 2.0f * mu * sqrtf(r0 * r0 + r1 * r1 + r2 * r2);
    const float dgam = edn - s.Y / (2.0f * mu);
    const float K = lam + 2.0f * mu / 3.0f;
    const float res2 = K * fabsf((q[0] + q[1] + q[2]) - s.esum - 3.0f * 0.3f * al * dgam);
    const float lim = tol * (ys > 1.0f ? ys : 1.0f);
    return res1 <= lim && res2 <= lim;
}

// Harvest: features, exact targets (s, 20 (alpha - 0.3), 0) and the state of every
// particle (1 plastic MPM, 0 elastic MPM, -1 not an active MPM particle);
// e_trial gets the principal trial log strains.  Reads the state, writes
// nothing of it.
__global__ void lrm_harvest_kernel(mpmx_particles P, float* __restrict__ X, float* __restrict__ Tg,
                                   float* __restrict__ E3, int32_t* __restrict__ flag) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= P.n) return;
    if (P.activeLabel[p] != 1 || P.materialLabel[p] > 1) {
        flag[p] = -1;
        return;
    }
    float Ft[9];
    ld9(P.F_trial, P.n, p, Ft);
    Svd3 sv;
    svd3(Ft, sv);
    float eps[3];
    for (int k = 0; k < 3; k++) eps[k] = log_clamped(sv.s[k], sv.lam[k], 0.01f);
    const float mu = P.mu[p], lam = P.lam[p], ys = P.ys[p], al = P.alpha[p], D = P.damage[p];
    const Principal s = principal_state(eps, mu, lam, ys, al, D);
    float x[10];
    lrm_features(s, mu, lam, ys, al, P.acc_strain[p], D, x);
    for (int k = 0; k < 10; k++) X[p * 10 + k] = x[k];
    Tg[p * 3 + 0] = s.plastic ? s.Y / (2.0f * mu * (s.dn + 1.0e-6f)) : 0.0f;
    Tg[p * 3 + 1] = s.plastic ? 20.0f * (al - 0.3f) : 0.0f;
    Tg[p * 3 + 2] = 0.0f;
    for (int k = 0; k < 3; k++) E3[p * 3 + k] = eps[k];
    flag[p] = s.plastic ? 1 : 0;
}

// One wavefront handles 64 consecutive particles per step.  Every particle
// runs the plain update's first phase (cu_begin: invariants and branch, as
// compute_stress_from_F_trial decides them).  Plastic particles that stay MPM
// are the network's candidates: they take the SVD of F_trial (if the series
// route skipped it), their principal state and features; the network runs on
// the matrix cores for the wavefront's 4 tiles of 16 whenever one lane is a
// candidate.  A judged prediction is the projection cu_finish uses; every other
// particle finishes exactly as the plain update, bit for bit.  counters:
// [0] += plastic states, [1] += accepted predictions, [2] += particles updated.
__global__ void __launch_bounds__(kBlock) stress_learned_kernel(mpmx_particles P, float eff, const float* __restrict__ prm,
                                                                float tol, int use_net, int32_t* __restrict__ counters) {
    __shared__ MlpWeights w;
    if (use_net) load_weights(prm, w);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * 4;
    int n_plastic = 0, n_accept = 0, n_done = 0;
    for (long base = wave * 64; base < P.n; base += nwaves * 64) {  // uniform per wavefront
        const long p = base + lane;
        const bool mpm = p < P.n && P.activeLabel[p] == 1 && P.materialLabel[p] <= 1;
        CuState c;
        Principal s;
        float eps[3] = {0.0f, 0.0f, 0.0f}, x[10];
        for (int k = 0; k < 10; k++) x[k] = 0.0f;
        bool cand = false;
        if (mpm) {
            cu_begin(P, p, c);
            if (use_net && c.branch == 1) {
                if (c.fast) svd3(c.Ft, c.sv);
                for (int k = 0; k < 3; k++) eps[k] = log_clamped(c.sv.s[k], c.sv.lam[k], 0.01f);
                s = principal_state(eps, c.mu, c.lam, c.ys, c.al, c.D);
                cand = s.plastic && s.Y >= 0.0f;
                if (cand) lrm_features(s, c.mu, c.lam, c.ys, c.al, P.acc_strain[p], c.D, x);
            }
        }
        float yo[3] = {0.0f, 0.0f, 0.0f};
        if (__ballot(cand)) {  // uniform: every lane runs the matrix-core tiles
#pragma unroll
            for (int t = 0; t < 4; t++) {
                // B fragment: lane l takes features 4(l/16)..+3 of particle t*16 + l%16
                const int src = t * 16 + (lane & 15);
                const int f0 = 4 * (lane >> 4);
                float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
                for (int q = 0; q < 10; q++) {
                    const float v = __shfl(x[q], src);
#pragma unroll
                    for (int j = 0; j < 4; j++)
                        if (f0 + j == q) xv[j] = v;
                }
                MlpTile tt;
                mlp_tile_forward(w, pack_bf16(xv[0], xv[1], xv[2], xv[3]), pack_bf16_lo(xv[0], xv[1], xv[2], xv[3]),
                                 lane, tt);
                // lanes 0..15 hold the outputs of particles t*16 + 0..15
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const float v = __shfl(tt.y[j], lane & 15);
                    if ((lane >> 4) == t) yo[j] = v;
                }
            }
        }
        if (mpm) {
            n_done++;
            if (c.branch != 0) n_plastic++;
            float q[3];
            bool ok = false;
            if (cand) {
                const float mu = c.mu, ys = c.ys;
                const float cc = ys > 1.0f ? ys : 1.0f;
                const float er = cc / (2.0f * mu), kap = 2.0f * mu * 1.0e-6f / cc;
                const float dil = 0.3f * (0.3f + 0.05f * yo[1]) * (er * ((2.0f * mu * s.dn - s.Y) / cc + kap));
                const float shr = s.dn > 0.0f ? er * kap * yo[0] / s.dn : 0.0f;
                radial_return(s, mu, q);
                for (int k = 0; k < 3; k++) q[k] = (q[k] - shr * s.d[k]) + dil;
                ok = lrm_judge(s, q, mu, c.lam, ys, c.al, tol);
                n_accept += ok ? 1 : 0;
            }
            float T[9];
            cu_finish(P, p, eff, c, T, ok ? q : nullptr);
        }
    }
    // one atomic per wavefront and counter
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        n_plastic += __shfl_xor(n_plastic, o);
        n_accept += __shfl_xor(n_accept, o);
        n_done += __shfl_xor(n_done, o);
    }
    if (lane == 0 && counters) {
        atomicAdd(&counters[0], n_plastic);
        atomicAdd(&counters[1], n_accept);
        atomicAdd(&counters[2], n_done);
    }
}

static int check_particles(const mpmx_particles* P) {
    MPMX_REQUIRE(P && P->n >= 0, "learned: bad particle struct");
    MPMX_REQUIRE(P->n == 0 || (P->activeLabel && P->materialLabel && P->F_trial && P->mu && P->lam && P->ys &&
                               P->alpha && P->damage && P->acc_strain),
                 "learned: null particle array");
    return 0;
}

int launch_stress_learned(const mpmx_particles& P, float eff, const float* params, float tol, int32_t* counters,
                          hipStream_t s) {
    if (P.n == 0) return 0;
    const long waves = ((long)P.n + 63) / 64;
    long nb = (waves + 3) / 4;
    if (nb > 8192) nb = 8192;
    MPMX_LAUNCH(stress_learned_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, P, eff, params, tol, params ? 1 : 0,
                counters);
    return check_launch("stress_learned", s);
}

}  // namespace mpmx

using namespace mpmx;

extern "C" {

int mpmx_lrm_harvest(const mpmx_particles* P, float* X, float* T, float* e_trial, int32_t* flag, void* stream) {
    int rc = check_particles(P);
    if (rc) return rc;
    if (P->n == 0) return 0;
    MPMX_REQUIRE(X && T && e_trial && flag, "mpmx_lrm_harvest: null output");
    hipStream_t s = as_stream(stream);
    MPMX_LAUNCH(lrm_harvest_kernel, dim3(grid_for(P->n, kBlock)), dim3(kBlock), 0, s, *P, X, T, e_trial, flag);
    return check_launch("lrm_harvest", s);
}

int mpmx_compute_stress_learned(const mpmx_particles* P, float eff, const float* params, float tol,
                                int32_t* counters, void* stream) {
    int rc = check_particles(P);
    if (rc) return rc;
    if (P->n == 0) return 0;
    MPMX_REQUIRE(P->x && P->v && P->x_init_xpbd && P->v_init_xpbd && P->F && P->C && P->stress && P->hardening &&
                     P->softening && P->density && P->strainCriteria && P->eta_shear && P->eta_bulk,
                 "mpmx_compute_stress_learned: null particle array");
    MPMX_REQUIRE(tol >= 0.0f, "mpmx_compute_stress_learned: negative tolerance");
    return launch_stress_learned(*P, eff, params, tol, counters, as_stream(stream));
}

}  // extern "C"
