// The per-particle constitutive update of compute_stress_from_F_trial
// (mpmRoutines.py:229-312) with drucker_prager_return_mapping (:314-453) and
// kirchoff_stress (:456-471), as one device function shared by the standalone
// stress kernel and the fused return-map + P2G tile kernel.
//
// One SVD per particle: the reference's second svd3 of the returned F
// (mpmRoutines.py:288) is the first one again in the elastic branch and
// U diag(exp eps) V^T in the plastic branch, and U diag(c) V^T F^T with
// F = U diag(s) V^T equals U diag(c_k s_k) U^T, so the Kirchhoff stress is
// assembled directly as U diag(d) U^T (symmetric, as the reference's
// (tau + tau^T)/2 makes it).
#pragma once
#include "mpmx_device.hpp"

namespace mpmx {

// Warp's scalar min/max builtins are a<b?a:b and a>b?a:b (NaN falls to b).
__device__ __forceinline__ float wmin(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float wmax(float a, float b) { return a > b ? a : b; }

// Runs the update for particle p, which the caller has checked to be active
// with materialLabel <= 1.  Writes F, stress, accumulated strain, damage, ys
// and, on a phase transition, the label, velocity, XPBD birth state and zeroed
// material.  Returns the stress in T (row-major) and whether the particle is
// still MPM (label <= 1) afterwards.
__device__ __forceinline__ bool constitutive_update(const mpmx_particles& P, long p, float eff, float T[9]) {
    const long n = P.n;
    const int label = P.materialLabel[p];
    float Ft[9];
    ld9(P.F_trial, n, p, Ft);
    Svd3 sv;
    svd3(Ft, sv);

    float mu = P.mu[p], lam = P.lam[p], ys = P.ys[p], al = P.alpha[p], D = P.damage[p];
    float eps[3];
#pragma unroll
    for (int k = 0; k < 3; k++) eps[k] = log_clamped(sv.s[k], sv.lam[k], 0.01f);
    float esum = eps[0] + eps[1] + eps[2];
    float mean_eps = esum / 3.0f;
    float edev[3] = {eps[0] - mean_eps, eps[1] - mean_eps, eps[2] - mean_eps};
    float tau[3];
#pragma unroll
    for (int k = 0; k < 3; k++) tau[k] = 2.0f * mu * eps[k] + lam * esum;
    float p_stress = (tau[0] + tau[1] + tau[2]) / 3.0f;
    float td0 = tau[0] - p_stress, td1 = tau[1] - p_stress, td2 = tau[2] - p_stress;
    float sigma_eq = sqrtf(td0 * td0 + td1 * td1 + td2 * td2);
    float yield_eff = (1.0f - D) * (ys - al * p_stress);

    int branch = 0;  // 0 elastic, 1 plastic projection, 2 phase transition
    float epsn[3] = {0.0f, 0.0f, 0.0f};
    if (sigma_eq > yield_eff || ys < 0.0f || yield_eff < 0.0f) {
        float edn = sqrtf(edev[0] * edev[0] + edev[1] * edev[1] + edev[2] * edev[2]) + 1.0e-6f;
        float dgam = edn - yield_eff / (2.0f * mu);
        float dep = sqrtf(2.0f / 3.0f) * dgam;
        P.acc_strain[p] = P.acc_strain[p] + dep;
        float sc = P.strainCriteria[p];
        float Dn = (sc > 0.0f) ? wmin(1.0f, D + dep / sc) : 1.0f;
        if (ys < 0.0f || yield_eff < 0.0f) Dn = 1.0f;
        P.damage[p] = Dn;
        if (Dn >= 1.0f && label == 1) {
            // phase transition to XPBD (mpmRoutines.py:379-408)
            branch = 2;
            P.materialLabel[p] = 2;
            f3 xp = ld3(P.x, n, p);
            st3(P.x_init_xpbd, n, p, xp);
            float rho = P.density[p];
            float u = 0.5f * (tau[0] * eps[0] + tau[1] * eps[1] + tau[2] * eps[2]);
            float v_expected = sqrtf(2.0f * eff * u / rho);
            f3 vp = ld3(P.v, n, p);
            f3 vdir = normalize3(f3{vp.x + 1.0e-12f, vp.y + 1.0e-12f, vp.z + 1.0e-12f});
            f3 vn = f3{vp.x + v_expected * vdir.x, vp.y + v_expected * vdir.y, vp.z + v_expected * vdir.z};
            st3(P.v, n, p, vn);
            st3(P.v_init_xpbd, n, p, vn);
            P.ys[p] = 0.0f;
            P.mu[p] = 0.0f;
            P.lam[p] = 0.0f;
            P.alpha[p] = 0.0f;
            mu = 0.0f;
            lam = 0.0f;
        } else {
            branch = 1;
            float r = dgam / edn;
            float volc = (0.3f * al) * dgam;
#pragma unroll
            for (int k = 0; k < 3; k++) epsn[k] = (eps[k] - r * edev[k]) + volc;
            P.ys[p] = wmax(0.0f, ys + 2.0f * mu * (P.hardening[p] - P.softening[p]) * dgam);
        }
    }

    // ---- deformation gradient and principal Kirchhoff stresses ----
    float d[3];
    if (branch == 1) {
        float e[3], l[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            e[k] = expf(epsn[k]);
            l[k] = (e[k] >= 1.0e-6f) ? epsn[k] : logf(1.0e-6f);
        }
        float ls = l[0] + l[1] + l[2];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float dk = 2.0f * mu * l[k] + lam * ls;
            d[k] = (e[k] >= 1.0e-6f) ? dk : dk * (e[k] / 1.0e-6f);
        }
        float Fo[9];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                Fo[i * 3 + j] = sv.U[i * 3 + 0] * e[0] * sv.V[j * 3 + 0] + sv.U[i * 3 + 1] * e[1] * sv.V[j * 3 + 1] +
                                sv.U[i * 3 + 2] * e[2] * sv.V[j * 3 + 2];
        st9(P.F, n, p, Fo);
    } else {
        float l[3];
#pragma unroll
        for (int k = 0; k < 3; k++) l[k] = log_clamped(sv.s[k], sv.lam[k], 1.0e-6f);
        float ls = l[0] + l[1] + l[2];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float dk = 2.0f * mu * l[k] + lam * ls;
            d[k] = (sv.s[k] >= 1.0e-6f) ? dk : dk * (sv.s[k] / 1.0e-6f);
        }
        st9(P.F, n, p, Ft);
    }

    // tau = U diag(d) U^T, plus Kelvin-Voigt with the APIC C as grad v (mpmRoutines.py:294-310)
    float G[9];
    ld9(P.C, n, p, G);
    float s00 = G[0], s11 = G[4], s22 = G[8];
    float s01 = 0.5f * (G[1] + G[3]), s02 = 0.5f * (G[2] + G[6]), s12 = 0.5f * (G[5] + G[7]);
    float vr = s00 + s11 + s22;
    float es = 2.0f * P.eta_shear[p], eb = P.eta_bulk[p];
    float third = vr / 3.0f;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++) {
            float t = sv.U[i * 3 + 0] * d[0] * sv.U[j * 3 + 0] + sv.U[i * 3 + 1] * d[1] * sv.U[j * 3 + 1] +
                      sv.U[i * 3 + 2] * d[2] * sv.U[j * 3 + 2];
            T[i * 3 + j] = t;
            T[j * 3 + i] = t;
        }
    T[0] += es * (s00 - third) + eb * vr;
    T[4] += es * (s11 - third) + eb * vr;
    T[8] += es * (s22 - third) + eb * vr;
    T[1] += es * s01; T[3] += es * s01;
    T[2] += es * s02; T[6] += es * s02;
    T[5] += es * s12; T[7] += es * s12;
    st9(P.stress, n, p, T);
    return branch != 2;
}

}  // namespace mpmx
