// The per-particle constitutive update of compute_stress_from_F_trial
// (mpmRoutines.py:229-312) with drucker_prager_return_mapping (:314-453) and
// kirchoff_stress (:456-471), as one device function shared by the standalone
// stress kernel and the fused return-map + P2G tile kernel.
//
// The reference takes two SVDs per particle: one of F_trial for the principal
// log strains eps = ln max(s, 0.01), one of the returned F for the Kirchhoff
// stress tau = U diag(2 mu ln s + lam sum ln s) U^T.  Both maps are isotropic
// functions of the left stretch, so they are evaluated here without a
// decomposition whenever F_trial is close to a rotation:
//
//   b - I = F F^T - I                          (compensated FP32 dot products)
//   eps_m = 1/2 log(I + (b - I))               (Horner series, terms chosen
//                                                from |b - I|_F)
//   elastic:  F = F_trial,  tau = 2 mu eps_m + lam tr(eps_m) I
//   plastic:  eps_m' = eps_m - r dev(eps_m) + c I  (the reference's projection
//             e' = e - (dgam/edn) dev e + 0.3 alpha dgam, eigenvector by
//             eigenvector), and because eps_m' and eps_m share eigenvectors,
//             F = U exp(e') V^T = exp(eps_m' - eps_m) F_trial
//               = e^c exp(-r dev eps_m) F_trial   (series for the exponential)
//
// The invariants the return map needs (tr eps, |dev eps|, tau : eps) do not
// depend on the eigenvectors.  The fast route is taken when det F_trial > 0
// and |b - I|_F <= 0.25 (principal stretches within about 12% of 1), which
// keeps the logarithm clamps of the reference (s < 0.01, s < 1e-6) out of
// reach; every other particle takes the SVD route (one SVD: the second one of
// the reference is the first again in the elastic branch and U diag(exp e') V^T
// in the plastic branch, and U diag(c) V^T F^T = U diag(c_k s_k) U^T).
#pragma once
#include "mpmx_device.hpp"

namespace mpmx {

// Warp's scalar min/max builtins are a<b?a:b and a>b?a:b (NaN falls to b).
MPMX_HD float wmin(float a, float b) { return a < b ? a : b; }
MPMX_HD float wmax(float a, float b) { return a > b ? a : b; }

// symmetric 3x3 matrices as {xx, yy, zz, xy, xz, yz}
// C = A B for commuting symmetric A, B (polynomials in one matrix), so C is symmetric
MPMX_HD void sym_mul(const float A[6], const float B[6], float C[6]) {
    C[0] = A[0] * B[0] + A[3] * B[3] + A[4] * B[4];
    C[1] = A[3] * B[3] + A[1] * B[1] + A[5] * B[5];
    C[2] = A[4] * B[4] + A[5] * B[5] + A[2] * B[2];
    C[3] = A[0] * B[3] + A[3] * B[1] + A[4] * B[5];
    C[4] = A[0] * B[4] + A[3] * B[5] + A[4] * B[2];
    C[5] = A[3] * B[4] + A[1] * B[5] + A[5] * B[2];
}

MPMX_HD float sym_norm(const float A[6]) {
    return sqrtf(A[0] * A[0] + A[1] * A[1] + A[2] * A[2] + 2.0f * (A[3] * A[3] + A[4] * A[4] + A[5] * A[5]));
}

// log(I + B) = sum_k (-1)^(k+1) B^k / k for |B|_F = nb <= 0.25, by Horner's rule
// with the number of terms m such that the first omitted term, nb^m / (m + 1)
// relative to B, is below 2^-26.
MPMX_HD void sym_log1p(const float B[6], float nb, float L[6]) {
    const int m = nb <= 1.0e-4f ? 2 : nb <= 1.0e-3f ? 3 : nb <= 1.0e-2f ? 4 : nb <= 3.0e-2f ? 5 : nb <= 0.1f ? 8 : 13;
    float X[6] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 13; k >= 1; k--) {
        if (k > m) continue;
        const float ck = ((k & 1) ? 1.0f : -1.0f) / (float)k;
        if (k == m) {
            X[0] = ck; X[1] = ck; X[2] = ck;
        } else {
            float Y[6];
            sym_mul(B, X, Y);
            X[0] = Y[0] + ck; X[1] = Y[1] + ck; X[2] = Y[2] + ck;
            X[3] = Y[3]; X[4] = Y[4]; X[5] = Y[5];
        }
    }
    sym_mul(B, X, L);
}

// exp(S) = I + S (I + S/2 (I + S/3 (...))) for |S|_F = ns <= 1 with the first
// omitted term ns^(m+1) / (m+1)! below 2^-26.
MPMX_HD void sym_exp(const float S[6], float ns, float X[6]) {
    const int m = ns <= 1.0e-4f ? 1 : ns <= 1.0e-3f ? 2 : ns <= 1.0e-2f ? 3 : ns <= 0.1f ? 5 : ns <= 0.5f ? 8 : 12;
    X[0] = 1.0f; X[1] = 1.0f; X[2] = 1.0f; X[3] = 0.0f; X[4] = 0.0f; X[5] = 0.0f;
#pragma unroll
    for (int k = 12; k >= 1; k--) {
        if (k > m) continue;
        const float ik = 1.0f / (float)k;
        float Y[6];
        sym_mul(S, X, Y);
        X[0] = 1.0f + Y[0] * ik; X[1] = 1.0f + Y[1] * ik; X[2] = 1.0f + Y[2] * ik;
        X[3] = Y[3] * ik; X[4] = Y[4] * ik; X[5] = Y[5] * ik;
    }
}

// b - I = F F^T - I (symmetric, {xx,yy,zz,xy,xz,yz}) and its Frobenius norm.
// Formed by the compensated dot product (dot3_plus), so the cancellation
// against I (F holds rotations of order 1) costs nothing and b - I keeps the
// relative precision of the strain, in FP32 arithmetic.  Returns whether the
// series route applies: det F > 0 and |b - I|_F <= 0.25.
MPMX_HD bool left_strain(const float Ft[9], float Bs[6], float& nb) {
    const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; q++) {
        const int i = ii[q], j = jj[q];
        Bs[q] = dot3_plus(Ft[i * 3 + 0], Ft[j * 3 + 0], Ft[i * 3 + 1], Ft[j * 3 + 1], Ft[i * 3 + 2], Ft[j * 3 + 2],
                          i == j ? -1.0f : 0.0f);
    }
    nb = sym_norm(Bs);
    const float det = Ft[0] * (Ft[4] * Ft[8] - Ft[5] * Ft[7]) - Ft[1] * (Ft[3] * Ft[8] - Ft[5] * Ft[6]) +
                      Ft[2] * (Ft[3] * Ft[7] - Ft[4] * Ft[6]);
    return det > 0.0f && nb <= 0.25f;
}

// State of one particle's update between its two phases: the strain
// invariants and the branch (cu_begin), then the writes (cu_finish).
struct CuState {
    float Ft[9];
    bool fast;     // series route (eps_m in E6) or SVD route (eps, sv)
    float E6[6];   // fast route: eps_m
    float eps[3];  // SVD route: principal log strains
    Svd3 sv;
    float esum, devn, enorm2;
    float mu, lam, ys, al, D;
    int label;
    int branch;  // 0 elastic, 1 plastic projection, 2 phase transition
    float yield_eff, dgam, Dn;
};

// First phase of the update of particle p (active, materialLabel <= 1): reads
// F_trial and the material, computes the strain invariants and decides the
// branch.  Writes nothing.
MPMX_HD void cu_begin(const mpmx_particles& P, long p, CuState& c) {
    const long n = P.n;
    c.label = P.materialLabel[p];
    ld9(P.F_trial, n, p, c.Ft);

    float Bs[6], nb;
    c.fast = left_strain(c.Ft, Bs, nb);

    c.mu = P.mu[p];
    c.lam = P.lam[p];
    c.ys = P.ys[p];
    c.al = P.alpha[p];
    c.D = P.damage[p];

    // strain invariants: trace, |dev eps| and |eps|^2
    if (c.fast) {
        float L[6];
        sym_log1p(Bs, nb, L);
#pragma unroll
        for (int q = 0; q < 6; q++) c.E6[q] = 0.5f * L[q];
        c.esum = c.E6[0] + c.E6[1] + c.E6[2];
        const float m3 = c.esum / 3.0f;
        const float a = c.E6[0] - m3, b = c.E6[1] - m3, d = c.E6[2] - m3;
        const float off = c.E6[3] * c.E6[3] + c.E6[4] * c.E6[4] + c.E6[5] * c.E6[5];
        c.devn = sqrtf(a * a + b * b + d * d + 2.0f * off);
        c.enorm2 = c.E6[0] * c.E6[0] + c.E6[1] * c.E6[1] + c.E6[2] * c.E6[2] + 2.0f * off;
    } else {
        svd3(c.Ft, c.sv);
#pragma unroll
        for (int k = 0; k < 3; k++) c.eps[k] = log_clamped(c.sv.s[k], c.sv.lam[k], 0.01f);
        c.esum = c.eps[0] + c.eps[1] + c.eps[2];
        const float m3 = c.esum / 3.0f;
        const float a = c.eps[0] - m3, b = c.eps[1] - m3, d = c.eps[2] - m3;
        c.devn = sqrtf(a * a + b * b + d * d);
        c.enorm2 = c.eps[0] * c.eps[0] + c.eps[1] * c.eps[1] + c.eps[2] * c.eps[2];
    }
    // mean Kirchhoff stress and its deviatoric norm: tau_k = 2 mu eps_k + lam sum(eps)
    const float p_stress = (2.0f * c.mu / 3.0f + c.lam) * c.esum;
    const float sigma_eq = 2.0f * c.mu * c.devn;
    c.yield_eff = (1.0f - c.D) * (c.ys - c.al * p_stress);

    c.branch = 0;
    c.dgam = 0.0f;
    c.Dn = c.D;
    if (sigma_eq > c.yield_eff || c.ys < 0.0f || c.yield_eff < 0.0f) {
        const float edn = c.devn + 1.0e-6f;
        c.dgam = edn - c.yield_eff / (2.0f * c.mu);
        const float dep = sqrtf(2.0f / 3.0f) * c.dgam;
        const float sc = P.strainCriteria[p];
        c.Dn = (sc > 0.0f) ? wmin(1.0f, c.D + dep / sc) : 1.0f;
        if (c.ys < 0.0f || c.yield_eff < 0.0f) c.Dn = 1.0f;
        c.branch = (c.Dn >= 1.0f && c.label == 1) ? 2 : 1;
    }
}

// Second phase: writes F, stress, accumulated strain, damage, ys and, on a
// phase transition, the label, velocity, XPBD birth state and zeroed material.
// The plastic branch projects e' = e - r dev(e) + volc with the reference's
// r = dgam / edn and volc = 0.3 alpha dgam, or with proj = {r, volc} when
// given (plastic branch only).  Returns the stress in T (row-major) and
// whether the particle is still MPM (label <= 1) afterwards.
MPMX_HD bool cu_finish(const mpmx_particles& P, long p, float eff, CuState& c, float T[9],
                       const float* proj = nullptr) {
    const long n = P.n;
    float mu = c.mu, lam = c.lam;
    const int branch = c.branch;
    float r = 0.0f, volc = 0.0f;
    if (branch != 0) {
        P.acc_strain[p] = P.acc_strain[p] + sqrtf(2.0f / 3.0f) * c.dgam;
        P.damage[p] = c.Dn;
        if (branch == 2) {
            // phase transition to XPBD (mpmRoutines.py:379-408); u = tau : eps / 2
            P.materialLabel[p] = 2;
            f3 xp = ld3(P.x, n, p);
            st3(P.x_init_xpbd, n, p, xp);
            const float rho = P.density[p];
            const float u = mu * c.enorm2 + 0.5f * lam * c.esum * c.esum;
            const float v_expected = sqrtf(2.0f * eff * u / rho);
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
            const float edn = c.devn + 1.0e-6f;
            r = c.dgam / edn;
            volc = (0.3f * c.al) * c.dgam;
            if (proj) {
                r = proj[0];
                volc = proj[1];
            }
            P.ys[p] = wmax(0.0f, c.ys + 2.0f * mu * (P.hardening[p] - P.softening[p]) * c.dgam);
        }
    }

    // ---- deformation gradient and Kirchhoff stress ----
    float tau[6];
    if (c.fast) {
        float* E6 = c.E6;
        float esum = c.esum;
        if (branch == 1) {
            const float m3 = esum / 3.0f;
            float S[6] = {-r * (E6[0] - m3), -r * (E6[1] - m3), -r * (E6[2] - m3), -r * E6[3], -r * E6[4], -r * E6[5]};
            float X[6];
            sym_exp(S, sym_norm(S), X);
            const float ec = expf(volc);
            const float Xf[9] = {X[0], X[3], X[4], X[3], X[1], X[5], X[4], X[5], X[2]};
            const float* Ft = c.Ft;
            float Fo[9];
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++)
                    Fo[i * 3 + j] = ec * (Xf[i * 3 + 0] * Ft[0 * 3 + j] + Xf[i * 3 + 1] * Ft[1 * 3 + j] +
                                          Xf[i * 3 + 2] * Ft[2 * 3 + j]);
            st9(P.F, n, p, Fo);
#pragma unroll
            for (int q = 0; q < 6; q++) E6[q] += S[q];
            E6[0] += volc; E6[1] += volc; E6[2] += volc;
            esum += 3.0f * volc;
        } else {
            st9(P.F, n, p, c.Ft);
        }
        const float ls = lam * esum;
        tau[0] = 2.0f * mu * E6[0] + ls;
        tau[1] = 2.0f * mu * E6[1] + ls;
        tau[2] = 2.0f * mu * E6[2] + ls;
        tau[3] = 2.0f * mu * E6[3];
        tau[4] = 2.0f * mu * E6[4];
        tau[5] = 2.0f * mu * E6[5];
    } else {
        const Svd3& sv = c.sv;
        float d[3];
        if (branch == 1) {
            float epsn[3], e[3], l[3];
            const float mean_eps = c.esum / 3.0f;
#pragma unroll
            for (int k = 0; k < 3; k++) epsn[k] = (c.eps[k] - r * (c.eps[k] - mean_eps)) + volc;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                e[k] = expf(epsn[k]);
                l[k] = (e[k] >= 1.0e-6f) ? epsn[k] : logf(1.0e-6f);
            }
            const float ls = l[0] + l[1] + l[2];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dk = 2.0f * mu * l[k] + lam * ls;
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
            const float ls = l[0] + l[1] + l[2];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float dk = 2.0f * mu * l[k] + lam * ls;
                d[k] = (sv.s[k] >= 1.0e-6f) ? dk : dk * (sv.s[k] / 1.0e-6f);
            }
            st9(P.F, n, p, c.Ft);
        }
        const int ii[6] = {0, 1, 2, 0, 0, 1}, jj[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
        for (int q = 0; q < 6; q++) {
            const int i = ii[q], j = jj[q];
            tau[q] = sv.U[i * 3 + 0] * d[0] * sv.U[j * 3 + 0] + sv.U[i * 3 + 1] * d[1] * sv.U[j * 3 + 1] +
                     sv.U[i * 3 + 2] * d[2] * sv.U[j * 3 + 2];
        }
    }

    // plus Kelvin-Voigt with the APIC C as grad v (mpmRoutines.py:294-310)
    float G[9];
    const float es = 2.0f * P.eta_shear[p], eb = P.eta_bulk[p];
    float s00 = 0.0f, s11 = 0.0f, s22 = 0.0f, s01 = 0.0f, s02 = 0.0f, s12 = 0.0f;
    if (es != 0.0f || eb != 0.0f) {
        // with both viscosities zero the viscous term is zero for any finite C, and C
        // (36 bytes per particle) is not read
        ld9(P.C, n, p, G);
        s00 = G[0];
        s11 = G[4];
        s22 = G[8];
        s01 = 0.5f * (G[1] + G[3]);
        s02 = 0.5f * (G[2] + G[6]);
        s12 = 0.5f * (G[5] + G[7]);
    }
    const float vr = s00 + s11 + s22;
    const float third = vr / 3.0f;
    T[0] = tau[0] + (es * (s00 - third) + eb * vr);
    T[4] = tau[1] + (es * (s11 - third) + eb * vr);
    T[8] = tau[2] + (es * (s22 - third) + eb * vr);
    T[1] = T[3] = tau[3] + es * s01;
    T[2] = T[6] = tau[4] + es * s02;
    T[5] = T[7] = tau[5] + es * s12;
    st9(P.stress, n, p, T);
    return branch != 2;
}

// The whole update of particle p, which the caller has checked to be active
// with materialLabel <= 1 (see cu_finish for what it writes).
MPMX_HD bool constitutive_update(const mpmx_particles& P, long p, float eff, float T[9]) {
    CuState c;
    cu_begin(P, p, c);
    return cu_finish(P, p, eff, c, T);
}

}  // namespace mpmx
