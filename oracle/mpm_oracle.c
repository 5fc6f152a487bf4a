/*
 * CPU oracle for the MPMWarpYDW per-substep hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, never by the
 * product path (mpmwarpydw_amd/), which runs the HIP kernels or fails.
 *
 * The same source is compiled twice:
 *   liboracle_f64.so  REAL=double  - the FP64 oracle of record
 *   liboracle_f32.so  REAL=float   - an FP32 restatement (reference semantics,
 *                                    sequential P2G in particle order), used to
 *                                    show how far the reference's own FP32
 *                                    arithmetic sits from FP64, and timed on one
 *                                    core as a CPU baseline.
 *   liboracle_f32_omp.so  the FP32 restatement with OpenMP over particles and
 *                                    grid nodes (P2G into one grid per thread,
 *                                    summed afterwards): the all-cores CPU
 *                                    baseline.  The pragmas are inert in the
 *                                    two builds above.
 * Every function restates one reference kernel; the citation in front of it
 * names the file:line under utils/ in the MPMWarpYDW reference.
 *
 * Layout: AoS, as the reference's Warp arrays (vec3 = 3 REAL, mat33 = 9 REAL
 * row-major, grid arrays [gx][gy][gz] with z fastest).  Labels are int32.
 *
 * Pinning: every function is checked in tests/test_oracle_*.py against closed
 * forms, invariants, numpy/LAPACK or brute force (see the header of each test).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef REAL
#define REAL double
#endif
typedef REAL R;

#define EXPORT __attribute__((visibility("default")))

/* the reference's 3x3 SVD runs in this mode: 1 = from F^T F (what an SVD on
 * F itself resolves), 0 = from F^T F - I formed through H = F - I (accurate
 * principal strains for near-identity F).  The FP64 oracle uses 0. */
static int g_svd_from_ftf = 0;
EXPORT void o_set_svd_mode(int from_ftf) { g_svd_from_ftf = from_ftf; }
EXPORT int o_real_size(void) { return (int)sizeof(R); }
/* threads the OpenMP build runs its loops on (1 in the other builds) */
EXPORT int o_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* small linear algebra                                                */
/* ------------------------------------------------------------------ */
static void mat_mul(const R *A, const R *B, R *C) {
    R T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            R s = 0;
            for (int k = 0; k < 3; k++) s += A[i * 3 + k] * B[k * 3 + j];
            T[i * 3 + j] = s;
        }
    memcpy(C, T, sizeof(T));
}
static void mat_T(const R *A, R *B) {
    R T[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) T[i * 3 + j] = A[j * 3 + i];
    memcpy(B, T, sizeof(T));
}
static R det3(const R *A) {
    return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
           A[2] * (A[3] * A[7] - A[4] * A[6]);
}
static R vlen(const R *v) { return (R)sqrt((double)(v[0] * v[0] + v[1] * v[1] + v[2] * v[2])); }

/* Warp's normalize: zero vector stays zero, otherwise divide by length. */
static void vnormalize(const R *a, R *out) {
    R l = vlen(a);
    if (l > (R)0) {
        out[0] = a[0] / l; out[1] = a[1] / l; out[2] = a[2] / l;
    } else {
        out[0] = out[1] = out[2] = 0;
    }
}

/*
 * 3x3 SVD, F = U diag(s) V^T with U, V proper rotations, |s| sorted
 * descending and the sign of det F carried by s[2] (the convention of the
 * McAdams et al. 2011 construction "Computing the singular value
 * decomposition of 3x3 matrices with minimal branching and elementary
 * floating point operations", which the reference's wp.svd3 follows;
 * call sites mpmRoutines.py:288,339).
 * Method: cyclic Jacobi on the symmetric A = F^T F (mode 1) or on
 * A = F^T F - I = H + H^T + H^T H with H = F - I (mode 0), then a Givens QR of
 * B = F V for U and the signs.  Also returns lam_out[k] = s_k^2 - 1.
 */
EXPORT void o_svd3(const R *F, R *U, R *s, R *V, R *lam_out) {
    R A[9];
    if (g_svd_from_ftf) {
        R Ft[9];
        mat_T(F, Ft);
        mat_mul(Ft, F, A);
        A[0] -= 1; A[4] -= 1; A[8] -= 1;
    } else {
        R H[9];
        memcpy(H, F, sizeof(H));
        H[0] -= 1; H[4] -= 1; H[8] -= 1;
        R Ht[9], HtH[9];
        mat_T(H, Ht);
        mat_mul(Ht, H, HtH);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) A[i * 3 + j] = H[i * 3 + j] + H[j * 3 + i] + HtH[i * 3 + j];
    }
    R Vm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    /* cyclic Jacobi, exact rotations (Golub & Van Loan, Alg. 8.4.2) */
    for (int sweep = 0; sweep < 30; sweep++) {
        R off = A[1] * A[1] + A[2] * A[2] + A[5] * A[5];
        R dia = A[0] * A[0] + A[4] * A[4] + A[8] * A[8];
        if (off <= (R)1e-40 * dia || off == 0) break;
        static const int P[3] = {0, 0, 1}, Q[3] = {1, 2, 2};
        for (int r = 0; r < 3; r++) {
            int p = P[r], q = Q[r];
            R apq = A[p * 3 + q];
            if (apq == 0) continue;
            R theta = (A[q * 3 + q] - A[p * 3 + p]) / (2 * apq);
            R t;
            if (fabs((double)theta) > 1e150)
                t = (R)0.5 / theta;
            else
                t = (theta >= 0 ? (R)1 : (R)-1) / ((R)fabs((double)theta) + (R)sqrt((double)(1 + theta * theta)));
            R c = (R)1 / (R)sqrt((double)(1 + t * t));
            R sn = t * c;
            /* A <- J^T A J with J the (p,q) rotation */
            for (int k = 0; k < 3; k++) {
                R akp = A[k * 3 + p], akq = A[k * 3 + q];
                A[k * 3 + p] = c * akp - sn * akq;
                A[k * 3 + q] = sn * akp + c * akq;
            }
            for (int k = 0; k < 3; k++) {
                R apk = A[p * 3 + k], aqk = A[q * 3 + k];
                A[p * 3 + k] = c * apk - sn * aqk;
                A[q * 3 + k] = sn * apk + c * aqk;
            }
            for (int k = 0; k < 3; k++) {
                R vkp = Vm[k * 3 + p], vkq = Vm[k * 3 + q];
                Vm[k * 3 + p] = c * vkp - sn * vkq;
                Vm[k * 3 + q] = sn * vkp + c * vkq;
            }
        }
    }
    R lam[3] = {A[0], A[4], A[8]};
    /* sort descending, swapping columns with a sign flip to stay in SO(3) */
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2 - a; b++)
            if (lam[b] < lam[b + 1]) {
                R tl = lam[b]; lam[b] = lam[b + 1]; lam[b + 1] = tl;
                for (int k = 0; k < 3; k++) {
                    R tv = Vm[k * 3 + b];
                    Vm[k * 3 + b] = Vm[k * 3 + b + 1];
                    Vm[k * 3 + b + 1] = -tv;
                }
            }
    /* U from a Givens QR of B = F V: R00, R11 >= 0, sign(R22) = sign(det F) */
    R B[9];
    mat_mul(F, Vm, B);
    R Qm[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    static const int GP[3] = {0, 0, 1}, GQ[3] = {1, 2, 2}, GC[3] = {0, 0, 1};
    for (int g = 0; g < 3; g++) {
        int p = GP[g], q = GQ[g], col = GC[g];
        R a = B[p * 3 + col], b = B[q * 3 + col];
        R rr = (R)sqrt((double)(a * a + b * b));
        R c = 1, sn = 0;
        if (rr > 0) { c = a / rr; sn = b / rr; }
        for (int k = 0; k < 3; k++) {
            R bp = B[p * 3 + k], bq = B[q * 3 + k];
            B[p * 3 + k] = c * bp + sn * bq;
            B[q * 3 + k] = -sn * bp + c * bq;
        }
        for (int k = 0; k < 3; k++) {
            R qp = Qm[k * 3 + p], qq = Qm[k * 3 + q];
            Qm[k * 3 + p] = c * qp + sn * qq;
            Qm[k * 3 + q] = -sn * qp + c * qq;
        }
    }
    for (int k = 0; k < 3; k++) {
        R l = lam[k];
        if (l < -1) l = -1;
        R mag = (R)sqrt((double)(1 + l));
        s[k] = (k == 2 && B[8] < 0) ? -mag : mag;
        if (lam_out) lam_out[k] = lam[k];
    }
    memcpy(U, Qm, sizeof(Qm));
    memcpy(V, Vm, sizeof(Vm));
}

/* ln(max(s, floor)), evaluated as 0.5*log1p(s^2-1) where s >= floor, so that
 * near-identity stretches keep their precision. */
static R log_clamped(R s, R lam, R floorv) {
    if (s >= floorv) return (R)(0.5 * log1p((double)lam));
    return (R)log((double)floorv);
}

/* ------------------------------------------------------------------ */
/* A1 compute_mu_lam_bulk_from_E_nu  (mpmRoutines.py:206-226)          */
/* ------------------------------------------------------------------ */
EXPORT void o_compute_mu_lam_bulk(int n, const R *E, const R *nu, R *mu, R *lam, R *bulk) {
    for (int p = 0; p < n; p++) {
        mu[p] = E[p] / ((R)2 * ((R)1 + nu[p]));
        lam[p] = E[p] * nu[p] / (((R)1 + nu[p]) * ((R)1 - (R)2 * nu[p]));
        bulk[p] = lam[p] + (R)(2.0 / 3.0) * mu[p];
    }
}

/* ------------------------------------------------------------------ */
/* initialize_geostatic_F (mpmRoutines.py:52-124) and                  */
/* initialize_geostatic_F_planestrain (:128-203); F is AoS 9 per particle */
/* ------------------------------------------------------------------ */
EXPORT void o_initialize_geostatic_F(int n, const R *x, R *F, const R *mu, const R *lam, const R *density, R g,
                                     R z_top, R K0, int plane_strain) {
    for (int p = 0; p < n; p++) {
        R depth = z_top - x[3 * p + 2];
        R e[3] = {0, 0, 0};
        if (!plane_strain) {
            /* solve [2l+2m, l; 2l, l+2m] (eps_h, eps_v) = (sigma_h, sigma_v) by Cramer (:104-110) */
            R sv = -density[p] * g * depth, sh = K0 * sv;
            R a11 = (R)2 * lam[p] + (R)2 * mu[p], a12 = lam[p], a21 = (R)2 * lam[p], a22 = lam[p] + (R)2 * mu[p];
            R det = a11 * a22 - a12 * a21;
            if (fabs((double)det) > 1e-12) {
                e[0] = e[1] = (sh * a22 - sv * a12) / det;
                e[2] = (a11 * sv - a21 * sh) / det;
            }
        } else if (depth > 0) {
            /* Hooke's law with eps_yy = 0, so sigma_yy = nu (sigma_xx + sigma_zz) (:172-191) */
            R nu = lam[p] / ((R)2 * (lam[p] + mu[p]));
            R E = mu[p] * ((R)3 * lam[p] + (R)2 * mu[p]) / (lam[p] + mu[p]);
            R sv = -density[p] * g * depth, sh = K0 * sv, syy = nu * (sh + sv);
            e[0] = (sh - nu * (syy + sv)) / E;
            e[2] = (sv - nu * (sh + syy)) / E;
        }
        R *Fp = F + 9 * p;
        for (int k = 0; k < 9; k++) Fp[k] = 0;
        for (int k = 0; k < 3; k++) Fp[4 * k] = (R)exp((double)e[k]);
    }
}

/* deactivate_and_teleport_xpbd_below_datum (mpmRoutines.py:1446-1475) */
EXPORT void o_deactivate_and_teleport(int n, R *x, R *v, int *active, const int *label, R z_datum, R teleport) {
    for (int p = 0; p < n; p++) {
        if (active[p] != 1 || label[p] != 2 || !(x[3 * p + 2] < z_datum)) continue;
        active[p] = 0;
        x[3 * p + 2] -= teleport;
        v[3 * p] = v[3 * p + 1] = v[3 * p + 2] = 0;
    }
}

/* ------------------------------------------------------------------ */
/* A4 kirchoff_stress (mpmRoutines.py:456-471):  U diag(c) V^T F^T      */
/* ------------------------------------------------------------------ */
static void kirchoff_stress(const R *F, const R *U, const R *V, const R *sig, const R *lam_sq,
                            R mu, R lam, R *out) {
    R sc[3], ls[3];
    for (int k = 0; k < 3; k++) {
        sc[k] = sig[k] > (R)1e-6 ? sig[k] : (R)1e-6;
        ls[k] = log_clamped(sig[k], lam_sq[k], (R)1e-6);
    }
    R lsum = ls[0] + ls[1] + ls[2];
    R c[3];
    for (int k = 0; k < 3; k++) c[k] = (R)2 * mu * ls[k] * ((R)1 / sc[k]) + lam * lsum * ((R)1 / sc[k]);
    R Ud[9], Vt[9], Ft[9], T1[9];
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) Ud[i * 3 + k] = U[i * 3 + k] * c[k];
    mat_T(V, Vt);
    mat_T(F, Ft);
    mat_mul(Ud, Vt, T1);
    mat_mul(T1, Ft, out);
}

/* ------------------------------------------------------------------ */
/* A3 drucker_prager_return_mapping (mpmRoutines.py:314-453)           */
/* ------------------------------------------------------------------ */
static void dp_return_map(const R *F_trial, const R *x, R *v, R *x_init, R *v_init, int *matLabel,
                          R *acc, R *damage, R *mu, R *lam, R *ys, R *alpha, const R *hard,
                          const R *soft, const R *density, const R *sc, R eff, int p, R *F_out) {
    R U[9], V[9], sig_old[3], lsq[3];
    o_svd3(F_trial, U, sig_old, V, lsq);
    R eps[3];
    for (int k = 0; k < 3; k++) eps[k] = log_clamped(sig_old[k], lsq[k], (R)0.01);
    R mean_eps = (eps[0] + eps[1] + eps[2]) / (R)3;
    R edev[3] = {eps[0] - mean_eps, eps[1] - mean_eps, eps[2] - mean_eps};
    R esum = eps[0] + eps[1] + eps[2];
    R tau[3];
    for (int k = 0; k < 3; k++) tau[k] = (R)2 * mu[p] * eps[k] + lam[p] * esum;
    R p_stress = (tau[0] + tau[1] + tau[2]) / (R)3;
    R tdev[3] = {tau[0] - p_stress, tau[1] - p_stress, tau[2] - p_stress};
    R sigma_eq = vlen(tdev);
    R yield_eff = ((R)1 - damage[p]) * (ys[p] - alpha[p] * p_stress);

    if (sigma_eq > yield_eff || ys[p] < 0 || yield_eff < 0) {
        R edn = vlen(edev) + (R)1e-6;
        R dgam = edn - yield_eff / ((R)2 * mu[p]);
        R dep = (R)sqrt(2.0 / 3.0) * dgam;
        acc[p] += dep;
        if (sc[p] > 0) {
            R d = damage[p] + dep / sc[p];
            damage[p] = (R)1 < d ? (R)1 : d; /* wp.min(1, d) = 1 < d ? 1 : d */
        } else {
            damage[p] = 1;
        }
        if (ys[p] < 0 || yield_eff < 0) damage[p] = 1;
        if (damage[p] >= (R)1 && matLabel[p] == 1) {
            matLabel[p] = 2;
            for (int k = 0; k < 3; k++) x_init[p * 3 + k] = x[p * 3 + k];
            R rho = density[p];
            R u = (R)0.5 * (tau[0] * eps[0] + tau[1] * eps[1] + tau[2] * eps[2]);
            R v_expected = (R)sqrt((double)((R)2 * eff * u / rho));
            R vv[3] = {v[p * 3 + 0] + (R)1e-12, v[p * 3 + 1] + (R)1e-12, v[p * 3 + 2] + (R)1e-12};
            R vdir[3];
            vnormalize(vv, vdir);
            for (int k = 0; k < 3; k++) {
                v[p * 3 + k] = v[p * 3 + k] + v_expected * vdir[k];
                v_init[p * 3 + k] = v[p * 3 + k];
            }
            ys[p] = 0; mu[p] = 0; lam[p] = 0; alpha[p] = 0;
            memcpy(F_out, F_trial, 9 * sizeof(R));
            return;
        }
        R beta = (R)0.3 * alpha[p];
        R volc = beta * dgam;
        for (int k = 0; k < 3; k++) eps[k] = eps[k] - (dgam / edn) * edev[k];
        for (int k = 0; k < 3; k++) eps[k] = eps[k] + volc;
        R dsy = (R)2 * mu[p] * (hard[p] - soft[p]) * dgam;
        R ny = ys[p] + dsy;
        ys[p] = (R)0 > ny ? (R)0 : ny; /* wp.max(0, ny) */
        /* the ys < 0 failure branch (mpmRoutines.py:438-441) is unreachable
         * after the clamp above and is not restated. */
        R Ud[9], Vt[9];
        for (int i = 0; i < 3; i++)
            for (int k = 0; k < 3; k++) Ud[i * 3 + k] = U[i * 3 + k] * (R)exp((double)eps[k]);
        mat_T(V, Vt);
        mat_mul(Ud, Vt, F_out);
        return;
    }
    memcpy(F_out, F_trial, 9 * sizeof(R));
}

/* ------------------------------------------------------------------ */
/* A2 compute_stress_from_F_trial (mpmRoutines.py:229-312)             */
/* ------------------------------------------------------------------ */
EXPORT void o_compute_stress_from_F_trial(int n, const int *activeLabel, int *materialLabel,
                                          const R *x, R *v, R *x_init, R *v_init, R *F,
                                          const R *F_trial, R *mu, R *lam, R *ys, const R *hard,
                                          const R *soft, const R *density, const R *sc, R eff,
                                          const R *eta_shear, const R *eta_bulk, const R *C,
                                          R *stress, R *acc, R *damage, int constitutive_model,
                                          R *alpha) {
    (void)constitutive_model; /* passed but never read by the reference kernel */
#pragma omp parallel for schedule(dynamic, 1024)
    for (int p = 0; p < n; p++) {
        if (materialLabel[p] > 1 || activeLabel[p] != 1) continue;
        dp_return_map(F_trial + 9 * p, x, v, x_init, v_init, materialLabel, acc, damage, mu, lam,
                      ys, alpha, hard, soft, density, sc, eff, p, F + 9 * p);
        R U[9], V[9], sig[3], lsq[3], tau[9];
        o_svd3(F + 9 * p, U, sig, V, lsq);
        kirchoff_stress(F + 9 * p, U, V, sig, lsq, mu[p], lam[p], tau);
        R st[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) st[i * 3 + j] = (tau[i * 3 + j] + tau[j * 3 + i]) / (R)2;
        const R *g = C + 9 * p;
        R sr[9];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) sr[i * 3 + j] = (R)0.5 * (g[i * 3 + j] + g[j * 3 + i]);
        R vr = sr[0] + sr[4] + sr[8];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                R dev = sr[i * 3 + j] - (i == j ? vr / (R)3 : (R)0);
                R visc = ((R)2 * eta_shear[p]) * dev + (i == j ? eta_bulk[p] * vr : (R)0);
                stress[9 * p + i * 3 + j] = st[i * 3 + j] + visc;
            }
    }
}

/* ------------------------------------------------------------------ */
/* quadratic B-spline stencil shared by P2G and G2P                    */
/* (mpmRoutines.py:514-530 and :733-749; compute_dweight :474-483)     */
/* ------------------------------------------------------------------ */
static int stencil(const R *xp, const R *minB, R inv_dx, int *base, R *fx, R w[3][3], R dw[3][3]) {
    for (int a = 0; a < 3; a++) {
        R gp = (xp[a] - minB[a]) * inv_dx;
        base[a] = (int)(gp - (R)0.5); /* truncation toward zero, as wp.int */
        fx[a] = gp - (R)base[a];
        R wa = (R)1.5 - fx[a], wb = fx[a] - (R)1, wc = fx[a] - (R)0.5;
        w[a][0] = wa * wa * (R)0.5;
        w[a][1] = (R)0 - wb * wb + (R)0.75;
        w[a][2] = wc * wc * (R)0.5;
        dw[a][0] = fx[a] - (R)1.5;
        dw[a][1] = (R)-2 * (fx[a] - (R)1);
        dw[a][2] = fx[a] - (R)0.5;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* A6 p2g_apic_with_stress (mpmRoutines.py:485-569), sequential order  */
/* returns the number of particles whose stencil left the grid (the    */
/* reference has no bounds check; they are skipped here).              */
/* ------------------------------------------------------------------ */
static int p2g_particle(int p, const int *activeLabel, const int *materialLabel, const R *stress, const R *x,
                        const R *v, const R *C, const R *vol, const R *mass, R dx, R inv_dx, const R *minB, R rpic,
                        int gx, int gy, int gz, R *grid_m, R *grid_v_in, R dt) {
    if (!(activeLabel[p] == 1 && materialLabel[p] <= 1)) return 0;
    int base[3];
    R fx[3], w[3][3], dw[3][3];
    stencil(x + 3 * p, minB, inv_dx, base, fx, w, dw);
    if (base[0] < 0 || base[1] < 0 || base[2] < 0 || base[0] + 2 >= gx || base[1] + 2 >= gy || base[2] + 2 >= gz)
        return 1;
    const R *Cp = C + 9 * p, *S = stress + 9 * p;
    R Cm[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            Cm[i * 3 + j] = ((R)1 - rpic) * Cp[i * 3 + j] + rpic / (R)2 * (Cp[i * 3 + j] - Cp[j * 3 + i]);
    if (rpic < (R)-0.001) memset(Cm, 0, sizeof(Cm));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                R dpos[3] = {((R)i - fx[0]) * dx, ((R)j - fx[1]) * dx, ((R)k - fx[2]) * dx};
                R weight = w[0][i] * w[1][j] * w[2][k];
                R dwt[3] = {dw[0][i] * w[1][j] * w[2][k] * inv_dx, w[0][i] * dw[1][j] * w[2][k] * inv_dx,
                            w[0][i] * w[1][j] * dw[2][k] * inv_dx};
                R add[3];
                for (int a = 0; a < 3; a++) {
                    R cd = Cm[a * 3 + 0] * dpos[0] + Cm[a * 3 + 1] * dpos[1] + Cm[a * 3 + 2] * dpos[2];
                    R ef = -vol[p] * (S[a * 3 + 0] * dwt[0] + S[a * 3 + 1] * dwt[1] + S[a * 3 + 2] * dwt[2]);
                    add[a] = weight * mass[p] * (v[3 * p + a] + cd) + dt * ef;
                }
                long idx = ((long)(base[0] + i) * gy + (base[1] + j)) * gz + (base[2] + k);
                for (int a = 0; a < 3; a++) grid_v_in[3 * idx + a] += add[a];
                grid_m[idx] += weight * mass[p];
            }
    return 0;
}

EXPORT int o_p2g(int n, const int *activeLabel, const int *materialLabel, const R *stress,
                 const R *x, const R *v, const R *C, const R *vol, const R *mass, R dx, R inv_dx,
                 const R *minB, R rpic, int gx, int gy, int gz, R *grid_m, R *grid_v_in, R dt) {
    int bad = 0;
#ifdef _OPENMP
    /* the all-cores baseline: each thread scatters a contiguous range of
     * particles into a grid of its own, then the grids are summed node by node */
    const long G = (long)gx * gy * gz;
    const int T = omp_get_max_threads();
    R *priv = (R *)calloc((size_t)T * 4 * (size_t)G, sizeof(R));
#pragma omp parallel reduction(+ : bad)
    {
        const int t = omp_get_thread_num();
        R *gm = priv + (size_t)t * 4 * G, *gv = gm + G;
#pragma omp for schedule(static)
        for (int p = 0; p < n; p++)
            bad += p2g_particle(p, activeLabel, materialLabel, stress, x, v, C, vol, mass, dx, inv_dx, minB, rpic,
                                gx, gy, gz, gm, gv, dt);
#pragma omp for schedule(static)
        for (long i = 0; i < G; i++) {
            R m = grid_m[i], a0 = grid_v_in[3 * i], a1 = grid_v_in[3 * i + 1], a2 = grid_v_in[3 * i + 2];
            for (int u = 0; u < T; u++) {
                const R *um = priv + (size_t)u * 4 * G, *uv = um + G;
                m += um[i];
                a0 += uv[3 * i];
                a1 += uv[3 * i + 1];
                a2 += uv[3 * i + 2];
            }
            grid_m[i] = m;
            grid_v_in[3 * i] = a0;
            grid_v_in[3 * i + 1] = a1;
            grid_v_in[3 * i + 2] = a2;
        }
    }
    free(priv);
#else
    for (int p = 0; p < n; p++)
        bad += p2g_particle(p, activeLabel, materialLabel, stress, x, v, C, vol, mass, dx, inv_dx, minB, rpic, gx,
                            gy, gz, grid_m, grid_v_in, dt);
#endif
    return bad;
}

/* A7 grid_normalization_and_gravity (mpmRoutines.py:691-704); v_out must
 * hold zeros where m <= 1e-15 (simulationRoutines.py:13 zeroes it). */
EXPORT void o_grid_normalization_and_gravity(long G, const R *grid_m, const R *v_in, R *v_out,
                                             const R *gravity, R dt) {
#pragma omp parallel for schedule(static)
    for (long i = 0; i < G; i++) {
        if (grid_m[i] > (R)1e-15) {
            R inv = (R)1 / grid_m[i];
            for (int a = 0; a < 3; a++) v_out[3 * i + a] = v_in[3 * i + a] * inv + dt * gravity[a];
        }
    }
}

/* A8 add_damping_via_grid (mpmRoutines.py:706-711) */
EXPORT void o_add_damping(long G, R *v_out, R scale) {
    for (long i = 0; i < 3 * G; i++) v_out[i] = v_out[i] * scale;
}

/* apply_coulomb_friction (mpmRoutines.py:915-940) */
static R coulomb(R v, R mu, R dt, R g) {
    if (fabs((double)v) < 1e-12) return 0;
    R dv = mu * g * dt;
    R s = v < 0 ? (R)-1 : (R)1;
    R m = (R)fabs((double)v) - dv;
    return s * (m > 0 ? m : (R)0);
}

/* A9 collideBounds (mpmRoutines.py:942-993): v read once, each matching face
 * overwrites the whole vector, the last match wins. */
EXPORT void o_collide_bounds(R *v_out, int gx, int gy, int gz, R mu, R dt, R gmag, int pad) {
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                long idx = ((long)ix * gy + iy) * gz + iz;
                R *o = v_out + 3 * idx;
                R v[3] = {o[0], o[1], o[2]};
                if (ix <= pad && v[0] < 0) { o[0] = 0; o[1] = coulomb(v[1], mu, dt, gmag); o[2] = coulomb(v[2], mu, dt, gmag); }
                if (ix >= gx - pad && v[0] > 0) { o[0] = 0; o[1] = coulomb(v[1], mu, dt, gmag); o[2] = coulomb(v[2], mu, dt, gmag); }
                if (iy <= pad && v[1] < 0) { o[0] = coulomb(v[0], mu, dt, gmag); o[1] = 0; o[2] = coulomb(v[2], mu, dt, gmag); }
                if (iy >= gy - pad && v[1] > 0) { o[0] = coulomb(v[0], mu, dt, gmag); o[1] = 0; o[2] = coulomb(v[2], mu, dt, gmag); }
                if (iz <= pad && v[2] < 0) { o[0] = coulomb(v[0], mu, dt, gmag); o[1] = coulomb(v[1], mu, dt, gmag); o[2] = 0; }
                if (iz >= gz - pad && v[2] > 0) { o[0] = coulomb(v[0], mu, dt, gmag); o[1] = coulomb(v[1], mu, dt, gmag); o[2] = 0; }
            }
}

/* the other boundary conditions of the grid update (simulationRoutines.py:
 * 141-228), each kernel restated per node */

/* collideBoundsPlaneStrainY (mpmRoutines.py:996-1062): x and z faces as
 * collideBounds; on the y faces v_y is zeroed whatever its sign, from a fresh
 * read of the node (which also becomes the v the z faces test) */
EXPORT void o_collide_bounds_plane_strain_y(R *v_out, int gx, int gy, int gz, R mu, R dt, R g, int pad) {
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                R *o = v_out + 3 * (((long)ix * gy + iy) * gz + iz);
                R v[3] = {o[0], o[1], o[2]};
                if (ix <= pad && v[0] < 0) { o[0] = 0; o[1] = coulomb(v[1], mu, dt, g); o[2] = coulomb(v[2], mu, dt, g); }
                if (ix >= gx - pad && v[0] > 0) { o[0] = 0; o[1] = coulomb(v[1], mu, dt, g); o[2] = coulomb(v[2], mu, dt, g); }
                if (iy <= pad) { v[0] = o[0]; v[1] = o[1]; v[2] = o[2]; o[1] = 0; }
                if (iy >= gy - pad) { v[0] = o[0]; v[1] = o[1]; v[2] = o[2]; o[1] = 0; }
                if (iz <= pad && v[2] < 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = coulomb(v[1], mu, dt, g); o[2] = 0; }
                if (iz >= gz - pad && v[2] > 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = coulomb(v[1], mu, dt, g); o[2] = 0; }
            }
}

/* collideBoundsPlaneStrainYGlobal (mpmRoutines.py:1065-1123): v_y = 0 on every
 * node first, then x and z faces on that velocity */
EXPORT void o_collide_bounds_plane_strain_y_global(R *v_out, int gx, int gy, int gz, R mu, R dt, R g, int pad) {
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                R *o = v_out + 3 * (((long)ix * gy + iy) * gz + iz);
                o[1] = 0;
                R v[3] = {o[0], o[1], o[2]};
                if (ix <= pad && v[0] < 0) { o[0] = 0; o[1] = 0; o[2] = coulomb(v[2], mu, dt, g); }
                if (ix >= gx - pad && v[0] > 0) { o[0] = 0; o[1] = 0; o[2] = coulomb(v[2], mu, dt, g); }
                if (iz <= pad && v[2] < 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = 0; o[2] = 0; }
                if (iz >= gz - pad && v[2] > 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = 0; o[2] = 0; }
            }
}

/* collideBoundsAbsorbing (mpmRoutines.py:1126-1176): the three outer node
 * layers of each face (a literal 3, whatever the padding) scale v by
 * 1 - (pad - distance) / pad; every face scales the node's original v and the
 * last matching face wins */
EXPORT void o_collide_bounds_absorbing(R *v_out, int gx, int gy, int gz, int pad) {
    const R P = (R)pad;
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                R *o = v_out + 3 * (((long)ix * gy + iy) * gz + iz);
                R v[3] = {o[0], o[1], o[2]};
                int dist[6] = {ix, gx - 1 - ix, iy, gy - 1 - iy, iz, gz - 1 - iz};
                int hit[6] = {ix < 3, ix >= gx - 3, iy < 3, iy >= gy - 3, iz < 3, iz >= gz - 3};
                for (int f = 0; f < 6; f++) {
                    if (!hit[f]) continue;
                    R lf = (P - (R)dist[f]) / P;
                    R d = (R)0 * lf + (R)1 * ((R)1 - lf);
                    o[0] = v[0] * d; o[1] = v[1] * d; o[2] = v[2] * d;
                }
            }
}

/* collideBoundsRestitution (mpmRoutines.py:1179-1247): the normal component of
 * an approaching node is reversed and scaled by e, the tangential ones get
 * Coulomb friction; last matching face wins, all from the original v */
EXPORT void o_collide_bounds_restitution(R *v_out, int gx, int gy, int gz, R e, R mu, R dt, R g, int pad) {
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                R *o = v_out + 3 * (((long)ix * gy + iy) * gz + iz);
                R v[3] = {o[0], o[1], o[2]};
                if (ix <= pad && v[0] < 0) { o[0] = -e * v[0]; o[1] = coulomb(v[1], mu, dt, g); o[2] = coulomb(v[2], mu, dt, g); }
                if (ix >= gx - pad && v[0] > 0) { o[0] = -e * v[0]; o[1] = coulomb(v[1], mu, dt, g); o[2] = coulomb(v[2], mu, dt, g); }
                if (iy <= pad && v[1] < 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = -e * v[1]; o[2] = coulomb(v[2], mu, dt, g); }
                if (iy >= gy - pad && v[1] > 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = -e * v[1]; o[2] = coulomb(v[2], mu, dt, g); }
                if (iz <= pad && v[2] < 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = coulomb(v[1], mu, dt, g); o[2] = -e * v[2]; }
                if (iz >= gz - pad && v[2] > 0) { o[0] = coulomb(v[0], mu, dt, g); o[1] = coulomb(v[1], mu, dt, g); o[2] = -e * v[2]; }
            }
}

/* collideBoundsGradualFriction (mpmRoutines.py:1249-1340): inside the padding
 * (index < pad, or > dim - 1 - pad) an approaching node keeps fade = distance
 * / pad of its normal velocity and gets friction mu (1 - fade) on the tangential
 * ones; per axis the lower face is tested first and the upper one only when the
 * lower does not apply (elif), the three axes in turn from the original v */
EXPORT void o_collide_bounds_gradual(R *v_out, int gx, int gy, int gz, R mu, R dt, R g, int pad) {
    const R P = (R)pad;
    for (int ix = 0; ix < gx; ix++)
        for (int iy = 0; iy < gy; iy++)
            for (int iz = 0; iz < gz; iz++) {
                R *o = v_out + 3 * (((long)ix * gy + iy) * gz + iz);
                R v[3] = {o[0], o[1], o[2]};
                int idx[3] = {ix, iy, iz}, dim[3] = {gx, gy, gz};
                for (int a = 0; a < 3; a++) {
                    R fade;
                    if (idx[a] < pad && v[a] < 0) fade = (R)idx[a] / P;
                    else if (idx[a] >= dim[a] - pad && v[a] > 0) fade = (R)(dim[a] - 1 - idx[a]) / P;
                    else continue;
                    R m = mu * ((R)1 - fade);
                    for (int b = 0; b < 3; b++)
                        o[b] = (b == a) ? (R)0 * ((R)1 - fade) + v[a] * fade : coulomb(v[b], m, dt, g);
                }
            }
}

/* Volumetric locking correction (mpmRoutines.py:805-911, run after G2P by
 * simulationRoutines.py:255-293): per cell int((x - minB) inv_dx) the sums of
 * tr(C) V and V over active MPM particles, then C += (div_cell - tr C) / 3 I
 * with div_cell = sum(tr C V) / sum(V) (0 when sum(V) <= 1e-20).  Particles
 * whose cell lies outside the grid are skipped (the reference would index out
 * of bounds).  cell_div and cell_vol [gx*gy*gz] are overwritten. */
EXPORT void o_vlc(int n, const int *active, const int *label, const R *x, R *C, const R *vol, R inv_dx,
                  const R *minB, int gx, int gy, int gz, R *cell_div, R *cell_vol) {
    long G = (long)gx * gy * gz;
    memset(cell_div, 0, sizeof(R) * G);
    memset(cell_vol, 0, sizeof(R) * G);
    for (int pass = 0; pass < 2; pass++)
        for (int p = 0; p < n; p++) {
            if (!(active[p] == 1 && label[p] <= 1)) continue;
            /* the cell is an index, fixed by the reference's float32 arithmetic
             * ((x - minBounds) * inv_dx in wp.vec3, :840-844): particles placed
             * exactly at grid_padding cells from minBounds sit on a cell face,
             * where FP64 would pick the other side */
            int c[3];
            for (int a = 0; a < 3; a++) {
                volatile float d = (float)x[3 * p + a] - (float)minB[a];
                volatile float g = d * (float)inv_dx;
                c[a] = (int)g;
            }
            if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= gx || c[1] >= gy || c[2] >= gz) continue;
            long k = ((long)c[0] * gy + c[1]) * gz + c[2];
            R *Cp = C + 9 * p;
            R div = Cp[0] + Cp[4] + Cp[8];
            if (pass == 0) {
                cell_div[k] += div * vol[p];
                cell_vol[k] += vol[p];
            } else {
                R dc = cell_vol[k] > (R)1e-20 ? cell_div[k] / cell_vol[k] : (R)0;
                R corr = (dc - div) / (R)3;
                Cp[0] += corr; Cp[4] += corr; Cp[8] += corr;
            }
        }
}

/* ------------------------------------------------------------------ */
/* A10 g2p + update_cov (mpmRoutines.py:715-802)                       */
/* ------------------------------------------------------------------ */
EXPORT int o_g2p(R dt, int n, const int *activeLabel, const int *materialLabel, R *x, R *v, R *C,
                 const R *F, R *F_trial, R *cov, R inv_dx, const R *v_out, int gx, int gy, int gz,
                 int update_cov, const R *minB) {
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int p = 0; p < n; p++) {
        if (activeLabel[p] != 1 || materialLabel[p] > 1) continue;
        int base[3];
        R fx[3], w[3][3], dw[3][3];
        stencil(x + 3 * p, minB, inv_dx, base, fx, w, dw);
        if (base[0] < 0 || base[1] < 0 || base[2] < 0 || base[0] + 2 >= gx || base[1] + 2 >= gy ||
            base[2] + 2 >= gz) {
            bad++;
            continue;
        }
        R nv[3] = {0, 0, 0}, nC[9] = {0}, nF[9] = {0};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    long idx = ((long)(base[0] + i) * gy + (base[1] + j)) * gz + (base[2] + k);
                    const R *gv = v_out + 3 * idx;
                    R dpos[3] = {(R)i - fx[0], (R)j - fx[1], (R)k - fx[2]};
                    R weight = w[0][i] * w[1][j] * w[2][k];
                    R dwt[3] = {dw[0][i] * w[1][j] * w[2][k] * inv_dx, w[0][i] * dw[1][j] * w[2][k] * inv_dx,
                                w[0][i] * w[1][j] * dw[2][k] * inv_dx};
                    for (int a = 0; a < 3; a++) {
                        nv[a] += gv[a] * weight;
                        for (int b = 0; b < 3; b++) {
                            nC[a * 3 + b] += gv[a] * dpos[b] * (weight * inv_dx * (R)4);
                            nF[a * 3 + b] += gv[a] * dwt[b];
                        }
                    }
                }
        for (int a = 0; a < 3; a++) {
            v[3 * p + a] = nv[a];
            x[3 * p + a] = x[3 * p + a] + dt * nv[a];
        }
        memcpy(C + 9 * p, nC, sizeof(nC));
        R M[9];
        for (int i = 0; i < 9; i++) M[i] = nF[i] * dt;
        M[0] += 1; M[4] += 1; M[8] += 1;
        mat_mul(M, F + 9 * p, F_trial + 9 * p);
        if (update_cov) {
            R *c6 = cov + 6 * p;
            R cn[9] = {c6[0], c6[1], c6[2], c6[1], c6[3], c6[4], c6[2], c6[4], c6[5]};
            R GC[9], CGt[9], Gt[9];
            mat_mul(nF, cn, GC);
            mat_T(nF, Gt);
            mat_mul(cn, Gt, CGt);
            R c1[9];
            for (int i = 0; i < 9; i++) c1[i] = cn[i] + dt * (GC[i] + CGt[i]);
            c6[0] = c1[0]; c6[1] = c1[1]; c6[2] = c1[2]; c6[3] = c1[4]; c6[4] = c1[5]; c6[5] = c1[8];
        }
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* XPBD (xpbdRoutines.py)                                              */
/* ------------------------------------------------------------------ */
/* X4 reset_mpm_contact_labels (xpbdRoutines.py:3-21) */
EXPORT void o_reset_mpm_contact_labels(int n, const int *active, int *mat) {
    for (int p = 0; p < n; p++)
        if (active[p] == 1 && mat[p] == 0) mat[p] = 1;
}

/* X3 integrateParticlesXPBD (xpbdRoutines.py:23-48) */
EXPORT void o_integrate_particles_xpbd(int n, const int *active, const int *mat, const R *x,
                                       const R *v, const R *g, R dt, R *x_new, R *v_new, R v_max) {
#pragma omp parallel for schedule(static)
    for (int p = 0; p < n; p++) {
        if (active[p] != 1 || mat[p] != 2) continue;
        R v1[3] = {v[3 * p] + g[0] * dt, v[3 * p + 1] + g[1] * dt, v[3 * p + 2] + g[2] * dt};
        R m = vlen(v1);
        if (m > v_max) {
            R s = v_max / m;
            v1[0] *= s; v1[1] *= s; v1[2] *= s;
        }
        R x1[3] = {x[3 * p] + v1[0] * dt, x[3 * p + 1] + v1[1] * dt, x[3 * p + 2] + v1[2] * dt};
        for (int a = 0; a < 3; a++) { x_new[3 * p + a] = x1[a]; v_new[3 * p + a] = v1[a]; }
    }
}

/* X5 my_solve_particle_bound_contacts (xpbdRoutines.py:49-136) */
EXPORT void o_solve_particle_bound_contacts(int n, const int *active, const int *mat, const R *x,
                                            const R *v, const R *mass, const R *radius, R mu,
                                            R thr, R static_mu, const R *bmin, const R *bmax, R dt,
                                            R relax, R *delta) {
#pragma omp parallel for schedule(static)
    for (int tid = 0; tid < n; tid++) {
        if (active[tid] != 1 || mat[tid] != 2 || mass[tid] == 0) continue;
        R wi = (R)1 / mass[tid];
        const R *xp = x + 3 * tid, *vp = v + 3 * tid;
        R r = radius[tid];
        for (int ax = 0; ax < 3; ax++) {
            for (int side = 0; side < 2; side++) {
                R nrm[3] = {0, 0, 0};
                nrm[ax] = side == 0 ? (R)1 : (R)-1;
                const R *b = side == 0 ? bmin : bmax;
                R dd = nrm[0] * (xp[0] - b[0]) + nrm[1] * (xp[1] - b[1]) + nrm[2] * (xp[2] - b[2]);
                R c = dd - r < 0 ? dd - r : (R)0; /* wp.min(., 0) */
                if (c < 0) {
                    R lambda_n = c;
                    R dn[3] = {nrm[0] * lambda_n, nrm[1] * lambda_n, nrm[2] * lambda_n};
                    R vn = nrm[0] * vp[0] + nrm[1] * vp[1] + nrm[2] * vp[2];
                    R fmu = mu;
                    if (fabs((double)vn) < thr * r) fmu = static_mu;
                    R vt[3] = {vp[0] - nrm[0] * vn, vp[1] - nrm[1] * vn, vp[2] - nrm[2] * vn};
                    R vtl = vlen(vt);
                    R a_ = fmu * lambda_n, b_ = -vtl * dt;
                    R lf = a_ > b_ ? a_ : b_; /* wp.max(a, b) = a > b ? a : b */
                    R df[3] = {0, 0, 0};
                    if (vtl > (R)1e-8) {
                        R nt[3];
                        vnormalize(vt, nt);
                        df[0] = nt[0] * lf; df[1] = nt[1] * lf; df[2] = nt[2] * lf;
                    }
                    for (int a = 0; a < 3; a++) delta[3 * tid + a] += (df[a] - dn[a]) / wi * relax * wi;
                }
            }
        }
    }
}

/* ---- X2: spatial hash grid with the layout of wp.HashGrid --------- */
/* cell coordinates int(p / width) (truncation), offset by 2^20 and wrapped
 * modulo the hash dims; key = cz*dx*dy + cy*dx + cx.  Points are sorted by
 * key with a stable sort, so ids inside a cell ascend. */
static int hash_index(int dimx, int dimy, int dimz, int x, int y, int z) {
    const int origin = 1 << 20;
    x += origin; y += origin; z += origin;
    int cx = ((x % dimx) + dimx) % dimx, cy = ((y % dimy) + dimy) % dimy, cz = ((z % dimz) + dimz) % dimz;
    return cz * (dimx * dimy) + cy * dimx + cx;
}

EXPORT void o_hashgrid_build(int n, const R *x, R cell_width, int dimx, int dimy, int dimz,
                             int *point_ids, int *cell_starts, int *cell_ends) {
    R inv = (R)1 / cell_width;
    long ncell = (long)dimx * dimy * dimz;
    int *keys = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    int *cnt = (int *)calloc(ncell + 1, sizeof(int));
    for (int p = 0; p < n; p++) {
        keys[p] = hash_index(dimx, dimy, dimz, (int)(x[3 * p] * inv), (int)(x[3 * p + 1] * inv),
                             (int)(x[3 * p + 2] * inv));
        cnt[keys[p] + 1]++;
    }
    for (long c = 0; c < ncell; c++) cnt[c + 1] += cnt[c];
    for (long c = 0; c < ncell; c++) {
        cell_starts[c] = cnt[c];
        cell_ends[c] = cnt[c + 1];
    }
    int *fill = (int *)malloc(sizeof(int) * ncell);
    memcpy(fill, cnt, sizeof(int) * ncell);
    for (int p = 0; p < n; p++) point_ids[fill[keys[p]]++] = p; /* stable */
    free(fill);
    free(cnt);
    free(keys);
}

/* wp.hash_grid_query + hash_grid_query_next for one point: writes the
 * candidate ids in visiting order into out[0..cap), returns their number. */
EXPORT int o_hashgrid_query(const R *pos, R radius, const int *point_ids, const int *cell_starts,
                            const int *cell_ends, R cell_width, int dimx, int dimy, int dimz, int *out,
                            int cap) {
    R inv = (R)1 / cell_width;
    int xs = (int)((pos[0] - radius) * inv), ys = (int)((pos[1] - radius) * inv), zs = (int)((pos[2] - radius) * inv);
    int xe = (int)((pos[0] + radius) * inv), ye = (int)((pos[1] + radius) * inv), ze = (int)((pos[2] + radius) * inv);
    if (xe > xs + dimx - 1) xe = xs + dimx - 1;
    if (ye > ys + dimy - 1) ye = ys + dimy - 1;
    if (ze > zs + dimz - 1) ze = zs + dimz - 1;
    int k = 0;
    for (int cz = zs; cz <= ze; cz++)
        for (int cy = ys; cy <= ye; cy++)
            for (int cx = xs; cx <= xe; cx++) {
                int cell = hash_index(dimx, dimy, dimz, cx, cy, cz);
                for (int s = cell_starts[cell]; s < cell_ends[cell]; s++) {
                    if (k < cap) out[k] = point_ids[s];
                    k++;
                }
            }
    return k;
}

/* X6 my_solve_particle_particle_contacts (xpbdRoutines.py:138-219),
 * candidates in wp.hash_grid_query order: x fastest, then y, then z, and the
 * sorted ids of each cell (order 0).  The order is not pinned by any test or
 * fixture of the reference; order 1 visits the same candidates in reverse
 * (cells z, y, x descending, ids descending) and order 2 with z fastest, so
 * that a test can bound what the order changes. */
static void solve_contacts(int n, const int *active, int *mat, const int *point_ids, const int *cell_starts,
                           const int *cell_ends, R cell_width, int dimx, int dimy, int dimz, const R *x,
                           const R *v, const R *mass, const R *radius, R k_mu_in, R thr, R static_mu,
                           R k_cohesion, R max_radius, R dt, R relax, int lock, R *deltas, int order) {
    R inv = (R)1 / cell_width;
    /* every thread adds to its own particle's delta only; the lock's label
     * writes go to neighbours, so with the lock on the loop stays serial */
#pragma omp parallel for schedule(dynamic, 256) if (lock == 0)
    for (int tid = 0; tid < n; tid++) {
        if (active[tid] != 1 || mat[tid] < 0) continue;
        int i = point_ids[tid];
        if (i == -1) continue;
        const R *xi = x + 3 * i, *vi = v + 3 * i;
        R radius_i = radius[i];
        if (mass[tid] == 0) continue; /* the reference tests tid here */
        R w1 = (R)1 / mass[i];
        R qr = radius_i + max_radius + k_cohesion * radius_i;
        int xs = (int)((xi[0] - qr) * inv), ys = (int)((xi[1] - qr) * inv), zs = (int)((xi[2] - qr) * inv);
        int xe = (int)((xi[0] + qr) * inv), ye = (int)((xi[1] + qr) * inv), ze = (int)((xi[2] + qr) * inv);
        if (xe > xs + dimx - 1) xe = xs + dimx - 1;
        if (ye > ys + dimy - 1) ye = ys + dimy - 1;
        if (ze > zs + dimz - 1) ze = zs + dimz - 1;
        R k_mu = k_mu_in;
        R delta[3] = {0, 0, 0};
        const R alpha = 0;
        const int nx = xe - xs + 1, ny = ye - ys + 1, nz = ze - zs + 1;
        const int ncell = (nx > 0 && ny > 0 && nz > 0) ? nx * ny * nz : 0;
        for (int q = 0; q < ncell; q++) {
            int cx, cy, cz;
            if (order == 2) {
                cz = zs + q % nz; cy = ys + (q / nz) % ny; cx = xs + q / (nz * ny);
            } else {
                const int r = order == 1 ? ncell - 1 - q : q;
                cx = xs + r % nx; cy = ys + (r / nx) % ny; cz = zs + r / (nx * ny);
            }
            int cell = hash_index(dimx, dimy, dimz, cx, cy, cz);
            const int c0 = cell_starts[cell], c1 = cell_ends[cell];
            for (int u = c0; u < c1; u++) {
                const int s = order == 1 ? c0 + c1 - 1 - u : u;
                int j = point_ids[s];
                if (!(active[j] == 1 && j != i && mass[j] > 0)) continue;
                if (mat[i] <= 1 && mat[j] <= 1) continue;
                R nv[3] = {xi[0] - x[3 * j], xi[1] - x[3 * j + 1], xi[2] - x[3 * j + 2]};
                R d = vlen(nv);
                R err = d - radius_i - radius[j];
                R w2 = (R)1 / mass[j];
                R denom = w1 + w2;
                if (err <= k_cohesion && denom > 0) {
                    if (lock == 1) {
                        if (mat[i] <= 1 && mat[j] == 2) mat[i] = 0;
                        if (mat[i] == 2 && mat[j] <= 1) mat[j] = 0;
                    }
                }
                if (err <= k_cohesion && denom > 0) {
                    R nn[3] = {nv[0] / d, nv[1] / d, nv[2] / d};
                    R vrel[3] = {vi[0] - v[3 * j], vi[1] - v[3 * j + 1], vi[2] - v[3 * j + 2]};
                    R lambda_n = err / ((R)1 + (alpha * dt * dt));
                    R dn[3] = {nn[0] * lambda_n, nn[1] * lambda_n, nn[2] * lambda_n};
                    R vn = nn[0] * vrel[0] + nn[1] * vrel[1] + nn[2] * vrel[2];
                    if (fabs((double)vn) < thr * radius_i) k_mu = static_mu;
                    R vt[3] = {vrel[0] - nn[0] * vn, vrel[1] - nn[1] * vn, vrel[2] - nn[2] * vn};
                    R a_ = k_mu * lambda_n, b_ = -vlen(vt) * dt;
                    R lf = a_ > b_ ? a_ : b_;
                    R nt[3];
                    vnormalize(vt, nt);
                    for (int a = 0; a < 3; a++) delta[a] += (nt[a] * lf - dn[a]) / denom * w1;
                }
            }
        }
        for (int a = 0; a < 3; a++) deltas[3 * i + a] += delta[a] * relax;
    }
}

EXPORT void o_solve_particle_particle_contacts(int n, const int *active, int *mat,
                                               const int *point_ids, const int *cell_starts,
                                               const int *cell_ends, R cell_width, int dimx,
                                               int dimy, int dimz, const R *x, const R *v,
                                               const R *mass, const R *radius, R k_mu_in, R thr,
                                               R static_mu, R k_cohesion, R max_radius, R dt,
                                               R relax, int lock, R *deltas) {
    solve_contacts(n, active, mat, point_ids, cell_starts, cell_ends, cell_width, dimx, dimy, dimz, x, v, mass,
                   radius, k_mu_in, thr, static_mu, k_cohesion, max_radius, dt, relax, lock, deltas, 0);
}

EXPORT void o_solve_particle_particle_contacts_ordered(int n, const int *active, int *mat,
                                                       const int *point_ids, const int *cell_starts,
                                                       const int *cell_ends, R cell_width, int dimx,
                                                       int dimy, int dimz, const R *x, const R *v,
                                                       const R *mass, const R *radius, R k_mu_in, R thr,
                                                       R static_mu, R k_cohesion, R max_radius, R dt,
                                                       R relax, int lock, R *deltas, int order) {
    solve_contacts(n, active, mat, point_ids, cell_starts, cell_ends, cell_width, dimx, dimy, dimz, x, v, mass,
                   radius, k_mu_in, thr, static_mu, k_cohesion, max_radius, dt, relax, lock, deltas, order);
}

/* X7 my_apply_particle_deltas (xpbdRoutines.py:221-257) */
EXPORT void o_apply_particle_deltas(int n, const int *active, const int *mat, const R *x_orig,
                                    const R *x_pred, const R *delta, R dt, R v_max, R *x_out,
                                    R *v_out) {
    (void)mat;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; t++) {
        if (active[t] != 1) continue;
        R d[3] = {delta[3 * t], delta[3 * t + 1], delta[3 * t + 2]};
        if (isnan((double)d[0]) || isnan((double)d[1]) || isnan((double)d[2])) d[0] = d[1] = d[2] = 0;
        R xn[3], vn[3];
        for (int a = 0; a < 3; a++) {
            xn[a] = x_pred[3 * t + a] + d[a];
            vn[a] = (xn[a] - x_orig[3 * t + a]) / dt;
        }
        R m = vlen(vn);
        if (m > v_max) {
            R s = v_max / m;
            vn[0] *= s; vn[1] *= s; vn[2] *= s;
        }
        for (int a = 0; a < 3; a++) { x_out[3 * t + a] = xn[a]; v_out[3 * t + a] = vn[a]; }
    }
}

/* X9 apply_mpm_contact_impulses (xpbdRoutines.py:259-310) */
EXPORT void o_apply_mpm_contact_impulses(int n, const int *active, const int *mat, const R *x_orig,
                                         const R *v_orig, R *x_final, const R *v_final, R dt,
                                         R v_max, R k, int xpbd_only, R *v_out) {
    (void)v_final;
#pragma omp parallel for schedule(static)
    for (int t = 0; t < n; t++) {
        if (!(active[t] == 1 && mat[t] <= 1)) continue;
        for (int a = 0; a < 3; a++) x_final[3 * t + a] = x_orig[3 * t + a];
        if (xpbd_only == 0) {
            R vc[3];
            for (int a = 0; a < 3; a++)
                vc[a] = v_orig[3 * t + a] + (x_final[3 * t + a] - x_orig[3 * t + a]) / dt * k;
            R m = vlen(vc);
            if (m > v_max) {
                R s = v_max / m;
                vc[0] *= s; vc[1] *= s; vc[2] *= s;
            }
            for (int a = 0; a < 3; a++) v_out[3 * t + a] = vc[a];
        }
    }
}

/* X8 sleepParticles (xpbdRoutines.py:313-334); counters[0]=total, [1]=sleeping */
EXPORT void o_sleep_particles(int n, const int *active, const int *mat, R thr, const R *radius,
                              const R *x_prev, R *x_after, R *dist_total, R dt, int *counters) {
    for (int t = 0; t < n; t++) {
        if (active[t] != 1 || mat[t] != 2) continue;
        counters[0]++;
        R dv[3] = {x_after[3 * t] - x_prev[3 * t], x_after[3 * t + 1] - x_prev[3 * t + 1],
                   x_after[3 * t + 2] - x_prev[3 * t + 2]};
        R d = vlen(dv);
        if (d / dt < thr * radius[t]) {
            for (int a = 0; a < 3; a++) x_after[3 * t + a] = x_prev[3 * t + a];
            counters[1]++;
        }
        dist_total[t] = dist_total[t] + d;
    }
}

/* X10 swellParticlesType2 (xpbdRoutines.py:354-376) */
EXPORT void o_swell_particles_type2(int n, const int *active, const int *mat, const R *x_init,
                                    const R *x, R *radius, R factor_in, R factor_max_in,
                                    const R *swell_r, const R *base_r) {
    for (int t = 0; t < n; t++) {
        if (active[t] != 1 || mat[t] != 2) continue;
        R dv[3] = {x[3 * t] - x_init[3 * t], x[3 * t + 1] - x_init[3 * t + 1], x[3 * t + 2] - x_init[3 * t + 2]};
        R dxl = vlen(dv) * base_r[t];
        R factor = factor_in * base_r[t];
        R factor_max = factor_max_in * base_r[t];
        if (dxl > factor) {
            R nr = base_r[t] + (swell_r[t] - base_r[t]) * (dxl - factor) / (factor_max - factor);
            if (nr < radius[t]) nr = radius[t];
            if (nr > swell_r[t]) nr = swell_r[t];
            radius[t] = nr;
        }
    }
}

/* ------------------------------------------------------------------ */
/* A12 reductions (simulationRoutines.py:538-610), accumulated in the  */
/* oracle's REAL type in particle order                                */
/* ------------------------------------------------------------------ */
EXPORT double o_velocity_convergence(int n, const int *active, const R *v, const R *radius) {
    R s = 0;
    for (int p = 0; p < n; p++)
        if (active[p] == 1) s += vlen(v + 3 * p) / radius[p];
    return (double)s;
}
EXPORT int o_count_active(int n, const int *active) {
    int c = 0;
    for (int p = 0; p < n; p++) c += active[p] == 1;
    return c;
}
EXPORT void o_count_by_type(int n, const int *active, const int *mat, int *out4) {
    out4[0] = out4[1] = out4[2] = out4[3] = 0; /* activeMPM, activeXPBD, inactiveMPM, inactiveXPBD */
    for (int p = 0; p < n; p++) {
        int a = active[p] == 1, xp = mat[p] == 2;
        if (xp) { if (a) out4[1]++; else out4[3]++; }
        else { if (a) out4[0]++; else out4[2]++; }
    }
}
EXPORT double o_sum_active(int n, const R *arr, const int *active, int *count) {
    R s = 0;
    int c = 0;
    for (int p = 0; p < n; p++)
        if (active[p] == 1) { s += arr[p]; c++; }
    *count = c;
    return (double)s;
}
