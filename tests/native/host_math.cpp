// Host build of the device math of libmpmx (the constitutive update, the SVD
// and the stencil helpers are __host__ __device__), so that the CPU test suite
// can run the exact code the gfx950 kernels run against the oracle.
#include "../../mpmwarpydw_amd/csrc/mpmx_constitutive.hpp"

extern "C" {

// the per-particle loop of compute_stress_from_F_trial (mpmRoutines.py:229-312)
void host_constitutive(const mpmx_particles* P, float eff) {
    for (long p = 0; p < P->n; p++) {
        if (P->activeLabel[p] != 1 || P->materialLabel[p] > 1) continue;
        float T[9];
        mpmx::constitutive_update(*P, p, eff, T);
    }
}

// which route the update takes for each F (1 = series, 0 = SVD); F is [9][n]
void host_route(long n, const float* F, int* fast) {
    for (long p = 0; p < n; p++) {
        float Ft[9];
        mpmx::ld9(F, n, p, Ft);
        float Bs[6], nb;
        fast[p] = mpmx::left_strain(Ft, Bs, nb);
    }
}

// log(I + B) and exp(S) of symmetric matrices given as [6] rows {xx,yy,zz,xy,xz,yz}
void host_sym_log1p(long n, const float* B, float* L) {
    for (long p = 0; p < n; p++) mpmx::sym_log1p(B + 6 * p, mpmx::sym_norm(B + 6 * p), L + 6 * p);
}
void host_sym_exp(long n, const float* S, float* X) {
    for (long p = 0; p < n; p++) mpmx::sym_exp(S + 6 * p, mpmx::sym_norm(S + 6 * p), X + 6 * p);
}

// the G2P sums of particles with fractional offsets fx [n][3] over node
// velocities v [n][27][3] (node (i, j, k) at 9 i + 3 j + k): nv [n][3],
// nB [n][9], nG [n][9]
void host_g2p_gather(long n, const float* fx, const float* v, float inv_dx, float* nv, float* nB, float* nG) {
    for (long p = 0; p < n; p++) {
        const float* vp = v + 81 * p;
        mpmx::g2p_gather(
            fx + 3 * p, inv_dx,
            [&](int i, int j, int k) {
                const float* q = vp + 3 * (9 * i + 3 * j + k);
                return mpmx::f3{q[0], q[1], q[2]};
            },
            nv + 3 * p, nB + 9 * p, nG + 9 * p);
    }
}

}  // extern "C"
