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

// the 10 features of a plastic trial state from its invariants: |dev e| = dn,
// the yield value Y and the unit principal deviator dh (descending)
MPMX_HD void lrm_features(float dn, float Y, const float dh[3], float mu, float lam, float ys, float al, float D,
                          float x[10]) {
    const float E = mu * (3.0f * lam + 2.0f * mu) / (lam + mu);
    const float nu = lam / (2.0f * (lam + mu));
    const float c = ys > 1.0f ? ys : 1.0f;
    const float kap = 2.0f * mu * 1.0e-6f / c, y0 = Y / c;
    const float pk = (2.0f * mu * dn - Y) / c + kap;
    x[0] = 0.1f * logf(pk > 1.0e-30f ? pk : 1.0e-30f);
    x[1] = 0.1f * logf(y0 > 1.0e-12f ? y0 : 1.0e-12f);
    x[2] = 0.1f * logf(kap);
    for (int k = 0; k < 3; k++) x[3 + k] = dh[k];
    x[6] = lrm_tan_of_alpha(al);
    x[7] = 4.0f * nu - 1.0f;
    x[8] = log10f(E) - 8.0f;
    x[9] = D;
}

// Eigenvalues, descending, of the unit deviator A = dev(E) / |dev E| of a
// symmetric E = {xx, yy, zz, xy, xz, yz} with trace esum and |dev E| = devn > 0
// (trigonometric solution of the characteristic cubic, Smith 1961: with
// J2 = 1/2 and cos 3t = 3 sqrt6 det A, the roots are sqrt(2/3) cos(t - 2 pi k / 3)).
MPMX_HD void unit_deviator_eigenvalues(const float E[6], float esum, float devn, float dh[3]) {
    const float m = esum / 3.0f, inv = 1.0f / devn;
    const float a = (E[0] - m) * inv, b = (E[1] - m) * inv, c = (E[2] - m) * inv;
    const float d = E[3] * inv, e = E[4] * inv, f = E[5] * inv;
    const float det = a * (b * c - f * f) - d * (d * c - f * e) + e * (d * f - b * e);
    float x = 7.348469228349534f * det;  // 3 sqrt6
    x = x > 1.0f ? 1.0f : (x < -1.0f ? -1.0f : x);
    const float t = acosf(x) / 3.0f;
    const float r = 0.816496580927726f;  // sqrt(2/3)
    dh[0] = r * cosf(t);
    dh[2] = r * cosf(t + 2.0943951023931953f);
    dh[1] = -dh[0] - dh[2];
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
    float x[10], dh[3];
    for (int k = 0; k < 3; k++) dh[k] = s.dn > 0.0f ? s.d[k] / s.dn : 0.0f;
    lrm_features(s.dn, s.Y, dh, mu, lam, ys, al, D, x);
    for (int k = 0; k < 10; k++) X[p * 10 + k] = x[k];
    Tg[p * 3 + 0] = s.plastic ? s.Y / (2.0f * mu * (s.dn + 1.0e-6f)) : 0.0f;
    Tg[p * 3 + 1] = s.plastic ? 20.0f * (al - 0.3f) : 0.0f;
    Tg[p * 3 + 2] = 0.0f;
    for (int k = 0; k < 3; k++) E3[p * 3 + k] = eps[k];
    flag[p] = s.plastic ? 1 : 0;
}

// a plastic state the network may project: stays MPM, Y >= 0 (not past the
// cone's apex) and a deviator to scale
MPMX_HD bool lrm_candidate(const CuState& c) { return c.branch == 1 && c.yield_eff >= 0.0f && c.devn > 0.0f; }

// wave-level sums of the counters, one 64-bit atomic per wavefront and counter
__device__ __forceinline__ void add_counters(int64_t* counters, int lane, int a, int b, int c) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        a += __shfl_xor(a, o);
        b += __shfl_xor(b, o);
        c += __shfl_xor(c, o);
    }
    if (lane == 0 && counters) {
        unsigned long long* cu = reinterpret_cast<unsigned long long*>(counters);
        if (a) atomicAdd(&cu[0], (unsigned long long)a);
        if (b) atomicAdd(&cu[1], (unsigned long long)b);
        if (c) atomicAdd(&cu[2], (unsigned long long)c);
    }
}

// Pass 1, the plain update for every particle that is not the network's: each
// active MPM particle runs cu_begin (invariants and branch, as
// compute_stress_from_F_trial decides them); with use_net the candidates are
// flagged and left for pass 2, everything else finishes here.  Wavefronts
// take 64 consecutive particles per round.  counters: [0] += plastic states,
// [2] += particles updated.
__global__ void __launch_bounds__(kBlock) stress_learned_pass1_kernel(mpmx_particles P, float eff, int use_net,
                                                                      uint8_t* __restrict__ flag,
                                                                      int64_t* __restrict__ counters) {
    const int lane = threadIdx.x & 63;
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * 4;
    int n_plastic = 0, n_done = 0;
    for (long base = wave * 64; base < P.n; base += nwaves * 64) {
        const long p = base + lane;
        if (p >= P.n) continue;
        uint8_t f = 0;
        if (P.activeLabel[p] == 1 && P.materialLabel[p] <= 1) {
            CuState c;
            cu_begin(P, p, c);
            if (use_net && lrm_candidate(c)) {
                f = 1;
            } else {
                n_done++;
                n_plastic += c.branch != 0 ? 1 : 0;
                float T[9];
                cu_finish(P, p, eff, c, T);
            }
        }
        if (use_net) flag[p] = f;
    }
    add_counters(counters, lane, n_plastic, 0, n_done);
}

// The k <= 64 candidates q[0..k) of one wavefront, one per lane: features,
// the network on the matrix cores (ceil(k / 16) tiles of 16), the judge and
// the finish.  The update state of cu_begin stays in registers across the
// tiles.
__device__ __forceinline__ void lrm_process(const mpmx_particles& P, float eff, const MlpWeights& w, const int* q,
                                            int k, int lane, float tol, int& n_plastic, int& n_accept, int& n_done) {
    const bool act = lane < k;
    const long p = act ? q[lane] : 0;
    float x[10];
    for (int j = 0; j < 10; j++) x[j] = 0.0f;
    CuState c;
    if (act) {
        cu_begin(P, p, c);
        if (lrm_candidate(c)) {
            float dh[3];
            if (c.fast) {
                unit_deviator_eigenvalues(c.E6, c.esum, c.devn, dh);
            } else {
                const float m = c.esum / 3.0f;
                for (int j = 0; j < 3; j++) dh[j] = (c.eps[j] - m) / c.devn;
            }
            lrm_features(c.devn, c.yield_eff, dh, c.mu, c.lam, c.ys, c.al, c.D, x);
        }
    }
    float yo[3] = {0.0f, 0.0f, 0.0f};
    const int ntiles = (k + 15) >> 4;  // uniform over the wavefront
    for (int t = 0; t < ntiles; t++) {
        // B fragment: lane l takes features 4(l/16)..+3 of candidate t*16 + l%16
        const int src = t * 16 + (lane & 15);
        const int f0 = 4 * (lane >> 4);
        float xv[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 10; j++) {
            const float v = __shfl(x[j], src);
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (f0 + i == j) xv[i] = v;
        }
        MlpTile tt;
        mlp_tile_forward(w, pack_bf16(xv[0], xv[1], xv[2], xv[3]), pack_bf16_lo(xv[0], xv[1], xv[2], xv[3]), lane, tt);
        // lanes 0..15 hold the outputs of candidates t*16 + 0..15
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const float v = __shfl(tt.y[j], lane & 15);
            if ((lane >> 4) == t) yo[j] = v;
        }
    }
    if (!act) return;
    n_done++;
    n_plastic += c.branch != 0 ? 1 : 0;
    float T[9];
    if (!lrm_candidate(c)) {  // pass 1 flagged it with the same state: not reached
        cu_finish(P, p, eff, c, T);
        return;
    }
    // e' = e - r dev(e) + volc with the network's flow split and cone coefficient
    const float mu = c.mu, Y = c.yield_eff, dn = c.devn;
    const float cc = c.ys > 1.0f ? c.ys : 1.0f;
    const float er = cc / (2.0f * mu), kap = 2.0f * mu * 1.0e-6f / cc;
    float rv[2];
    rv[0] = (1.0f - Y / (2.0f * mu * dn)) + er * kap * yo[0] / dn;
    rv[1] = 0.3f * (0.3f + 0.05f * yo[1]) * (er * ((2.0f * mu * dn - Y) / cc + kap));
    // the judge: dev e' against Y / (2 mu edn) dev e, tr e' against tr e + 0.9 alpha dgam
    const float edn = dn + 1.0e-6f;
    const float res1 = 2.0f * mu * fabsf((1.0f - rv[0]) - Y / (2.0f * mu * edn)) * dn;
    const float K = c.lam + 2.0f * mu / 3.0f;
    const float res2 = K * fabsf(3.0f * rv[1] - 0.9f * c.al * c.dgam);
    const float lim = tol * cc;
    const bool ok = res1 <= lim && res2 <= lim;
    n_accept += ok ? 1 : 0;
    cu_finish(P, p, eff, c, T, ok ? rv : nullptr);
}

// Pass 2, the flagged candidates: each wavefront scans its rounds of 64
// particles and queues the flagged ones in LDS; every 64 queued (and the rest
// at the end) go through lrm_process with one candidate per lane, so the
// network's tiles are full.  counters: [0] += plastic, [1] += accepted,
// [2] += updated.
__global__ void __launch_bounds__(kBlock) stress_learned_pass2_kernel(mpmx_particles P, float eff,
                                                                      const float* __restrict__ prm, float tol,
                                                                      const uint8_t* __restrict__ flag,
                                                                      int64_t* __restrict__ counters) {
    __shared__ MlpWeights w;
    __shared__ int queue[4][128];
    load_weights(prm, w);
    __syncthreads();
    const int lane = threadIdx.x & 63;
    int* q = queue[threadIdx.x >> 6];
    const long wave = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    const long nwaves = (long)gridDim.x * 4;
    const long nchunks = ((long)P.n + 63) / 64;
    int n_plastic = 0, n_accept = 0, n_done = 0;
    int qn = 0;  // queued candidates, uniform over the wavefront, at most 63 between rounds
    for (long ch = wave;; ch += nwaves) {
        const bool more = ch < nchunks;  // uniform
        if (more) {
            const long p = ch * 64 + lane;
            const bool f = p < P.n && flag[p];
            const unsigned long long m = __ballot(f);
            if (f) q[qn + __popcll(m & ((1ull << lane) - 1ull))] = (int)p;
            qn += __popcll(m);
            __builtin_amdgcn_wave_barrier();
        }
        while (qn >= 64 || (!more && qn > 0)) {
            const int k = qn < 64 ? qn : 64;
            lrm_process(P, eff, w, q, k, lane, tol, n_plastic, n_accept, n_done);
            const int rest = qn - k;  // move the rest to the front
            const int v = lane < rest ? q[k + lane] : 0;
            __builtin_amdgcn_wave_barrier();
            if (lane < rest) q[lane] = v;
            __builtin_amdgcn_wave_barrier();
            qn = rest;
        }
        if (!more) break;
    }
    add_counters(counters, lane, n_plastic, n_accept, n_done);
}

static int check_particles(const mpmx_particles* P) {
    MPMX_REQUIRE(P && P->n >= 0, "learned: bad particle struct");
    MPMX_REQUIRE(P->n == 0 || (P->activeLabel && P->materialLabel && P->F_trial && P->mu && P->lam && P->ys &&
                               P->alpha && P->damage && P->acc_strain),
                 "learned: null particle array");
    return 0;
}

int launch_stress_learned(const mpmx_particles& P, float eff, const float* params, float tol, int64_t* counters,
                          uint8_t* flag, hipStream_t s) {
    if (P.n == 0) return 0;
    const long waves = ((long)P.n + 63) / 64;
    long nb = (waves + 3) / 4;
    if (nb > 8192) nb = 8192;
    const int use_net = params ? 1 : 0;
    MPMX_REQUIRE(!use_net || flag, "learned stress update: no flag buffer");
    MPMX_LAUNCH(stress_learned_pass1_kernel, dim3((unsigned)nb), dim3(kBlock), 0, s, P, eff, use_net, flag, counters);
    int rc = check_launch("stress_learned(plain)", s);
    if (rc || !use_net) return rc;
    // pass 2 with a quarter of the waves per SIMD of pass 1 at most: every wavefront then
    // scans enough rounds to fill its batches of 64 candidates (with 8192 workgroups a
    // wavefront met about 15 candidates, and three quarters of its lanes idled)
    const long nb2 = nb < 1024 ? nb : 1024;
    MPMX_LAUNCH(stress_learned_pass2_kernel, dim3((unsigned)nb2), dim3(kBlock), 0, s, P, eff, params, tol, flag,
                counters);
    return check_launch("stress_learned(network)", s);
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

size_t mpmx_lrm_workspace_bytes(int n) { return n > 0 ? (((size_t)n + 255) / 256) * 256 : 256; }

int mpmx_compute_stress_learned(const mpmx_particles* P, float eff, const float* params, float tol,
                                int64_t* counters, void* workspace, size_t ws_bytes, void* stream) {
    int rc = check_particles(P);
    if (rc) return rc;
    if (P->n == 0) return 0;
    MPMX_REQUIRE(P->x && P->v && P->x_init_xpbd && P->v_init_xpbd && P->F && P->C && P->stress && P->hardening &&
                     P->softening && P->density && P->strainCriteria && P->eta_shear && P->eta_bulk,
                 "mpmx_compute_stress_learned: null particle array");
    MPMX_REQUIRE(tol >= 0.0f, "mpmx_compute_stress_learned: negative tolerance");
    MPMX_REQUIRE(!params || (workspace && ws_bytes >= mpmx_lrm_workspace_bytes(P->n)),
                 "mpmx_compute_stress_learned: workspace smaller than mpmx_lrm_workspace_bytes(n)");
    return launch_stress_learned(*P, eff, params, tol, counters, static_cast<uint8_t*>(workspace),
                                 as_stream(stream));
}

}  // extern "C"
