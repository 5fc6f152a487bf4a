/*
 * CPU oracle of the learned return map (TEST INFRASTRUCTURE; see mpm_oracle.c).
 *
 * The network is a 10 -> 64 -> 64 -> 3 MLP with GELU activations (tanh form,
 * Hendrycks & Gimpel 2016: 0.5 x (1 + tanh(sqrt(2/pi) (x + 0.044715 x^3)))).
 * Parameters are one flat vector: W1[64][10], b1[64], W2[64][64], b2[64],
 * W3[3][64], b3[3] (row-major, out x in).
 *
 * With quant = 1 the oracle rounds to bfloat16 (round to nearest even) at the
 * points where the GPU kernels feed their MFMA units: the weights and, in the
 * backward pass, the back-propagated gradients; the inputs and the hidden
 * activations enter as the sum of two bfloat16 values, hi = bf16(x) and
 * lo = bf16(x - hi) (two MFMA passes), which keeps 16 bits of each value.
 * Sums and everything else stay in REAL.  The loss is the mean squared error
 * over all B*3 outputs.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifndef REAL
#define REAL double
#endif
typedef REAL R;
#define EXPORT __attribute__((visibility("default")))

#define NI 10
#define NH 64
#define NO 3
#define NPARAM (NH * NI + NH + NH * NH + NH + NO * NH + NO)

EXPORT int o_mlp_nparams(void) { return NPARAM; }

/* round a float32 value to the nearest bfloat16 (ties to even), back in R */
static R bf16(R v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) != 0x7f800000u) { /* finite */
        uint32_t lsb = (u >> 16) & 1u;
        u += 0x7fffu + lsb;
    }
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (R)f;
}
EXPORT double o_bf16(double v) { return (double)bf16((R)v); }

static R q(R v, int quant) { return quant ? bf16(v) : v; }
/* an input as the GPU feeds it: bf16(x) + bf16(x - bf16(x)) */
static R qin(R v, int quant) {
    if (!quant) return v;
    const R hi = bf16(v);
    return hi + bf16(v - hi);
}
EXPORT double o_bf16_split(double v) { return (double)qin((R)v, 1); }

static R gelu(R x) {
    const R c = (R)0.7978845608028654; /* sqrt(2/pi) */
    R u = c * (x + (R)0.044715 * x * x * x);
    return (R)0.5 * x * ((R)1 + (R)tanh((double)u));
}
static R gelu_grad(R x) {
    const R c = (R)0.7978845608028654;
    R u = c * (x + (R)0.044715 * x * x * x);
    R t = (R)tanh((double)u);
    R du = c * ((R)1 + (R)3 * (R)0.044715 * x * x);
    return (R)0.5 * ((R)1 + t) + (R)0.5 * x * ((R)1 - t * t) * du;
}
EXPORT double o_gelu(double x) { return (double)gelu((R)x); }
EXPORT double o_gelu_grad(double x) { return (double)gelu_grad((R)x); }

struct Params {
    const R *W1, *b1, *W2, *b2, *W3, *b3;
};
static struct Params split(const R *p) {
    struct Params P;
    P.W1 = p;
    P.b1 = P.W1 + NH * NI;
    P.W2 = P.b1 + NH;
    P.b2 = P.W2 + NH * NH;
    P.W3 = P.b2 + NH;
    P.b3 = P.W3 + NO * NH;
    return P;
}

/* forward of one sample; z1, z2 pre-activations, h1, h2 activations (as fed on) */
static void forward1(const struct Params *P, const R *x, R *z1, R *h1, R *z2, R *h2, R *y, int quant) {
    R xq[NI];
    for (int i = 0; i < NI; i++) xq[i] = qin(x[i], quant);
    for (int o = 0; o < NH; o++) {
        R s = 0;
        for (int i = 0; i < NI; i++) s += q(P->W1[o * NI + i], quant) * xq[i];
        z1[o] = s + P->b1[o];
        h1[o] = qin(gelu(z1[o]), quant);
    }
    for (int o = 0; o < NH; o++) {
        R s = 0;
        for (int i = 0; i < NH; i++) s += q(P->W2[o * NH + i], quant) * h1[i];
        z2[o] = s + P->b2[o];
        h2[o] = qin(gelu(z2[o]), quant);
    }
    for (int o = 0; o < NO; o++) {
        R s = 0;
        for (int i = 0; i < NH; i++) s += q(P->W3[o * NH + i], quant) * h2[i];
        y[o] = s + P->b3[o];
    }
}

EXPORT void o_mlp_forward(int B, const R *params, const R *X, R *Y, int quant) {
    struct Params P = split(params);
    R z1[NH], h1[NH], z2[NH], h2[NH];
    for (int b = 0; b < B; b++) forward1(&P, X + (long)b * NI, z1, h1, z2, h2, Y + (long)b * NO, quant);
}

/* mean squared error over B*NO outputs and its gradient w.r.t. every parameter */
EXPORT double o_mlp_loss_grad(int B, const R *params, const R *X, const R *T, R *grad, int quant) {
    struct Params P = split(params);
    memset(grad, 0, sizeof(R) * NPARAM);
    R *gW1 = grad, *gb1 = gW1 + NH * NI, *gW2 = gb1 + NH, *gb2 = gW2 + NH * NH, *gW3 = gb2 + NH,
      *gb3 = gW3 + NO * NH;
    R z1[NH], h1[NH], z2[NH], h2[NH], y[NO];
    double loss = 0;
    const R scale = (R)2 / (R)((long)B * NO);
    for (int b = 0; b < B; b++) {
        const R *x = X + (long)b * NI;
        forward1(&P, x, z1, h1, z2, h2, y, quant);
        R g3[NO];
        for (int o = 0; o < NO; o++) {
            R d = y[o] - T[(long)b * NO + o];
            loss += (double)(d * d);
            g3[o] = d * scale;
        }
        R g3q[NO];
        for (int o = 0; o < NO; o++) {
            g3q[o] = q(g3[o], quant);
            gb3[o] += g3[o];
            for (int i = 0; i < NH; i++) gW3[o * NH + i] += g3q[o] * h2[i];
        }
        R g2[NH], g2q[NH];
        for (int i = 0; i < NH; i++) {
            R s = 0;
            for (int o = 0; o < NO; o++) s += q(P.W3[o * NH + i], quant) * g3q[o];
            g2[i] = s * gelu_grad(z2[i]);
            g2q[i] = q(g2[i], quant);
            gb2[i] += g2[i];
        }
        for (int o = 0; o < NH; o++)
            for (int i = 0; i < NH; i++) gW2[o * NH + i] += g2q[o] * h1[i];
        R g1[NH];
        for (int i = 0; i < NH; i++) {
            R s = 0;
            for (int o = 0; o < NH; o++) s += q(P.W2[o * NH + i], quant) * g2q[o];
            g1[i] = s * gelu_grad(z1[i]);
            gb1[i] += g1[i];
        }
        for (int o = 0; o < NH; o++) {
            R g1q = q(g1[o], quant);
            for (int i = 0; i < NI; i++) gW1[o * NI + i] += g1q * qin(x[i], quant);
        }
    }
    return loss / (double)((long)B * NO);
}

/* Adam step on the flat parameter vector (Kingma & Ba 2015), t >= 1 */
EXPORT void o_adam_step(int n, R *params, const R *grad, R *m, R *v, R lr, R beta1, R beta2, R eps, int t) {
    R c1 = (R)1 - (R)pow((double)beta1, t), c2 = (R)1 - (R)pow((double)beta2, t);
    for (int i = 0; i < n; i++) {
        m[i] = beta1 * m[i] + ((R)1 - beta1) * grad[i];
        v[i] = beta2 * v[i] + ((R)1 - beta2) * grad[i] * grad[i];
        R mh = m[i] / c1, vh = v[i] / c2;
        params[i] -= lr * mh / ((R)sqrt((double)vh) + eps);
    }
}

/* ------------------------------------------------------------------------ */
/* Learned return map: features, exact labels and the judge.                 */
/*                                                                          */
/* The return map of compute_stress_from_F_trial (mpmRoutines.py:298-441)    */
/* acts on the principal trial log strains e (descending, as the SVD orders  */
/* them).  With d = e - mean(e), edn = |d| + 1e-6, p = mean(2 mu e + lam     */
/* sum(e)) and Y = (1 - D)(ys - alpha p), a trial state is plastic when      */
/* 2 mu |d| > Y or ys < 0 or Y < 0 (mpmRoutines.py:357), and the projection  */
/* is e' = e - (dgam / edn) d + 0.3 alpha dgam with dgam = edn - Y / (2 mu)  */
/* (mpmRoutines.py:360-413).                                                 */
/*                                                                          */
/* With c = max(ys, 1), eref = c / (2 mu) (the yield strain), the overshoot  */
/* phi = (2 mu |d| - Y) / c, y0 = Y / c and kap = 2 mu 1e-6 / c, the exact   */
/* projection is the classical radial return e0' = e - (1 - Y/(2 mu |d|)) d  */
/* (von Mises' projection onto the yield surface at the trial pressure)      */
/* plus two corrections, in units of eref:                                   */
/*   e' = e0' - eref kap s d/|d| + 0.3 alpha eref (phi + kap) (1, 1, 1),     */
/*   s = y0 / (phi + y0 + kap)   (the part of the deviator that the 1e-6 in  */
/*                                edn keeps: s = Y / (2 mu edn)),            */
/*   alpha                       (the Drucker-Prager cone coefficient, which */
/*                                sets the dilatancy).                       */
/* The network predicts y = (s, 20 (alpha - 0.3), 0) (alpha centred and     */
/* scaled, so that the loss weighs its error as the dilatancy needs) from 10 */
/* features: 0.1 ln(phi+kap),                                                */
/* 0.1 ln(max(y0, 1e-12)), 0.1 ln kap, d/|d| (3), tan(friction angle),       */
/* 4 nu - 1, log10 E - 8, damage.  The tolerance of the judge then asks for  */
/* kap |ds| and 0.9 (phi + kap) |dalpha| (2 mu / K) below tol: loose where   */
/* the overshoot is small, as it is within one substep.  States with Y < 0   */
/* (tension past the cone's apex) are left to the exact projection.          */
static R eref_of(R mu, R ys) { return (ys > (R)1 ? ys : (R)1) / ((R)2 * mu); }

static R tan_of_alpha(R a) {
    /* invert alpha = 2 s / (sqrt3 (3 - s)):  s = 3 sqrt3 a / (2 + sqrt3 a) */
    const R r3 = (R)1.7320508075688772;
    R s = (R)3 * r3 * a / ((R)2 + r3 * a);
    if (s > (R)0.999999) s = (R)0.999999;
    return s / (R)sqrt((double)((R)1 - s * s));
}

struct Trial {
    R d[3], dn, sum, Y;
};
static struct Trial trial(const R *e, R mu, R lam, R ys, R alpha, R D) {
    struct Trial t;
    R mean = (e[0] + e[1] + e[2]) / (R)3;
    for (int k = 0; k < 3; k++) t.d[k] = e[k] - mean;
    t.sum = e[0] + e[1] + e[2];
    R p = ((R)2 * mu * e[0] + lam * t.sum + (R)2 * mu * e[1] + lam * t.sum + (R)2 * mu * e[2] + lam * t.sum) / (R)3;
    t.dn = (R)sqrt((double)(t.d[0] * t.d[0] + t.d[1] * t.d[1] + t.d[2] * t.d[2]));
    t.Y = ((R)1 - D) * (ys - alpha * p);
    return t;
}

/* exact principal projection; returns 1 when the trial state is plastic */
static int exact_principal(const R *e, R mu, R lam, R ys, R alpha, R D, R *e_new, R *dgam_out) {
    struct Trial t = trial(e, mu, lam, ys, alpha, D);
    for (int k = 0; k < 3; k++) e_new[k] = e[k];
    *dgam_out = 0;
    if (!((R)2 * mu * t.dn > t.Y || ys < 0 || t.Y < 0)) return 0;
    R edn = t.dn + (R)1e-6;
    R dgam = edn - t.Y / ((R)2 * mu);
    R beta = (R)0.3 * alpha;
    for (int k = 0; k < 3; k++) e_new[k] = e[k] - (dgam / edn) * t.d[k] + beta * dgam;
    *dgam_out = dgam;
    return 1;
}

/* the radial return e0' = e - r0 d, r0 = 1 - Y / (2 mu |d|) (0 for |d| = 0) */
static void radial_return(const R *e, const struct Trial *t, R mu, R *e0) {
    R r0 = t->dn > 0 ? (R)1 - t->Y / ((R)2 * mu * t->dn) : (R)0;
    for (int k = 0; k < 3; k++) e0[k] = e[k] - r0 * t->d[k];
}

EXPORT void o_lrm_features(int n, const R *e, const R *mu, const R *lam, const R *ys, const R *alpha,
                           const R *acc, const R *damage, R *X) {
    (void)acc;
    for (int p = 0; p < n; p++) {
        R m = mu[p], l = lam[p];
        R E = m * ((R)3 * l + (R)2 * m) / (l + m);
        R nu = l / ((R)2 * (l + m));
        R c = ys[p] > (R)1 ? ys[p] : (R)1;
        struct Trial t = trial(e + 3L * p, m, l, ys[p], alpha[p], damage[p]);
        R phi = ((R)2 * m * t.dn - t.Y) / c, y0 = t.Y / c, kap = (R)2 * m * (R)1e-6 / c;
        R *x = X + (long)p * NI;
        R pk = phi + kap;
        x[0] = (R)0.1 * (R)log((double)(pk > (R)1e-30 ? pk : (R)1e-30));
        x[1] = (R)0.1 * (R)log((double)(y0 > (R)1e-12 ? y0 : (R)1e-12));
        x[2] = (R)0.1 * (R)log((double)kap);
        for (int k = 0; k < 3; k++) x[3 + k] = t.dn > 0 ? t.d[k] / t.dn : (R)0;
        x[6] = tan_of_alpha(alpha[p]);
        x[7] = (R)4 * nu - (R)1;
        x[8] = (R)log10((double)E) - (R)8;
        x[9] = damage[p];
    }
}

/* exact targets (s, 20 (alpha - 0.3), 0) of plastic states (zeros for elastic ones) and
 * the plastic flag */
EXPORT void o_lrm_targets(int n, const R *e, const R *mu, const R *lam, const R *ys, const R *alpha,
                          const R *damage, R *T, int *plastic) {
    for (int p = 0; p < n; p++) {
        R en[3], dg;
        const R *ep = e + 3L * p;
        plastic[p] = exact_principal(ep, mu[p], lam[p], ys[p], alpha[p], damage[p], en, &dg);
        struct Trial t = trial(ep, mu[p], lam[p], ys[p], alpha[p], damage[p]);
        R *tp = T + 3L * p;
        tp[0] = tp[1] = tp[2] = 0;
        if (plastic[p]) {
            tp[0] = t.Y / ((R)2 * mu[p] * (t.dn + (R)1e-6));
            tp[1] = (R)20 * (alpha[p] - (R)0.3);
        }
    }
}

/* the prediction e0' - eref kap y0 d/|d| + 0.3 (0.3 + y1 / 20) eref (phi + kap)
 * from network outputs y */
EXPORT void o_lrm_apply(int n, const R *e, const R *mu, const R *lam, const R *ys, const R *alpha,
                        const R *damage, const R *Y, R *e_pred) {
    for (int p = 0; p < n; p++) {
        const R *ep = e + 3L * p;
        struct Trial t = trial(ep, mu[p], lam[p], ys[p], alpha[p], damage[p]);
        R e0[3];
        radial_return(ep, &t, mu[p], e0);
        R er = eref_of(mu[p], ys[p]);
        R c = ys[p] > (R)1 ? ys[p] : (R)1;
        R kap = (R)2 * mu[p] * (R)1e-6 / c, phi = ((R)2 * mu[p] * t.dn - t.Y) / c;
        const R *y = Y + 3L * p;
        for (int k = 0; k < 3; k++) {
            R dh = t.dn > 0 ? t.d[k] / t.dn : (R)0;
            e_pred[3L * p + k] = e0[k] - er * kap * y[0] * dh + (R)0.3 * ((R)0.3 + y[1] / (R)20) * er * (phi + kap);
        }
    }
}

/* Judge of a predicted projection e_pred of a plastic trial state, in stress
 * units against tol * max(cohesion, 1): the exact projection has
 * dev e' = Y / (2 mu edn) d (edn = |d| + 1e-6), i.e. it lies on the yield
 * surface 2 mu |dev e'| = Y |d| / (|d| + 1e-6) at the trial pressure and old
 * damage, along the trial deviator, and tr e' = tr e + 3 * 0.3 alpha dgam.
 * The deviatoric residual 2 mu |dev e' - Y/(2 mu edn) d| (which holds both the
 * yield-surface identity and the flow direction) and the volumetric residual
 * K |tr e' - tr e - 0.9 alpha dgam| must both be within the tolerance. */
EXPORT void o_lrm_judge(int n, const R *e, const R *e_pred, const R *mu, const R *lam, const R *ys,
                        const R *alpha, const R *damage, R tol, int *accept) {
    for (int p = 0; p < n; p++) {
        const R *t = e + 3L * p, *q = e_pred + 3L * p;
        R m = mu[p], l = lam[p];
        R mean = (t[0] + t[1] + t[2]) / (R)3;
        R d[3] = {t[0] - mean, t[1] - mean, t[2] - mean};
        R sum = t[0] + t[1] + t[2];
        R pr = ((R)2 * m * t[0] + l * sum + (R)2 * m * t[1] + l * sum + (R)2 * m * t[2] + l * sum) / (R)3;
        R dn = (R)sqrt((double)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
        R edn = dn + (R)1e-6;
        R Y = ((R)1 - damage[p]) * (ys[p] - alpha[p] * pr);
        R c = Y / ((R)2 * m * edn);
        R qm = (q[0] + q[1] + q[2]) / (R)3;
        R rd[3] = {q[0] - qm - c * d[0], q[1] - qm - c * d[1], q[2] - qm - c * d[2]};
        R r1 = (R)2 * m * (R)sqrt((double)(rd[0] * rd[0] + rd[1] * rd[1] + rd[2] * rd[2]));
        R dgam = edn - Y / ((R)2 * m);
        R K = l + (R)2 * m / (R)3;
        R r2 = K * (R)fabs((double)((q[0] + q[1] + q[2]) - sum - (R)3 * (R)0.3 * alpha[p] * dgam));
        R lim = tol * (ys[p] > (R)1 ? ys[p] : (R)1);
        accept[p] = (r1 <= lim && r2 <= lim) ? 1 : 0;
    }
}
