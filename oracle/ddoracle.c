/* ddoracle: FP64 CPU restatement of jaxdd's steady-state J-V solve (test infrastructure).
 *
 * Not product code: the checker that the HIP path in jaxdd_amd/ is compared with. See the
 * header for the reference functions each part follows; every function below names its
 * reference file:line. Plain C99, one device per call, straightforward two-pass structure
 * (assemble all blocks, then block Thomas), as the reference computes it.
 */
#include "ddoracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Q_E 1.6022e-19 /* src/jaxdd/constants.py:8 */

/* Convergence constants, bit-equal to the reference's. */
#define ROUND_OFF_FLOOR 1e-6  /* solver.py:46 */
#define STALL_STEPS 3         /* solver.py:47 */
#define NOISE_FLOOR_CAP 1e-3  /* solver.py:51 */
#define DENSITY_FLOOR 1e-12   /* solver.py:54 */
#define CAP_V 10.0            /* solver.py:59 STEP_CAPS */
#define CAP_N 5.0
#define ION_LOG_CAP 5.0       /* solver.py:61 */
#define FLOOR_DROP_CAP 50.0   /* solver.py:71 */
#define SERIES_BELOW 1e-2     /* solver.py:92 */
#define CURRENT_TOL 1e-5      /* batch.py:81 */
#define CURRENT_STEP_CAP 1e-3 /* batch.py:83 */
#define QUADRATIC_TOL 1e-5    /* batch.py:86 */
#define QUADRATIC_RATIO 1e-2  /* batch.py:87 */
#define POISSON_STOP 1e-8     /* batch.py:500 */

static const double ION_STAGES[4] = {1.0 / 64, 1.0 / 16, 1.0 / 4, 1.0}; /* solver.py:2009 */

int ddo_abi_version(void) { return 1; }

/* ---- Bernoulli: solver.py:95-121 --------------------------------------------------------- */
double ddo_bernoulli(double x, int kind, double *dB) {
    double B, d;
    if (kind == 0) {
        if (fabs(x) < SERIES_BELOW) {
            double x4 = (x * x) * (x * x);
            B = 1.0 - x / 2.0 + x * x / 12.0 - x4 / 720.0;
            d = -0.5 + 2.0 * x / 12.0 - 4.0 * (x * x * x) / 720.0;
        } else {
            double e = expm1(x);
            B = x / e;
            /* d/dx [x / expm1(x)] = 1/expm1(x) - x*exp(x)/expm1(x)^2 */
            d = 1.0 / e - x * exp(x) / (e * e);
        }
    } else {
        double ax = fabs(x);
        if (ax < 0.05) {
            B = 1.0 - 0.5 * x;
            d = -0.5;
        } else if (ax < 0.45) {
            B = 1.0 - x * (0.5 - x / 12.0);
            d = -0.5 + x / 6.0;
        } else if (ax < 1.5) {
            B = 1.0 + x * (-0.5 + x * (1.0 / 12.0 - x * x / 720.0));
            d = -0.5 + x / 6.0 - 4.0 * x * x * x / 720.0;
        } else {
            double e = expm1(x);
            B = x / e;
            d = 1.0 / e - x * exp(x) / (e * e);
        }
    }
    if (dB) *dB = d;
    return B;
}

/* ---- the problem one Newton step sees -------------------------------------------------- */
typedef struct {
    const ddo_device *dev;
    int NP, bern;
    /* stencils on the interior points, solver.py:210-249 */
    double *fac_l, *fac_m, *fac_u, *hm, *hp, *fac, *dop, *kd, *ni2, *Nt, *Cn, *Cp, *nt0, *pt0;
    double *Gs, *am, *cm, *afix, *cfix, *sP, *sC;
    double n_floor;
    /* ions over the full grid (one slot per species): masks, totals, hL = h*L per interval */
    double *amask_full, *cmask_full, *hL;
    double a_total, c_total;
    /* recombination arrays over the full grid (batch._device_tree "rec") */
    double *ni2_full;
    double *work;
} problem;

static double *alloc_d(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* jnp.clip and jnp.maximum keep a NaN; C's fmin/fmax drop it. */
static double clipd(double x, double lo, double hi) { return x < lo ? lo : (x > hi ? hi : x); }
static double nmax(double a, double b) { return (isnan(a) || isnan(b)) ? NAN : (a > b ? a : b); }

static void problem_init(problem *P, const ddo_device *d, int bern) {
    int NP = d->NP, i;
    double L = d->L, L2 = L * L, ncmax = 0.0;
    memset(P, 0, sizeof(*P));
    P->dev = d;
    P->NP = NP;
    P->bern = bern;
    double **fields[] = {&P->fac_l, &P->fac_m, &P->fac_u, &P->hm, &P->hp, &P->fac, &P->dop,
                         &P->kd,    &P->ni2,   &P->Nt,    &P->Cn, &P->Cp, &P->nt0, &P->pt0,
                         &P->Gs,    &P->am,    &P->cm,    &P->afix, &P->cfix, &P->sP, &P->sC};
    for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); ++f) *fields[f] = alloc_d(NP);
    P->amask_full = alloc_d(NP + 2);
    P->cmask_full = alloc_d(NP + 2);
    P->hL = alloc_d(NP + 1);
    P->ni2_full = alloc_d(NP + 2);
    for (i = 0; i < NP + 2; ++i) {
        if (d->N_c[i] > ncmax) ncmax = d->N_c[i];
        P->amask_full[i] = d->a_mask[i];
        P->cmask_full[i] = d->c_mask[i];
        P->ni2_full[i] = d->ni[i] * d->ni[i];
    }
    for (i = 0; i < NP + 1; ++i) P->hL[i] = d->h[i] * L;
    for (int r = 0; r < NP; ++r) {
        i = r + 1;
        const double *h = d->h, *eps = d->eps;
        double hs = h[i] + h[i - 1];
        P->fac_m[r] = (eps[i] - eps[i - 1]) / (h[i - 1] * h[i - 1] * L2) -
                      2 * eps[i] * (1 / (h[i] * hs * L2) + 1 / (h[i - 1] * hs * L2));
        P->fac_l[r] = 2 * eps[i] / (h[i - 1] * hs * L2) - (eps[i] - eps[i - 1]) / (h[i - 1] * h[i - 1] * L2);
        P->fac_u[r] = 2 * eps[i] / (h[i] * hs * L2);
        P->hm[r] = h[i - 1];
        P->hp[r] = h[i];
        P->fac[r] = 0.5 * L2 * h[i] * h[i - 1] * hs;
        P->dop[r] = d->N_A[i] - d->N_D[i];
        P->kd[r] = d->k_direct[i];
        P->ni2[r] = d->ni[i] * d->ni[i];
        P->Nt[r] = d->Nt[i];
        P->Cn[r] = d->Cn[i];
        P->Cp[r] = d->Cp[i];
        P->nt0[r] = d->nt0[i];
        P->pt0[r] = d->pt0[i];
        P->Gs[r] = d->G[i];
        P->am[r] = d->a_mask[i];
        P->cm[r] = d->c_mask[i];
        P->afix[r] = d->a_fixed[i];
        P->cfix[r] = d->c_fixed[i];
        P->sP[r] = 1.0 / (Q_E * d->N_c[i]);
        double mu = d->mu_n[i] > d->mu_p[i] ? d->mu_n[i] : d->mu_p[i];
        P->sC[r] = 1.0 / (mu * d->Vt * d->N_c[i] * h[i]);
    }
    P->n_floor = DENSITY_FLOOR * ncmax;
    P->a_total = d->a_total;
    P->c_total = d->c_total;
    P->work = alloc_d((size_t)(NP + 2) * 64);
}

static void problem_free(problem *P) {
    double *fields[] = {P->fac_l, P->fac_m, P->fac_u, P->hm,  P->hp,  P->fac,  P->dop,
                        P->kd,    P->ni2,   P->Nt,    P->Cn,  P->Cp,  P->nt0,  P->pt0,
                        P->Gs,    P->am,    P->cm,    P->afix, P->cfix, P->sP, P->sC,
                        P->amask_full, P->cmask_full, P->hL, P->ni2_full, P->work};
    for (size_t f = 0; f < sizeof(fields) / sizeof(fields[0]); ++f) free(fields[f]);
}

/* _scaled_ions (solver.py:2012-2015): fixed densities and totals times `share`. */
typedef struct {
    double share;
    double a_total, c_total;
} ionscale;

/* ---- recombination: solver.py:734-745, and its derivatives in (ln n, ln p) --------------- */
static double recomb(double n, double p, double kd, double ni2, double Nt, double Cn, double Cp,
                     double nt0, double pt0, double *dln_n, double *dln_p) {
    double np_ = n * p;
    double R = kd * (np_ - ni2);
    double dn = kd * p, dp = kd * n;
    if (Nt > 0) {
        double num = Nt * Cn * Cp * (np_ - nt0 * pt0);
        double den = Cn * (n + nt0) + Cp * (p + pt0);
        R += num / den;
        double base = Nt * Cn * Cp;
        dn += (base * p * den - num * Cn) / (den * den);
        dp += (base * n * den - num * Cp) / (den * den);
    }
    if (dln_n) *dln_n = dn * n;
    if (dln_p) *dln_p = dp * p;
    return R;
}

/* ---- assembly: solver.py:1021-1187 (Boltzmann, no optional modules) --------------------- */
/* full: (NP+2) x 3. F: NP x 3. A, B, C: NP x 9. U: NP x 6 (3 rows x 2 slots). */
/* la_pt, lc_pt: each interior point's ion logarithms (the 5x5 form, _assemble_ions passing
   `la, lc` per point, solver.py:1393-1395), or NULL for the slots' la, lc at every point. */
static void assemble_pt(const problem *P, const double *full, double g_frac, double la,
                        double lc, const double *la_pt, const double *lc_pt, double share,
                        double *F, double *A, double *B, double *C, double *U) {
    const ddo_device *d = P->dev;
    int NP = P->NP;
    double Vt = d->Vt;
    double *Fn = P->work, *Fp = Fn + (NP + 1);
    double *dFn0 = Fp + (NP + 1), *dFn1 = dFn0 + 3 * (NP + 1);
    double *dFp0 = dFn1 + 3 * (NP + 1), *dFp1 = dFp0 + 3 * (NP + 1);
    for (int e = 0; e <= NP; ++e) {
        double V0 = full[3 * e], V1 = full[3 * (e + 1)];
        double n0 = exp(full[3 * e + 1]), n1 = exp(full[3 * (e + 1) + 1]);
        double p0 = exp(full[3 * e + 2]), p1 = exp(full[3 * (e + 1) + 2]);
        double cn = Vt * d->mu_n[e], cp = Vt * d->mu_p[e];
        double xn = ((V1 + d->offset_n[e + 1]) - (V0 + d->offset_n[e])) / Vt;
        double xp = ((V1 + d->offset_p[e + 1]) - (V0 + d->offset_p[e])) / Vt;
        double dBnf, dBnb, dBpf, dBpb;
        double Bnf = ddo_bernoulli(xn, P->bern, &dBnf), Bnb = ddo_bernoulli(-xn, P->bern, &dBnb);
        double Bpf = ddo_bernoulli(xp, P->bern, &dBpf), Bpb = ddo_bernoulli(-xp, P->bern, &dBpb);
        Fn[e] = cn * (Bnf * n1 - Bnb * n0);
        double gn = cn * (dBnf * n1 + dBnb * n0) / Vt;
        dFn0[3 * e] = -gn; dFn0[3 * e + 1] = -cn * Bnb * n0; dFn0[3 * e + 2] = 0;
        dFn1[3 * e] = gn;  dFn1[3 * e + 1] = cn * Bnf * n1;  dFn1[3 * e + 2] = 0;
        Fp[e] = cp * (Bpf * p0 - Bpb * p1);
        double gp = cp * (dBpf * p0 + dBpb * p1) / Vt;
        dFp0[3 * e] = -gp; dFp0[3 * e + 1] = 0; dFp0[3 * e + 2] = cp * Bpf * p0;
        dFp1[3 * e] = gp;  dFp1[3 * e + 1] = 0; dFp1[3 * e + 2] = -cp * Bpb * p1;
    }
    for (int r = 0; r < NP; ++r) {
        int i = r + 1;
        double Vm = full[3 * i], nm = exp(full[3 * i + 1]), pm = exp(full[3 * i + 2]);
        double dRn, dRp;
        double R = recomb(nm, pm, P->kd[r], P->ni2[r], P->Nt[r], P->Cn[r], P->Cp[r], P->nt0[r],
                          P->pt0[r], &dRn, &dRp);
        double gen = P->Gs[r] * g_frac;
        double hm = P->hm[r], hp = P->hp[r], fac = P->fac[r], sP = P->sP[r], sC = P->sC[r];
        double elec = hm * Fn[i] - hp * Fn[i - 1] + fac * (gen - R);
        double hole = hp * Fp[i - 1] - hm * Fp[i] + fac * (gen - R);
        double anion = P->am[r] * exp((la_pt ? la_pt[r] : la) + Vm / Vt);
        double cation = P->cm[r] * exp((lc_pt ? lc_pt[r] : lc) - Vm / Vt);
        double afix = P->afix[r] * share, cfix = P->cfix[r] * share;
        double poisson = P->fac_l[r] * full[3 * (i - 1)] + P->fac_m[r] * Vm +
                         P->fac_u[r] * full[3 * (i + 1)] -
                         Q_E * (nm - pm + P->dop[r] + afix + anion - cfix - cation);
        double *a = A + 9 * r, *b = B + 9 * r, *c = C + 9 * r;
        const double *l0 = dFn0 + 3 * (i - 1), *l1 = dFn1 + 3 * (i - 1);
        const double *r0 = dFn0 + 3 * i, *r1 = dFn1 + 3 * i;
        const double *q0 = dFp0 + 3 * (i - 1), *q1 = dFp1 + 3 * (i - 1);
        const double *s0 = dFp0 + 3 * i, *s1 = dFp1 + 3 * i;
        double dnet[3] = {0.0, -dRn, -dRp};
        a[0] = P->fac_l[r] * sP; a[1] = 0; a[2] = 0;
        c[0] = P->fac_u[r] * sP; c[1] = 0; c[2] = 0;
        b[0] = (P->fac_m[r] - Q_E * (anion + cation) / Vt) * sP;
        b[1] = -Q_E * nm * sP;
        b[2] = Q_E * pm * sP;
        for (int col = 0; col < 3; ++col) {
            a[3 + col] = -hp * l0[col] * sC;
            a[6 + col] = hp * q0[col] * sC;
            b[3 + col] = (-hp * l1[col] + hm * r0[col] + fac * dnet[col]) * sC;
            b[6 + col] = (hp * q1[col] - hm * s0[col] + fac * dnet[col]) * sC;
            c[3 + col] = hm * r1[col] * sC;
            c[6 + col] = -hm * s1[col] * sC;
        }
        F[3 * r] = poisson * sP;
        F[3 * r + 1] = elec * sC;
        F[3 * r + 2] = hole * sC;
        if (U) {
            double *u = U + 6 * r;
            u[0] = -Q_E * anion * sP; u[1] = Q_E * cation * sP;
            u[2] = u[3] = u[4] = u[5] = 0;
        }
    }
}

static void assemble(const problem *P, const double *full, double g_frac, double la, double lc,
                     double share, double *F, double *A, double *B, double *C, double *U) {
    assemble_pt(P, full, g_frac, la, lc, NULL, NULL, share, F, A, B, C, U);
}

int ddo_assemble(const ddo_device *dev, const double *full, double g_frac, double la, double lc,
                 int bernoulli, double *F, double *A, double *B, double *C, double *U) {
    problem P;
    problem_init(&P, dev, bernoulli);
    assemble(&P, full, g_frac, la, lc, 1.0, F, A, B, C, U);
    problem_free(&P);
    return 0;
}

/* ---- the 5x5 steady system: solver.py:1364-1473 (`_assemble_ions`, dti = 0) ---------------- */
/* full5: (NP+2) x 5 (V, ln n, ln p, ln a, ln c); logs: the anion and cation slots' logarithms.
   F: NP x 5. A, B, C: NP x 25. U: NP x 10 (5 rows x 2 slots). The carriers' rows are the 3x3
   assembly's at each point's ion logarithms ln a - V/V_T, ln c + V/V_T (solver.py:1393-1395);
   Poisson's loses the ions' V dependence and gains their columns (:1405-1409); each ion's row is
   Boltzmann's in its region, its unknown outside (`steady` of `species`, :1448-1451). */
int ddo_assemble_ions(const ddo_device *dev, const double *full5, double g_frac,
                      const double *logs, int bernoulli, double *F, double *A, double *B,
                      double *C, double *U) {
    problem P;
    problem_init(&P, dev, bernoulli);
    const int NP = P.NP;
    const double Vt = dev->Vt;
    double *full3 = alloc_d((size_t)(NP + 2) * 3);
    double *la_pt = alloc_d((size_t)NP), *lc_pt = alloc_d((size_t)NP);
    double *F3 = alloc_d((size_t)NP * 3), *U3 = alloc_d((size_t)NP * 6);
    double *A3 = alloc_d((size_t)NP * 9), *B3 = alloc_d((size_t)NP * 9), *C3 = alloc_d((size_t)NP * 9);
    for (int i = 0; i < NP + 2; ++i)
        for (int k = 0; k < 3; ++k) full3[3 * i + k] = full5[5 * i + k];
    for (int r = 0; r < NP; ++r) {
        const double *y = full5 + 5 * (r + 1);
        la_pt[r] = y[3] - y[0] / Vt;
        lc_pt[r] = y[4] + y[0] / Vt;
    }
    assemble_pt(&P, full3, g_frac, 0.0, 0.0, la_pt, lc_pt, 1.0, F3, A3, B3, C3, U3);
    for (int r = 0; r < NP; ++r) {
        const double *y = full5 + 5 * (r + 1);
        double *f = F + 5 * r, *a = A + 25 * r, *b = B + 25 * r, *c = C + 25 * r, *u = U + 10 * r;
        for (int k = 0; k < 25; ++k) a[k] = b[k] = c[k] = 0.0;
        for (int k = 0; k < 10; ++k) u[k] = 0.0;
        for (int row = 0; row < 3; ++row) {
            for (int col = 0; col < 3; ++col) {
                a[5 * row + col] = A3[9 * r + 3 * row + col];
                b[5 * row + col] = B3[9 * r + 3 * row + col];
                c[5 * row + col] = C3[9 * r + 3 * row + col];
            }
            f[row] = F3[3 * r + row];
        }
        const double a_m = P.am[r] * exp(y[3]), c_m = P.cm[r] * exp(y[4]);
        b[0] += Q_E * (a_m + c_m) / Vt * P.sP[r];
        b[3] = U3[6 * r + 0];
        b[4] = U3[6 * r + 1];
        const int in_a = P.am[r] > 0, in_c = P.cm[r] > 0;
        f[3] = in_a ? y[3] - logs[0] + (-1.0 * y[0]) / Vt : y[3];
        f[4] = in_c ? y[4] - logs[1] + (1.0 * y[0]) / Vt : y[4];
        b[15] = in_a ? -1.0 / Vt : 0.0;
        b[18] = 1.0;
        b[20] = in_c ? 1.0 / Vt : 0.0;
        b[24] = 1.0;
        u[6] = in_a ? -1.0 : 0.0;
        u[9] = in_c ? -1.0 : 0.0;
    }
    free(full3); free(la_pt); free(lc_pt); free(F3); free(U3); free(A3); free(B3); free(C3);
    problem_free(&P);
    return 0;
}

/* ---- small dense algebra: solver.py:1481-1531 --------------------------------------------- */
static void inverse3(const double *m, double *out) {
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    double co[9] = {e * i - f * h, c * h - b * i, b * f - c * e,
                    f * g - d * i, a * i - c * g, c * d - a * f,
                    d * h - e * g, b * g - a * h, a * e - b * d};
    double det = a * co[0] + b * co[3] + c * co[6];
    for (int k = 0; k < 9; ++k) out[k] = co[k] / det;
}

/* Inverse of a b x b block by Gauss-Jordan with partial pivoting (b != 3). */
static void inverse_general(int b, const double *m, double *out) {
    double t[64], inv[64];
    for (int i = 0; i < b * b; ++i) t[i] = m[i];
    for (int i = 0; i < b; ++i)
        for (int j = 0; j < b; ++j) inv[i * b + j] = i == j;
    for (int c = 0; c < b; ++c) {
        int piv = c;
        for (int r = c + 1; r < b; ++r)
            if (fabs(t[r * b + c]) > fabs(t[piv * b + c])) piv = r;
        if (piv != c)
            for (int j = 0; j < b; ++j) {
                double s = t[c * b + j]; t[c * b + j] = t[piv * b + j]; t[piv * b + j] = s;
                s = inv[c * b + j]; inv[c * b + j] = inv[piv * b + j]; inv[piv * b + j] = s;
            }
        double p = t[c * b + c];
        for (int j = 0; j < b; ++j) { t[c * b + j] /= p; inv[c * b + j] /= p; }
        for (int r = 0; r < b; ++r) {
            if (r == c) continue;
            double f = t[r * b + c];
            for (int j = 0; j < b; ++j) { t[r * b + j] -= f * t[c * b + j]; inv[r * b + j] -= f * inv[c * b + j]; }
        }
    }
    for (int i = 0; i < b * b; ++i) out[i] = inv[i];
}

/* out (r x k) = a (r x m) @ b (m x k) */
static void mm(int r, int m, int k, const double *a, const double *b, double *out) {
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < k; ++j) {
            double s = 0.0;
            for (int l = 0; l < m; ++l) s += a[i * m + l] * b[l * k + j];
            out[i * k + j] = s;
        }
}

/* _block_thomas (solver.py:1534-1560). cp, dp: scratch NP*b*b, NP*b*k. */
static void block_thomas(int NP, int b, int k, const double *A, const double *B, const double *C,
                         const double *d, double *x, double *cp, double *dp) {
    double M[64], Mi[64], t[64], rhs[64];
    int bb = b * b, bk = b * k;
    for (int i = 0; i < NP; ++i) {
        const double *a = A + (size_t)bb * i;
        for (int q = 0; q < bb; ++q) M[q] = B[(size_t)bb * i + q];
        for (int q = 0; q < bk; ++q) rhs[q] = d[(size_t)bk * i + q];
        if (i > 0) {
            mm(b, b, b, a, cp + (size_t)bb * (i - 1), t);
            for (int q = 0; q < bb; ++q) M[q] -= t[q];
            mm(b, b, k, a, dp + (size_t)bk * (i - 1), t);
            for (int q = 0; q < bk; ++q) rhs[q] -= t[q];
        }
        if (b == 3) inverse3(M, Mi); else inverse_general(b, M, Mi);
        mm(b, b, b, Mi, C + (size_t)bb * i, cp + (size_t)bb * i);
        mm(b, b, k, Mi, rhs, dp + (size_t)bk * i);
    }
    for (int i = NP - 1; i >= 0; --i) {
        double *xi = x + (size_t)bk * i;
        for (int q = 0; q < bk; ++q) xi[q] = dp[(size_t)bk * i + q];
        if (i < NP - 1) {
            mm(b, b, k, cp + (size_t)bb * i, x + (size_t)bk * (i + 1), t);
            for (int q = 0; q < bk; ++q) xi[q] -= t[q];
        }
    }
}

int ddo_block_thomas(int NP, int b, int k, const double *A, const double *B, const double *C,
                     const double *d, double *x) {
    if (b < 1 || b > 8 || k < 1 || k > 8 || NP < 1) return -1;
    double *cp = alloc_d((size_t)NP * b * b), *dp = alloc_d((size_t)NP * b * k);
    block_thomas(NP, b, k, A, B, C, d, x, cp, dp);
    free(cp);
    free(dp);
    return 0;
}

/* K x K solve (K <= 2) by elimination with partial pivoting, as LAPACK's gesv would. */
static void solve_small(int K, double *S, double *rhs) {
    if (K == 1) { rhs[0] /= S[0]; return; }
    if (fabs(S[2]) > fabs(S[0])) { /* swap rows */
        double t;
        t = S[0]; S[0] = S[2]; S[2] = t;
        t = S[1]; S[1] = S[3]; S[3] = t;
        t = rhs[0]; rhs[0] = rhs[1]; rhs[1] = t;
    }
    double l = S[2] / S[0];
    double u11 = S[3] - l * S[1];
    double r1 = rhs[1] - l * rhs[0];
    rhs[1] = r1 / u11;
    rhs[0] = (rhs[0] - S[1] * rhs[1]) / S[0];
}

/* ---- ion constraints: solver.py:1241-1265 (trapezoidal number over the region) ---------- */
/* g[k] = number_k / total_k - 1; W (NP x 2) = dg/dV at the interior; Ddiag = dg/dlogs. */
static void ion_constraints(const problem *P, const double *Vfull, const double *logs,
                            const ionscale *sc, double *g, double *W, double *Ddiag) {
    int NP = P->NP;
    double Vt = P->dev->Vt;
    const double *masks[2] = {P->amask_full, P->cmask_full};
    double totals[2] = {sc->a_total, sc->c_total};
    double sign[2] = {1.0, -1.0};
    double *dens = P->work; /* NP + 2 */
    for (int k = 0; k < 2; ++k) {
        for (int i = 0; i < NP + 2; ++i) dens[i] = masks[k][i] * exp(logs[k] + sign[k] * Vfull[i] / Vt);
        double number = 0.0;
        for (int e = 0; e <= NP; ++e) {
            double pair = masks[k][e + 1] * masks[k][e];
            number += 0.5 * (dens[e + 1] + dens[e]) * pair * P->hL[e];
        }
        int has = totals[k] > 0;
        double tot = has ? totals[k] : 1.0;
        if (g) g[k] = has ? number / tot - 1.0 : 0.0;
        if (Ddiag) Ddiag[k] = has ? number / tot : 0.0;
        if (W)
            for (int r = 0; r < NP; ++r) {
                int i = r + 1;
                double wsum = masks[k][i] * masks[k][i - 1] * P->hL[i - 1] + masks[k][i + 1] * masks[k][i] * P->hL[i];
                W[2 * r + k] = has ? 0.5 * wsum * dens[i] * (sign[k] / Vt) / tot : 0.0;
            }
    }
}

/* _ion_logs, Boltzmann (solver.py:1695-1712): the normalisations that put each total into
 * its region at the potential Vfull. */
static void ion_logs(const problem *P, const double *Vfull, const ionscale *sc, double *logs) {
    int NP = P->NP;
    double Vt = P->dev->Vt;
    const double *masks[2] = {P->amask_full, P->cmask_full};
    double totals[2] = {sc->a_total, sc->c_total};
    double sign[2] = {1.0, -1.0};
    for (int k = 0; k < 2; ++k) {
        if (!(totals[k] > 0)) { logs[k] = 0.0; continue; }
        double number = 0.0;
        for (int e = 0; e <= NP; ++e) {
            double pair = masks[k][e + 1] * masks[k][e];
            double d0 = masks[k][e] * exp(sign[k] * Vfull[e] / Vt);
            double d1 = masks[k][e + 1] * exp(sign[k] * Vfull[e + 1] / Vt);
            number += 0.5 * (d1 + d0) * pair * P->hL[e];
        }
        logs[k] = log(totals[k] / number);
    }
}

/* ---- the integral current: batch.py:302-403 (steady state) ------------------------------ */
static double current(const problem *P, const double *full, double g_frac) {
    const ddo_device *d = P->dev;
    int NP = P->NP;
    double Vt = d->Vt, L = d->L;
    double *U = P->work, *ev = U + (NP + 2), *iu = ev + (NP + 2), *J = iu + (NP + 2);
    double *n = J + (NP + 2), *p = n + (NP + 2), *Jsum = p + (NP + 2);
    for (int i = 0; i < NP + 2; ++i) {
        n[i] = exp(full[3 * i + 1]);
        p[i] = exp(full[3 * i + 2]);
        double R = recomb(n[i], p[i], d->k_direct[i], P->ni2_full[i], d->Nt[i], d->Cn[i], d->Cp[i],
                          d->nt0[i], d->pt0[i], NULL, NULL);
        U[i] = R - d->G[i] * g_frac;
        Jsum[i] = 0.0;
    }
    U[0] = U[NP + 1] = 0.0;
    for (int c = 0; c < 2; ++c) {
        double sn = c == 0 ? -1.0 : 1.0;
        const double *off = c == 0 ? d->offset_n : d->offset_p;
        const double *mu = c == 0 ? d->mu_n : d->mu_p;
        const double *dens = c == 0 ? n : p;
        for (int i = 0; i <= NP; ++i) ev[i] = exp(sn * (full[3 * i] + off[i]) / Vt);
        iu[0] = 0.0;
        for (int j = 0; j < NP; ++j) iu[j + 1] = iu[j] + 0.5 * (d->x[j + 1] - d->x[j]) * (U[j] + U[j + 1]);
        double single = 0.0, dbl = 0.0;
        for (int j = 0; j < NP; ++j) {
            double dx = d->x[j + 1] - d->x[j];
            single += 0.5 * dx * (ev[j] / mu[j] + ev[j + 1] / mu[j + 1]);
            dbl += 0.5 * dx * (ev[j] * iu[j] / mu[j] + ev[j + 1] * iu[j + 1] / mu[j + 1]);
        }
        double VgN = full[3 * (NP + 1)] + off[NP + 1], Vg0 = full[0] + off[0];
        double K = (Vt * (sn * dens[NP + 1] * exp(sn * VgN / Vt) - sn * dens[0] * exp(sn * Vg0 / Vt)) - sn * dbl) / single;
        J[0] = Q_E * K;
        double acc = 0.0;
        for (int e = 1; e <= NP; ++e) {
            acc += sn * Q_E * L * d->h[e] * U[e];
            J[e] = Q_E * K + acc;
        }
        /* across an interface, the current before it (no interface traps on this path) */
        for (int e = NP; e >= 1; --e)
            if (d->iface[e]) J[e] = J[e - 1];
        double sign = -d->left_elec;
        for (int e = 0; e <= NP; ++e) Jsum[e] += sign * J[e];
    }
    double num = 0.0, den = 0.0;
    for (int e = 0; e <= NP; ++e) { num += d->h[e] * Jsum[e]; den += d->h[e]; }
    return num / den;
}

double ddo_current(const ddo_device *dev, const double *full, double g_frac) {
    problem P;
    problem_init(&P, dev, 0);
    double J = current(&P, full, g_frac);
    problem_free(&P);
    return J;
}

/* ---- one Newton step: batch.py:759-785 with _system (:518-549), _bordered_solve (:595-603) */
typedef struct {
    double *A, *B, *C, *F, *U, *rhs, *X, *cp, *dp, *W, *full, *Vf;
} scratch;

static void scratch_init(scratch *s, int NP) {
    s->A = alloc_d((size_t)NP * 9); s->B = alloc_d((size_t)NP * 9); s->C = alloc_d((size_t)NP * 9);
    s->F = alloc_d((size_t)NP * 3); s->U = alloc_d((size_t)NP * 6);
    s->rhs = alloc_d((size_t)NP * 9); s->X = alloc_d((size_t)NP * 9);
    s->cp = alloc_d((size_t)NP * 9); s->dp = alloc_d((size_t)NP * 9);
    s->W = alloc_d((size_t)NP * 2); s->full = alloc_d((size_t)(NP + 2) * 3);
    s->Vf = alloc_d((size_t)NP + 2);
}
static void scratch_free(scratch *s) {
    double *f[] = {s->A, s->B, s->C, s->F, s->U, s->rhs, s->X, s->cp, s->dp, s->W, s->full, s->Vf};
    for (int i = 0; i < 12; ++i) free(f[i]);
}

static void bounds_of(const ddo_device *d, double Vint, double *b /* 6 */) {
    b[0] = d->V0 + 0.5 * d->left_elec * Vint; b[1] = d->ln_n_left; b[2] = d->ln_p_left;
    b[3] = d->VL - 0.5 * d->left_elec * Vint; b[4] = d->ln_n_right; b[5] = d->ln_p_right;
}

static void make_full(int NP, const double *bnd, const double *y, double *full) {
    memcpy(full, bnd, 3 * sizeof(double));
    memcpy(full + 3, y, (size_t)NP * 3 * sizeof(double));
    memcpy(full + 3 * (NP + 1), bnd + 3, 3 * sizeof(double));
}

/* Newton step in place on (y, logs); returns the step size (batch.py:759-785). */
static double newton_step(const problem *P, scratch *s, double *y, double *logs, double Vint,
                          double g_frac, const ionscale *sc) {
    const ddo_device *d = P->dev;
    int NP = P->NP;
    double Vt = d->Vt, bnd[6], g[2], Dd[2];
    bounds_of(d, Vint, bnd);
    make_full(NP, bnd, y, s->full);
    assemble(P, s->full, g_frac, logs[0], logs[1], sc->share, s->F, s->A, s->B, s->C, s->U);
    for (int i = 0; i < NP + 2; ++i) s->Vf[i] = s->full[3 * i];
    ion_constraints(P, s->Vf, logs, sc, g, s->W, Dd);
    double held[2] = {sc->a_total > 0 ? 0.0 : 1.0, sc->c_total > 0 ? 0.0 : 1.0};
    /* rhs columns: -F, U_anion, U_cation (slot columns: U times the region's mask) */
    for (int r = 0; r < NP; ++r)
        for (int row = 0; row < 3; ++row) {
            s->rhs[9 * r + 3 * row + 0] = -s->F[3 * r + row];
            s->rhs[9 * r + 3 * row + 1] = s->U[6 * r + 2 * row + 0] * P->am[r];
            s->rhs[9 * r + 3 * row + 2] = s->U[6 * r + 2 * row + 1] * P->cm[r];
        }
    block_thomas(NP, 3, 3, s->A, s->B, s->C, s->rhs, s->X, s->cp, s->dp);
    double S[4] = {Dd[0] + held[0], 0.0, 0.0, Dd[1] + held[1]};
    double z[2] = {-g[0], -g[1]};
    for (int r = 0; r < NP; ++r)
        for (int k = 0; k < 2; ++k) { /* W only has the potential's row */
            double w = s->W[2 * r + k];
            S[2 * k + 0] -= w * s->X[9 * r + 1];
            S[2 * k + 1] -= w * s->X[9 * r + 2];
            z[k] -= w * s->X[9 * r + 0];
        }
    solve_small(2, S, z);
    double capv = CAP_V * Vt, lnfloor = log(P->n_floor), size = 0.0;
    for (int r = 0; r < NP; ++r) {
        double delta[3], step[3];
        for (int c = 0; c < 3; ++c)
            delta[c] = s->X[9 * r + 3 * c + 0] - (s->X[9 * r + 3 * c + 1] * z[0] + s->X[9 * r + 3 * c + 2] * z[1]);
        step[0] = clipd(delta[0], -capv, capv);
        for (int c = 1; c < 3; ++c) {
            step[c] = clipd(delta[c], -CAP_N, CAP_N);
            /* density_steps (solver.py:74-88): below n_floor, the step in the density */
            if (y[3 * r + c] < lnfloor) {
                double in_n = delta[c] > -1 ? log1p(delta[c]) : -CAP_N;
                step[c] = clipd(in_n, -FLOOR_DROP_CAP, CAP_N);
            }
        }
        for (int c = 0; c < 3; ++c) y[3 * r + c] += step[c];
        size = nmax(size, fabs(step[0]) / Vt);
        for (int c = 1; c < 3; ++c) {
            double dens = exp(y[3 * r + c]);
            size = nmax(size, fabs(step[c]) * (dens / (dens + P->n_floor)));
        }
    }
    double dl0 = clipd(z[0], -ION_LOG_CAP, ION_LOG_CAP);
    double dl1 = clipd(z[1], -ION_LOG_CAP, ION_LOG_CAP);
    logs[0] += dl0;
    logs[1] += dl1;
    return nmax(size, nmax(fabs(dl0), fabs(dl1)));
}

/* The current of state y at Vint. */
static double current_at(const problem *P, scratch *s, const double *y, double Vint, double g_frac) {
    double bnd[6];
    bounds_of(P->dev, Vint, bnd);
    make_full(P->NP, bnd, y, s->full);
    return current(P, s->full, g_frac);
}

static int all_finite(const double *a, int n) {
    for (int i = 0; i < n; ++i)
        if (!isfinite(a[i])) return 0;
    return 1;
}

/* _newton_loop (batch.py:740-848). Returns converged; *it the steps, *J the last current. */
static int newton_loop(const problem *P, scratch *s, double *y, double *logs, double Vint,
                       double g_frac, const ionscale *sc, double tol, int max_iter, int strict,
                       int skip, double *J_out, int *it_out) {
    double hist[STALL_STEPS + 1], Jh[STALL_STEPS + 1];
    for (int k = 0; k <= STALL_STEPS; ++k) hist[k] = Jh[k] = INFINITY;
    int it = 0, done = skip ? 1 : 0;
    double J_scale = P->dev->J_scale;
    while (!done && it < max_iter && !isnan(hist[STALL_STEPS])) {
        double size = newton_step(P, s, y, logs, Vint, g_frac, sc);
        for (int k = 0; k < STALL_STEPS; ++k) { hist[k] = hist[k + 1]; Jh[k] = Jh[k + 1]; }
        hist[STALL_STEPS] = size;
        Jh[STALL_STEPS] = current_at(P, s, y, Vint, g_frac);
        ++it;
        double hmax = -INFINITY, hmin_tail = INFINITY;
        for (int k = 0; k <= STALL_STEPS; ++k) hmax = fmax(hmax, hist[k]);
        for (int k = 1; k <= STALL_STEPS; ++k) hmin_tail = fmin(hmin_tail, hist[k]);
        int anynan = 0;
        for (int k = 0; k <= STALL_STEPS; ++k) anynan |= isnan(hist[k]);
        int stalled = it > STALL_STEPS && !anynan && hmax < ROUND_OFF_FLOOR && hmin_tail >= 0.1 * hist[0];
        double Jn = Jh[STALL_STEPS];
        double scale = J_scale + fabs(Jn), jdev = 0.0;
        int jnan = 0;
        for (int k = 0; k <= STALL_STEPS; ++k) {
            double v = fabs(Jh[k] - Jn);
            if (isnan(v)) jnan = 1;
            jdev = fmax(jdev, v);
        }
        int settled = it > STALL_STEPS && size < CURRENT_STEP_CAP && !jnan && jdev < CURRENT_TOL * scale;
        double prev = hist[STALL_STEPS - 1];
        int quadratic = it > 1 && size < QUADRATIC_TOL && size < QUADRATIC_RATIO * prev;
        if (strict) done = size < tol || stalled;
        else done = size < tol || stalled || settled || quadratic;
    }
    *J_out = Jh[STALL_STEPS];
    *it_out = it;
    return done;
}

/* ---- Poisson alone: batch.py:486-503 with solver._poisson_delta (:1757-1800) ------------ */
static void poisson_solve(const problem *P, scratch *s, double *y, double Vint, const ionscale *sc,
                          int max_iter) {
    const ddo_device *d = P->dev;
    int NP = P->NP;
    double Vt = d->Vt, bnd[6];
    double *lo = s->A, *mid = s->B, *up = s->C, *rhs = s->rhs, *c2 = s->cp, *d2 = s->dp;
    double *Vf = s->W; /* NP + 2 fits in 2*NP for NP >= 2 */
    double *XU = s->X; /* NP x 3 (dV, XU_a, XU_c) */
    int mobile = sc->a_total > 0 || sc->c_total > 0;
    double size = INFINITY;
    int it = 0;
    bounds_of(d, Vint, bnd);
    while (size > POISSON_STOP && it < max_iter) {
        make_full(NP, bnd, y, s->full);
        for (int i = 0; i < NP + 2; ++i) Vf[i] = s->full[3 * i];
        double logs[2];
        ion_logs(P, Vf, sc, logs);
        /* the residual with the quasi-Fermi levels held, and its tridiagonal Jacobian */
        for (int r = 0; r < NP; ++r) {
            int i = r + 1;
            double Vm = Vf[i], n = exp(s->full[3 * i + 1]), p = exp(s->full[3 * i + 2]);
            double anion = P->am[r] * exp(logs[0] + Vm / Vt), cation = P->cm[r] * exp(logs[1] - Vm / Vt);
            double afix = P->afix[r] * sc->share, cfix = P->cfix[r] * sc->share;
            double F = P->fac_l[r] * Vf[i - 1] + P->fac_m[r] * Vm + P->fac_u[r] * Vf[i + 1] -
                       Q_E * (n - p + P->dop[r] + afix + anion - cfix - cation);
            rhs[3 * r] = -F * P->sP[r];
            lo[r] = r == 0 ? 0.0 : P->fac_l[r] * P->sP[r];
            up[r] = r == NP - 1 ? 0.0 : P->fac_u[r] * P->sP[r];
            mid[r] = (P->fac_m[r] - Q_E * (n / Vt + p / Vt + anion / Vt + cation / Vt)) * P->sP[r];
            rhs[3 * r + 1] = -Q_E * anion * P->sP[r] * P->am[r];
            rhs[3 * r + 2] = Q_E * cation * P->sP[r] * P->cm[r];
        }
        /* tridiagonal solve of the three columns (Thomas, no pivoting) */
        for (int r = 0; r < NP; ++r) {
            double m = mid[r] - (r > 0 ? lo[r] * c2[r - 1] : 0.0);
            c2[r] = up[r] / m;
            for (int k = 0; k < 3; ++k)
                d2[3 * r + k] = (rhs[3 * r + k] - (r > 0 ? lo[r] * d2[3 * (r - 1) + k] : 0.0)) / m;
        }
        for (int r = NP - 1; r >= 0; --r)
            for (int k = 0; k < 3; ++k)
                XU[3 * r + k] = d2[3 * r + k] - (r < NP - 1 ? c2[r] * XU[3 * (r + 1) + k] : 0.0);
        double cap = CAP_V * Vt, m = 0.0;
        if (mobile) {
            double g[2], Dd[2], held[2] = {sc->a_total > 0 ? 0.0 : 1.0, sc->c_total > 0 ? 0.0 : 1.0};
            double *W = s->F; /* NP x 2 */
            ion_constraints(P, Vf, logs, sc, g, W, Dd);
            double S[4] = {Dd[0] + held[0], 0.0, 0.0, Dd[1] + held[1]}, z[2] = {-g[0], -g[1]};
            for (int r = 0; r < NP; ++r)
                for (int k = 0; k < 2; ++k) {
                    S[2 * k + 0] -= W[2 * r + k] * XU[3 * r + 1];
                    S[2 * k + 1] -= W[2 * r + k] * XU[3 * r + 2];
                    z[k] -= W[2 * r + k] * XU[3 * r];
                }
            solve_small(2, S, z);
            for (int r = 0; r < NP; ++r) {
                double st = XU[3 * r] - (XU[3 * r + 1] * z[0] + XU[3 * r + 2] * z[1]);
                XU[3 * r] = st;
                m = nmax(m, fabs(st));
            }
            m /= Vt;
            double scale = m > 0 ? fmin(log1p(m), CAP_V) / m : 1.0;
            for (int r = 0; r < NP; ++r) XU[3 * r] *= scale;
        } else {
            for (int r = 0; r < NP; ++r) XU[3 * r] = clipd(XU[3 * r], -cap, cap);
        }
        size = 0.0;
        for (int r = 0; r < NP; ++r) {
            double dV = XU[3 * r], shift = dV / Vt;
            y[3 * r] += dV;
            y[3 * r + 1] += shift;
            y[3 * r + 2] -= shift;
            size = nmax(size, fabs(shift));
        }
        ++it;
        if (isnan(size)) break;
    }
}

/* _initial (batch.py:276-299, no free ends). */
static void initial(const ddo_device *d, double Vint, double *y /* NP x 3 */) {
    double e = d->left_elec;
    for (int r = 0; r < d->NP; ++r) {
        int i = r + 1;
        double x = d->x[i];
        double V = d->V0 + 0.5 * e * Vint + x * (d->VL - d->V0 - e * Vint) / d->L;
        double EF = d->W_L - x * (d->W_L - d->W_R) / d->L;
        double ln_n = log(d->N_c[i]) - (EF - d->E_c[i]) / d->Vt;
        y[3 * r] = V;
        y[3 * r + 1] = ln_n;
        y[3 * r + 2] = 2 * log(d->ni[i]) - ln_n;
    }
}

static void full_V(const ddo_device *d, double Vint, const double *y, double *Vf) {
    double bnd[6];
    bounds_of(d, Vint, bnd);
    Vf[0] = bnd[0];
    for (int r = 0; r < d->NP; ++r) Vf[r + 1] = y[3 * r];
    Vf[d->NP + 1] = bnd[3];
}

/* _freeze (batch.py:859-872): the mobile ions of y as fixed densities. */
static void freeze(problem *P, const double *y, const double *logs, ionscale *sc) {
    double Vt = P->dev->Vt;
    for (int r = 0; r < P->NP; ++r) {
        double V = y[3 * r];
        P->afix[r] = P->afix[r] + P->am[r] * exp(logs[0] + V / Vt);
        P->cfix[r] = P->cfix[r] + P->cm[r] * exp(logs[1] - V / Vt);
        P->am[r] = 0.0;
        P->cm[r] = 0.0;
    }
    for (int i = 0; i < P->NP + 2; ++i) P->amask_full[i] = P->cmask_full[i] = 0.0;
    sc->a_total = sc->c_total = 0.0;
}

/* ---- _solve_one (batch.py:937-1006) with _start (:875-934) ------------------------------ */
int ddo_solve_one_states(const ddo_device *d, const double *Vs, int nV, const ddo_opts *o,
                         double *Jout, uint8_t *okout, int32_t *itsout, double *states) {
    int NP = d->NP;
    if (NP < 2 || NP > 4096 || nV < 1) return -1;
    problem P;
    problem_init(&P, d, o->bernoulli);
    scratch s;
    scratch_init(&s, NP);
    size_t ny = (size_t)NP * 3;
    double *y = alloc_d(ny), *yf = alloc_d(ny), *yc = alloc_d(ny), *yk = alloc_d(ny);
    double *Vf = alloc_d(NP + 2);
    double g_frac = d->g_frac, Jdummy;
    int itdummy;
    ionscale full_sc = {1.0, d->a_total, d->c_total};
    double first = d->pre_cond ? d->V_pre : Vs[0];

    /* Poisson's start, then Boltzmann's ions at once, else through the stages */
    initial(d, first, y);
    poisson_solve(&P, &s, y, first, &full_sc, o->start_iter);
    double logs[2];
    full_V(d, first, y, Vf);
    ion_logs(&P, Vf, &full_sc, logs);
    memcpy(yf, y, ny * sizeof(double));
    double logs_f[2] = {logs[0], logs[1]};
    int ok_f = newton_loop(&P, &s, yf, logs_f, first, g_frac, &full_sc, o->tol, o->start_iter,
                           o->strict, 0, &Jdummy, &itdummy);
    int mobile = d->a_total > 0 || d->c_total > 0;
    int done = ok_f || !mobile;
    memcpy(yc, y, ny * sizeof(double));
    double logs_c[2] = {logs[0], logs[1]}, before = 1.0;
    int ok_c = 1;
    for (int k = 0; k < 4; ++k) {
        double share = ION_STAGES[k];
        ionscale sk = {share, d->a_total * share, d->c_total * share};
        logs_c[0] += log(share / before);
        logs_c[1] += log(share / before);
        memcpy(yk, yc, ny * sizeof(double));
        double logs_k[2] = {logs_c[0], logs_c[1]};
        ok_c = newton_loop(&P, &s, yk, logs_k, first, g_frac, &sk, o->tol, o->start_iter, o->strict,
                           done, &Jdummy, &itdummy);
        if (all_finite(yk, (int)ny) && isfinite(logs_k[0]) && isfinite(logs_k[1])) {
            memcpy(yc, yk, ny * sizeof(double));
            logs_c[0] = logs_k[0];
            logs_c[1] = logs_k[1];
        }
        before = share;
    }
    int staged = !done && ok_c;
    int kept = all_finite(yf, (int)ny) && isfinite(logs_f[0]) && isfinite(logs_f[1]);
    if (staged) { memcpy(y, yc, ny * sizeof(double)); logs[0] = logs_c[0]; logs[1] = logs_c[1]; }
    else if (kept) { memcpy(y, yf, ny * sizeof(double)); logs[0] = logs_f[0]; logs[1] = logs_f[1]; }
    int solved = ok_f || staged;
    /* the pre-bias (or first voltage) solve, skipped where the start is the solution */
    int pre_ok = newton_loop(&P, &s, y, logs, first, g_frac, &full_sc, o->tol, o->start_iter,
                             o->strict, solved, &Jdummy, &itdummy);
    int frozen = d->pre_cond && d->fix_ions;
    ionscale sc = full_sc;
    if (frozen) freeze(&P, y, logs, &sc);
    pre_ok = pre_ok || !frozen;

    /* the scan outwards from `first`: up, then down from the start again */
    int below = 0;
    while (below < nV && Vs[below] < first) ++below;
    double *ys = alloc_d(ny), *yprev = alloc_d(ny), *ycur = alloc_d(ny), *guess = alloc_d(ny);
    memcpy(ys, y, ny * sizeof(double));
    double logs_s[2] = {logs[0], logs[1]};
    memcpy(ycur, y, ny * sizeof(double));
    memcpy(yprev, y, ny * sizeof(double));
    double lc_[2] = {logs[0], logs[1]}, lp_[2] = {logs[0], logs[1]};
    double V_last = first, V_before = first;
    double capv = CAP_V * d->Vt;
    for (int j = 0; j < nV; ++j) {
        int idx = j < nV - below ? below + j : nV - 1 - j;
        double Vint = Vs[idx];
        if (j == nV - below) { /* turn: from the start again */
            memcpy(ycur, ys, ny * sizeof(double));
            memcpy(yprev, ys, ny * sizeof(double));
            lc_[0] = lp_[0] = logs_s[0];
            lc_[1] = lp_[1] = logs_s[1];
            V_last = V_before = first;
        }
        double ratio = V_last != V_before ? (Vint - V_last) / (V_last - V_before) : 0.0;
        for (size_t q = 0; q < ny; ++q) {
            double cap = (q % 3 == 0) ? capv : CAP_N;
            guess[q] = ycur[q] + clipd(ratio * (ycur[q] - yprev[q]), -cap, cap);
        }
        double lg[2];
        for (int k = 0; k < 2; ++k)
            lg[k] = lc_[k] + clipd(ratio * (lc_[k] - lp_[k]), -ION_LOG_CAP, ION_LOG_CAP);
        double J;
        int it;
        int ok = newton_loop(&P, &s, guess, lg, Vint, g_frac, &sc, o->tol, o->max_iter, o->strict, 0, &J, &it);
        int keep = ok && all_finite(guess, (int)ny) && isfinite(lg[0]) && isfinite(lg[1]);
        if (keep) {
            memcpy(yprev, ycur, ny * sizeof(double));
            memcpy(ycur, guess, ny * sizeof(double));
            lp_[0] = lc_[0]; lp_[1] = lc_[1];
            lc_[0] = lg[0]; lc_[1] = lg[1];
            V_before = V_last;
            V_last = Vint;
        }
        Jout[idx] = J;
        okout[idx] = (uint8_t)(ok && pre_ok);
        itsout[idx] = it;
        if (states) memcpy(states + (size_t)idx * ny, guess, ny * sizeof(double));
    }
    free(ys); free(yprev); free(ycur); free(guess);
    free(y); free(yf); free(yc); free(yk); free(Vf);
    scratch_free(&s);
    problem_free(&P);
    return 0;
}

int ddo_solve_one(const ddo_device *d, const double *Vs, int nV, const ddo_opts *o, double *Jout,
                  uint8_t *okout, int32_t *itsout) {
    return ddo_solve_one_states(d, Vs, nV, o, Jout, okout, itsout, NULL);
}

/* ---- solve_jv (solver.py:2041-2245) --------------------------------------------------- */
/* _noise_step (solver.py:1728-1751): the step the residual's round-off alone gives. */
static double noise_step_eps(const problem *P, scratch *s, const double *y, const double *logs,
                             double Vint, double g_frac, const ionscale *sc, double eps);

static double noise_step(const problem *P, scratch *s, const double *y, const double *logs,
                         double Vint, double g_frac, const ionscale *sc) {
    return noise_step_eps(P, s, y, logs, Vint, g_frac, sc, DBL_EPSILON);
}

/* The same with the round-off unit eps in place of the double's (the step is linear in it). */
static double noise_step_eps(const problem *P, scratch *s, const double *y, const double *logs,
                             double Vint, double g_frac, const ionscale *sc, double eps) {
    int NP = P->NP;
    double bnd[6], Vt = P->dev->Vt;
    bounds_of(P->dev, Vint, bnd);
    make_full(NP, bnd, y, s->full);
    assemble(P, s->full, g_frac, logs[0], logs[1], sc->share, s->F, s->A, s->B, s->C, NULL);
    for (int r = 0; r < NP; ++r)
        for (int row = 0; row < 3; ++row) {
            double acc = 0.0;
            for (int c = 0; c < 3; ++c) {
                acc += fabs(s->A[9 * r + 3 * row + c]) * fabs(s->full[3 * r + c]);
                acc += fabs(s->B[9 * r + 3 * row + c]) * fabs(s->full[3 * (r + 1) + c]);
                acc += fabs(s->C[9 * r + 3 * row + c]) * fabs(s->full[3 * (r + 2) + c]);
            }
            s->rhs[3 * r + row] = eps * acc;
        }
    block_thomas(NP, 3, 1, s->A, s->B, s->C, s->rhs, s->X, s->cp, s->dp);
    double m = 0.0;
    for (int r = 0; r < NP; ++r) {
        m = nmax(m, fabs(s->X[3 * r]) / Vt);
        for (int c = 1; c < 3; ++c) {
            double dens = exp(y[3 * r + c]);
            m = nmax(m, fabs(s->X[3 * r + c]) * (dens / (dens + P->n_floor)));
        }
    }
    return m;
}

double ddo_noise_step(const ddo_device *d, const double *y, const double *logs, double Vint,
                      double g_frac, int bernoulli, double eps) {
    problem P;
    problem_init(&P, d, bernoulli);
    scratch s;
    scratch_init(&s, d->NP);
    ionscale sc = {1.0, d->a_total, d->c_total};
    double out = noise_step_eps(&P, &s, y, logs, Vint, g_frac, &sc, eps);
    scratch_free(&s);
    problem_free(&P);
    return out;
}

/* settle (solver.py:2132-2175, no contacts with finite surface recombination). */
static int settle(const problem *P, scratch *s, double *y, double *logs, double Vint,
                  const ionscale *sc, double tol, int max_iter, int *steps_out) {
    double g_frac = P->dev->g_frac, sizes[8];
    int count = 0, steps = 0, have_floor = 0;
    double floor_ = 0.0;
    while (steps < max_iter) {
        double size = newton_step(P, s, y, logs, Vint, g_frac, sc);
        ++steps;
        if (count == STALL_STEPS + 1) {
            for (int k = 0; k < STALL_STEPS; ++k) sizes[k] = sizes[k + 1];
            sizes[STALL_STEPS] = size;
        } else {
            sizes[count++] = size;
        }
        if (isnan(size)) { *steps_out = steps; return 0; }
        int full4 = count > STALL_STEPS;
        double mx = 0.0, mn_tail = INFINITY;
        for (int k = 0; k < count; ++k) mx = fmax(mx, sizes[k]);
        for (int k = 1; k < count; ++k) mn_tail = fmin(mn_tail, sizes[k]);
        int done = size < tol || (full4 && mx < ROUND_OFF_FLOOR && mn_tail >= 0.1 * sizes[0]);
        if (!done && full4 && mx < NOISE_FLOOR_CAP) {
            if (!have_floor) { floor_ = noise_step(P, s, y, logs, Vint, g_frac, sc); have_floor = 1; }
            done = mx < floor_;
        }
        if (done) { *steps_out = steps; return 1; }
    }
    *steps_out = steps;
    return 0;
}

int ddo_solve_jv(const ddo_device *d, const double *Vs, int nV, double tol, int max_iter,
                 int bernoulli, double *Jout, uint8_t *okout, int32_t *itsout) {
    int NP = d->NP;
    if (NP < 2 || NP > 4096 || nV < 1) return -1;
    problem P;
    problem_init(&P, d, bernoulli);
    scratch s;
    scratch_init(&s, NP);
    size_t ny = (size_t)NP * 3;
    double *y = alloc_d(ny), *yc = alloc_d(ny), *yn = alloc_d(ny), *Vf = alloc_d(NP + 2);
    double g_frac = d->g_frac;
    ionscale sc = {1.0, d->a_total, d->c_total};
    double first = d->pre_cond ? d->V_pre : Vs[0];
    int mobile = d->a_total > 0 || d->c_total > 0, steps, rc = 0;
    double logs[2];
    initial(d, first, y);
    poisson_solve(&P, &s, y, first, &sc, max_iter);
    full_V(d, first, y, Vf);
    ion_logs(&P, Vf, &sc, logs);
    if (mobile) {
        memcpy(yn, y, ny * sizeof(double));
        double lf[2] = {logs[0], logs[1]};
        int ok = settle(&P, &s, yn, lf, first, &sc, tol, max_iter, &steps);
        if (!ok) {
            memcpy(yc, y, ny * sizeof(double));
            double lcx[2] = {logs[0], logs[1]}, before = 1.0;
            for (int k = 0; k < 4; ++k) {
                double share = ION_STAGES[k];
                ionscale sk = {share, d->a_total * share, d->c_total * share};
                lcx[0] += log(share / before);
                lcx[1] += log(share / before);
                ok = settle(&P, &s, yc, lcx, first, &sk, tol, max_iter, &steps);
                before = share;
                if (!ok) break;
            }
            if (ok) { memcpy(yn, yc, ny * sizeof(double)); lf[0] = lcx[0]; lf[1] = lcx[1]; }
        }
        if (ok) { memcpy(y, yn, ny * sizeof(double)); logs[0] = lf[0]; logs[1] = lf[1]; }
    }
    if (d->pre_cond) {
        int ok = settle(&P, &s, y, logs, d->V_pre, &sc, tol, max_iter, &steps);
        if (!ok) rc = -2;
        else if (d->fix_ions) freeze(&P, y, logs, &sc);
    }
    for (int j = 0; j < nV && rc == 0; ++j) {
        double Vint = Vs[j];
        memcpy(yn, y, ny * sizeof(double));
        double ln_[2] = {logs[0], logs[1]};
        int ok = settle(&P, &s, yn, ln_, Vint, &sc, tol, max_iter, &steps);
        double J = current_at(&P, &s, yn, Vint, g_frac);
        if (d->G_shunt > 0) J += Vint * d->G_shunt;
        if (ok && all_finite(yn, (int)ny) && isfinite(ln_[0]) && isfinite(ln_[1])) {
            memcpy(y, yn, ny * sizeof(double));
            logs[0] = ln_[0];
            logs[1] = ln_[1];
        }
        Jout[j] = J;
        okout[j] = (uint8_t)ok;
        itsout[j] = steps;
    }
    free(y); free(yc); free(yn); free(Vf);
    scratch_free(&s);
    problem_free(&P);
    return rc;
}
