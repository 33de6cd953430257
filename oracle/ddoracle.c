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
    double *Gs, *am, *cm, *afix, *cfix, *sP, *sC, *tt;
    int charged;  /* some point holds charged bulk traps */
    int neutral;  /* their charge left out (the start's first solve, batch.py:898-900) */
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
                         &P->Gs,    &P->am,    &P->cm,    &P->afix, &P->cfix, &P->sP, &P->sC,
                         &P->tt};
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
        P->tt[r] = d->tt ? d->tt[i] : 0.0;
        if (P->tt[r] != 0.0) P->charged = 1;
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
                        P->Gs,    P->am,    P->cm,    P->afix, P->cfix, P->sP, P->sC, P->tt,
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

/* The charged bulk traps' charge density (m^-3) at a point, and its derivatives in ln n and
   ln p (solver.py:762-773 `_trap_charge`, one level; device.py:940-948): Nt (cwe - f), cwe 1
   for donors and 0 for acceptors, f = (Cn n + Cp pt0) / (Cn (n + nt0) + Cp (p + pt0)) the
   level's steady occupation; 0 for neutral traps and where there are none. */
static double trap_charge(double tt, double n, double p, double Nt, double Cn, double Cp,
                          double nt0, double pt0, double *dln_n, double *dln_p) {
    *dln_n = *dln_p = 0.0;
    if (tt == 0.0 || !(Nt > 0)) return 0.0;
    double num = Cn * n + Cp * pt0, den = Cn * (n + nt0) + Cp * (p + pt0);
    double f = num / den, cwe = tt > 0 ? 1.0 : 0.0;
    *dln_n = -Nt * n * Cn * (Cn * nt0 + Cp * p) / (den * den);
    *dln_p = Nt * p * Cp * num / (den * den);
    return Nt * (cwe - f);
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
        double dQn, dQp;
        double tq = P->neutral ? 0.0 : P->tt[r];
        double Qt = trap_charge(tq, nm, pm, P->Nt[r], P->Cn[r], P->Cp[r], P->nt0[r], P->pt0[r],
                                &dQn, &dQp);
        if (tq != 0.0) poisson = poisson + Q_E * Qt;  /* solver.py:1136-1140 */
        double *a = A + 9 * r, *b = B + 9 * r, *c = C + 9 * r;
        const double *l0 = dFn0 + 3 * (i - 1), *l1 = dFn1 + 3 * (i - 1);
        const double *r0 = dFn0 + 3 * i, *r1 = dFn1 + 3 * i;
        const double *q0 = dFp0 + 3 * (i - 1), *q1 = dFp1 + 3 * (i - 1);
        const double *s0 = dFp0 + 3 * i, *s1 = dFp1 + 3 * i;
        double dnet[3] = {0.0, -dRn, -dRp};
        a[0] = P->fac_l[r] * sP; a[1] = 0; a[2] = 0;
        c[0] = P->fac_u[r] * sP; c[1] = 0; c[2] = 0;
        b[0] = (P->fac_m[r] - Q_E * (anion + cation) / Vt) * sP;
        b[1] = (-Q_E * nm + Q_E * dQn) * sP;
        b[2] = (Q_E * pm + Q_E * dQp) * sP;
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
            double dQn, dQp;  /* the traps' charge, the densities moving with V (solver.py:1672) */
            double Qt = trap_charge(P->tt[r], n, p, P->Nt[r], P->Cn[r], P->Cp[r], P->nt0[r],
                                    P->pt0[r], &dQn, &dQp);
            if (P->tt[r] != 0.0) F = F + Q_E * Qt;
            rhs[3 * r] = -F * P->sP[r];
            lo[r] = r == 0 ? 0.0 : P->fac_l[r] * P->sP[r];
            up[r] = r == NP - 1 ? 0.0 : P->fac_u[r] * P->sP[r];
            mid[r] = (P->fac_m[r] - Q_E * (n / Vt + p / Vt + anion / Vt + cation / Vt) +
                      Q_E * (dQn - dQp) / Vt) * P->sP[r];
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

/* _start (batch.py:875-934): Poisson's start at `first`, then Boltzmann's ions at once, else
 * through the ion density stages. y (NP x 3) and logs receive the start; returns whether it
 * is a solution (the solve that follows may then be skipped). */
static int start_state(problem *P, scratch *s, const ddo_device *d, double first, double g_frac,
                       const ddo_opts *o, double *y, double *logs) {
    int NP = d->NP;
    size_t ny = (size_t)NP * 3;
    double *yf = alloc_d(ny), *yc = alloc_d(ny), *yk = alloc_d(ny), *Vf = alloc_d(NP + 2);
    double Jdummy;
    int itdummy;
    ionscale full_sc = {1.0, d->a_total, d->c_total};
    /* Poisson's start, then Boltzmann's ions at once, else through the stages */
    initial(d, first, y);
    poisson_solve(P, s, y, first, &full_sc, o->start_iter);
    full_V(d, first, y, Vf);
    ion_logs(P, Vf, &full_sc, logs);
    if (P->charged) {  /* charged bulk traps: from the solution with them neutral (batch.py:898) */
        double logs_n[2] = {logs[0], logs[1]};
        memcpy(yf, y, ny * sizeof(double));
        P->neutral = 1;
        int ok_n = newton_loop(P, s, yf, logs_n, first, g_frac, &full_sc, o->tol, o->start_iter,
                               o->strict, 0, &Jdummy, &itdummy);
        P->neutral = 0;
        if (ok_n) { memcpy(y, yf, ny * sizeof(double)); logs[0] = logs_n[0]; logs[1] = logs_n[1]; }
    }
    memcpy(yf, y, ny * sizeof(double));
    double logs_f[2] = {logs[0], logs[1]};
    int ok_f = newton_loop(P, s, yf, logs_f, first, g_frac, &full_sc, o->tol, o->start_iter,
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
        ok_c = newton_loop(P, s, yk, logs_k, first, g_frac, &sk, o->tol, o->start_iter, o->strict,
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
    free(yf); free(yc); free(yk); free(Vf);
    return solved;
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

    double logs[2];
    int solved = start_state(&P, &s, d, first, g_frac, o, y, logs);
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
    if (P.charged) {  /* charged bulk traps: from the solution with them neutral (solver.py:2192) */
        memcpy(yn, y, ny * sizeof(double));
        double ln_[2] = {logs[0], logs[1]};
        P.neutral = 1;
        int ok = settle(&P, &s, yn, ln_, first, &sc, tol, max_iter, &steps);
        P.neutral = 0;
        if (ok) { memcpy(y, yn, ny * sizeof(double)); logs[0] = ln_[0]; logs[1] = ln_[1]; }
    }
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

/* ==== transients: src/jaxdd/transient.py:270-464 (`_solve_one`), the subset without an open
 * circuit, a series resistance, interface traps, free ends, finite surface recombination or
 * ZimT's automatic stop (Track 0): a steady start, then one implicit-Euler Newton solve per
 * step (`solver._time_terms` :573-591), mobile ions as unknowns of their own (5 x 5 blocks,
 * `solver._assemble_ions` :1364-1473, `_ion_constraints_time` :1349-1361), the bulk traps'
 * occupations carried from step to step (`_trap_step` :594-620), the transient current
 * (`batch._current_parts` with a time step :316-418, `_ion_currents` :433-483). ========== */

/* Forward-mode numbers with NDUAL partial derivatives: the rows' derivatives follow their
 * values exactly, as the reference's jvp does. */
#define NDUAL 7
typedef struct {
    double v, d[NDUAL];
} dual;

static dual dc(double v) { dual r; r.v = v; for (int k = 0; k < NDUAL; ++k) r.d[k] = 0.0; return r; }
static dual dv(double v, int k) { dual r = dc(v); r.d[k] = 1.0; return r; }
static dual dadd(dual a, dual b) { for (int k = 0; k < NDUAL; ++k) a.d[k] += b.d[k]; a.v += b.v; return a; }
static dual dsub(dual a, dual b) { for (int k = 0; k < NDUAL; ++k) a.d[k] -= b.d[k]; a.v -= b.v; return a; }
static dual dsc(dual a, double c) { for (int k = 0; k < NDUAL; ++k) a.d[k] *= c; a.v *= c; return a; }
static dual dadds(dual a, double c) { a.v += c; return a; }
static dual dmul(dual a, dual b) {
    dual r; r.v = a.v * b.v;
    for (int k = 0; k < NDUAL; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
static dual ddiv(dual a, dual b) {
    dual r; r.v = a.v / b.v;
    for (int k = 0; k < NDUAL; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
    return r;
}
static dual dfun(dual a, double f, double df) {
    dual r; r.v = f;
    for (int k = 0; k < NDUAL; ++k) r.d[k] = df * a.d[k];
    return r;
}
static dual dexp_(dual a) { double e = exp(a.v); return dfun(a, e, e); }
static dual dbern(dual x, int kind) { double dB, B = ddo_bernoulli(x.v, kind, &dB); return dfun(x, B, dB); }
/* -expm1(-x)/x, the trap occupation's mean factor over a step of a x Delta t = x */
static dual dmean(dual x) {
    double e = expm1(-x.v), m = -e / x.v;
    double dm = (exp(-x.v) * x.v + e) / (x.v * x.v); /* d/dx of -expm1(-x)/x */
    return dfun(x, m, dm);
}

/* What a time step holds fixed (`_time_terms` :573-591 and the current's `tr`). */
typedef struct {
    double dti;
    double *n_prev, *p_prev;   /* NP + 2: the densities before the step */
    double *a_prev, *c_prev;   /* NP + 2: the ions before the step (0 outside their regions) */
    double *V_prev;            /* NP + 2 */
    double *fb;                /* NP: the bulk traps' occupation before the step */
    double *width;             /* NP + 1: each point's width in the integral current's sums */
    int zimt;                  /* ZimT's sums (the left contact's change counted too) */
} tstep;

/* The carriers' rates over the step (`_bulk_rates` :680-700 with `_trap_step`'s mean and
 * vacant share): direct recombination, and each carrier's capture by the bulk traps at the
 * occupation's mean over the step (the steady SRH rate when dti = 0). */
static void step_rates(double kd, double ni2, double Nt, double Cn, double Cp, double nt0,
                       double pt0, dual n, dual p, double f_old, double dti, dual *Rn, dual *Rp) {
    dual direct = dsc(dadds(dmul(n, p), -ni2), kd);
    *Rn = *Rp = direct;
    if (!(Nt > 0)) return;
    dual a = dadd(dsc(dadds(n, nt0), Cn), dsc(dadds(p, pt0), Cp));
    dual steady = ddiv(dadds(dsc(n, Cn), Cp * pt0), a);
    dual vac_s = ddiv(dadds(dsc(p, Cp), Cn * nt0), a);
    dual fbar = steady, vbar = vac_s;
    if (dti > 0) {
        dual M = dmean(dsc(a, 1.0 / dti));
        fbar = dadd(steady, dmul(dsub(dc(f_old), steady), M));
        vbar = dadd(vac_s, dmul(dsub(dc(1.0 - f_old), vac_s), M));
    }
    *Rn = dadd(direct, dsc(dsub(dmul(n, vbar), dsc(fbar, nt0)), Nt * Cn));
    *Rp = dadd(direct, dsc(dsub(dmul(p, fbar), dsc(vbar, pt0)), Nt * Cp));
}

/* The traps' occupation at the step's end (`_bulk_occupation` :711-716). */
static double trap_end(double Nt, double Cn, double Cp, double nt0, double pt0, double n,
                       double p, double f_old, double dti) {
    if (!(Nt > 0)) return 0.0;
    double a = Cn * (n + nt0) + Cp * (p + pt0);
    double steady = (Cn * n + Cp * pt0) / a;
    if (!(dti > 0)) return steady;
    return steady + (f_old - steady) * exp(-a / dti);
}

/* `_ion_full` (:1315-1346): each contact's ion densities where a region reaches it, from its
 * neighbour (no flux through the contact's interval); 0 elsewhere. full: (NP + 2) x 5. */
static void ion_full(const problem *P, double *full, double dti) {
    int NP = P->NP, nc = 5;
    double Vt = P->dev->Vt;
    double x0 = (full[nc * 1] - full[0]) / Vt, xN = (full[nc * (NP + 1)] - full[nc * NP]) / Vt;
    double r0 = dti > 0 ? log(ddo_bernoulli(x0, P->bern, NULL)) - log(ddo_bernoulli(-x0, P->bern, NULL)) : -x0;
    double rN = dti > 0 ? log(ddo_bernoulli(xN, P->bern, NULL)) - log(ddo_bernoulli(-xN, P->bern, NULL)) : -xN;
    double *L0 = full, *R0 = full + nc * (NP + 1);
    L0[3] = P->amask_full[0] > 0 ? full[nc + 3] + r0 : 0.0;
    L0[4] = P->cmask_full[0] > 0 ? full[nc + 4] - r0 : 0.0;
    R0[3] = P->amask_full[NP + 1] > 0 ? full[nc * NP + 3] - rN : 0.0;
    R0[4] = P->cmask_full[NP + 1] > 0 ? full[nc * NP + 4] + rN : 0.0;
}

/* One time step's residual F (NP x nc, rows scaled), blocks A, B, C (NP x nc x nc) and the
 * slots' columns U (NP x nc x 2) at full ((NP + 2) x nc; nc = 5 with the ions' logarithms as
 * unknowns, its contacts' ions from ion_full), the slots' logarithms logs. The carriers' rows
 * are `_assemble`'s (:1021-1187) with each carrier's rate over the step and the implicit
 * Euler term -fac dti (n - n_prev); each ion's row (nc = 5) is `_assemble_ions`' `species`
 * (:1414-1451): Scharfetter-Gummel fluxes inside its region with the density before the step
 * scaled by exp(its slot's log), Boltzmann's at dti = 0, its unknown held outside. */
static void assemble_step(const problem *P, const double *full, int nc, double g_frac,
                          const tstep *ts, const double *logs, double *F, double *A, double *B,
                          double *C, double *U) {
    const ddo_device *d = P->dev;
    int NP = P->NP, bb = nc * nc;
    double Vt = d->Vt, dti = ts->dti;
    dual *Fn = (dual *)calloc((size_t)(NP + 1), sizeof(dual));
    dual *Fp = (dual *)calloc((size_t)(NP + 1), sizeof(dual));
    /* per edge e (points e, e + 1): partials 0 V_e, 1 V_e+1, 2 ln n_e, 3 ln n_e+1, 4 ln p_e,
       5 ln p_e+1 */
    for (int e = 0; e <= NP; ++e) {
        dual V0 = dv(full[nc * e], 0), V1 = dv(full[nc * (e + 1)], 1);
        dual n0 = dexp_(dv(full[nc * e + 1], 2)), n1 = dexp_(dv(full[nc * (e + 1) + 1], 3));
        dual p0 = dexp_(dv(full[nc * e + 2], 4)), p1 = dexp_(dv(full[nc * (e + 1) + 2], 5));
        dual xn = dsc(dadds(dsub(V1, V0), d->offset_n[e + 1] - d->offset_n[e]), 1.0 / Vt);
        dual xp = dsc(dadds(dsub(V1, V0), d->offset_p[e + 1] - d->offset_p[e]), 1.0 / Vt);
        double cn = Vt * d->mu_n[e], cp = Vt * d->mu_p[e];
        Fn[e] = dsc(dsub(dmul(dbern(xn, P->bern), n1), dmul(dbern(dsc(xn, -1.0), P->bern), n0)), cn);
        Fp[e] = dsc(dsub(dmul(dbern(xp, P->bern), p0), dmul(dbern(dsc(xp, -1.0), P->bern), p1)), cp);
    }
    for (int r = 0; r < NP; ++r) {
        int i = r + 1;
        double *a = A + (size_t)bb * r, *b = B + (size_t)bb * r, *c = C + (size_t)bb * r;
        double *f = F + (size_t)nc * r, *u = U + (size_t)nc * 2 * r;
        for (int k = 0; k < bb; ++k) a[k] = b[k] = c[k] = 0.0;
        for (int k = 0; k < 2 * nc; ++k) u[k] = 0.0;
        double hm = P->hm[r], hp = P->hp[r], fac = P->fac[r], sP = P->sP[r], sC = P->sC[r];
        /* the rates at the point: partials 0 ln n, 1 ln p */
        dual n = dexp_(dv(full[nc * i + 1], 0)), p = dexp_(dv(full[nc * i + 2], 1)), Rn, Rp;
        step_rates(P->kd[r], P->ni2[r], P->Nt[r], P->Cn[r], P->Cp[r], P->nt0[r], P->pt0[r], n,
                   p, ts->fb ? ts->fb[r] : 0.0, ts->fb ? dti : 0.0, &Rn, &Rp);
        double gen = P->Gs[r] * g_frac;
        /* electrons: hm Fn[i] - hp Fn[i-1] + fac (gen - Rn) - fac dti (n - n_prev) */
        double elec = hm * Fn[i].v - hp * Fn[i - 1].v + fac * (gen - Rn.v) - fac * dti * (n.v - ts->n_prev[i]);
        double hole = hp * Fp[i - 1].v - hm * Fp[i].v + fac * (gen - Rp.v) - fac * dti * (p.v - ts->p_prev[i]);
        /* columns: left point (edge i-1's partials 0, 2, 4), own (edge i-1's 1, 3, 5 and edge
           i's 0, 2, 4), right (edge i's 1, 3, 5) */
        for (int q = 0; q < 3; ++q) {
            a[nc * 1 + q] = -hp * Fn[i - 1].d[2 * q] * sC;
            a[nc * 2 + q] = hp * Fp[i - 1].d[2 * q] * sC;
            b[nc * 1 + q] = (hm * Fn[i].d[2 * q] - hp * Fn[i - 1].d[2 * q + 1]) * sC;
            b[nc * 2 + q] = (hp * Fp[i - 1].d[2 * q + 1] - hm * Fp[i].d[2 * q]) * sC;
            c[nc * 1 + q] = hm * Fn[i].d[2 * q + 1] * sC;
            c[nc * 2 + q] = -hm * Fp[i].d[2 * q + 1] * sC;
        }
        b[nc * 1 + 1] += (-fac * Rn.d[0] - fac * dti * n.v) * sC;
        b[nc * 1 + 2] += -fac * Rn.d[1] * sC;
        b[nc * 2 + 1] += -fac * Rp.d[0] * sC;
        b[nc * 2 + 2] += (-fac * Rp.d[1] - fac * dti * p.v) * sC;
        /* Poisson's: the mobile ions at their own densities (nc = 5) */
        double an = nc == 5 ? P->am[r] * exp(full[nc * i + 3]) : 0.0;
        double ca = nc == 5 ? P->cm[r] * exp(full[nc * i + 4]) : 0.0;
        double poisson = P->fac_l[r] * full[nc * (i - 1)] + P->fac_m[r] * full[nc * i] +
                         P->fac_u[r] * full[nc * (i + 1)] -
                         Q_E * (n.v - p.v + P->dop[r] + P->afix[r] + an - P->cfix[r] - ca);
        a[0] = P->fac_l[r] * sP;
        c[0] = P->fac_u[r] * sP;
        b[0] = P->fac_m[r] * sP;
        b[1] = -Q_E * n.v * sP;
        b[2] = Q_E * p.v * sP;
        f[0] = poisson * sP;
        f[1] = elec * sC;
        f[2] = hole * sC;
        if (nc == 3) continue;
        b[3] = -Q_E * an * sP;
        b[4] = Q_E * ca * sP;
        /* the ions' rows: partials 0 V_l, 1 V_c, 2 V_r, 3 ln_l, 4 ln_c, 5 ln_r, 6 the slot's log */
        for (int sp = 0; sp < 2; ++sp) {
            int col = 3 + sp;
            double sn = sp == 0 ? -1.0 : 1.0;
            const double *mfull = sp == 0 ? P->amask_full : P->cmask_full;
            const double *mu = sp == 0 ? d->mu_a : d->mu_c;
            const double *prevs = sp == 0 ? ts->a_prev : ts->c_prev;
            double m_i = mfull[i];
            double el = r == 0 ? 0.0 : mfull[i - 1] * m_i, er = r == NP - 1 ? 0.0 : m_i * mfull[i + 1];
            double cl = Vt * mu[i - 1], cr = Vt * mu[i];
            dual Vl = dv(full[nc * (i - 1)], 0), Vc = dv(full[nc * i], 1), Vr = dv(full[nc * (i + 1)], 2);
            dual ll = dv(full[nc * (i - 1) + col], 3), lc = dv(full[nc * i + col], 4);
            dual lr = dv(full[nc * (i + 1) + col], 5), ls = dv(logs[sp], 6);
            dual val;
            if (!(m_i > 0)) {
                val = lc;
            } else if (!(dti > 0)) {
                val = dadd(dsub(lc, ls), dsc(Vc, sn / Vt));
            } else {
                dual xl = dsc(dsub(Vc, Vl), 1.0 / Vt), xr = dsc(dsub(Vr, Vc), 1.0 / Vt);
                dual nl = dexp_(ll), nm_ = dexp_(lc), nr = dexp_(lr);
                dual Fl, Fr, flux;
                if (sn < 0) { /* anions, as electrons */
                    Fl = dsc(dsub(dmul(dbern(xl, P->bern), nm_), dmul(dbern(dsc(xl, -1.0), P->bern), nl)), cl);
                    Fr = dsc(dsub(dmul(dbern(xr, P->bern), nr), dmul(dbern(dsc(xr, -1.0), P->bern), nm_)), cr);
                    flux = dsub(dsc(Fr, er * hm), dsc(Fl, el * hp));
                } else { /* cations, as holes */
                    Fl = dsc(dsub(dmul(dbern(xl, P->bern), nl), dmul(dbern(dsc(xl, -1.0), P->bern), nm_)), cl);
                    Fr = dsc(dsub(dmul(dbern(xr, P->bern), nm_), dmul(dbern(dsc(xr, -1.0), P->bern), nr)), cr);
                    flux = dsub(dsc(Fl, el * hp), dsc(Fr, er * hm));
                }
                double before = prevs[i], prev = before > 0 ? before : 1.0;
                double scale = 1.0 / (prev * (fac * dti + hm * cr + hp * cl));
                dual stored = dsc(dsub(dsc(dexp_(ls), before), nm_), fac * dti);
                val = dsc(dadd(flux, stored), scale);
            }
            f[col] = val.v;
            a[nc * col + 0] = val.d[0];
            b[nc * col + 0] = val.d[1];
            c[nc * col + 0] = val.d[2];
            a[nc * col + col] = val.d[3];
            b[nc * col + col] = val.d[4];
            c[nc * col + col] = val.d[5];
            u[2 * col + sp] = val.d[6] * m_i;
        }
    }
    free(Fn);
    free(Fp);
}

/* `_ion_constraints_time` (:1349-1361): each slot's number of ions from their own densities
 * (trapezoidal over its region, the contacts' from ion_full) over its total, less 1; W (NP x 2)
 * its derivative in each interior point's own ion logarithm. full: (NP + 2) x 5. */
static void ion_constraints_time(const problem *P, const double *full, double *g, double *W) {
    int NP = P->NP;
    const double *masks[2] = {P->amask_full, P->cmask_full};
    double totals[2] = {P->a_total, P->c_total};
    for (int k = 0; k < 2; ++k) {
        double number = 0.0;
        for (int e = 0; e <= NP; ++e) {
            double pair = masks[k][e] * masks[k][e + 1];
            double d0 = masks[k][e] * exp(full[5 * e + 3 + k]);
            double d1 = masks[k][e + 1] * exp(full[5 * (e + 1) + 3 + k]);
            number += 0.5 * (d1 + d0) * pair * P->hL[e];
        }
        int has = totals[k] > 0;
        g[k] = has ? number / totals[k] - 1.0 : 0.0;
        for (int r = 0; r < NP; ++r) {
            int i = r + 1;
            double w = masks[k][i - 1] * masks[k][i] * P->hL[i - 1] + masks[k][i] * masks[k][i + 1] * P->hL[i];
            W[2 * r + k] = has ? 0.5 * w * masks[k][i] * exp(full[5 * i + 3 + k]) / totals[k] : 0.0;
        }
    }
}

typedef struct {
    double *full, *F, *A, *B, *C, *U, *rhs, *X, *cp, *dp, *W;
} tscratch;

static void tscratch_init(tscratch *s, int NP) {
    size_t P2 = (size_t)(NP + 2);
    s->full = alloc_d(P2 * 5); s->F = alloc_d((size_t)NP * 5);
    s->A = alloc_d((size_t)NP * 25); s->B = alloc_d((size_t)NP * 25); s->C = alloc_d((size_t)NP * 25);
    s->U = alloc_d((size_t)NP * 10); s->rhs = alloc_d((size_t)NP * 15); s->X = alloc_d((size_t)NP * 15);
    s->cp = alloc_d((size_t)NP * 25); s->dp = alloc_d((size_t)NP * 15); s->W = alloc_d((size_t)NP * 2);
}
static void tscratch_free(tscratch *s) {
    double *f[] = {s->full, s->F, s->A, s->B, s->C, s->U, s->rhs, s->X, s->cp, s->dp, s->W};
    for (int i = 0; i < 11; ++i) free(f[i]);
}

/* `_full` (batch.py:262-267) with the contacts at Vint, the ions' contacts by ion_full. */
static void full_of(const problem *P, const double *y, int nc, double Vint, double dti, double *full) {
    int NP = P->NP;
    double bnd[6];
    bounds_of(P->dev, Vint, bnd);
    for (int k = 0; k < nc; ++k) {
        full[k] = k < 3 ? bnd[k] : 0.0;
        full[nc * (NP + 1) + k] = k < 3 ? bnd[3 + k] : 0.0;
    }
    memcpy(full + nc, y, (size_t)NP * nc * sizeof(double));
    if (nc == 5) ion_full(P, full, dti);
}

/* The transient current (`batch._current` with `tr`, :302-418): each carrier's integral form
 * with the densities' change in the integrand, the mobile ions' currents of their regions
 * (`_ion_currents` :433-483) and the displacement current, averaged over the intervals. */
static double current_step(const problem *P, const double *full, int nc, double g_frac,
                           const tstep *ts) {
    const ddo_device *d = P->dev;
    int NP = P->NP;
    double Vt = d->Vt, L = d->L, dti = ts->dti;
    double *w = alloc_d((size_t)(NP + 2) * 8);
    double *U = w, *ev = U + (NP + 2), *iu = ev + (NP + 2), *J = iu + (NP + 2);
    double *n = J + (NP + 2), *p = n + (NP + 2), *Jsum = p + (NP + 2), *dens = Jsum + (NP + 2);
    double sign = -d->left_elec;
    for (int i = 0; i < NP + 2; ++i) {
        n[i] = exp(full[nc * i + 1]);
        p[i] = exp(full[nc * i + 2]);
        Jsum[i] = 0.0;
    }
    for (int c = 0; c < 2; ++c) {
        double sn = c == 0 ? -1.0 : 1.0;
        const double *off = c == 0 ? d->offset_n : d->offset_p;
        const double *mu = c == 0 ? d->mu_n : d->mu_p;
        const double *dn = c == 0 ? n : p;
        const double *before = c == 0 ? ts->n_prev : ts->p_prev;
        for (int i = 0; i < NP + 2; ++i) {
            double Ri = 0.0;
            if (i > 0 && i <= NP) {
                dual Rn, Rp;
                step_rates(d->k_direct[i], P->ni2_full[i], d->Nt[i], d->Cn[i], d->Cp[i], d->nt0[i],
                           d->pt0[i], dc(n[i]), dc(p[i]), ts->fb ? ts->fb[i - 1] : 0.0,
                           ts->fb ? dti : 0.0, &Rn, &Rp);
                Ri = c == 0 ? Rn.v : Rp.v;
            }
            U[i] = (i == 0 || i == NP + 1) ? 0.0 : Ri - d->G[i] * g_frac;
            double change = dti * (dn[i] - before[i]);
            if (i > 0 || ts->zimt) U[i] += change;
        }
        for (int i = 0; i <= NP; ++i) ev[i] = exp(sn * (full[nc * i] + off[i]) / Vt);
        iu[0] = 0.0;
        for (int j = 0; j < NP; ++j) iu[j + 1] = iu[j] + 0.5 * (d->x[j + 1] - d->x[j]) * (U[j] + U[j + 1]);
        double single = 0.0, dbl = 0.0;
        for (int j = 0; j < NP; ++j) {
            double dx = d->x[j + 1] - d->x[j];
            single += 0.5 * dx * (ev[j] / mu[j] + ev[j + 1] / mu[j + 1]);
            dbl += 0.5 * dx * (ev[j] * iu[j] / mu[j] + ev[j + 1] * iu[j + 1] / mu[j + 1]);
        }
        double VgN = full[nc * (NP + 1)] + off[NP + 1], Vg0 = full[0] + off[0];
        double K = (Vt * (sn * dn[NP + 1] * exp(sn * VgN / Vt) - sn * dn[0] * exp(sn * Vg0 / Vt)) - sn * dbl) / single;
        J[0] = Q_E * K;
        double acc = 0.0;
        for (int e = 1; e <= NP; ++e) {
            acc += sn * Q_E * L * ts->width[e] * U[e];
            J[e] = Q_E * K + acc;
        }
        for (int e = NP; e >= 1; --e)
            if (d->iface[e]) J[e] = J[e - 1];
        for (int e = 0; e <= NP; ++e) Jsum[e] += sign * J[e];
    }
    if (nc == 5) { /* the mobile ions, one region per species */
        const double *masks[2] = {P->amask_full, P->cmask_full};
        double totals[2] = {P->a_total, P->c_total};
        for (int k = 0; k < 2; ++k) {
            if (!(totals[k] > 0)) continue;
            double sn = k == 0 ? -1.0 : 1.0;
            const double *m = masks[k], *mu = k == 0 ? d->mu_a : d->mu_c;
            const double *before = k == 0 ? ts->a_prev : ts->c_prev;
            int first = -1, last = -1;
            for (int i = 0; i < NP + 2; ++i)
                if (m[i] > 0) { if (first < 0) first = i; last = i; }
            if (first < 0) continue;
            for (int i = 0; i < NP + 2; ++i) {
                dens[i] = exp(full[nc * i + 3 + k]);
                U[i] = dti * (dens[i] - before[i]);
                ev[i] = exp(sn * full[nc * i] / Vt);
            }
            iu[0] = 0.0;
            double single = 0.0, dbl = 0.0;
            for (int e = 0; e <= NP; ++e) {
                double pair = m[e] * m[e + 1], dx = d->x[e + 1] - d->x[e];
                iu[e + 1] = iu[e] + 0.5 * dx * (U[e] + U[e + 1]) * pair;
            }
            for (int e = 0; e <= NP; ++e) {
                double pair = m[e] * m[e + 1], dx = d->x[e + 1] - d->x[e];
                double m0 = mu[e] > 0 ? mu[e] : 1.0, m1 = mu[e + 1] > 0 ? mu[e + 1] : 1.0;
                single += 0.5 * dx * (ev[e] / m0 + ev[e + 1] / m1) * pair;
                dbl += 0.5 * dx * (ev[e] * iu[e] / m0 + ev[e + 1] * iu[e + 1] / m1) * pair;
            }
            if (!(single > 0)) single = 1.0;
            double ends = sn * (dens[last] * ev[last] - dens[first] * ev[first]);
            double K = (Vt * ends - sn * dbl) / single;
            double acc = 0.0;
            for (int e = 0; e <= NP; ++e) {
                double pair = m[e] * m[e + 1];
                if (e > first) acc += sn * Q_E * L * ts->width[e] * U[e] * pair;
                J[e] = Q_E * K + acc;
            }
            for (int e = NP; e >= 1; --e)
                if (d->iface[e]) J[e] = J[e - 1];
            for (int e = 0; e <= NP; ++e)
                if (m[e] * m[e + 1] > 0) Jsum[e] += sign * J[e];
        }
    }
    /* the displacement current on each interval */
    for (int e = 0; e <= NP; ++e) {
        double dV = full[nc * (e + 1)] - full[nc * e], dV0 = ts->V_prev[e + 1] - ts->V_prev[e];
        Jsum[e] += sign * d->eps[e] * (dV - dV0) * dti / (L * d->h[e]);
    }
    double num = 0.0, den = 0.0;
    for (int e = 0; e <= NP; ++e) { num += d->h[e] * Jsum[e]; den += d->h[e]; }
    free(w);
    return num / den;
}

/* One Newton step of a time step in place on (y, logs) (batch.py:759-785 with `_system`'s 5 x 5
 * branch :518-536 and `_bordered_solve` :595-603); returns the step's size. nc = 3: no border
 * (a device without mobile ions holds both slots). */
static double newton_step_time(const problem *P, tscratch *s, double *y, double *logs, int nc,
                               double Vint, double g_frac, const tstep *ts) {
    int NP = P->NP;
    double Vt = P->dev->Vt;
    full_of(P, y, nc, Vint, ts->dti, s->full);
    assemble_step(P, s->full, nc, g_frac, ts, logs, s->F, s->A, s->B, s->C, s->U);
    int kr = nc == 5 ? 3 : 1; /* right-hand sides: -F, and the slots' columns */
    for (int r = 0; r < NP; ++r)
        for (int row = 0; row < nc; ++row) {
            s->rhs[(size_t)kr * (nc * r + row)] = -s->F[nc * r + row];
            if (nc == 5) {
                s->rhs[(size_t)kr * (nc * r + row) + 1] = s->U[2 * (nc * r + row)];
                s->rhs[(size_t)kr * (nc * r + row) + 2] = s->U[2 * (nc * r + row) + 1];
            }
        }
    block_thomas(NP, nc, kr, s->A, s->B, s->C, s->rhs, s->X, s->cp, s->dp);
    double z[2] = {0.0, 0.0};
    if (nc == 5) {
        double g[2];
        ion_constraints_time(P, s->full, g, s->W);
        double held[2] = {P->a_total > 0 ? 0.0 : 1.0, P->c_total > 0 ? 0.0 : 1.0};
        double S[4] = {held[0], 0.0, 0.0, held[1]};
        z[0] = -g[0] - held[0] * logs[0];
        z[1] = -g[1] - held[1] * logs[1];
        for (int r = 0; r < NP; ++r)
            for (int k = 0; k < 2; ++k) { /* W lives in the ions' own columns */
                double w = s->W[2 * r + k];
                const double *x = s->X + (size_t)kr * (nc * r + 3 + k);
                S[2 * k + 0] -= w * x[1];
                S[2 * k + 1] -= w * x[2];
                z[k] -= w * x[0];
            }
        solve_small(2, S, z);
    }
    double capv = CAP_V * Vt, lnfloor = log(P->n_floor), size = 0.0;
    for (int r = 0; r < NP; ++r) {
        double step[5];
        for (int c = 0; c < nc; ++c) {
            const double *x = s->X + (size_t)kr * (nc * r + c);
            double delta = nc == 5 ? x[0] - (x[1] * z[0] + x[2] * z[1]) : x[0];
            if (c == 0) {
                step[c] = clipd(delta, -capv, capv);
            } else {
                step[c] = clipd(delta, -CAP_N, CAP_N);
                if (c < 3 && y[nc * r + c] < lnfloor) { /* density_steps (solver.py:74-88) */
                    double in_n = delta > -1 ? log1p(delta) : -CAP_N;
                    step[c] = clipd(in_n, -FLOOR_DROP_CAP, CAP_N);
                }
            }
        }
        for (int c = 0; c < nc; ++c) y[nc * r + c] += step[c];
        size = nmax(size, fabs(step[0]) / Vt);
        for (int c = 1; c < nc; ++c) {
            double dens = exp(y[nc * r + c]);
            size = nmax(size, fabs(step[c]) * (dens / (dens + P->n_floor)));
        }
    }
    if (nc == 5) {
        double dl0 = clipd(z[0], -ION_LOG_CAP, ION_LOG_CAP), dl1 = clipd(z[1], -ION_LOG_CAP, ION_LOG_CAP);
        logs[0] += dl0;
        logs[1] += dl1;
        size = nmax(size, nmax(fabs(dl0), fabs(dl1)));
    }
    return size;
}

/* `_newton_loop` (batch.py:740-848) of a time step: the four stop rules, the current of each
 * iterate the transient's. Returns converged; *J the last iterate's current, *it the steps. */
static int newton_loop_time(const problem *P, tscratch *s, double *y, double *logs, int nc,
                            double Vint, double g_frac, const tstep *ts, double tol, int max_iter,
                            double *J_out, int *it_out) {
    double hist[STALL_STEPS + 1], Jh[STALL_STEPS + 1];
    for (int k = 0; k <= STALL_STEPS; ++k) hist[k] = Jh[k] = INFINITY;
    int it = 0, done = 0;
    double J_scale = P->dev->J_scale;
    while (!done && it < max_iter && !isnan(hist[STALL_STEPS])) {
        double size = newton_step_time(P, s, y, logs, nc, Vint, g_frac, ts);
        for (int k = 0; k < STALL_STEPS; ++k) { hist[k] = hist[k + 1]; Jh[k] = Jh[k + 1]; }
        hist[STALL_STEPS] = size;
        full_of(P, y, nc, Vint, ts->dti, s->full);
        Jh[STALL_STEPS] = current_step(P, s->full, nc, g_frac, ts);
        ++it;
        double hmax = -INFINITY, hmin_tail = INFINITY;
        int anynan = 0;
        for (int k = 0; k <= STALL_STEPS; ++k) { hmax = fmax(hmax, hist[k]); anynan |= isnan(hist[k]); }
        for (int k = 1; k <= STALL_STEPS; ++k) hmin_tail = fmin(hmin_tail, hist[k]);
        int stalled = it > STALL_STEPS && !anynan && hmax < ROUND_OFF_FLOOR && hmin_tail >= 0.1 * hist[0];
        double Jn = Jh[STALL_STEPS], scale = J_scale + fabs(Jn), jdev = 0.0;
        int jnan = 0;
        for (int k = 0; k <= STALL_STEPS; ++k) {
            double v = fabs(Jh[k] - Jn);
            if (isnan(v)) jnan = 1;
            jdev = fmax(jdev, v);
        }
        int settled = it > STALL_STEPS && size < CURRENT_STEP_CAP && !jnan && jdev < CURRENT_TOL * scale;
        double prev = hist[STALL_STEPS - 1];
        int quadratic = it > 1 && size < QUADRATIC_TOL && size < QUADRATIC_RATIO * prev;
        done = size < tol || stalled || settled || quadratic;
    }
    *J_out = Jh[STALL_STEPS];
    *it_out = it;
    return done;
}

/* `solve_transient` for one device (transient.py:270-464 `_solve_one`, init None): the steady
 * state at the first step's voltage from the start (`batch._start`), then every step from the
 * one before, the first guess extrapolated linearly in time from the last two kept states
 * (held to one Newton step's caps; none from or into a steady state), each step's dt from the
 * last step kept; failureMode 0 stops at a failed step, 1 keeps it, 2 skips it (Track 0
 * throughout). Outputs per step: J (internal), J_ext (with the shunt), converged, iterations,
 * kept (a row of ZimT's tJFile). Returns 0, -1 for a bad shape, -3 for a device outside this
 * subset (a pre-bias; an ion region reaching a contact). */
int ddo_transient(const ddo_device *d, const double *t, const double *V, const double *G,
                  int nt, const ddo_topts *o, double *Jout, double *Jext, uint8_t *okout,
                  int32_t *itsout, uint8_t *keptout) {
    int NP = d->NP;
    if (NP < 2 || NP > 4096 || nt < 1) return -1;
    if (d->pre_cond) return -3;
    if (d->tt)  /* charged traps' occupations in time are outside the restatement */
        for (int i = 0; i < NP + 2; ++i)
            if (d->tt[i] != 0.0) return -5;
    problem P;
    problem_init(&P, d, o->bernoulli);
    int ions = d->a_total > 0 || d->c_total > 0;
    if (ions && (P.amask_full[0] > 0 || P.cmask_full[0] > 0 || P.amask_full[NP + 1] > 0 ||
                 P.cmask_full[NP + 1] > 0)) {
        problem_free(&P);
        return -3;
    }
    int nc = ions ? 5 : 3;
    size_t ny = (size_t)NP * nc;
    scratch s3;
    scratch_init(&s3, NP);
    tscratch s;
    tscratch_init(&s, NP);
    double Vt = d->Vt;
    /* the steady start at the first voltage (transient.py:307-313) */
    ddo_opts so = {o->tol, o->max_iter, o->max_iter, 0, o->bernoulli};
    double *y3 = alloc_d((size_t)NP * 3), logs[2];
    double V0 = V[0];
    start_state(&P, &s3, d, V0, G[0], &so, y3, logs);
    {
        double *yn = alloc_d((size_t)NP * 3), ln[2] = {logs[0], logs[1]}, Jd;
        int itd;
        ionscale sc = {1.0, d->a_total, d->c_total};
        memcpy(yn, y3, (size_t)NP * 3 * sizeof(double));
        int ok_s = newton_loop(&P, &s3, yn, ln, V0, G[0], &sc, o->tol, o->max_iter, 0, 0, &Jd, &itd);
        if (ok_s) { memcpy(y3, yn, (size_t)NP * 3 * sizeof(double)); logs[0] = ln[0]; logs[1] = ln[1]; }
        free(yn);
    }
    double *y = alloc_d(ny), *y_prev = alloc_d(ny), *guess = alloc_d(ny), *after = alloc_d((size_t)(NP + 2) * nc);
    for (int r = 0; r < NP; ++r) {
        for (int c = 0; c < 3; ++c) y[nc * r + c] = y3[3 * r + c];
        if (ions) { /* the ions' densities as unknowns: Boltzmann's profiles, 0 outside */
            double Vm = y3[3 * r];
            y[nc * r + 3] = P.am[r] > 0 ? logs[0] + Vm / Vt : 0.0;
            y[nc * r + 4] = P.cm[r] > 0 ? logs[1] - Vm / Vt : 0.0;
        }
    }
    memcpy(y_prev, y, ny * sizeof(double));
    double logs_prev[2] = {logs[0], logs[1]};
    double *fb = alloc_d(NP), *fb_new = alloc_d(NP);
    tstep ts;
    ts.n_prev = alloc_d(NP + 2); ts.p_prev = alloc_d(NP + 2); ts.a_prev = alloc_d(NP + 2);
    ts.c_prev = alloc_d(NP + 2); ts.V_prev = alloc_d(NP + 2); ts.fb = fb;
    ts.width = alloc_d(NP + 1);
    ts.zimt = o->zimt_current;
    for (int e = 0; e <= NP; ++e) /* `_sum_widths` (batch.py:405-418) */
        ts.width[e] = o->zimt_current || e == 0 ? d->h[e] : 0.5 * (d->h[e - 1] + d->h[e]);
    double V_now = V0, V_before = V0, dti_now = 0.0, t_last = 0.0;
    int first = 1, stopped = 0;
    double caps[5] = {CAP_V * Vt, CAP_N, CAP_N, CAP_N, CAP_N};
    for (int k = 0; k < nt; ++k) {
        double dti = t[k] == 0 ? 0.0 : 1.0 / (t[k] > t_last ? t[k] - t_last : 1.0);
        /* the state before the step and its time terms (`_time_terms`) */
        full_of(&P, y, nc, V_now, dti_now, after);
        ts.dti = dti;
        for (int i = 0; i < NP + 2; ++i) {
            ts.n_prev[i] = exp(after[nc * i + 1]);
            ts.p_prev[i] = exp(after[nc * i + 2]);
            ts.V_prev[i] = after[nc * i];
            ts.a_prev[i] = ions ? exp(after[nc * i + 3]) * P.amask_full[i] : 0.0;
            ts.c_prev[i] = ions ? exp(after[nc * i + 4]) * P.cmask_full[i] : 0.0;
        }
        int moving = dti > 0 && dti_now > 0 && !first;
        double rt = moving ? dti_now / dti : 0.0;
        for (size_t q = 0; q < ny; ++q) {
            double cap = caps[q % nc];
            guess[q] = y[q] + clipd(rt * (y[q] - y_prev[q]), -cap, cap);
        }
        double lg[2];
        for (int j = 0; j < 2; ++j) lg[j] = logs[j] + clipd(rt * (logs[j] - logs_prev[j]), -ION_LOG_CAP, ION_LOG_CAP);
        double Vk = V[k];
        if (ions) { /* in a steady state Boltzmann's normalisations, over a step the scales */
            if (dti > 0) {
                if (!moving) lg[0] = lg[1] = 0.0;
            } else {
                double *Vf = alloc_d(NP + 2), bnd[6];
                ionscale sc = {1.0, d->a_total, d->c_total};
                bounds_of(d, Vk, bnd);
                Vf[0] = bnd[0]; Vf[NP + 1] = bnd[3];
                for (int r = 0; r < NP; ++r) Vf[r + 1] = guess[nc * r];
                ion_logs(&P, Vf, &sc, lg);
                free(Vf);
            }
        }
        double J;
        int it;
        int ok = newton_loop_time(&P, &s, guess, lg, nc, Vk, G[k], &ts, o->tol, o->max_iter, &J, &it);
        int finite = all_finite(guess, (int)ny) && isfinite(lg[0]) && isfinite(lg[1]);
        int running = !stopped;
        int failed = running && !ok && o->failure_mode == 0;
        int keep = running && !failed && finite &&
                   (ok || o->failure_mode == 1 || (o->failure_mode == 2 && dti == 0));
        Jout[k] = J;
        Jext[k] = J + d->G_shunt * Vk;
        okout[k] = (uint8_t)ok;
        itsout[k] = it;
        keptout[k] = (uint8_t)keep;
        if (keep) {
            full_of(&P, guess, nc, Vk, dti, after);
            for (int r = 0; r < NP; ++r) {
                int i = r + 1;
                fb_new[r] = trap_end(P.Nt[r], P.Cn[r], P.Cp[r], P.nt0[r], P.pt0[r],
                                     exp(after[nc * i + 1]), exp(after[nc * i + 2]), fb[r], dti);
            }
            memcpy(fb, fb_new, (size_t)NP * sizeof(double));
            memcpy(y_prev, y, ny * sizeof(double));
            memcpy(y, guess, ny * sizeof(double));
            logs_prev[0] = logs[0]; logs_prev[1] = logs[1];
            logs[0] = lg[0]; logs[1] = lg[1];
            V_before = V_now;
            V_now = Vk;
            dti_now = dti;
            t_last = t[k];
        }
        if (failed) stopped = 1;
        first = 0;
    }
    (void)V_before;
    free(y3); free(y); free(y_prev); free(guess); free(after); free(fb); free(fb_new);
    free(ts.n_prev); free(ts.p_prev); free(ts.a_prev); free(ts.c_prev); free(ts.V_prev); free(ts.width);
    tscratch_free(&s);
    scratch_free(&s3);
    problem_free(&P);
    return 0;
}

/* One time step's system at a state, for the tests: F, A, B, C, U as assemble_step writes them
 * (nc = 3 or 5), the current, and the constraints g (nc = 5). The step's fixed part: dti, the
 * state before it (full_prev, (NP + 2) x nc, its ions' contacts set by the caller's ion_full
 * rule) and the traps' occupation fb (NP). */
int ddo_assemble_step(const ddo_device *d, const double *full, int nc, double g_frac, double dti,
                      const double *full_prev, const double *fb, const double *logs, int bernoulli,
                      int zimt, double *F, double *A, double *B, double *C, double *U, double *J,
                      double *g, double *Wout) {
    int NP = d->NP;
    if (nc != 3 && nc != 5) return -1;
    problem P;
    problem_init(&P, d, bernoulli);
    tstep ts;
    ts.dti = dti;
    ts.n_prev = alloc_d(NP + 2); ts.p_prev = alloc_d(NP + 2); ts.a_prev = alloc_d(NP + 2);
    ts.c_prev = alloc_d(NP + 2); ts.V_prev = alloc_d(NP + 2); ts.width = alloc_d(NP + 1);
    ts.fb = (double *)fb;
    ts.zimt = zimt;
    for (int e = 0; e <= NP; ++e)
        ts.width[e] = zimt || e == 0 ? d->h[e] : 0.5 * (d->h[e - 1] + d->h[e]);
    for (int i = 0; i < NP + 2; ++i) {
        ts.n_prev[i] = exp(full_prev[nc * i + 1]);
        ts.p_prev[i] = exp(full_prev[nc * i + 2]);
        ts.V_prev[i] = full_prev[nc * i];
        ts.a_prev[i] = nc == 5 ? exp(full_prev[nc * i + 3]) * P.amask_full[i] : 0.0;
        ts.c_prev[i] = nc == 5 ? exp(full_prev[nc * i + 4]) * P.cmask_full[i] : 0.0;
    }
    assemble_step(&P, full, nc, g_frac, &ts, logs, F, A, B, C, U);
    if (J) *J = current_step(&P, full, nc, g_frac, &ts);
    if ((g || Wout) && nc == 5) {
        double *W = alloc_d((size_t)NP * 2), gl[2];
        ion_constraints_time(&P, full, gl, W);
        if (g) { g[0] = gl[0]; g[1] = gl[1]; }
        if (Wout) for (int k = 0; k < 2 * NP; ++k) Wout[k] = W[k];
        free(W);
    }
    free(ts.n_prev); free(ts.p_prev); free(ts.a_prev); free(ts.c_prev); free(ts.V_prev); free(ts.width);
    problem_free(&P);
    return 0;
}
