// The impedance operator (src/jaxdd/impedance.py:230-360 `_admittance`, devices without mobile
// ions): at a steady state a small internal voltage dV e^(iwt) drives
// (J + lap M) dy = -(dF/dV_int) dV, one complex block-tridiagonal system per frequency, and the
// admittance is the transient current's linearisation dJ = c . dy + J_V (impedance.py:316-358).
//
// Two launches. `impedance_prep_lane`, one lane per device, assembles the steady Jacobian's
// blocks at the state (the Newton sweeps' row arithmetic, row_blocks_s), the time terms M, the
// bulk traps' occupation terms and the coefficients of the linearised current, which do not
// depend on the frequency, into a per-device table. `impedance_lane`, one lane per (device,
// frequency), eliminates its complex system from the first row to the last and contracts the
// current's coefficients on the way: with x_i = d'_i - C'_i x_(i+1), sum_i c_i x_i =
// sum_i g_i d'_i for g_0 = c_0, g_i = c_i - g_(i-1) C'_(i-1), so the solve needs no
// back-substitution and stores no factors: the lanes of one device read its table together
// and keep everything else in registers.
//
// The linearised current (batch.py:314-403 over a time step, each point's rate over its control
// volume): per carrier (sign sn, -1 electrons, +1 holes) the integral form
// J_j = q K + sum_(1<=i<=j) sn q L w_i U_i, K = (V_T sn (dens_(N+1) E_(N+1) - dens_0 E_0)
// - sn D)/S, S = sum_j c_j, D = sum_j c_j iu_j, c_j the trapezoid weight of g_j = e^(sn Vg_j/V_T)/mu_j,
// iu_j = int_0^x_j U; averaged over the intervals (weights h_j/H; across an interface the
// interval takes the one before it) plus the displacement current eps (dV_(j+1) - dV_j) lap/(L h_j).
// Its derivatives: dK/dV_k = -c_k (iu_k + sn K)/(V_T S), dK/dU_k = -sn (dx_k P_(k+1) +
// dx_(k-1) P_k)/(2 S) with P_k = sum_(j>=k) c_j, and the direct part sn q L w_k T_k with
// T_k = (sum_(j>=k) h_j - h_k [interface k])/H. The rates move by
// dU = dR/dln(n,p) . dy (each carrier's own density's change lap dens dln dens in addition,
// and the traps' occupation beyond its steady response in both).
#pragma once

#include "ddhip_core.h"

namespace ddh {

// A complex number for the small-signal systems.
struct cplx {
    double re, im;
    DDH_MFN cplx() : re(0.0), im(0.0) {}
    DDH_MFN cplx(double r, double i = 0.0) : re(r), im(i) {}
};
DDH_MFN cplx operator+(cplx a, cplx b) { return cplx(a.re + b.re, a.im + b.im); }
DDH_MFN cplx operator-(cplx a, cplx b) { return cplx(a.re - b.re, a.im - b.im); }
DDH_MFN cplx operator-(cplx a) { return cplx(-a.re, -a.im); }
DDH_MFN cplx operator*(cplx a, cplx b) {
    return cplx(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
DDH_MFN cplx operator*(cplx a, double b) { return cplx(a.re * b, a.im * b); }
DDH_MFN cplx operator*(double b, cplx a) { return cplx(a.re * b, a.im * b); }
DDH_MFN cplx operator/(cplx a, cplx b) {
    const double r = 1.0 / (b.re * b.re + b.im * b.im);
    return cplx((a.re * b.re + a.im * b.im) * r, (a.im * b.re - a.re * b.im) * r);
}

// Per point of the per-device table [n][NP][IM_N]: the steady blocks (row-major 3x3, the last
// row's C zero: its coupling is the contact's), the time terms, the traps' occupation terms
// and the current's coefficients.
enum ImpField {
    IM_A = 0, IM_B = 9, IM_C = 18,
    IM_MN = 27, IM_MP,     // -fac sC n, -fac sC p: lap times these on the continuity diagonals
    IM_RATE,               // the trap level's rate a (1 without traps)
    IM_E, IM_M,            // e^(-a dt), (1 - e^(-a dt))/(a dt) (implicit Euler's steps)
    IM_UN, IM_UP,          // the steady occupation's drive per unit of ln n, ln p, over a
    IM_TN, IM_TP,          // -fac sC dR_n/df, -fac sC dR_p/df: the occupation in the rows
    IM_KV, IM_KD,          // the current's coefficient of dV: KV + lap KD
    IM_G0N, IM_G0P,        // ... of dln n: G0N + lap GLN + h UN GT; of dln p likewise
    IM_GLN, IM_GLP, IM_GT,
    IM_DRN, IM_DRP,        // dR_n/df, dR_p/df
    IM_NN, IM_PP,          // n, p
    IM_N
};
// Per device [n][IS_N]: the drive (-dF/dV_int) of the first and the last row, the current's
// direct dependence on V_int (IS_JV + lap IS_JD).
enum ImpScalar { IS_D0 = 0, IS_DN = 3, IS_JV = 6, IS_JD, IS_N };

struct ImpArgs {
    int n, NP, nseg, bern, nf;
    const double *pt;
    const uint8_t *sid;
    const double *seg;
    const double *sc;
    const double *y;        // [NP][3][n] the steady interior state
    const double *Vint;     // [n]
    const uint8_t *iface;   // [NP + 1][n] 1 where the interval after the point crosses layers
    const double *laps;     // [nf][2] lap (re, im)
    double dt;              // implicit Euler's step (0: the continuous response)
    double *tab;            // [n][NP][IM_N]
    double *dsc;            // [n][IS_N]
    double *Y;              // [n][nf][2] dJ/dV_int
    int32_t *status;        // [n] 0, or 1: mobile ions (not handled; Y = NaN)
};

#define ISC(f) a.sc[(size_t)(f) * a.n + d]
#define IPT(j, f) a.pt[((size_t)(j) * PT_N + (f)) * a.n + d]
#define ISID(j) ((int)a.sid[(size_t)(j) * a.n + d])
#define ISEG(id, f) a.seg[((size_t)(id) * SEG_N + (f)) * a.n + d]
#define IY(j, f) a.y[(((size_t)(j) - 1) * 3 + (f)) * a.n + d]

DDH_FN void impedance_prep_lane(const ImpArgs &a, int d) {
    const int NP = a.NP;
    double *tab = a.tab + (size_t)d * NP * IM_N;
    double *ds = a.dsc + (size_t)d * IS_N;
    if (ISC(SC_AT) > 0 || ISC(SC_CT) > 0) {
        a.status[d] = 1;
        return;
    }
    a.status[d] = 0;
    const double Vt = ISC(SC_VT), iVt = 1.0 / Vt, L = ISC(SC_L), L2 = L * L;
    const double gfr = ISC(SC_GFRAC), elec = ISC(SC_ELEC), Vint = a.Vint[d];
    const double half = 0.5 * elec;
    const double V0 = ISC(SC_V0) + half * Vint, VN = ISC(SC_VL) - half * Vint;
    auto point = [&](int j, Pt &q) {
        q.sid = ISID(j);
        q.offn = ISEG(q.sid, SEG_OFFN); q.offp = ISEG(q.sid, SEG_OFFP);
        q.mun = ISEG(q.sid, SEG_MUN); q.mup = ISEG(q.sid, SEG_MUP);
        q.am = q.cm = q.an = q.ca = 0.0;
        q.h = IPT(j, PT_H);
        if (j == 0 || j == NP + 1) {
            q.V = j == 0 ? V0 : VN;
            q.n = exp(j == 0 ? ISC(SC_LNL) : ISC(SC_LNR));
            q.p = exp(j == 0 ? ISC(SC_LPL) : ISC(SC_LPR));
            q.G = q.fix = q.R = q.dRn = q.dRp = 0.0;
        } else {
            q.V = IY(j, 0);
            q.n = exp(IY(j, 1));
            q.p = exp(IY(j, 2));
            q.G = IPT(j, PT_G);
            q.fix = IPT(j, PT_FIX);
            q.R = recombination(q.n, q.p, ISEG(q.sid, SEG_KD), ISEG(q.sid, SEG_NI2),
                                ISEG(q.sid, SEG_NT), ISEG(q.sid, SEG_CN), ISEG(q.sid, SEG_CP),
                                ISEG(q.sid, SEG_NT0), ISEG(q.sid, SEG_PT0), q.dRn, q.dRp);
        }
    };
    // pass 1: the blocks, the time and trap terms; the current's sums (S, D per carrier), each
    // point's trapezoid weights c_j, iu_j and the prefixes sum_(j<i) c_j in the table's
    // coefficient slots until pass 2
    Pt cur, mid, rgt;
    point(0, cur);
    point(1, mid);
    Edge eL = edge_flux_pt(cur, mid, Vt, iVt, a.bern);
    double H = 0.0;
    for (int j = 0; j <= NP; ++j) H += IPT(j, PT_H);
    const double dx0 = IPT(1, PT_X) - IPT(0, PT_X);
    const double gn0 = exp(-(V0 + cur.offn) * iVt) * ISEG(cur.sid, SEG_IMUN);
    const double gp0 = exp((V0 + cur.offp) * iVt) * ISEG(cur.sid, SEG_IMUP);
    double cpn = 0.5 * dx0 * gn0, cpp = 0.5 * dx0 * gp0;  // c_0
    double Sn = cpn, Sp = cpp, Dn = 0.0, Dp = 0.0;        // iu_0 = 0
    double U_prev = 0.0, iu = 0.0, x_prev = IPT(0, PT_X);
    for (int i = 1; i <= NP; ++i) {
        point(i + 1, rgt);
        const Edge eR = edge_flux_pt(mid, rgt, Vt, iVt, a.bern);
        RowSeg g;
        g.meps = ISEG(mid.sid, SEG_EPS);
        g.ceps = ISEG(cur.sid, SEG_EPS);
        g.sP = ISEG(mid.sid, SEG_SP); g.scs = ISEG(mid.sid, SEG_SCS); g.dop = ISEG(mid.sid, SEG_DOP);
        double Ab[9], Bb[9], Cb[9], r[3];
        row_blocks_s<1>(g, cur, mid, rgt.V, eL, eR, false, L2, iVt, 1.0, gfr, Ab, Bb, Cb, r);
        double *t = tab + (size_t)(i - 1) * IM_N;
        if (i == 1) {
            for (int row = 0; row < 3; ++row) ds[IS_D0 + row] = -Ab[3 * row] * half;
        }
        if (i == NP) {
            for (int row = 0; row < 3; ++row) ds[IS_DN + row] = Cb[3 * row] * half;
            for (int q = 0; q < 9; ++q) Cb[q] = 0.0;
        }
        for (int q = 0; q < 9; ++q) {
            t[IM_A + q] = Ab[q];
            t[IM_B + q] = Bb[q];
            t[IM_C + q] = Cb[q];
        }
        const double hm = cur.h, hs = mid.h + cur.h;
        const double fs = -0.5 * L2 * hm * hs * g.scs;  // -fac sC
        t[IM_MN] = fs * mid.n;
        t[IM_MP] = fs * mid.p;
        // the bulk traps' occupation (impedance.py:63-87), one level
        const double Nt = ISEG(mid.sid, SEG_NT), Cn = ISEG(mid.sid, SEG_CN), Cp = ISEG(mid.sid, SEG_CP);
        const double nt0 = ISEG(mid.sid, SEG_NT0), pt0 = ISEG(mid.sid, SEG_PT0);
        double rate = 1.0, un = 0.0, up = 0.0, drn = 0.0, drp = 0.0;
        if (Nt > 0) {
            rate = Cn * (mid.n + nt0) + Cp * (mid.p + pt0);
            const double f = (Cn * mid.n + Cp * pt0) / rate;
            un = Cn * mid.n * (1.0 - f) / rate;
            up = -f * Cp * mid.p / rate;
            drn = -Nt * Cn * (mid.n + nt0);
            drp = Nt * Cp * (mid.p + pt0);
        }
        t[IM_RATE] = rate;
        if (a.dt > 0) {
            const double x = rate * a.dt;
            t[IM_E] = exp(-x);
            t[IM_M] = -expm1(-x) / x;
        } else {
            t[IM_E] = t[IM_M] = 0.0;
        }
        t[IM_UN] = un; t[IM_UP] = up;
        t[IM_TN] = fs * drn; t[IM_TP] = fs * drp;
        t[IM_DRN] = drn; t[IM_DRP] = drp;
        t[IM_NN] = mid.n; t[IM_PP] = mid.p;
        // the current's sums at point i (the steady rates: U = R - G g_frac, 0 at the contacts)
        const double Uu = mid.R - mid.G * gfr;
        const double xj = IPT(i, PT_X), dxl = xj - x_prev;
        iu += 0.5 * dxl * (U_prev + Uu);
        const double dxr = i < NP ? IPT(i + 1, PT_X) - xj : 0.0;
        const double wgt = 0.5 * (dxl + dxr);
        const double gn = exp(-(mid.V + mid.offn) * iVt) * ISEG(mid.sid, SEG_IMUN);
        const double gp = exp((mid.V + mid.offp) * iVt) * ISEG(mid.sid, SEG_IMUP);
        const double cn = wgt * gn, cp = wgt * gp;
        t[IM_KV] = cn; t[IM_KD] = cp; t[IM_G0N] = iu;
        t[IM_G0P] = Sn; t[IM_GLN] = Sp;  // prefixes sum_(j<i) c_j
        t[IM_GLP] = mid.dRn; t[IM_GT] = mid.dRp;
        Sn += cn; Sp += cp;
        Dn += cn * iu; Dp += cp * iu;
        U_prev = Uu; x_prev = xj;
        cur = mid; mid = rgt; eL = eR;
    }
    // K per carrier (sn = -1 electrons, +1 holes)
    const int sid0 = ISID(0), sidN = ISID(NP + 1);
    const double En0 = exp(ISC(SC_LNL) - (V0 + ISEG(sid0, SEG_OFFN)) * iVt);   // n_0 E_0
    const double Ep0 = exp(ISC(SC_LPL) + (V0 + ISEG(sid0, SEG_OFFP)) * iVt);
    const double EnN = exp(ISC(SC_LNR) - (VN + ISEG(sidN, SEG_OFFN)) * iVt);
    const double EpN = exp(ISC(SC_LPR) + (VN + ISEG(sidN, SEG_OFFP)) * iVt);
    const double Kn = (-Vt * (EnN - En0) + Dn) / Sn;
    const double Kp = (Vt * (EpN - Ep0) - Dp) / Sp;
    const double sign = -elec, qs = sign * Q_E;
    // the contacts: dK/dV_0 (through e^(sn Vg_0/V_T) and g_0), dK/dV_(N+1)
    const double dKn0 = (-En0 - Kn * (-1.0) * cpn * iVt) / Sn, dKnN = EnN / Sn;
    const double dKp0 = (-Ep0 - Kp * cpp * iVt) / Sp, dKpN = EpN / Sp;
    ds[IS_JV] = qs * ((dKn0 + dKp0) * half - (dKnN + dKpN) * half);
    const double iLH = 1.0 / (L * H);
    const double eps0 = ISEG(sid0, SEG_EPS), epsN = ISEG(ISID(NP), SEG_EPS);
    ds[IS_JD] = sign * (-half) * (epsN + eps0) * iLH;
    // pass 2: the coefficients
    double hsum = 0.0;  // sum_(j<i) h_j
    double eps_prev = eps0;
    x_prev = IPT(0, PT_X);
    hsum = IPT(0, PT_H);
    for (int i = 1; i <= NP; ++i) {
        double *t = tab + (size_t)(i - 1) * IM_N;
        const double cn = t[IM_KV], cp = t[IM_KD], iu_i = t[IM_G0N];
        const double Pn = Sn - t[IM_G0P], Pp = Sp - t[IM_GLN];  // sum_(j>=i) c_j
        const double dRln = t[IM_GLP], dRlp = t[IM_GT];
        const double xj = IPT(i, PT_X), dxl = xj - x_prev;
        const double dxr = i < NP ? IPT(i + 1, PT_X) - xj : 0.0;
        const double dKnV = -cn * (iu_i - Kn) * iVt / Sn;
        const double dKpV = -cp * (iu_i + Kp) * iVt / Sp;
        // dD/dU_i = dx_i P_(i+1)/2 + dx_(i-1) P_i/2
        const double dDn = 0.5 * dxr * (Pn - cn) + 0.5 * dxl * Pn;
        const double dDp = 0.5 * dxr * (Pp - cp) + 0.5 * dxl * Pp;
        const double hi = IPT(i, PT_H), hl = IPT(i - 1, PT_H);
        const double width = 0.5 * (hl + hi);
        const double T = (H - hsum - (a.iface[(size_t)i * a.n + d] ? hi : 0.0)) / H;
        const double direct = Q_E * L * width * T;
        const double wn = qs * (dDn / Sn) - sign * direct;   // dK_n/dU = +dD_n/S_n (sn = -1)
        const double wp = qs * (-dDp / Sp) + sign * direct;
        const double eps_i = ISEG(ISID(i), SEG_EPS);
        t[IM_KV] = qs * (dKnV + dKpV);
        t[IM_KD] = sign * (eps_prev - eps_i) * iLH;
        t[IM_G0N] = (wn + wp) * dRln;
        t[IM_G0P] = (wn + wp) * dRlp;
        t[IM_GLN] = wn * t[IM_NN];
        t[IM_GLP] = wp * t[IM_PP];
        t[IM_GT] = wn * t[IM_DRN] + wp * t[IM_DRP];
        eps_prev = eps_i;
        hsum += hi;
        x_prev = xj;
    }
}

// One (device, frequency): the complex system eliminated row by row, the current's
// coefficients contracted on the way.
DDH_FN void impedance_lane(const ImpArgs &a, int d, int f) {
    double *Yo = a.Y + ((size_t)d * a.nf + f) * 2;
    if (a.status[d] != 0) {
        Yo[0] = Yo[1] = NAN;
        return;
    }
    const int NP = a.NP;
    const double *tab = a.tab + (size_t)d * NP * IM_N;
    const double *ds = a.dsc + (size_t)d * IS_N;
    const cplx lap(a.laps[2 * f], a.laps[2 * f + 1]);
    const bool steps = a.dt > 0;
    const cplx ld = lap * a.dt;
    cplx Cp[9], dp[3], g[3];
    cplx acc(0.0, 0.0);
    for (int r = 0; r < NP; ++r) {
        const double *t = tab + (size_t)r * IM_N;
        double A[9], C[9];
        cplx B[9];
        for (int q = 0; q < 9; ++q) {
            A[q] = t[IM_A + q];
            B[q] = cplx(t[IM_B + q]);
            C[q] = t[IM_C + q];
        }
        // the occupation beyond its steady response, per unit of it (impedance.py:104-117)
        cplx h;
        if (steps) {
            const double E = t[IM_E], m = t[IM_M];
            h = -(m * ld) / (cplx(1.0 - E) + E * ld);
        } else {
            h = -lap / (lap + cplx(t[IM_RATE]));
        }
        const double un = t[IM_UN], up = t[IM_UP], tn = t[IM_TN], tp = t[IM_TP];
        B[4] = B[4] + lap * t[IM_MN] + h * (tn * un);
        B[5] = B[5] + h * (tn * up);
        B[7] = B[7] + h * (tp * un);
        B[8] = B[8] + lap * t[IM_MP] + h * (tp * up);
        cplx rhs[3];
        if (r > 0) {
            for (int i = 0; i < 3; ++i) {
                for (int k = 0; k < 3; ++k) {
                    B[3 * i + k] = B[3 * i + k] - (A[3 * i] * Cp[k] + A[3 * i + 1] * Cp[3 + k] +
                                                   A[3 * i + 2] * Cp[6 + k]);
                }
                rhs[i] = -(A[3 * i] * dp[0] + A[3 * i + 1] * dp[1] + A[3 * i + 2] * dp[2]);
            }
        }
        if (r == 0)
            for (int i = 0; i < 3; ++i) rhs[i] = cplx(ds[IS_D0 + i]);
        if (r == NP - 1)
            for (int i = 0; i < 3; ++i) rhs[i] = rhs[i] + cplx(ds[IS_DN + i]);
        cplx inv[9];
        inverse3<cplx>(B, inv);
        for (int i = 0; i < 3; ++i)
            dp[i] = inv[3 * i] * rhs[0] + inv[3 * i + 1] * rhs[1] + inv[3 * i + 2] * rhs[2];
        // the current's coefficients of this point's dy
        const double gt = t[IM_GT];
        cplx c[3];
        c[0] = cplx(t[IM_KV]) + lap * t[IM_KD];
        c[1] = cplx(t[IM_G0N]) + lap * t[IM_GLN] + h * (un * gt);
        c[2] = cplx(t[IM_G0P]) + lap * t[IM_GLP] + h * (up * gt);
        if (r > 0) {
            cplx gn[3];
            for (int k = 0; k < 3; ++k)
                gn[k] = c[k] - (g[0] * Cp[k] + g[1] * Cp[3 + k] + g[2] * Cp[6 + k]);
            for (int k = 0; k < 3; ++k) g[k] = gn[k];
        } else {
            for (int k = 0; k < 3; ++k) g[k] = c[k];
        }
        for (int i = 0; i < 3; ++i)
            for (int k = 0; k < 3; ++k)
                Cp[3 * i + k] = inv[3 * i] * C[k] + inv[3 * i + 1] * C[3 + k] + inv[3 * i + 2] * C[6 + k];
        acc = acc + g[0] * dp[0] + g[1] * dp[1] + g[2] * dp[2];
    }
    const cplx Y = cplx(ds[IS_JV]) + lap * ds[IS_JD] + acc;
    Yo[0] = Y.re;
    Yo[1] = Y.im;
}

#undef ISC
#undef IPT
#undef ISID
#undef ISEG
#undef IY

}  // namespace ddh
