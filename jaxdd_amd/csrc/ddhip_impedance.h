// The impedance operator (src/jaxdd/impedance.py:230-360 `_admittance`, devices without mobile
// ions): at a steady state a small internal voltage dV e^(iwt) drives
// (J + lap M) dy = -(dF/dV_int) dV, one complex block-tridiagonal system per frequency, and the
// admittance is the transient current's linearisation dJ = c . dy + J_V (impedance.py:316-358).
//
// Two launches. The first, one wavefront per device, assembles the steady Jacobian's blocks at
// the state (the Newton sweeps' row arithmetic, row_blocks_s), the time terms M, the bulk traps'
// occupation terms and the coefficients of the linearised current, which do not depend on the
// frequency, into a per-device table. `impedance_lane`, one lane per (device, frequency),
// eliminates its complex system from the first row to the last and contracts the
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

// The per-device table [n][NP][IM_ROW] (a point's fields together, 384 bytes: three cache lines
// that the lanes of its device read at once): the steady blocks (row-major 3x3, the last row's C zero: its coupling is the
// contact's), the time terms, the traps' occupation terms and the current's coefficients.
enum ImpField {
    // the blocks: A and C keep their five entries that are not structural zeros (each
    // continuity row couples to V and to its own carrier only: [[0,0,0],[3,4,0],[6,0,8]]),
    // B all nine
    IM_A = 0, IM_B = 5, IM_C = 14,
    IM_MN = 19, IM_MP,     // -fac sC n, -fac sC p: lap times these on the continuity diagonals
    IM_RATE,               // the trap level's rate a (1 without traps)
    IM_E, IM_M,            // e^(-a dt), (1 - e^(-a dt))/(a dt) (implicit Euler's steps)
    IM_UN, IM_UP,          // the steady occupation's drive per unit of ln n, ln p, over a
    IM_TN, IM_TP,          // -fac sC dR_n/df, -fac sC dR_p/df: the occupation in the rows
    IM_KV, IM_KD,          // the current's coefficient of dV: KV + lap KD
    IM_G0N, IM_G0P,        // ... of dln n: G0N + lap GLN + h UN GT; of dln p likewise
    IM_GLN, IM_GLP, IM_GT,
    // the assembly's own: dR_n/df, dR_p/df, n, p, the sum of h before the point
    IM_DRN, IM_DRP, IM_NN, IM_PP, IM_TMP,
    IM_N,
    IM_ROW = 48            // doubles per point: three 128-byte lines
};
// With mobile ions (impedance.py:242-262) each point's row of the table carries the ions'
// fields after these (IMI_ROW doubles, four 128-byte lines): their rows of the 5 x 5 blocks (each
// couples to the potential and to its own species only), Poisson's derivatives in their
// logarithms, their time terms, the slots' constraint rows and the current's coefficients.
enum ImpIonField {
    IM_B03 = IM_N, IM_B04,            // Poisson's derivatives in ln a, ln c
    IM_IA,                            // the anions' row: A30, A33, B30, B33 (no time term), C30, C33
    IM_IC = IM_IA + 6,                // the cations' row likewise (column 4)
    IM_MA = IM_IC + 6, IM_MC,         // their time terms per unit of lap; the slots' columns are -M
    IM_WA, IM_WC,                     // the slots' constraint rows (d number/total / d ln)
    IM_CA0, IM_CAL, IM_CC0, IM_CCL,   // the current's coefficients of dln a: CA0 + lap CAL; dln c
    IM_LA, IM_LC,                     // ln a, ln c at the point (the assembly's own)
    IMI_N,
    IMI_ROW = 64
};
static_assert(IMI_N <= IMI_ROW, "the ion table's row holds its fields");

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
    double *tab;            // [n][NP][IM_ROW]
    double *dsc;            // [n][IS_N]
    double *Y;              // [n][nf][2] dJ/dV_int
    int32_t *status;        // [n] 0; 1: mobile ions on the 3 x 3 path; 2: ions outside the
                            // 5 x 5 path (a region at a contact, several of a species); Y = NaN
    const double *imu;      // [nseg][2][n] the ions' mobilities (the 5 x 5 path)
    int row;                // doubles per point of the table: IM_ROW, or IMI_ROW with ions
};

#define ISC(f) a.sc[(size_t)(f) * a.n + d]
#define IPT(j, f) a.pt[((size_t)(j) * PT_N + (f)) * a.n + d]
#define ISID(j) ((int)a.sid[(size_t)(j) * a.n + d])
#define ISEG(id, f) a.seg[((size_t)(id) * SEG_N + (f)) * a.n + d]
#define IY(j, f) a.y[(((size_t)(j) - 1) * 3 + (f)) * a.n + d]
#define ITAB(f, r) a.tab[((size_t)d * a.NP + (r)) * a.row + (f)]
#define IMU(id, q) a.imu[((size_t)(id) * 2 + (q)) * a.n + d]

// The table is assembled by one wavefront per device in three phases: (A) each point's blocks,
// time and trap terms and its parts of the current's integrals, lanes over the points; (B) the
// integrals' running sums, each lane over a contiguous segment of points, the segments' totals
// scanned across the wavefront; (C) the current's coefficients, lanes over the points again.
struct ImpDev {
    double Vt, iVt, L, L2, gfr, elec, half, V0, VN;
    double cn0, cp0, h0;   // the left contact's trapezoid weights and spacing
};

DDH_FN bool imp_mobile_ions(const ImpArgs &a, int d) { return ISC(SC_AT) > 0 || ISC(SC_CT) > 0; }

DDH_FN ImpDev imp_device(const ImpArgs &a, int d) {
    ImpDev k;
    k.Vt = ISC(SC_VT); k.iVt = 1.0 / k.Vt; k.L = ISC(SC_L); k.L2 = k.L * k.L;
    k.gfr = ISC(SC_GFRAC); k.elec = ISC(SC_ELEC); k.half = 0.5 * k.elec;
    const double Vint = a.Vint[d];
    k.V0 = ISC(SC_V0) + k.half * Vint;
    k.VN = ISC(SC_VL) - k.half * Vint;
    const int s0 = ISID(0);
    const double dx0 = IPT(1, PT_X) - IPT(0, PT_X);
    k.cn0 = 0.5 * dx0 * exp(-(k.V0 + ISEG(s0, SEG_OFFN)) * k.iVt) * ISEG(s0, SEG_IMUN);
    k.cp0 = 0.5 * dx0 * exp((k.V0 + ISEG(s0, SEG_OFFP)) * k.iVt) * ISEG(s0, SEG_IMUP);
    k.h0 = IPT(0, PT_H);
    return k;
}

DDH_FN void imp_point_at(const ImpArgs &a, int d, const ImpDev &k, int j, Pt &q) {
    const int NP = a.NP;
    q.sid = ISID(j);
    q.offn = ISEG(q.sid, SEG_OFFN); q.offp = ISEG(q.sid, SEG_OFFP);
    q.mun = ISEG(q.sid, SEG_MUN); q.mup = ISEG(q.sid, SEG_MUP);
    q.am = q.cm = q.an = q.ca = 0.0;
    q.h = IPT(j, PT_H);
    if (j == 0 || j == NP + 1) {
        q.V = j == 0 ? k.V0 : k.VN;
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
}

// (A) row r (point i = r + 1): blocks, time and trap terms, and in the coefficients' slots the
// point's trapezoid weights c_n, c_p, its step of the integral of U, 0.5 dx_(i-1) (U_(i-1) +
// U_i), and dR/dln n, dR/dln p.
DDH_FN void imp_assemble_row(const ImpArgs &a, int d, const ImpDev &k, int r) {
    const int NP = a.NP, i = r + 1;
    Pt cur, mid, rgt;
    imp_point_at(a, d, k, i - 1, cur);
    imp_point_at(a, d, k, i, mid);
    imp_point_at(a, d, k, i + 1, rgt);
    const Edge eL = edge_flux_pt(cur, mid, k.Vt, k.iVt, a.bern);
    const Edge eR = edge_flux_pt(mid, rgt, k.Vt, k.iVt, a.bern);
    RowSeg g;
    g.meps = ISEG(mid.sid, SEG_EPS);
    g.ceps = ISEG(cur.sid, SEG_EPS);
    g.sP = ISEG(mid.sid, SEG_SP); g.scs = ISEG(mid.sid, SEG_SCS); g.dop = ISEG(mid.sid, SEG_DOP);
    double Ab[9], Bb[9], Cb[9], rr[3];
    row_blocks_s<1>(g, cur, mid, rgt.V, eL, eR, false, k.L2, k.iVt, 1.0, k.gfr, Ab, Bb, Cb, rr);
    double *ds = a.dsc + (size_t)d * IS_N;
    if (r == 0)
        for (int row = 0; row < 3; ++row) ds[IS_D0 + row] = -Ab[3 * row] * k.half;
    if (r == NP - 1) {
        for (int row = 0; row < 3; ++row) ds[IS_DN + row] = Cb[3 * row] * k.half;
        for (int q = 0; q < 9; ++q) Cb[q] = 0.0;
    }
    for (int q = 0; q < 9; ++q) ITAB(IM_B + q, r) = Bb[q];
    ITAB(IM_A, r) = Ab[0]; ITAB(IM_A + 1, r) = Ab[3]; ITAB(IM_A + 2, r) = Ab[4];
    ITAB(IM_A + 3, r) = Ab[6]; ITAB(IM_A + 4, r) = Ab[8];
    ITAB(IM_C, r) = Cb[0]; ITAB(IM_C + 1, r) = Cb[3]; ITAB(IM_C + 2, r) = Cb[4];
    ITAB(IM_C + 3, r) = Cb[6]; ITAB(IM_C + 4, r) = Cb[8];
    const double fs = -0.5 * k.L2 * cur.h * (mid.h + cur.h) * g.scs;  // -fac sC
    ITAB(IM_MN, r) = fs * mid.n;
    ITAB(IM_MP, r) = fs * mid.p;
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
    ITAB(IM_RATE, r) = rate;
    if (a.dt > 0) {
        const double x = rate * a.dt;
        ITAB(IM_E, r) = exp(-x);
        ITAB(IM_M, r) = -expm1(-x) / x;
    } else {
        ITAB(IM_E, r) = 0.0;
        ITAB(IM_M, r) = 0.0;
    }
    ITAB(IM_UN, r) = un; ITAB(IM_UP, r) = up;
    ITAB(IM_TN, r) = fs * drn; ITAB(IM_TP, r) = fs * drp;
    ITAB(IM_DRN, r) = drn; ITAB(IM_DRP, r) = drp;
    ITAB(IM_NN, r) = mid.n; ITAB(IM_PP, r) = mid.p;
    // the current's parts (the steady rates: U = R - G g_frac, 0 at the contacts)
    const double U_i = mid.R - mid.G * k.gfr;
    const double U_l = i > 1 ? cur.R - cur.G * k.gfr : 0.0;
    const double xj = IPT(i, PT_X), dxl = xj - IPT(i - 1, PT_X);
    const double dxr = i < NP ? IPT(i + 1, PT_X) - xj : 0.0;
    const double wgt = 0.5 * (dxl + dxr);
    ITAB(IM_KV, r) = wgt * exp(-(mid.V + mid.offn) * k.iVt) * ISEG(mid.sid, SEG_IMUN);
    ITAB(IM_KD, r) = wgt * exp((mid.V + mid.offp) * k.iVt) * ISEG(mid.sid, SEG_IMUP);
    ITAB(IM_G0N, r) = 0.5 * dxl * (U_l + U_i);
    ITAB(IM_GLP, r) = mid.dRn;
    ITAB(IM_GT, r) = mid.dRp;
}

// (B) the rows of one segment [r0, r1): their totals of the integral's steps, c_n, c_p and h
DDH_FN void imp_segment_totals(const ImpArgs &a, int d, int r0, int r1, double tot[4]) {
    tot[0] = tot[1] = tot[2] = tot[3] = 0.0;
    for (int r = r0; r < r1; ++r) {
        tot[0] += ITAB(IM_G0N, r);
        tot[1] += ITAB(IM_KV, r);
        tot[2] += ITAB(IM_KD, r);
        tot[3] += IPT(r + 1, PT_H);
    }
}

// (B) the segment's running sums from the totals of the segments before it (`off`, with the
// left contact's parts): iu_i, sum_(j<i) c_n,j, sum_(j<i) c_p,j, sum_(j<i) h_j; returns the
// segment's parts of D_n, D_p
DDH_FN void imp_segment_prefix(const ImpArgs &a, int d, int r0, int r1, const double off[4],
                               double &Dn, double &Dp) {
    double iu = off[0], pn = off[1], pp = off[2], hs = off[3];
    Dn = Dp = 0.0;
    for (int r = r0; r < r1; ++r) {
        const double cn = ITAB(IM_KV, r), cp = ITAB(IM_KD, r);
        iu += ITAB(IM_G0N, r);
        ITAB(IM_G0N, r) = iu;
        ITAB(IM_G0P, r) = pn;
        ITAB(IM_GLN, r) = pp;
        ITAB(IM_TMP, r) = hs;
        Dn += cn * iu;
        Dp += cp * iu;
        pn += cn;
        pp += cp;
        hs += IPT(r + 1, PT_H);
    }
}

// The integrals' totals of a device, after (B): S_n, S_p (with the left contact's weight),
// D_n, D_p, H; K per carrier; the contacts' and the displacement current's direct terms.
struct ImpTot {
    double Sn, Sp, Dn, Dp, H, Kn, Kp;
};

DDH_FN ImpTot imp_totals(const ImpArgs &a, int d, const ImpDev &k, double Sn, double Sp,
                         double Dn, double Dp, double H, bool write) {
    const int NP = a.NP;
    ImpTot t;
    t.Sn = Sn; t.Sp = Sp; t.Dn = Dn; t.Dp = Dp; t.H = H;
    const int sid0 = ISID(0), sidN = ISID(NP + 1);
    const double En0 = exp(ISC(SC_LNL) - (k.V0 + ISEG(sid0, SEG_OFFN)) * k.iVt);  // n_0 E_0
    const double Ep0 = exp(ISC(SC_LPL) + (k.V0 + ISEG(sid0, SEG_OFFP)) * k.iVt);
    const double EnN = exp(ISC(SC_LNR) - (k.VN + ISEG(sidN, SEG_OFFN)) * k.iVt);
    const double EpN = exp(ISC(SC_LPR) + (k.VN + ISEG(sidN, SEG_OFFP)) * k.iVt);
    t.Kn = (-k.Vt * (EnN - En0) + Dn) / Sn;
    t.Kp = (k.Vt * (EpN - Ep0) - Dp) / Sp;
    if (write) {
        double *ds = a.dsc + (size_t)d * IS_N;
        const double qs = -k.elec * Q_E;
        // dK/dV_0 through e^(sn Vg_0/V_T) and the left contact's weight; dK/dV_(N+1)
        const double dKn0 = (-En0 + t.Kn * k.cn0 * k.iVt) / Sn, dKnN = EnN / Sn;
        const double dKp0 = (-Ep0 - t.Kp * k.cp0 * k.iVt) / Sp, dKpN = EpN / Sp;
        ds[IS_JV] = qs * ((dKn0 + dKp0) * k.half - (dKnN + dKpN) * k.half);
        const double eps0 = ISEG(sid0, SEG_EPS), epsN = ISEG(ISID(NP), SEG_EPS);
        ds[IS_JD] = -k.elec * (-k.half) * (epsN + eps0) / (k.L * H);
    }
    return t;
}

// (C) row r's coefficients of the linearised current
DDH_FN void imp_coefficients(const ImpArgs &a, int d, const ImpDev &k, const ImpTot &t, int r) {
    const int NP = a.NP, i = r + 1;
    const double sign = -k.elec, qs = sign * Q_E;
    const double cn = ITAB(IM_KV, r), cp = ITAB(IM_KD, r), iu = ITAB(IM_G0N, r);
    const double Pn = t.Sn - ITAB(IM_G0P, r), Pp = t.Sp - ITAB(IM_GLN, r);  // sum_(j>=i) c_j
    const double hsum = ITAB(IM_TMP, r);
    const double dRln = ITAB(IM_GLP, r), dRlp = ITAB(IM_GT, r);
    const double xj = IPT(i, PT_X), dxl = xj - IPT(i - 1, PT_X);
    const double dxr = i < NP ? IPT(i + 1, PT_X) - xj : 0.0;
    const double dKnV = -cn * (iu - t.Kn) * k.iVt / t.Sn;
    const double dKpV = -cp * (iu + t.Kp) * k.iVt / t.Sp;
    // dD/dU_i = dx_i P_(i+1)/2 + dx_(i-1) P_i/2
    const double dDn = 0.5 * dxr * (Pn - cn) + 0.5 * dxl * Pn;
    const double dDp = 0.5 * dxr * (Pp - cp) + 0.5 * dxl * Pp;
    const double hi = IPT(i, PT_H), hl = IPT(i - 1, PT_H);
    const double width = 0.5 * (hl + hi);
    const double T = (t.H - hsum - (a.iface[(size_t)i * a.n + d] ? hi : 0.0)) / t.H;
    const double direct = Q_E * k.L * width * T;
    const double wn = qs * (dDn / t.Sn) - sign * direct;  // dK_n/dU = +dD_n/S_n (sn = -1)
    const double wp = qs * (-dDp / t.Sp) + sign * direct;
    const double eps_l = ISEG(ISID(i - 1), SEG_EPS), eps_i = ISEG(ISID(i), SEG_EPS);
    ITAB(IM_KV, r) = qs * (dKnV + dKpV);
    ITAB(IM_KD, r) = sign * (eps_l - eps_i) / (k.L * t.H);
    ITAB(IM_G0N, r) = (wn + wp) * dRln;
    ITAB(IM_G0P, r) = (wn + wp) * dRlp;
    ITAB(IM_GLN, r) = wn * ITAB(IM_NN, r);
    ITAB(IM_GLP, r) = wp * ITAB(IM_PP, r);
    ITAB(IM_GT, r) = wn * ITAB(IM_DRN, r) + wp * ITAB(IM_DRP, r);
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
    const double *tab = a.tab + (size_t)d * NP * a.row;
    const double *ds = a.dsc + (size_t)d * IS_N;
    const cplx lap(a.laps[2 * f], a.laps[2 * f + 1]);
    const bool steps = a.dt > 0;
    const cplx ld = lap * a.dt;
    cplx Cp[9], dp[3], g[3];
    cplx acc(0.0, 0.0);
    // the row's fields in registers, the next row's loads in flight during this row's work
    double cur[IM_DRN], nxt[IM_DRN];
#pragma unroll
    for (int q = 0; q < IM_DRN; ++q) cur[q] = tab[q];
    for (int r = 0; r < NP; ++r) {
        if (r + 1 < NP) {
            const double *tn_ = tab + (size_t)(r + 1) * a.row;
#pragma unroll
            for (int q = 0; q < IM_DRN; ++q) nxt[q] = tn_[q];
        }
        auto T = [&](int f) { return cur[f]; };
        // A = [[a0,0,0],[a3,a4,0],[a6,0,a8]], C likewise
        const double a0 = T(IM_A), a3 = T(IM_A + 1), a4 = T(IM_A + 2), a6 = T(IM_A + 3),
                     a8 = T(IM_A + 4);
        const double c0 = T(IM_C), c3 = T(IM_C + 1), c4 = T(IM_C + 2), c6 = T(IM_C + 3),
                     c8 = T(IM_C + 4);
        cplx B[9];
        for (int q = 0; q < 9; ++q) B[q] = cplx(T(IM_B + q));
        // the occupation beyond its steady response, per unit of it (impedance.py:104-117)
        cplx h;
        if (steps) {
            const double E = T(IM_E), m = T(IM_M);
            h = -(m * ld) / (cplx(1.0 - E) + E * ld);
        } else {
            h = -lap / (lap + cplx(T(IM_RATE)));
        }
        const double un = T(IM_UN), up = T(IM_UP), tn = T(IM_TN), tp = T(IM_TP);
        B[4] = B[4] + lap * T(IM_MN) + h * (tn * un);
        B[5] = B[5] + h * (tn * up);
        B[7] = B[7] + h * (tp * un);
        B[8] = B[8] + lap * T(IM_MP) + h * (tp * up);
        cplx rhs[3];
        if (r > 0) {
            for (int k = 0; k < 3; ++k) {
                B[k] = B[k] - a0 * Cp[k];
                B[3 + k] = B[3 + k] - (a3 * Cp[k] + a4 * Cp[3 + k]);
                B[6 + k] = B[6 + k] - (a6 * Cp[k] + a8 * Cp[6 + k]);
            }
            rhs[0] = -(a0 * dp[0]);
            rhs[1] = -(a3 * dp[0] + a4 * dp[1]);
            rhs[2] = -(a6 * dp[0] + a8 * dp[2]);
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
        const double gt = T(IM_GT);
        cplx c[3];
        c[0] = cplx(T(IM_KV)) + lap * T(IM_KD);
        c[1] = cplx(T(IM_G0N)) + lap * T(IM_GLN) + h * (un * gt);
        c[2] = cplx(T(IM_G0P)) + lap * T(IM_GLP) + h * (up * gt);
        if (r > 0) {
            cplx gn[3];
            for (int k = 0; k < 3; ++k)
                gn[k] = c[k] - (g[0] * Cp[k] + g[1] * Cp[3 + k] + g[2] * Cp[6 + k]);
            for (int k = 0; k < 3; ++k) g[k] = gn[k];
        } else {
            for (int k = 0; k < 3; ++k) g[k] = c[k];
        }
        for (int i = 0; i < 3; ++i) {
            Cp[3 * i] = inv[3 * i] * c0 + inv[3 * i + 1] * c3 + inv[3 * i + 2] * c6;
            Cp[3 * i + 1] = inv[3 * i + 1] * c4;
            Cp[3 * i + 2] = inv[3 * i + 2] * c8;
        }
        acc = acc + g[0] * dp[0] + g[1] * dp[1] + g[2] * dp[2];
#pragma unroll
        for (int q = 0; q < IM_DRN; ++q) cur[q] = nxt[q];
    }
    const cplx Y = cplx(ds[IS_JV]) + lap * ds[IS_JD] + acc;
    Yo[0] = Y.re;
    Yo[1] = Y.im;
}

// ---- mobile ions: the ions' logarithms as unknowns of their own (impedance.py:242-262) ------
//
// The 5 x 5 system of `_assemble_ions` (solver.py:1364-1473) at a step of dti = 1 from the
// steady state itself, with dti's time terms taken back out (`shift`, impedance.py:261): the
// carriers' 3 x 3 rows as above, Poisson's row with the ions' densities as unknowns, and each
// ion's row inside its region Scharfetter-Gummel's flux with implicit Euler's time term, scaled
// by 1/(a_prev (fac + h_(i-1) V_T mu_i + h_i V_T mu_(i-1))). Each species' number is kept by a
// slot: the column U = -M of its rows and its trapezoidal number as a constraint W, closed by
// the 2 x 2 Schur complement of the bordered system (impedance.py:301-308).
//
// The ions' current (batch.py:421-485 `_ion_currents`, one region [f, l] per species): J_e =
// q K + sum_(f<k<=e) sn q L w_k U_k on the region's intervals, K = (V_T sn (G_l - G_f) - sn D)/S
// with G = dens e^(sn V/V_T), S = sum c_j, D = sum c_j iu_j, U = dti (dens - dens_prev). At the
// steady state U = 0 and G is constant over the region, so its linearisation is
// dK = (V_T sn (dG_l - dG_f) - sn dD - K dS)/S, dG = G (dln + sn dV/V_T), dS = sum c_j sn dV_j/V_T,
// dD = sum_k (dx_k P_(k+1) + dx_(k-1) P_k)/2 dU_k (P_k = sum_(j>=k) c_j over the region),
// dU = dti dens dln, and the device average takes K over the region's share of the intervals.

// Each species' normalisation in closed form from the steady potential (`_ion_logs`
// solver.py:1695-1712, trapezoidal number :1241-1252) and its region's first and last point.
struct ImpIon {
    double la[2];
    int first[2], last[2];
    bool ok;  // one region per species, reaching no contact
};

DDH_FN double imp_V(const ImpArgs &a, int d, const ImpDev &k, int j) {
    return j == 0 ? k.V0 : (j == a.NP + 1 ? k.VN : IY(j, 0));
}

DDH_FN double imp_mask(const ImpArgs &a, int d, int sid, int q) {
    return ISEG(sid, q == 0 ? SEG_AM : SEG_CM);
}

DDH_FN ImpIon imp_ion_setup(const ImpArgs &a, int d, const ImpDev &k) {
    ImpIon s;
    const double tot[2] = {ISC(SC_AT), ISC(SC_CT)};
    double num[2] = {0.0, 0.0}, dprev[2] = {0.0, 0.0}, mprev[2] = {0.0, 0.0};
    int count[2] = {0, 0};
    s.first[0] = s.first[1] = s.last[0] = s.last[1] = -1;
    double hprev = 0.0;
    for (int j = 0; j <= a.NP + 1; ++j) {
        const double V = imp_V(a, d, k, j);
        const int sid = ISID(j);
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const double m = imp_mask(a, d, sid, q);
            const double dq = m * exp((q == 0 ? 1.0 : -1.0) * V * k.iVt);
            if (j >= 1) num[q] += 0.5 * (dq + dprev[q]) * (m * mprev[q]) * hprev * k.L;
            dprev[q] = dq;
            mprev[q] = m;
            if (m > 0) {
                if (s.first[q] < 0) s.first[q] = j;
                s.last[q] = j;
                ++count[q];
            }
        }
        hprev = IPT(j, PT_H);
    }
    s.ok = true;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        if (tot[q] > 0) {
            s.la[q] = log(tot[q] / num[q]);
            if (s.first[q] < 1 || s.last[q] > a.NP || count[q] != s.last[q] - s.first[q] + 1)
                s.ok = false;
        } else {
            s.la[q] = 0.0;
            s.first[q] = s.last[q] = -1;
        }
    }
    return s;
}

// ln of species q's density at point j (0 outside its region)
DDH_FN double imp_ion_log(const ImpArgs &a, int d, const ImpDev &k, const ImpIon &s, int j, int q) {
    if (s.first[q] < 0 || j < s.first[q] || j > s.last[q]) return 0.0;
    return s.la[q] + (q == 0 ? 1.0 : -1.0) * imp_V(a, d, k, j) * k.iVt;
}

// Row r's ion fields (after imp_assemble_row): each species' row, its time term and constraint
// row, Poisson's derivatives in the ions' logarithms; the current's ion coefficients zeroed.
DDH_FN void imp_ion_row(const ImpArgs &a, int d, const ImpDev &k, const ImpIon &s, int r) {
    const int NP = a.NP, i = r + 1;
    const int sl = ISID(i - 1), sm = ISID(i), sr = ISID(i + 1);
    const double Vl = imp_V(a, d, k, i - 1), Vc = IY(i, 0), Vr = imp_V(a, d, k, i + 1);
    const double hl = IPT(i - 1, PT_H), hi = IPT(i, PT_H);
    const double fac = 0.5 * k.L2 * hi * hl * (hi + hl);
    const double sP = ISEG(sm, SEG_SP);
    const double tot[2] = {ISC(SC_AT), ISC(SC_CT)};
    const double xl = (Vc - Vl) * k.iVt, xr = (Vr - Vc) * k.iVt;
    double Blf, Blb, dBlf, dBlb, Brf, Brb, dBrf, dBrb;
    bernoulli_pair(xl, a.bern, Blf, Blb, dBlf, dBlb);
    bernoulli_pair(xr, a.bern, Brf, Brb, dBrf, dBrb);
    double dens[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int base = q == 0 ? IM_IA : IM_IC;
        const double mc = imp_mask(a, d, sm, q);
        const double lnc = imp_ion_log(a, d, k, s, i, q);
        dens[q] = mc > 0 ? exp(lnc) : 0.0;
        if (!(mc > 0) || s.first[q] < 0) {  // held: the row is the unknown itself
            ITAB(base + 0, r) = 0.0; ITAB(base + 1, r) = 0.0; ITAB(base + 2, r) = 0.0;
            ITAB(base + 3, r) = 1.0; ITAB(base + 4, r) = 0.0; ITAB(base + 5, r) = 0.0;
            ITAB(q == 0 ? IM_MA : IM_MC, r) = 0.0;
            ITAB(q == 0 ? IM_WA : IM_WC, r) = 0.0;
            dens[q] = 0.0;
            continue;
        }
        const double ml = imp_mask(a, d, sl, q), mr = imp_mask(a, d, sr, q);
        const double el = r == 0 ? 0.0 : ml * mc, er = r == NP - 1 ? 0.0 : mc * mr;
        const double cl = k.Vt * IMU(sl, q), cr = k.Vt * IMU(sm, q);
        const double nl = exp(imp_ion_log(a, d, k, s, i - 1, q)), nc = dens[q];
        const double nr = exp(imp_ion_log(a, d, k, s, i + 1, q));
        const double scale = 1.0 / (nc * (fac + hl * cr + hi * cl));
        double dVl, dVc, dVr, dll, dlc, dlr;
        if (q == 0) {  // anions as electrons: er h_(i-1) F_r - el h_i F_l
            const double dFl = cl * (dBlf * nc + dBlb * nl), dFr = cr * (dBrf * nr + dBrb * nc);
            dVl = el * hi * dFl * k.iVt;
            dVc = -er * hl * dFr * k.iVt - el * hi * dFl * k.iVt;
            dVr = er * hl * dFr * k.iVt;
            dll = el * hi * cl * Blb * nl;
            dlc = -er * hl * cr * Brb * nc - el * hi * cl * Blf * nc;
            dlr = er * hl * cr * Brf * nr;
        } else {  // cations as holes: el h_i F_l - er h_(i-1) F_r
            const double dFl = cl * (dBlf * nl + dBlb * nc), dFr = cr * (dBrf * nc + dBrb * nr);
            dVl = -el * hi * dFl * k.iVt;
            dVc = el * hi * dFl * k.iVt + er * hl * dFr * k.iVt;
            dVr = -er * hl * dFr * k.iVt;
            dll = el * hi * cl * Blf * nl;
            dlc = -el * hi * cl * Blb * nc - er * hl * cr * Brf * nc;
            dlr = er * hl * cr * Brb * nr;
        }
        ITAB(base + 0, r) = dVl * scale; ITAB(base + 1, r) = dll * scale;
        ITAB(base + 2, r) = dVc * scale; ITAB(base + 3, r) = dlc * scale;
        ITAB(base + 4, r) = dVr * scale; ITAB(base + 5, r) = dlr * scale;
        ITAB(q == 0 ? IM_MA : IM_MC, r) = -fac * nc * scale;
        const double w = ml * mc * hl * k.L + mc * mr * hi * k.L;
        ITAB(q == 0 ? IM_WA : IM_WC, r) = tot[q] > 0 ? 0.5 * w * mc * nc / tot[q] : 0.0;
    }
    ITAB(IM_B03, r) = -Q_E * dens[0] * sP;
    ITAB(IM_B04, r) = Q_E * dens[1] * sP;
    ITAB(IM_LA, r) = imp_ion_log(a, d, k, s, i, 0);
    ITAB(IM_LC, r) = imp_ion_log(a, d, k, s, i, 1);
    ITAB(IM_CA0, r) = 0.0; ITAB(IM_CAL, r) = 0.0;
    ITAB(IM_CC0, r) = 0.0; ITAB(IM_CCL, r) = 0.0;
}

// The ions' current coefficients (after imp_coefficients; one lane, the regions in turn): each
// point's dln coefficient (CA0/CC0 + lap CAL/CCL) and its dV coefficient added to IM_KV.
DDH_FN void imp_ion_current(const ImpArgs &a, int d, const ImpDev &k, const ImpIon &s, double H) {
    const double sign = -k.elec, qs = sign * Q_E;
    for (int q = 0; q < 2; ++q) {
        const int f = s.first[q], l = s.last[q];
        if (f < 0) continue;
        const double sn = q == 0 ? -1.0 : 1.0;
        const int c0 = q == 0 ? IM_CA0 : IM_CC0, cl_ = q == 0 ? IM_CAL : IM_CCL;
        // S, the region's share of the intervals, G at the ends
        double S = 0.0, Hreg = 0.0;
        for (int j = f; j <= l; ++j) {
            const double dxl = j > f ? IPT(j, PT_X) - IPT(j - 1, PT_X) : 0.0;
            const double dxr = j < l ? IPT(j + 1, PT_X) - IPT(j, PT_X) : 0.0;
            const double mu = IMU(ISID(j), q);
            S += 0.5 * (dxl + dxr) * exp(sn * IY(j, 0) * k.iVt) / (mu > 0 ? mu : 1.0);
            if (j < l) Hreg += IPT(j, PT_H);
        }
        const double Gf = exp(imp_ion_log(a, d, k, s, f, q) + sn * IY(f, 0) * k.iVt);
        const double Gl = exp(imp_ion_log(a, d, k, s, l, q) + sn * IY(l, 0) * k.iVt);
        const double K = k.Vt * sn * (Gl - Gf) / S;
        const double base = qs * (Hreg / H) / S;
        // from the last point back: P_j = sum_(i>=j) c_i, the suffix of h over the region
        double P = 0.0, hsuf = 0.0;
        for (int j = l; j >= f; --j) {
            const double xj = IPT(j, PT_X);
            const double dxl = j > f ? xj - IPT(j - 1, PT_X) : 0.0;
            const double dxr = j < l ? IPT(j + 1, PT_X) - xj : 0.0;
            const double mu = IMU(ISID(j), q);
            const double ev = exp(sn * IY(j, 0) * k.iVt);
            const double c = 0.5 * (dxl + dxr) * ev / (mu > 0 ? mu : 1.0);
            const double Pnext = P;  // sum over i > j
            P += c;
            const double hj = IPT(j, PT_H);
            if (j < l) hsuf += hj;
            const double ends = (j == l ? 1.0 : 0.0) - (j == f ? 1.0 : 0.0);
            const double dens = exp(imp_ion_log(a, d, k, s, j, q));
            const double G = dens * ev;
            const double dDdU = 0.5 * dxr * Pnext + 0.5 * dxl * P;
            double lapc = -base * sn * dDdU * dens;
            if (j > f && j < l) {
                const double w = 0.5 * (IPT(j - 1, PT_H) + hj);
                const double T = (hsuf - (a.iface[(size_t)j * a.n + d] ? hj : 0.0)) / H;
                lapc += qs * sn * k.L * w * T * dens;
            }
            const int r = j - 1;
            ITAB(c0, r) = base * k.Vt * sn * G * ends;
            ITAB(cl_, r) = lapc;
            ITAB(IM_KV, r) += base * (G * ends - K * c * sn * k.iVt);
        }
    }
}

// 5 x 5 inverses by their partition (solver.py:1505-1517 `_inverse5`): the carriers' 3 x 3
// block by its adjugate, the ions' 2 x 2 Schur complement by its own.
DDH_FN void inverse5c(const cplx *m, cplx *out) {
    cplx P[9], Pi[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) P[3 * i + j] = m[5 * i + j];
    inverse3<cplx>(P, Pi);
    cplx PiQ[6], RPi[6];  // 3 x 2, 2 x 3
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            PiQ[2 * i + j] = Pi[3 * i] * m[3 + j] + Pi[3 * i + 1] * m[8 + j] + Pi[3 * i + 2] * m[13 + j];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            RPi[3 * i + j] = m[5 * (3 + i)] * Pi[j] + m[5 * (3 + i) + 1] * Pi[3 + j] +
                             m[5 * (3 + i) + 2] * Pi[6 + j];
    cplx Sc[4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            Sc[2 * i + j] = m[5 * (3 + i) + 3 + j] -
                            (m[5 * (3 + i)] * PiQ[j] + m[5 * (3 + i) + 1] * PiQ[2 + j] +
                             m[5 * (3 + i) + 2] * PiQ[4 + j]);
    const cplx rd = cplx(1.0) / (Sc[0] * Sc[3] - Sc[1] * Sc[2]);
    const cplx Si[4] = {Sc[3] * rd, -Sc[1] * rd, -Sc[2] * rd, Sc[0] * rd};
    cplx SiR[6];  // Si RPi, 2 x 3
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) SiR[3 * i + j] = Si[2 * i] * RPi[j] + Si[2 * i + 1] * RPi[3 + j];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
            out[5 * i + j] = Pi[3 * i + j] + PiQ[2 * i] * SiR[j] + PiQ[2 * i + 1] * SiR[3 + j];
#pragma unroll
        for (int j = 0; j < 2; ++j)
            out[5 * i + 3 + j] = -(PiQ[2 * i] * Si[j] + PiQ[2 * i + 1] * Si[2 + j]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 3; ++j) out[5 * (3 + i) + j] = -SiR[3 * i + j];
#pragma unroll
        for (int j = 0; j < 2; ++j) out[5 * (3 + i) + 3 + j] = Si[2 * i + j];
    }
}

// One (device, frequency) with the ions as unknowns: the complex 5 x 5 system with three
// right-hand sides (the drive and the two slots' columns) eliminated row by row, three
// functionals (the current, the two constraints) contracted on the way (g_i = c_i -
// g_(i-1) C'_(i-1)), the border closed at the end by the slots' 2 x 2 Schur complement.
DDH_FN void impedance_ion_lane(const ImpArgs &a, int d, int f) {
    double *Yo = a.Y + ((size_t)d * a.nf + f) * 2;
    if (a.status[d] != 0) {
        Yo[0] = Yo[1] = NAN;
        return;
    }
    const int NP = a.NP;
    const double *tab = a.tab + (size_t)d * NP * a.row;
    const double *ds = a.dsc + (size_t)d * IS_N;
    const cplx lap(a.laps[2 * f], a.laps[2 * f + 1]);
    const bool steps = a.dt > 0;
    const cplx ld = lap * a.dt;
    const double held[2] = {a.sc[(size_t)SC_AT * a.n + d] > 0 ? 0.0 : 1.0,
                            a.sc[(size_t)SC_CT * a.n + d] > 0 ? 0.0 : 1.0};
    cplx Cp[25], dp[15], g[15], acc[9];  // dp [5][3]; g [3][5]; acc [3 functionals][3 columns]
#pragma unroll
    for (int q = 0; q < 9; ++q) acc[q] = cplx(0.0);
    for (int r = 0; r < NP; ++r) {
        const double *t = tab + (size_t)r * a.row;
        // the blocks' sparse A and C: column 0 of every row, the diagonal of rows 1-4
        const double A0[5] = {t[IM_A], t[IM_A + 1], t[IM_A + 3], t[IM_IA + 0], t[IM_IC + 0]};
        const double Ad[5] = {0.0, t[IM_A + 2], t[IM_A + 4], t[IM_IA + 1], t[IM_IC + 1]};
        const double C0[5] = {t[IM_C], t[IM_C + 1], t[IM_C + 3], t[IM_IA + 4], t[IM_IC + 4]};
        const double Cd[5] = {0.0, t[IM_C + 2], t[IM_C + 4], t[IM_IA + 5], t[IM_IC + 5]};
        cplx h;
        if (steps) {
            const double E = t[IM_E], m = t[IM_M];
            h = -(m * ld) / (cplx(1.0 - E) + E * ld);
        } else {
            h = -lap / (lap + cplx(t[IM_RATE]));
        }
        const double un = t[IM_UN], up = t[IM_UP], tn = t[IM_TN], tp = t[IM_TP];
        cplx B[25];
#pragma unroll
        for (int q = 0; q < 25; ++q) B[q] = cplx(0.0);
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) B[5 * i + j] = cplx(t[IM_B + 3 * i + j]);
        B[6] = B[6] + lap * t[IM_MN] + h * (tn * un);
        B[7] = B[7] + h * (tn * up);
        B[11] = B[11] + h * (tp * un);
        B[12] = B[12] + lap * t[IM_MP] + h * (tp * up);
        B[3] = cplx(t[IM_B03]);
        B[4] = cplx(t[IM_B04]);
        B[15] = cplx(t[IM_IA + 2]);
        B[18] = cplx(t[IM_IA + 3]) + lap * t[IM_MA];
        B[20] = cplx(t[IM_IC + 2]);
        B[24] = cplx(t[IM_IC + 3]) + lap * t[IM_MC];
        cplx rhs[15];
#pragma unroll
        for (int q = 0; q < 15; ++q) rhs[q] = cplx(0.0);
        if (r == 0)
            for (int i = 0; i < 3; ++i) rhs[3 * i] = cplx(ds[IS_D0 + i]);
        if (r == NP - 1)
            for (int i = 0; i < 3; ++i) rhs[3 * i] = rhs[3 * i] + cplx(ds[IS_DN + i]);
        rhs[3 * 3 + 1] = cplx(-t[IM_MA]);  // the slots' columns U = -M
        rhs[3 * 4 + 2] = cplx(-t[IM_MC]);
        // the current's and the constraints' coefficients of this point's dy
        const double gt = t[IM_GT];
        const cplx cc[5] = {cplx(t[IM_KV]) + lap * t[IM_KD],
                            cplx(t[IM_G0N]) + lap * t[IM_GLN] + h * (un * gt),
                            cplx(t[IM_G0P]) + lap * t[IM_GLP] + h * (up * gt),
                            cplx(t[IM_CA0]) + lap * t[IM_CAL], cplx(t[IM_CC0]) + lap * t[IM_CCL]};
        if (r > 0) {
#pragma unroll
            for (int i = 0; i < 5; ++i) {
#pragma unroll
                for (int j = 0; j < 5; ++j)
                    B[5 * i + j] = B[5 * i + j] - (A0[i] * Cp[j] + (i > 0 ? Ad[i] * Cp[5 * i + j] : cplx(0.0)));
#pragma unroll
                for (int c = 0; c < 3; ++c)
                    rhs[3 * i + c] = rhs[3 * i + c] - (A0[i] * dp[c] + (i > 0 ? Ad[i] * dp[3 * i + c] : cplx(0.0)));
            }
            cplx gn[15];
#pragma unroll
            for (int fn = 0; fn < 3; ++fn)
#pragma unroll
                for (int j = 0; j < 5; ++j) {
                    cplx v = fn == 0 ? cc[j] : cplx((fn == 1 && j == 3) ? t[IM_WA] : ((fn == 2 && j == 4) ? t[IM_WC] : 0.0));
#pragma unroll
                    for (int i = 0; i < 5; ++i) v = v - g[5 * fn + i] * Cp[5 * i + j];
                    gn[5 * fn + j] = v;
                }
#pragma unroll
            for (int q = 0; q < 15; ++q) g[q] = gn[q];
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) {
                g[j] = cc[j];
                g[5 + j] = cplx(j == 3 ? t[IM_WA] : 0.0);
                g[10 + j] = cplx(j == 4 ? t[IM_WC] : 0.0);
            }
        }
        cplx inv[25];
        inverse5c(B, inv);
#pragma unroll
        for (int i = 0; i < 5; ++i) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cplx v(0.0);
#pragma unroll
                for (int j = 0; j < 5; ++j) v = v + inv[5 * i + j] * rhs[3 * j + c];
                dp[3 * i + c] = v;
            }
            cplx v(0.0);
#pragma unroll
            for (int j = 0; j < 5; ++j) v = v + inv[5 * i + j] * C0[j];
            Cp[5 * i] = v;
#pragma unroll
            for (int j = 1; j < 5; ++j) Cp[5 * i + j] = inv[5 * i + j] * Cd[j];
        }
#pragma unroll
        for (int fn = 0; fn < 3; ++fn)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                cplx v = acc[3 * fn + c];
#pragma unroll
                for (int i = 0; i < 5; ++i) v = v + g[5 * fn + i] * dp[3 * i + c];
                acc[3 * fn + c] = v;
            }
    }
    // the border: S z = -W X_F with S = diag(held) - W X_U, then dy = X_F - X_U z
    const cplx s00 = cplx(held[0]) - acc[3 + 1], s01 = -acc[3 + 2];
    const cplx s10 = -acc[6 + 1], s11 = cplx(held[1]) - acc[6 + 2];
    const cplx b0 = -acc[3], b1 = -acc[6];
    const cplx det = s00 * s11 - s01 * s10;
    const cplx z0 = (b0 * s11 - s01 * b1) / det, z1 = (s00 * b1 - s10 * b0) / det;
    const cplx Y = cplx(ds[IS_JV]) + lap * ds[IS_JD] + acc[0] - (acc[1] * z0 + acc[2] * z1);
    Yo[0] = Y.re;
    Yo[1] = Y.im;
}

#undef IMU
#undef ISC
#undef IPT
#undef ISID
#undef ISEG
#undef IY

}  // namespace ddh
