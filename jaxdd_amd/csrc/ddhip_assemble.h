// The assembly operator: the scaled residual F, the Jacobian blocks A, B, C, the ion slots'
// columns U and the integral current J of every device at a given state
// (src/jaxdd/solver.py:1021-1187 `_assemble` with its ion columns, src/jaxdd/batch.py:518-549
// `_system`, and src/jaxdd/batch.py:302-403 `_current_parts`), one lane per device, the same
// row and edge arithmetic as the Newton sweeps (row_blocks_s, edge_flux_pt). Rows are the
// reference's: the first row's A and the last row's C couple to the contacts.
#pragma once

#include "ddhip_core.h"

namespace ddh {

struct AsmArgs {
    int n, NP, nseg, bern;
    const double *pt;
    const uint8_t *sid;
    const double *seg;
    const double *sc;
    const double *y;      // [NP][3][n] the interior state (V, ln n, ln p)
    const double *Vint;   // [n] the internal voltage
    const double *logs;   // [2][n] the ion slots' normalisations (NULL: no mobile ions)
    double *F;            // [NP][3][n]
    double *A, *B, *C;    // [NP][9][n] row-major 3x3
    double *U;            // [NP][6][n] [row][slot]
    double *J;            // [n] without the shunt
    // the 5x5 form (ions as unknowns): y5 [NP][5][n] (V, ln n, ln p, ln a, ln c) in place of y;
    // outputs F5 [NP][5][n], A5, B5, C5 [NP][25][n] row-major, U5 [NP][10][n] [row][slot]
    const double *y5;
    double *F5, *A5, *B5, *C5, *U5;
};

#define KSC(f) a.sc[(size_t)(f) * a.n + d]
#define KPT(j, f) a.pt[((size_t)(j) * PT_N + (f)) * a.n + d]
#define KSID(j) ((int)a.sid[(size_t)(j) * a.n + d])
#define KSEG(id, f) a.seg[((size_t)(id) * SEG_N + (f)) * a.n + d]
#define KY(j, f) (FIVE ? a.y5[(((size_t)(j) - 1) * 5 + (f)) * a.n + d] \
                      : a.y[(((size_t)(j) - 1) * 3 + (f)) * a.n + d])

// FIVE: the 5x5 steady form of src/jaxdd/solver.py:1364-1473 (`_assemble_ions` at dti = 0): the
// carriers' rows are `_assemble`'s with each point's ion logarithms ln a - V/V_T, ln c + V/V_T
// (its ion densities the unknowns' own), Poisson's row without the ions' V dependence and with
// their columns, each ion's row Boltzmann's in its region (ln a - la - V/V_T; ln c - lc + V/V_T)
// and its unknown held outside it; the slots enter only the ions' rows.
template <bool FIVE>
DDH_FN void assemble_lane(const AsmArgs &a, int d) {
    const int NP = a.NP;
    const size_t n = (size_t)a.n;
    const double Vt = KSC(SC_VT), iVt = 1.0 / Vt, L = KSC(SC_L), L2 = L * L;
    const double gfr = KSC(SC_GFRAC), elec = KSC(SC_ELEC), Vint = a.Vint[d];
    const double V0 = KSC(SC_V0) + 0.5 * elec * Vint, VN = KSC(SC_VL) - 0.5 * elec * Vint;
    const double la = a.logs ? a.logs[d] : 0.0, lc = a.logs ? a.logs[n + d] : 0.0;
    const bool ions = FIVE || a.logs != nullptr;
    // a point's ion logarithms: the slots' in the 3x3 form; in the 5x5 form its unknowns' (the
    // contacts take their neighbours', which no row reads)
    auto point_logs = [&](int j, double V, double &lj_a, double &lj_c) {
        if (FIVE) {
            const int k = j < 1 ? 1 : (j > NP ? NP : j);
            const double Vk = k == j ? V : KY(k, 0);
            lj_a = KY(k, 3) - Vk / Vt;
            lj_c = KY(k, 4) + Vk / Vt;
        } else {
            lj_a = la;
            lj_c = lc;
        }
    };
    auto point = [&](int j, Pt &q) {
        q.sid = KSID(j);
        q.offn = KSEG(q.sid, SEG_OFFN); q.offp = KSEG(q.sid, SEG_OFFP);
        q.mun = KSEG(q.sid, SEG_MUN); q.mup = KSEG(q.sid, SEG_MUP);
        q.am = ions ? KSEG(q.sid, SEG_AM) : 0.0;
        q.cm = ions ? KSEG(q.sid, SEG_CM) : 0.0;
        q.h = KPT(j, PT_H);
        if (j == 0 || j == NP + 1) {
            q.V = j == 0 ? V0 : VN;
            q.n = exp(j == 0 ? KSC(SC_LNL) : KSC(SC_LNR));
            q.p = exp(j == 0 ? KSC(SC_LPL) : KSC(SC_LPR));
            q.G = q.fix = q.R = q.dRn = q.dRp = 0.0;
        } else {
            q.V = KY(j, 0);
            q.n = exp(KY(j, 1));
            q.p = exp(KY(j, 2));
            q.G = KPT(j, PT_G);
            q.fix = KPT(j, PT_FIX);
            q.R = recombination(q.n, q.p, KSEG(q.sid, SEG_KD), KSEG(q.sid, SEG_NI2),
                                KSEG(q.sid, SEG_NT), KSEG(q.sid, SEG_CN), KSEG(q.sid, SEG_CP),
                                KSEG(q.sid, SEG_NT0), KSEG(q.sid, SEG_PT0), q.dRn, q.dRp);
        }
        double lj_a, lj_c;
        point_logs(j, q.V, lj_a, lj_c);
        q.an = q.am != 0.0 ? q.am * exp(lj_a + q.V * iVt) : 0.0;
        q.ca = q.cm != 0.0 ? q.cm * exp(lj_c - q.V * iVt) : 0.0;
    };
    // rows left to right; the current's sums on the way (the Newton sweeps' order)
    Pt cur, mid, rgt;
    point(0, cur);
    point(1, mid);
    Edge eL = edge_flux_pt(cur, mid, Vt, iVt, a.bern);
    const double Vg0n = V0 + cur.offn, Vg0p = V0 + cur.offp;
    double wn_prev = exp(-Vg0n * iVt) * KSEG(cur.sid, SEG_IMUN);
    double wp_prev = exp(Vg0p * iVt) * KSEG(cur.sid, SEG_IMUP);
    const double endn0 = exp(KSC(SC_LNL) - Vg0n * iVt), endp0 = exp(KSC(SC_LPL) + Vg0p * iVt);
    double U_prev = 0.0, iu_prev = 0.0, x_prev = KPT(0, PT_X);
    double Sn = 0.0, Dn = 0.0, Sp = 0.0, Dp = 0.0;
    for (int i = 1; i <= NP; ++i) {
        point(i + 1, rgt);
        const Edge eR = edge_flux_pt(mid, rgt, Vt, iVt, a.bern);
        RowSeg g;
        g.meps = KSEG(mid.sid, SEG_EPS);
        g.ceps = KSEG(cur.sid, SEG_EPS);
        g.sP = KSEG(mid.sid, SEG_SP); g.scs = KSEG(mid.sid, SEG_SCS); g.dop = KSEG(mid.sid, SEG_DOP);
        const double tt = KSEG(mid.sid, SEG_TT), Nt = KSEG(mid.sid, SEG_NT);
        if (tt != 0.0 && Nt > 0.0)  // charged bulk traps (solver.py:1136-1140)
            g.qt = trap_charge(tt, mid.n, mid.p, Nt, KSEG(mid.sid, SEG_CN), KSEG(mid.sid, SEG_CP),
                               KSEG(mid.sid, SEG_NT0), KSEG(mid.sid, SEG_PT0), g.dqn, g.dqp);
        double Ab[9], Bb[9], Cb[9], r[9];
        row_blocks_s<3>(g, cur, mid, rgt.V, eL, eR, false, L2, iVt, 1.0, gfr, Ab, Bb, Cb, r);
        const size_t o = (size_t)(i - 1);
        if (FIVE) {
            double A5[25], B5[25], C5[25], F5[5], U5[10];
#pragma unroll
            for (int q = 0; q < 25; ++q) A5[q] = B5[q] = C5[q] = 0.0;
#pragma unroll
            for (int q = 0; q < 10; ++q) U5[q] = 0.0;
#pragma unroll
            for (int row = 0; row < 3; ++row) {
#pragma unroll
                for (int col = 0; col < 3; ++col) {
                    A5[5 * row + col] = Ab[3 * row + col];
                    B5[5 * row + col] = Bb[3 * row + col];
                    C5[5 * row + col] = Cb[3 * row + col];
                }
                F5[row] = -r[row * 3];
            }
            // Poisson's: the ions are unknowns of their own, not Boltzmann's in V
            const double ya = KY(i, 3), yc = KY(i, 4), Vm = mid.V;
            const double a_m = mid.am * exp(ya), c_m = mid.cm * exp(yc);
            B5[0] += Q_E * (a_m + c_m) / Vt * g.sP;
            B5[3] = r[1];
            B5[4] = r[2];
            // the ions' rows: Boltzmann's in their regions, the unknown held outside
            const bool in_a = mid.am > 0, in_c = mid.cm > 0;
            F5[3] = in_a ? (ya - la) + (-1.0) * Vm / Vt : ya;
            F5[4] = in_c ? (yc - lc) + 1.0 * Vm / Vt : yc;
            B5[15] = in_a ? -1.0 / Vt : 0.0;
            B5[18] = 1.0;
            B5[20] = in_c ? 1.0 / Vt : 0.0;
            B5[24] = 1.0;
            U5[6] = in_a ? -1.0 : 0.0;
            U5[9] = in_c ? -1.0 : 0.0;
#pragma unroll
            for (int q = 0; q < 25; ++q) {
                a.A5[(o * 25 + q) * n + d] = A5[q];
                a.B5[(o * 25 + q) * n + d] = B5[q];
                a.C5[(o * 25 + q) * n + d] = C5[q];
            }
#pragma unroll
            for (int q = 0; q < 5; ++q) a.F5[(o * 5 + q) * n + d] = F5[q];
#pragma unroll
            for (int q = 0; q < 10; ++q) a.U5[(o * 10 + q) * n + d] = U5[q];
            cur = mid; mid = rgt; eL = eR;
            continue;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            a.A[(o * 9 + q) * n + d] = Ab[q];
            a.B[(o * 9 + q) * n + d] = Bb[q];
            a.C[(o * 9 + q) * n + d] = Cb[q];
        }
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            a.F[(o * 3 + row) * n + d] = -r[row * 3];
            a.U[(o * 6 + 2 * row) * n + d] = r[row * 3 + 1];
            a.U[(o * 6 + 2 * row + 1) * n + d] = r[row * 3 + 2];
        }
        // the current at point i (src/jaxdd/batch.py:321-373)
        const double R = recombination_value(mid.n, mid.p, KSEG(mid.sid, SEG_KD),
                                             KSEG(mid.sid, SEG_NI2), KSEG(mid.sid, SEG_NT),
                                             KSEG(mid.sid, SEG_CN), KSEG(mid.sid, SEG_CP),
                                             KSEG(mid.sid, SEG_NT0), KSEG(mid.sid, SEG_PT0));
        const double Uu = R - mid.G * gfr;
        const double wn = exp(-(mid.V + mid.offn) * iVt) * KSEG(mid.sid, SEG_IMUN);
        const double wp = exp((mid.V + mid.offp) * iVt) * KSEG(mid.sid, SEG_IMUP);
        const double xj = KPT(i, PT_X), dx = xj - x_prev;
        const double iu = iu_prev + 0.5 * dx * (U_prev + Uu);
        Sn += 0.5 * dx * (wn_prev + wn);
        Dn += 0.5 * dx * (wn_prev * iu_prev + wn * iu);
        Sp += 0.5 * dx * (wp_prev + wp);
        Dp += 0.5 * dx * (wp_prev * iu_prev + wp * iu);
        wn_prev = wn; wp_prev = wp; U_prev = Uu; iu_prev = iu; x_prev = xj;
        cur = mid; mid = rgt; eL = eR;
    }
    if (FIVE) return;  // the 5x5 form has no current of its own
    const int sidN = KSID(NP + 1);
    const double endnN = exp(KSC(SC_LNR) - (VN + KSEG(sidN, SEG_OFFN)) * iVt);
    const double endpN = exp(KSC(SC_LPR) + (VN + KSEG(sidN, SEG_OFFP)) * iVt);
    const double Kn = (Vt * (-endnN + endn0) + Dn) / Sn;
    const double Kp = (Vt * (endpN - endp0) - Dp) / Sp;
    a.J[d] = -elec * Q_E * (Kn + Kp);
}

#undef KSC
#undef KPT
#undef KSID
#undef KSEG
#undef KY

}  // namespace ddh
