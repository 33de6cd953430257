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
};

#define KSC(f) a.sc[(size_t)(f) * a.n + d]
#define KPT(j, f) a.pt[((size_t)(j) * PT_N + (f)) * a.n + d]
#define KSID(j) ((int)a.sid[(size_t)(j) * a.n + d])
#define KSEG(id, f) a.seg[((size_t)(id) * SEG_N + (f)) * a.n + d]
#define KY(j, f) a.y[(((size_t)(j) - 1) * 3 + (f)) * a.n + d]

DDH_FN void assemble_lane(const AsmArgs &a, int d) {
    const int NP = a.NP;
    const size_t n = (size_t)a.n;
    const double Vt = KSC(SC_VT), iVt = 1.0 / Vt, L = KSC(SC_L), L2 = L * L;
    const double gfr = KSC(SC_GFRAC), elec = KSC(SC_ELEC), Vint = a.Vint[d];
    const double V0 = KSC(SC_V0) + 0.5 * elec * Vint, VN = KSC(SC_VL) - 0.5 * elec * Vint;
    const double la = a.logs ? a.logs[d] : 0.0, lc = a.logs ? a.logs[n + d] : 0.0;
    const bool ions = a.logs != nullptr;
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
        q.an = q.am != 0.0 ? q.am * exp(la + q.V * iVt) : 0.0;
        q.ca = q.cm != 0.0 ? q.cm * exp(lc - q.V * iVt) : 0.0;
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
        double Ab[9], Bb[9], Cb[9], r[9];
        row_blocks_s<3>(g, cur, mid, rgt.V, eL, eR, false, L2, iVt, 1.0, gfr, Ab, Bb, Cb, r);
        const size_t o = (size_t)(i - 1);
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
