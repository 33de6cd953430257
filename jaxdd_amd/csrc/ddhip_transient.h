// Per-lane logic of the batched transient (implicit Euler in time, ZimT's steps) shared by the
// HIP kernel (ddhip_solve.hip `transient_kernel`) and the test-only CPU replay
// (tests/kernel_emu). Reference: src/jaxdd/transient.py:270-464 (`_solve_one`), with
// src/jaxdd/solver.py:573-620 (`_time_terms`, `_trap_step`), :680-700 (`_bulk_rates` over a
// step), :1315-1473 (`_ion_full`, `_ion_constraints_time`, `_assemble_ions`) and
// src/jaxdd/batch.py:302-483 (the transient current, `_ion_currents`).
//
// One lane runs one device through all of its steps. Its state machine takes one Newton step
// per call; when a step's solve ends, the lane records it and begins its next step at once,
// so no lane waits for another device's iterations. A Newton step is three passes over the
// grid:
//   forward (left to right): each row's residual and its blocks at the iterate, the mobile
//     ions as unknowns of their own (NC = 5), eliminated as they are made (M = B - A c',
//     c' = M^-1 C, d' = M^-1 (r - A d') for the right-hand sides r = -F and the two ion slots'
//     columns, by Gauss-Jordan on the augmented block), the factors stored; the ion slots'
//     constraints close in the same pass through the recursion gamma_i = W_i - c'_{i-1}^T
//     gamma_{i-1}, which gives W^T X = sum gamma_i^T d'_i without the solution X;
//   back-substitution (right to left): x_i = d'_i - c'_i x_{i+1} with the slots' step z, the
//     step caps and density steps, the new iterate and its step size;
//   current (left to right): the transient current of the new iterate (the densities' change
//     in the integral form's integrand, the ions' currents of their region, the displacement
//     current), each part's integration constant applied at the end, as the prefix sums are
//     linear in it.
// The first pass of a step reads its first guess (the last two kept states extrapolated in
// time, held to the caps) point by point and stores it, and advances the traps' occupation to
// the end of the step kept before.
#pragma once

#include <stdint.h>

#ifdef DDH_CPU_EMULATION
#include <cmath>
#define DDH_TFN inline
#else
#include <hip/hip_runtime.h>
#define DDH_TFN __device__ inline
#endif

#include "ddhip_layout.h"
#include "ddhip_math.h"

namespace ddh {

// Factors per grid point: c' (5 x 5) and d' (5 x 3 right-hand sides).
constexpr int TR_FAC = 40;
constexpr int TR_NC = 5;  // unknowns per point in the state buffers (3 used without ions)

struct TrArgs {
    int n, NP, nseg, nt;
    const double *pt;      // [NP + 2][PT_N][n]
    const uint8_t *sid;    // [NP + 2][n]
    const uint8_t *iface;  // [NP + 1][n] 1 where the interval after a point crosses a layer interface
    const double *seg;     // [nseg][SEG_N][n]
    const double *sc;      // [SC_N][n]
    const double *imu;     // [nseg][2][n] the mobile ions' mobility per segment
    const double *t, *V, *G;  // [n][nt] each step's time, internal voltage, generation fraction
    const double *y0;      // [NP][3][n] the steady state at each device's first voltage
    double *Y;             // [3][NP][TR_NC][n] state buffers (iterate, kept, kept before)
    double *fb;            // [NP][n] the bulk traps' occupation
    double *fac;           // [NP][TR_FAC][n]
    double *J, *Jext;      // [n][nt]
    uint8_t *ok, *kept;    // [n][nt]
    int32_t *its;          // [n][nt]
    int32_t *status;       // [n]
    double tol;
    int max_iter, bern, zimt, failure_mode;
};

struct TrLane {
    int k, it, first, stopped, fresh, pending, done;
    int bG, bY, bP;                      // state buffers: iterate, kept, kept before
    double dti, dti_now, dti_pend, t_last, V_now, Vk, gk, rt;
    double logs[2], logs_prev[2], lg[2];
    double hist[4], Jh[4];
    int finite;
};

// Forward-mode numbers with N partial derivatives.
template <int N>
struct Dual {
    double v, d[N];
};

template <int N>
DDH_TFN Dual<N> dconst(double v) {
    Dual<N> r; r.v = v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = 0.0;
    return r;
}
template <int N>
DDH_TFN Dual<N> dvar(double v, int i) { Dual<N> r = dconst<N>(v); r.d[i] = 1.0; return r; }
template <int N>
DDH_TFN Dual<N> operator+(Dual<N> a, const Dual<N> &b) {
    a.v += b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) a.d[k] += b.d[k];
    return a;
}
template <int N>
DDH_TFN Dual<N> operator-(Dual<N> a, const Dual<N> &b) {
    a.v -= b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) a.d[k] -= b.d[k];
    return a;
}
template <int N>
DDH_TFN Dual<N> operator*(Dual<N> a, double c) {
    a.v *= c;
#pragma unroll
    for (int k = 0; k < N; ++k) a.d[k] *= c;
    return a;
}
template <int N>
DDH_TFN Dual<N> operator+(Dual<N> a, double c) { a.v += c; return a; }
template <int N>
DDH_TFN Dual<N> operator*(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r; r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
DDH_TFN Dual<N> operator/(const Dual<N> &a, const Dual<N> &b) {
    Dual<N> r; const double ib = 1.0 / b.v; r.v = a.v * ib;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
DDH_TFN Dual<N> dapply(const Dual<N> &a, double f, double df) {
    Dual<N> r; r.v = f;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = df * a.d[k];
    return r;
}
template <int N>
DDH_TFN Dual<N> dexp(const Dual<N> &a) { const double e = exp(a.v); return dapply(a, e, e); }

// Each carrier's rate over a step (direct recombination, and capture by the bulk traps at the
// occupation's mean over the step, `_trap_step`'s mean and vacant share; the steady SRH rate at
// dti = 0), as duals in (ln n, ln p).
DDH_TFN void tr_rates(double kd, double ni2, double Nt, double Cn, double Cp, double nt0,
                      double pt0, const Dual<2> &n, const Dual<2> &p, double f_old, double dti,
                      Dual<2> &Rn, Dual<2> &Rp) {
    Dual<2> direct = (n * p + (-ni2)) * kd;
    Rn = Rp = direct;
    if (!(Nt > 0)) return;
    Dual<2> a = (n + nt0) * Cn + (p + pt0) * Cp;
    Dual<2> ia = dconst<2>(1.0) / a;
    Dual<2> steady = (n * Cn + Cp * pt0) * ia;
    Dual<2> vac = (p * Cp + Cn * nt0) * ia;
    Dual<2> fbar = steady, vbar = vac;
    if (dti > 0) {
        const Dual<2> x = a * (1.0 / dti);
        const double e = expm1(-x.v), m = -e / x.v;
        const double dm = (exp(-x.v) * x.v + e) / (x.v * x.v);
        const Dual<2> M = dapply(x, m, dm);
        fbar = steady + (dconst<2>(f_old) - steady) * M;
        vbar = vac + (dconst<2>(1.0 - f_old) - vac) * M;
    }
    Rn = direct + (n * vbar - fbar * nt0) * (Nt * Cn);
    Rp = direct + (p * fbar - vbar * pt0) * (Nt * Cp);
}

DDH_TFN double tr_rate_value(double kd, double ni2, double Nt, double Cn, double Cp, double nt0,
                             double pt0, double n, double p, double f_old, double dti, bool holes) {
    Dual<2> Rn, Rp;
    tr_rates(kd, ni2, Nt, Cn, Cp, nt0, pt0, dconst<2>(n), dconst<2>(p), f_old, dti, Rn, Rp);
    return holes ? Rp.v : Rn.v;
}

// The traps' occupation at the end of a step of 1/dti from f_old (`_bulk_occupation`).
DDH_TFN double tr_trap_end(double Nt, double Cn, double Cp, double nt0, double pt0, double n,
                           double p, double f_old, double dti) {
    if (!(Nt > 0)) return 0.0;
    const double a = Cn * (n + nt0) + Cp * (p + pt0);
    const double steady = (Cn * n + Cp * pt0) / a;
    return dti > 0 ? steady + (f_old - steady) * exp(-a / dti) : steady;
}

// Array accesses by (row, field, lane), each index within its array's extent (a CPU build
// with DDH_TRANSIENT_BOUNDS checks every one).
#ifdef DDH_TRANSIENT_BOUNDS
void tr_out_of_bounds(int line, size_t i, size_t lim);
#define TCHK(i, lim) \
    ((size_t)(i) < (size_t)(lim) ? (size_t)(i) : (tr_out_of_bounds(__LINE__, (size_t)(i), (size_t)(lim)), (size_t)0))
#else
#define TCHK(i, lim) ((size_t)(i))
#endif
#define TSC(f) a.sc[TCHK((size_t)(f) * a.n + lane, (size_t)SC_N * a.n)]
#define TPT(j, f) a.pt[TCHK(((size_t)(j) * PT_N + (f)) * a.n + lane, (size_t)(a.NP + 2) * PT_N * a.n)]
#define TSID(j) a.sid[TCHK((size_t)(j) * a.n + lane, (size_t)(a.NP + 2) * a.n)]
#define TY(b, j, c) a.Y[TCHK((((size_t)(b) * a.NP + ((j) - 1)) * TR_NC + (c)) * a.n + lane, (size_t)3 * a.NP * TR_NC * a.n)]
#define TFB(j) a.fb[TCHK((size_t)((j) - 1) * a.n + lane, (size_t)a.NP * a.n)]
#define TFAC(j, q) a.fac[TCHK(((size_t)((j) - 1) * TR_FAC + (q)) * a.n + lane, (size_t)a.NP * TR_FAC * a.n)]
#define TOUT(arr, k) arr[TCHK((size_t)lane * a.nt + (k), (size_t)a.n * a.nt)]
#define TIFACE(e) a.iface[TCHK((size_t)(e) * a.n + lane, (size_t)(a.NP + 1) * a.n)]

// The segment fields of a point (cached while the segment does not change).
struct TSeg {
    int id;
    double eps, offn, offp, Nc, dop, kd, ni2, Nt, Cn, Cp, nt0, pt0, am, cm, mun, mup, sP, sCs;
    double mua, muc;
};

DDH_TFN void tseg_at(const TrArgs &a, int lane, int j, TSeg &s) {
    const int id = TSID(j);
    if (id == s.id) return;
    s.id = id;
    (void)TCHK((size_t)id, (size_t)a.nseg);
    const double *g = a.seg + (size_t)id * SEG_N * a.n + lane;
    const size_t n = a.n;
    s.eps = g[SEG_EPS * n]; s.offn = g[SEG_OFFN * n]; s.offp = g[SEG_OFFP * n];
    s.Nc = g[SEG_NC * n]; s.dop = g[SEG_DOP * n]; s.kd = g[SEG_KD * n]; s.ni2 = g[SEG_NI2 * n];
    s.Nt = g[SEG_NT * n]; s.Cn = g[SEG_CN * n]; s.Cp = g[SEG_CP * n]; s.nt0 = g[SEG_NT0 * n];
    s.pt0 = g[SEG_PT0 * n]; s.am = g[SEG_AM * n]; s.cm = g[SEG_CM * n]; s.mun = g[SEG_MUN * n];
    s.mup = g[SEG_MUP * n]; s.sP = g[SEG_SP * n]; s.sCs = g[SEG_SCS * n];
    s.mua = a.imu[((size_t)id * 2 + 0) * n + lane];
    s.muc = a.imu[((size_t)id * 2 + 1) * n + lane];
}

// The mobile ions' masks (anions, cations) at point j.
DDH_TFN void tr_masks(const TrArgs &a, int lane, int j, double m[2]) {
    const size_t id = TSID(j), n = a.n;
    (void)TCHK(id, (size_t)a.nseg);
    m[0] = a.seg[(id * SEG_N + SEG_AM) * n + lane];
    m[1] = a.seg[(id * SEG_N + SEG_CM) * n + lane];
}

// The contacts' state at an internal voltage: V, ln n, ln p (the ions 0: no region reaches a
// contact on this path).
DDH_TFN void tr_contact(const TrArgs &a, int lane, double Vint, bool right, double *y) {
    const double e = TSC(SC_ELEC);
    y[0] = right ? TSC(SC_VL) - 0.5 * e * Vint : TSC(SC_V0) + 0.5 * e * Vint;
    y[1] = right ? TSC(SC_LNR) : TSC(SC_LNL);
    y[2] = right ? TSC(SC_LPR) : TSC(SC_LPL);
    y[3] = y[4] = 0.0;
}

// The iterate at interior point j: the stored one, or on a step's first pass its first guess
// from the kept states (written to the iterate's buffer by the caller).
template <int NC>
DDH_TFN void tr_iterate(const TrArgs &a, int lane, const TrLane &S, int j, double *y) {
    if (!S.fresh) {
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = TY(S.bG, j, c);
        return;
    }
    const double caps[5] = {CAP_V * TSC(SC_VT), CAP_N, CAP_N, CAP_N, CAP_N};
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const double yy = TY(S.bY, j, c), yp = TY(S.bP, j, c);
        y[c] = yy + clipd(S.rt * (yy - yp), -caps[c], caps[c]);
    }
}

template <int NC>
DDH_TFN void tr_point(const TrArgs &a, int lane, const TrLane &S, int j, double *y) {
    if (j == 0 || j == a.NP + 1) {
        double c[5];
        tr_contact(a, lane, S.Vk, j != 0, c);
#pragma unroll
        for (int q = 0; q < NC; ++q) y[q] = c[q];
        return;
    }
    tr_iterate<NC>(a, lane, S, j, y);
}

// Gauss-Jordan with partial pivoting on the augmented block [M | R] (5 x (5 + W)): R becomes
// M^-1 R.
template <int B, int W>
DDH_TFN void tr_solve_block(double (&M)[B][B], double (&R)[B][W]) {
#pragma unroll
    for (int c = 0; c < B; ++c) {
        int piv = c;
        double best = fabs(M[c][c]);
#pragma unroll
        for (int r = c + 1; r < B; ++r)
            if (fabs(M[r][c]) > best) { best = fabs(M[r][c]); piv = r; }
        if (piv != c) {
#pragma unroll
            for (int r = c + 1; r < B; ++r) {
                if (r != piv) continue;
#pragma unroll
                for (int q = 0; q < B; ++q) { const double t = M[c][q]; M[c][q] = M[r][q]; M[r][q] = t; }
#pragma unroll
                for (int q = 0; q < W; ++q) { const double t = R[c][q]; R[c][q] = R[r][q]; R[r][q] = t; }
            }
        }
        const double ip = 1.0 / M[c][c];
#pragma unroll
        for (int q = 0; q < B; ++q) M[c][q] *= ip;
#pragma unroll
        for (int q = 0; q < W; ++q) R[c][q] *= ip;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
#pragma unroll
            for (int q = 0; q < B; ++q) M[r][q] -= f * M[c][q];
#pragma unroll
            for (int q = 0; q < W; ++q) R[r][q] -= f * R[c][q];
        }
    }
}

// The carriers' fluxes over the interval from l to r and their derivatives in (V_l, V_r,
// ln n_l, ln n_r) / (V_l, V_r, ln p_l, ln p_r) (src/jaxdd/solver.py:1069-1083).
struct TEdge {
    double Fn, dn[4], Fp, dp[4];
};

DDH_TFN TEdge tr_edge(const double *yl, const double *yr, double offn_l, double offn_r,
                      double offp_l, double offp_r, double mun, double mup, double Vt, int bern) {
    TEdge e;
    const double iVt = 1.0 / Vt, cn = Vt * mun, cp = Vt * mup;
    const double nl = exp(yl[1]), nr = exp(yr[1]), pl = exp(yl[2]), pr = exp(yr[2]);
    double Bf, Bb, dBf, dBb;
    bernoulli_pair(((yr[0] + offn_r) - (yl[0] + offn_l)) * iVt, bern, Bf, Bb, dBf, dBb);
    e.Fn = cn * (Bf * nr - Bb * nl);
    const double gn = cn * (dBf * nr + dBb * nl) * iVt;
    e.dn[0] = -gn; e.dn[1] = gn; e.dn[2] = -cn * Bb * nl; e.dn[3] = cn * Bf * nr;
    bernoulli_pair(((yr[0] + offp_r) - (yl[0] + offp_l)) * iVt, bern, Bf, Bb, dBf, dBb);
    e.Fp = cp * (Bf * pl - Bb * pr);
    const double gp = cp * (dBf * pl + dBb * pr) * iVt;
    e.dp[0] = -gp; e.dp[1] = gp; e.dp[2] = cp * Bf * pl; e.dp[3] = -cp * Bb * pr;
    return e;
}

// One ion species' row at an interior point (`_assemble_ions`' `species`): value and its
// derivatives in (V_l, V_c, V_r, ln_l, ln_c, ln_r, the slot's log).
DDH_TFN Dual<7> tr_ion_row(double sn, double m_l, double m_c, double m_r, bool first_row,
                           bool last_row, double mu_l, double mu_c, double hm, double hp,
                           double fac, double dti, double before, const double *yl,
                           const double *yc, const double *yr, int col, double ls, double Vt,
                           int bern) {
    const Dual<7> Vl = dvar<7>(yl[0], 0), Vc = dvar<7>(yc[0], 1), Vr = dvar<7>(yr[0], 2);
    const Dual<7> ll = dvar<7>(yl[col], 3), lc = dvar<7>(yc[col], 4), lr = dvar<7>(yr[col], 5);
    const Dual<7> lsd = dvar<7>(ls, 6);
    if (!(m_c > 0)) return lc;  // outside the region: the unknown held
    if (!(dti > 0)) return lc - lsd + Vc * (sn / Vt);  // steady: Boltzmann's
    const double el = first_row ? 0.0 : m_l * m_c, er = last_row ? 0.0 : m_c * m_r;
    const double cl = Vt * mu_l, cr = Vt * mu_c;
    const Dual<7> xl = (Vc - Vl) * (1.0 / Vt), xr = (Vr - Vc) * (1.0 / Vt);
    double Bf, Bb, dBf, dBb;
    bernoulli_pair(xl.v, bern, Bf, Bb, dBf, dBb);
    const Dual<7> Blf = dapply(xl, Bf, dBf), Blb = dapply(xl, Bb, -dBb);
    bernoulli_pair(xr.v, bern, Bf, Bb, dBf, dBb);
    const Dual<7> Brf = dapply(xr, Bf, dBf), Brb = dapply(xr, Bb, -dBb);
    const Dual<7> nl = dexp(ll), nc = dexp(lc), nr = dexp(lr);
    Dual<7> flux;
    if (sn < 0) {  // anions, as electrons
        const Dual<7> Fl = (Blf * nc - Blb * nl) * cl, Fr = (Brf * nr - Brb * nc) * cr;
        flux = Fr * (er * hm) - Fl * (el * hp);
    } else {  // cations, as holes
        const Dual<7> Fl = (Blf * nl - Blb * nc) * cl, Fr = (Brf * nc - Brb * nr) * cr;
        flux = Fl * (el * hp) - Fr * (er * hm);
    }
    const double prev = before > 0 ? before : 1.0;
    const double scale = 1.0 / (prev * (fac * dti + hm * cr + hp * cl));
    const Dual<7> stored = (dexp(lsd) * before - nc) * (fac * dti);
    return (flux + stored) * scale;
}

// The forward pass: assembly, elimination, the slots' Schur complement through gamma. Returns
// the pass's g (the slots' constraint values) and sums of gamma^T d' in gd[2][3].
template <int NC>
DDH_TFN void tr_forward(const TrArgs &a, int lane, TrLane &S, double g[2], double gd[2][3]) {
    const int NP = a.NP;
    const int K = NC == 5 ? 3 : 1;
    const double Vt = TSC(SC_VT), L = TSC(SC_L), L2 = L * L;
    double yl[5], yc[5], yr[5];
    tr_point<NC>(a, lane, S, 0, yl);
    tr_point<NC>(a, lane, S, 1, yc);
    TSeg sl, sg, sr;
    sl.id = sg.id = sr.id = -1;
    tseg_at(a, lane, 0, sl);
    tseg_at(a, lane, 1, sg);
    double hl = TPT(0, PT_H), hc = TPT(1, PT_H);
    TEdge eL = tr_edge(yl, yc, sl.offn, sg.offn, sl.offp, sg.offp, sl.mun, sl.mup, Vt, a.bern);
    double cprev[5][5], dprev[5][3], gam[5][2];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
#pragma unroll
        for (int q = 0; q < 5; ++q) cprev[r][q] = 0.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) dprev[r][q] = 0.0;
        gam[r][0] = gam[r][1] = 0.0;
    }
    g[0] = g[1] = 0.0;
    for (int s = 0; s < 2; ++s)
        for (int q = 0; q < 3; ++q) gd[s][q] = 0.0;
    const double tot[2] = {TSC(SC_AT), TSC(SC_CT)};
    double dens_l[2] = {0.0, 0.0};
    for (int i = 1; i <= NP; ++i) {
        tr_point<NC>(a, lane, S, i + 1, yr);
        tseg_at(a, lane, i + 1, sr);
        const double hr = TPT(i + 1, PT_H);
        if (S.fresh) {  // the step's first guess, stored for the later passes
#pragma unroll
            for (int c = 0; c < NC; ++c) TY(S.bG, i, c) = yc[c];
        }
        // the state before the step at this point, and the traps' occupation over it
        double yb[5];
#pragma unroll
        for (int c = 0; c < NC; ++c) yb[c] = TY(S.bY, i, c);
        const double nb = exp(yb[1]), pb = exp(yb[2]);
        double fb = TFB(i);
        if (S.fresh && S.pending) {
            fb = tr_trap_end(sg.Nt, sg.Cn, sg.Cp, sg.nt0, sg.pt0, nb, pb, fb, S.dti_pend);
            TFB(i) = fb;
        }
        const TEdge eR = tr_edge(yc, yr, sg.offn, sr.offn, sg.offp, sr.offp, sg.mun, sg.mup, Vt, a.bern);
        // the row (scaled): Poisson, electrons, holes, anions, cations
        const double hm = hl, hp = hc, hs = hp + hm;
        const double fac = 0.5 * L2 * hp * hm * hs;
        const double epsm = sl.eps, epsc = sg.eps;
        const double fac_m = (epsc - epsm) / (hm * hm * L2) - 2 * epsc * (1 / (hp * hs * L2) + 1 / (hm * hs * L2));
        const double fac_l = 2 * epsc / (hm * hs * L2) - (epsc - epsm) / (hm * hm * L2);
        const double fac_u = 2 * epsc / (hp * hs * L2);
        const double sP = sg.sP, sC = sg.sCs / hp;
        double F[5], A[5][5], B[5][5], C[5][5], U[5][2];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            F[r] = 0.0; U[r][0] = U[r][1] = 0.0;
#pragma unroll
            for (int q = 0; q < 5; ++q) A[r][q] = B[r][q] = C[r][q] = 0.0;
        }
        const Dual<2> n = dexp(dvar<2>(yc[1], 0)), p = dexp(dvar<2>(yc[2], 1));
        Dual<2> Rn, Rp;
        tr_rates(sg.kd, sg.ni2, sg.Nt, sg.Cn, sg.Cp, sg.nt0, sg.pt0, n, p, fb, S.dti, Rn, Rp);
        const double gen = TPT(i, PT_G) * S.gk;
        const double dti = S.dti;
        F[1] = (hm * eR.Fn - hp * eL.Fn + fac * (gen - Rn.v) - fac * dti * (n.v - nb)) * sC;
        F[2] = (hp * eL.Fp - hm * eR.Fp + fac * (gen - Rp.v) - fac * dti * (p.v - pb)) * sC;
        // electrons' row: V, ln n columns of the left/own/right points
        A[1][0] = -hp * eL.dn[0] * sC; A[1][1] = -hp * eL.dn[2] * sC;
        B[1][0] = (hm * eR.dn[0] - hp * eL.dn[1]) * sC;
        B[1][1] = (hm * eR.dn[2] - hp * eL.dn[3] - fac * Rn.d[0] - fac * dti * n.v) * sC;
        B[1][2] = -fac * Rn.d[1] * sC;
        C[1][0] = hm * eR.dn[1] * sC; C[1][1] = hm * eR.dn[3] * sC;
        A[2][0] = hp * eL.dp[0] * sC; A[2][2] = hp * eL.dp[2] * sC;
        B[2][0] = (hp * eL.dp[1] - hm * eR.dp[0]) * sC;
        B[2][1] = -fac * Rp.d[0] * sC;
        B[2][2] = (hp * eL.dp[3] - hm * eR.dp[2] - fac * Rp.d[1] - fac * dti * p.v) * sC;
        C[2][0] = -hm * eR.dp[1] * sC; C[2][2] = -hm * eR.dp[3] * sC;
        const double an = NC == 5 ? sg.am * exp(yc[3]) : 0.0, ca = NC == 5 ? sg.cm * exp(yc[4]) : 0.0;
        const double fix = TPT(i, PT_FIX);
        F[0] = (fac_l * yl[0] + fac_m * yc[0] + fac_u * yr[0] -
                Q_E * (n.v - p.v + sg.dop + fix + an - ca)) * sP;
        A[0][0] = fac_l * sP; C[0][0] = fac_u * sP;
        B[0][0] = fac_m * sP; B[0][1] = -Q_E * n.v * sP; B[0][2] = Q_E * p.v * sP;
        double W[2] = {0.0, 0.0};  // the slots' constraints in this point's own ion unknown
        if (NC == 5) {
            B[0][3] = -Q_E * an * sP;
            B[0][4] = Q_E * ca * sP;
            const double befores[2] = {sg.am * exp(yb[3]), sg.cm * exp(yb[4])};
            const double masks_l[2] = {sl.am, sl.cm}, masks_c[2] = {sg.am, sg.cm}, masks_r[2] = {sr.am, sr.cm};
            const double mus_l[2] = {sl.mua, sl.muc}, mus_c[2] = {sg.mua, sg.muc};
#pragma unroll
            for (int sp = 0; sp < 2; ++sp) {
                const int col = 3 + sp;
                const Dual<7> v = tr_ion_row(sp == 0 ? -1.0 : 1.0, masks_l[sp], masks_c[sp],
                                             masks_r[sp], i == 1, i == NP, mus_l[sp], mus_c[sp],
                                             hm, hp, fac, dti, befores[sp], yl, yc, yr, col,
                                             S.lg[sp], Vt, a.bern);
                F[col] = v.v;
                A[col][0] = v.d[0]; B[col][0] = v.d[1]; C[col][0] = v.d[2];
                A[col][col] = v.d[3]; B[col][col] = v.d[4]; C[col][col] = v.d[5];
                U[col][sp] = v.d[6] * masks_c[sp];
                // the constraint: trapezoidal number over the region, from the ions' own density
                const double dens = masks_c[sp] * exp(yc[col]);
                const double pair = masks_l[sp] * masks_c[sp];
                const double hL = hm * L;
                if (tot[sp] > 0) {
                    g[sp] += 0.5 * (dens + dens_l[sp]) * pair * hL;
                    W[sp] = 0.5 * (pair * hm + masks_c[sp] * masks_r[sp] * hp) * L * dens / tot[sp];
                }
                dens_l[sp] = dens;
            }
        }
        // elimination: M = B - A c'_prev, rhs = [-F | U] - A d'_prev
        double M[NC][NC], R[NC][NC + K];
#pragma unroll
        for (int r = 0; r < NC; ++r) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                double s = B[r][q];
#pragma unroll
                for (int l = 0; l < NC; ++l) s -= A[r][l] * cprev[l][q];
                M[r][q] = s;
                R[r][q] = C[r][q];
            }
#pragma unroll
            for (int q = 0; q < K; ++q) {
                double s = q == 0 ? -F[r] : U[r][q - 1];
#pragma unroll
                for (int l = 0; l < NC; ++l) s -= A[r][l] * dprev[l][q];
                R[r][NC + q] = s;
            }
        }
        tr_solve_block<NC, NC + K>(M, R);
        // gamma_i = W_i - c'_{i-1}^T gamma_{i-1} (W in the ions' own columns)
        double gn[5][2];
#pragma unroll
        for (int r = 0; r < NC; ++r)
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                double v = (NC == 5 && r == 3 + s) ? W[s] : 0.0;
#pragma unroll
                for (int l = 0; l < NC; ++l) v -= cprev[l][r] * gam[l][s];
                gn[r][s] = v;
            }
#pragma unroll
        for (int r = 0; r < NC; ++r) {
#pragma unroll
            for (int q = 0; q < NC; ++q) {
                cprev[r][q] = R[r][q];
                TFAC(i, r * 5 + q) = R[r][q];
            }
#pragma unroll
            for (int q = 0; q < K; ++q) {
                dprev[r][q] = R[r][NC + q];
                TFAC(i, 25 + r * 3 + q) = R[r][NC + q];
            }
            gam[r][0] = gn[r][0];
            gam[r][1] = gn[r][1];
        }
        if (NC == 5) {
#pragma unroll
            for (int s = 0; s < 2; ++s)
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    double v = 0.0;
#pragma unroll
                    for (int r = 0; r < NC; ++r) v += gam[r][s] * dprev[r][q];
                    gd[s][q] += v;
                }
        }
        // shift the window
#pragma unroll
        for (int c = 0; c < 5; ++c) { yl[c] = yc[c]; yc[c] = yr[c]; }
        sl = sg; sg = sr;
        hl = hc; hc = hr;
        eL = eR;
    }
    if (NC == 5) {  // the last interval of each region (the right contact is outside)
        for (int sp = 0; sp < 2; ++sp) {
            const double pair = (sp == 0 ? sl.am * sg.am : sl.cm * sg.cm);
            const double dens = 0.0;
            g[sp] += 0.5 * (dens + dens_l[sp]) * pair * hl * L;
            g[sp] = tot[sp] > 0 ? g[sp] / tot[sp] - 1.0 : 0.0;
        }
    }
}

// The back-substitution: the step with the slots' z, its caps, the new iterate, its size.
template <int NC>
DDH_TFN double tr_backward(const TrArgs &a, int lane, TrLane &S, const double z[2]) {
    const int K = NC == 5 ? 3 : 1;
    const double Vt = TSC(SC_VT), nfloor = TSC(SC_NFLOOR), lnfloor = log(nfloor);
    const double caps[5] = {CAP_V * Vt, CAP_N, CAP_N, CAP_N, CAP_N};
    double xn[5][3];
#pragma unroll
    for (int r = 0; r < 5; ++r) xn[r][0] = xn[r][1] = xn[r][2] = 0.0;
    double size = 0.0;
    bool finite = true;
    for (int i = a.NP; i >= 1; --i) {
        double x[5][3];
#pragma unroll
        for (int r = 0; r < NC; ++r)
#pragma unroll
            for (int q = 0; q < K; ++q) {
                double v = TFAC(i, 25 + r * 3 + q);
#pragma unroll
                for (int l = 0; l < NC; ++l) v -= TFAC(i, r * 5 + l) * xn[l][q];
                x[r][q] = v;
            }
#pragma unroll
        for (int r = 0; r < NC; ++r)
#pragma unroll
            for (int q = 0; q < K; ++q) xn[r][q] = x[r][q];
        double y[5], step[5];
#pragma unroll
        for (int c = 0; c < NC; ++c) y[c] = TY(S.bG, i, c);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double delta = NC == 5 ? x[c][0] - (x[c][1] * z[0] + x[c][2] * z[1]) : x[c][0];
            step[c] = clipd(delta, -caps[c], caps[c]);
            if ((c == 1 || c == 2) && y[c] < lnfloor) {  // density_steps (solver.py:74-88)
                const double in_n = delta > -1 ? log1p(delta) : -CAP_N;
                step[c] = clipd(in_n, -FLOOR_DROP_CAP, CAP_N);
            }
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            const double v = y[c] + step[c];
            TY(S.bG, i, c) = v;
            finite = finite && isfinite(v);
        }
        size = nmax(size, fabs(step[0]) / Vt);
#pragma unroll
        for (int c = 1; c < NC; ++c) {
            const double dens = exp(y[c] + step[c]);
            size = nmax(size, fabs(step[c]) * (dens / (dens + nfloor)));
        }
    }
    S.finite = finite;
    return size;
}

// The transient current of the iterate (batch.py:302-418 with a step; `_ion_currents`
// :433-483; the displacement current), averaged over the intervals as `_current` does, in one
// pass from the left contact: each part's current on interval e is q K + acc_e (acc the prefix
// sum of sn q L width U), K known only at the end, so the h-weighted average keeps K's
// weight (sum of h) and acc's apart; across an interface an interval takes the one before.
template <int NC>
DDH_TFN double tr_current(const TrArgs &a, int lane, const TrLane &S) {
    const int NP = a.NP;
    const double Vt = TSC(SC_VT), iVt = 1.0 / Vt, L = TSC(SC_L), sign = -TSC(SC_ELEC);
    const double dti = S.dti, q = Q_E;
    // carriers c = 0 electrons (sn -1), 1 holes (sn +1)
    double iu[2] = {0, 0}, sgl[2] = {0, 0}, dbl[2] = {0, 0}, acc[2] = {0, 0}, hacc[2] = {0, 0};
    double Up[2] = {0, 0}, evp[2] = {0, 0}, imup[2] = {1, 1}, endl[2] = {0, 0}, endr[2] = {0, 0};
    // ions k = 0 anions (sn -1), 1 cations (sn +1): their region's sums
    double iui[2] = {0, 0}, sgli[2] = {0, 0}, dbli[2] = {0, 0}, acci[2] = {0, 0};
    double hacci[2] = {0, 0}, hq[2] = {0, 0}, Uip[2] = {0, 0}, evip[2] = {0, 0}, imuip[2] = {1, 1};
    double mp[2] = {0, 0}, dfirst[2] = {0, 0}, dlast[2] = {0, 0};
    int first[2] = {-1, -1};
    const double tot[2] = {TSC(SC_AT), TSC(SC_CT)};
    double H = 0.0, disp = 0.0, accp[2] = {0, 0}, accip[2] = {0, 0};
    TSeg sg;
    sg.id = -1;
    double xp = 0.0, hp = 0.0, Vp = 0.0, Vbp = 0.0, epsp = 0.0;
    for (int j = 0; j <= NP + 1; ++j) {
        const bool contact = j == 0 || j == NP + 1;
        double y[5], yb[5];
        if (contact) {
            tr_contact(a, lane, S.Vk, j != 0, y);
            tr_contact(a, lane, S.V_now, j != 0, yb);
        } else {
#pragma unroll
            for (int c = 0; c < NC; ++c) { y[c] = TY(S.bG, j, c); yb[c] = TY(S.bY, j, c); }
        }
        tseg_at(a, lane, j, sg);
        const double x = TPT(j, PT_X), h = TPT(j, PT_H);
        const double n = exp(y[1]), p = exp(y[2]);
        const double dens[2] = {n, p}, before[2] = {exp(yb[1]), exp(yb[2])};
        const double width = (a.zimt || j == 0) ? h : 0.5 * (hp + h);
        double U[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            double u = 0.0;
            if (!contact)
                u = tr_rate_value(sg.kd, sg.ni2, sg.Nt, sg.Cn, sg.Cp, sg.nt0, sg.pt0, n, p, TFB(j),
                                  dti, c == 1) - TPT(j, PT_G) * S.gk;
            if (j > 0 || a.zimt) u += dti * (dens[c] - before[c]);
            U[c] = u;
        }
        const double ev[2] = {exp(-(y[0] + sg.offn) * iVt), exp((y[0] + sg.offp) * iVt)};
        const double imu[2] = {1.0 / sg.mun, 1.0 / sg.mup};
        if (j == 0) { endl[0] = n * ev[0]; endl[1] = p * ev[1]; }
        if (j == NP + 1) { endr[0] = n * ev[0]; endr[1] = p * ev[1]; }
        if (j >= 1 && j <= NP) {  // the interval (j - 1, j) of the integrals
            const double dx = x - xp;
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const double iun = iu[c] + 0.5 * dx * (Up[c] + U[c]);
                sgl[c] += 0.5 * dx * (evp[c] * imup[c] + ev[c] * imu[c]);
                dbl[c] += 0.5 * dx * (evp[c] * iu[c] * imup[c] + ev[c] * iun * imu[c]);
                iu[c] = iun;
            }
        }
        if (j <= NP) {  // the interval after j: J_j = q K + acc_j
            const bool across = j >= 1 && TIFACE(j);
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const double sn = c == 0 ? -1.0 : 1.0;
                accp[c] = acc[c];
                if (j >= 1) acc[c] += sn * q * L * width * U[c];
                hacc[c] += h * (across ? accp[c] : acc[c]);
            }
            H += h;
        }
        if (j >= 1) {  // the displacement current on the interval (j - 1, j)
            disp += epsp * ((y[0] - Vp) - (yb[0] - Vbp));
        }
        if (NC == 5) {
            const double ms[2] = {sg.am, sg.cm};
            const double mus[2] = {sg.mua, sg.muc};
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const double sn = k == 0 ? -1.0 : 1.0;
                const double m = ms[k];
                const double d = exp(y[3 + k]);
                const double Ui = dti * (d - exp(yb[3 + k]) * m);
                const double evi = exp(sn * y[0] * iVt);
                const double imui = 1.0 / (mus[k] > 0 ? mus[k] : 1.0);
                if (m > 0) {
                    if (first[k] < 0) { first[k] = j; dfirst[k] = d * evi; }
                    dlast[k] = d * evi;
                }
                if (j >= 1) {  // the interval (j - 1, j)
                    const double pair = mp[k] * m, dx = x - xp;
                    const double iun = iui[k] + 0.5 * dx * (Uip[k] + Ui) * pair;
                    sgli[k] += 0.5 * dx * (evip[k] * imuip[k] + evi * imui) * pair;
                    dbli[k] += 0.5 * dx * (evip[k] * iui[k] * imuip[k] + evi * iun * imui) * pair;
                    iui[k] = iun;
                }
                Uip[k] = Ui; evip[k] = evi; imuip[k] = imui;
                mp[k] = m;
            }
            if (j <= NP) {  // the interval after j, inside a region
                const bool across = j >= 1 && TIFACE(j);
                // the next point's masks decide the pair
                double mn[2];
                tr_masks(a, lane, j + 1, mn);
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    const double sn = k == 0 ? -1.0 : 1.0;
                    const double pair = mp[k] * mn[k];
                    accip[k] = acci[k];
                    if (first[k] >= 0 && j > first[k]) acci[k] += sn * q * L * width * Uip[k] * pair;
                    if (pair > 0 && tot[k] > 0) {
                        hq[k] += h;
                        hacci[k] += h * (across ? accip[k] : acci[k]);
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 2; ++c) { Up[c] = U[c]; evp[c] = ev[c]; imup[c] = imu[c]; }
        xp = x; hp = h; Vp = y[0]; Vbp = yb[0]; epsp = sg.eps;
    }
    double total = 0.0;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const double sn = c == 0 ? -1.0 : 1.0;
        const double K = (Vt * (sn * endr[c] - sn * endl[c]) - sn * dbl[c]) / sgl[c];
        total += q * K * H + hacc[c];
    }
    if (NC == 5) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            if (!(tot[k] > 0) || first[k] < 0) continue;
            const double sn = k == 0 ? -1.0 : 1.0;
            const double sg_ = sgli[k] > 0 ? sgli[k] : 1.0;
            const double K = (Vt * sn * (dlast[k] - dfirst[k]) - sn * dbli[k]) / sg_;
            total += q * K * hq[k] + hacci[k];
        }
    }
    return sign * (total + dti * disp / L) / H;
}

// The Boltzmann normalisations (`_ion_logs`, solver.py:1695-1712) that put each species' total
// into its region at the potentials of state buffer b (the contacts at Vint).
DDH_TFN void tr_boltzmann_logs(const TrArgs &a, int lane, int b, double Vint, double out[2]) {
    const double Vt = TSC(SC_VT), L = TSC(SC_L);
    const double tot[2] = {TSC(SC_AT), TSC(SC_CT)};
    double num[2] = {0.0, 0.0}, dprev[2] = {0.0, 0.0}, mprev[2] = {0.0, 0.0};
    TSeg sg;
    sg.id = -1;
    double hprev = 0.0;
    for (int j = 0; j <= a.NP + 1; ++j) {
        double V;
        if (j == 0 || j == a.NP + 1) {
            double c[5];
            tr_contact(a, lane, Vint, j != 0, c);
            V = c[0];
        } else {
            V = TY(b, j, 0);
        }
        tseg_at(a, lane, j, sg);
        const double ms[2] = {sg.am, sg.cm};
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double d = ms[k] * exp((k == 0 ? 1.0 : -1.0) * V / Vt);
            if (j >= 1) num[k] += 0.5 * (d + dprev[k]) * (ms[k] * mprev[k]) * hprev * L;
            dprev[k] = d;
            mprev[k] = ms[k];
        }
        hprev = TPT(j, PT_H);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) out[k] = tot[k] > 0 ? log(tot[k] / num[k]) : 0.0;
}

// A step's beginning (transient.py:355-381): its dt from the last step kept, the first guess's
// extrapolation ratio, the slots' first guess.
template <int NC>
DDH_TFN void tr_begin_step(const TrArgs &a, int lane, TrLane &S) {
    const int k = S.k;
    const double tk = TOUT(a.t, k);
    S.dti = tk == 0 ? 0.0 : 1.0 / (tk > S.t_last ? tk - S.t_last : 1.0);
    const bool moving = S.dti > 0 && S.dti_now > 0 && !S.first;
    S.rt = moving ? S.dti_now / S.dti : 0.0;
    S.Vk = TOUT(a.V, k);
    S.gk = TOUT(a.G, k);
#pragma unroll
    for (int s = 0; s < 2; ++s)
        S.lg[s] = S.logs[s] + clipd(S.rt * (S.logs[s] - S.logs_prev[s]), -ION_LOG_CAP, ION_LOG_CAP);
    if (NC == 5) {
        if (S.dti > 0) {
            if (!moving) S.lg[0] = S.lg[1] = 0.0;
        } else {  // a steady state: Boltzmann's normalisations at the guess (the kept state)
            tr_boltzmann_logs(a, lane, S.bY, S.Vk, S.lg);
        }
    }
    S.fresh = 1;
    S.it = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) S.hist[q] = S.Jh[q] = INFINITY;
}

// A lane's setup: its steady start (the state the steady solve kept at its first voltage), the
// ions' densities as unknowns (Boltzmann's profiles, 0 outside their regions), the traps empty
// (their occupation does not enter a steady state), and its first step.
template <int NC>
DDH_TFN void tr_init(const TrArgs &a, int lane, TrLane &S) {
    S.done = 0;
    S.k = 0; S.first = 1; S.stopped = 0; S.pending = 0;
    S.bY = 0; S.bP = 0; S.bG = 1;
    S.dti_now = 0.0; S.dti_pend = 0.0; S.t_last = 0.0;
    S.V_now = TOUT(a.V, 0);
    S.finite = 1;
    double m0[2], m1[2];
    tr_masks(a, lane, 0, m0);
    tr_masks(a, lane, a.NP + 1, m1);
    const bool at_contact = NC == 5 && (m0[0] > 0 || m0[1] > 0 || m1[0] > 0 || m1[1] > 0);
    if (TSC(SC_PRE) > 0 || at_contact) {  // a pre-bias; ions reaching a contact: not on this path
        a.status[lane] = 3;
        for (int k = 0; k < a.nt; ++k) {
            TOUT(a.J, k) = NAN; TOUT(a.Jext, k) = NAN;
            TOUT(a.ok, k) = 0; TOUT(a.kept, k) = 0; TOUT(a.its, k) = 0;
        }
        S.done = 1;
        return;
    }
    a.status[lane] = 0;
    const double Vt = TSC(SC_VT);
    for (int j = 1; j <= a.NP; ++j) {
#pragma unroll
        for (int c = 0; c < 3; ++c)
            TY(S.bY, j, c) = a.y0[TCHK(((size_t)(j - 1) * 3 + c) * a.n + lane, (size_t)a.NP * 3 * a.n)];
        TFB(j) = 0.0;
    }
    S.logs[0] = S.logs[1] = 0.0;
    if (NC == 5) {
        tr_boltzmann_logs(a, lane, S.bY, S.V_now, S.logs);
        TSeg sg;
        sg.id = -1;
        for (int j = 1; j <= a.NP; ++j) {
            tseg_at(a, lane, j, sg);
            const double V = TY(S.bY, j, 0);
            TY(S.bY, j, 3) = sg.am > 0 ? S.logs[0] + V / Vt : 0.0;
            TY(S.bY, j, 4) = sg.cm > 0 ? S.logs[1] - V / Vt : 0.0;
        }
    }
    S.logs_prev[0] = S.logs[0];
    S.logs_prev[1] = S.logs[1];
    tr_begin_step<NC>(a, lane, S);
}

// A step's end: its record, and ZimT's main loop (failureMode; Track 0).
template <int NC>
DDH_TFN void tr_end_step(const TrArgs &a, int lane, TrLane &S, int ok) {
    const int k = S.k;
    const bool finite = S.finite && isfinite(S.lg[0]) && isfinite(S.lg[1]);
    const bool running = !S.stopped;
    const bool failed = running && !ok && a.failure_mode == 0;
    const bool keep = running && !failed && finite &&
                      (ok || a.failure_mode == 1 || (a.failure_mode == 2 && S.dti == 0));
    const double J = S.Jh[3];
    TOUT(a.J, k) = J;
    TOUT(a.Jext, k) = J + TSC(SC_GSH) * S.Vk;
    TOUT(a.ok, k) = (uint8_t)ok;
    TOUT(a.its, k) = S.it;
    TOUT(a.kept, k) = (uint8_t)keep;
    if (keep) {
        S.logs_prev[0] = S.logs[0]; S.logs_prev[1] = S.logs[1];
        S.logs[0] = S.lg[0]; S.logs[1] = S.lg[1];
        S.V_now = S.Vk;
        S.dti_now = S.dti;
        S.t_last = TOUT(a.t, k);
        S.pending = 1;
        S.dti_pend = S.dti;
        const int nP = S.bY, nY = S.bG;
        S.bP = nP; S.bY = nY; S.bG = 3 - nP - nY;
    }
    if (failed) S.stopped = 1;
    S.first = 0;
    S.k = k + 1;
    if (S.k >= a.nt) { S.done = 1; return; }
    tr_begin_step<NC>(a, lane, S);
}

// One Newton step of the lane's current time step (batch.py:759-839).
template <int NC>
DDH_TFN void tr_newton(const TrArgs &a, int lane, TrLane &S) {
    double g[2], gd[2][3];
    tr_forward<NC>(a, lane, S, g, gd);
    double z[2] = {0.0, 0.0};
    if (NC == 5) {
        const double held[2] = {TSC(SC_AT) > 0 ? 0.0 : 1.0, TSC(SC_CT) > 0 ? 0.0 : 1.0};
        z[0] = -g[0] - held[0] * S.lg[0] - gd[0][0];
        z[1] = -g[1] - held[1] * S.lg[1] - gd[1][0];
        solve2(held[0] - gd[0][1], -gd[0][2], -gd[1][1], held[1] - gd[1][2], z[0], z[1]);
    }
    double size = tr_backward<NC>(a, lane, S, z);
    if (NC == 5) {
        const double d0 = clipd(z[0], -ION_LOG_CAP, ION_LOG_CAP), d1 = clipd(z[1], -ION_LOG_CAP, ION_LOG_CAP);
        S.lg[0] += d0;
        S.lg[1] += d1;
        size = nmax(size, nmax(fabs(d0), fabs(d1)));
    }
    S.fresh = 0;
    S.pending = 0;
    const double J = tr_current<NC>(a, lane, S);
#pragma unroll
    for (int q = 0; q < 3; ++q) { S.hist[q] = S.hist[q + 1]; S.Jh[q] = S.Jh[q + 1]; }
    S.hist[3] = size;
    S.Jh[3] = J;
    S.it += 1;
    // the four stop rules (batch.py:793-818)
    double hmax = S.hist[0], hmin_tail = S.hist[1];
    bool anynan = S.hist[0] != S.hist[0];
#pragma unroll
    for (int q = 1; q < 4; ++q) {
        hmax = fmax(hmax, S.hist[q]);
        hmin_tail = fmin(hmin_tail, S.hist[q]);
        anynan = anynan || S.hist[q] != S.hist[q];
    }
    const bool stalled = S.it > STALL_STEPS && !anynan && hmax < ROUND_OFF_FLOOR && hmin_tail >= 0.1 * S.hist[0];
    const double scale = TSC(SC_JS) + fabs(J);
    double jdev = 0.0;
    bool jnan = false;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = fabs(S.Jh[q] - J);
        jnan = jnan || v != v;
        jdev = fmax(jdev, v);
    }
    const bool settled = S.it > STALL_STEPS && size < CURRENT_STEP_CAP && !jnan && jdev < CURRENT_TOL * scale;
    const bool quadratic = S.it > 1 && size < QUADRATIC_TOL && size < QUADRATIC_RATIO * S.hist[2];
    const bool done = size < a.tol || stalled || settled || quadratic;
    if (done || S.it >= a.max_iter || size != size) tr_end_step<NC>(a, lane, S, done ? 1 : 0);
}

}  // namespace ddh
