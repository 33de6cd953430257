// Two lanes per device: each Newton step's block-tridiagonal solve split at the middle of the
// grid (a "twisted" factorisation). The left lane eliminates rows 1..m from the left contact
// towards the middle (x_i = f'_i - e'_i x_{i+1}), the right lane rows NP..m+1 from the right
// contact (x_i = d'_i - c'_i x_{i-1}); the two meet in one 3x3 solve for x_m, x_{m+1} (and the
// ion border's Schur complement, whose adjoint recursions run from both ends), then each lane
// back-substitutes its half outwards and accumulates its part of the integral current. Both
// lanes run the same instructions (the direction is data, not control flow), so a wavefront
// of 32 devices keeps all 64 lanes busy and a batch fills twice the wavefronts of the one-lane
// kernel, each with half the serial chain.
//
// The lanes exchange two small records per step (register shuffles on the GPU). Every value
// that a decision depends on (step size, current, ion step) is combined in the same order on
// both lanes, so the pair's two state machines stay identical. The arithmetic is the one-lane
// kernel's except for the order of elimination and of the current's sums: results agree with
// it to round-off.
#pragma once

#include "ddhip_core.h"

namespace ddh {

// What one half's elimination hands the middle: the factor and right-hand sides of its row
// next to the middle (left: e'_m, f'_m; right: c'_{m+1}, d'_{m+1}), the border's adjoint
// recursion there and its sums, and the half's ion numbers.
struct PairElim {
    double cf[9];
    double df[9];        // [row][column], 3 columns (the residual, two ion columns)
    double gam[6];       // [row][slot]
    double acc[6];       // [slot][column]: W^T X of the half
    double num[2];
};

// What one half's back-substitution hands the combination: its step size and finiteness, its
// integral current parts, the values of its point next to the middle, its contact's terms.
struct PairBack {
    double size, finite;
    double Sn, Dn, Sp, Dp, iu;  // left: sums over intervals 1..m; right: over m+2..NP (local)
    double wn, wp, U, x;        // left: point m; right: point m+1
    double endn, endp;          // left: the left contact's; right: the right contact's
};

// The wavefront's segment tables in the local data share: 32 device slots, row stride 32 (a
// pair's two lanes read their device's column), 39 KB for seven segments.
constexpr int PAIR_LS = 32;

#ifdef DDH_CPU_EMULATION
#define DDH_PAIRFN inline
#else
#define DDH_PAIRFN __device__ __attribute__((noinline))
#endif

// The elimination of one half (src/jaxdd/solver.py:1021-1187 rows, :1534-1560 elimination),
// for the Bernoulli function of kind BK (0 exact, 1 SIMsalabim's; the launch's `bern`).
template <typename T, typename S, int K, int BK>
DDH_PAIRFN void elim_half(const Args &a, int lane, int half, const SweepCtl &cref, PairElim &out) {
    const SweepCtl c = cref;
    const int NP = DDH_UNIFORM(a.NP), bern = BK;
    const int m = NP / 2;
    const int dir = half ? -1 : 1;
    const int j0 = half ? NP + 1 : 0;
    const int nvis = half ? NP + 2 - m : m + 2;  // left: points 0..m+1; right: NP+1..m
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, L = SC(SC_L), L2 = L * L;
    const double tot0 = c.mobile ? SC(SC_AT) * c.share : 0.0;
    const double tot1 = c.mobile ? SC(SC_CT) * c.share : 0.0;
    const double itot0 = tot0 > 0 ? 1.0 / tot0 : 0.0, itot1 = tot1 > 0 ? 1.0 / tot1 : 0.0;
    const Lane Lv = lane_view(a, lane, c, PAIR_LS);
    const size_t n = Lv.n;
    DDH_GLOBAL S *fac = (DDH_GLOBAL S *)(factors<S>(a) + lane);
    gdouble *s1 = c.fresh ? Lv.ys1 : Lv.yw;
    gdouble *s2 = (c.fresh && c.ratio != 0.0) ? Lv.ys2 : s1;
    const Contact ct = contact_point(a, lane, Lv, c.Vint, half == 1);

    T cp[9], dp[3 * K];
#pragma unroll
    for (int q = 0; q < 9; ++q) cp[q] = T(0);
#pragma unroll
    for (int q = 0; q < 3 * K; ++q) dp[q] = T(0);
    double gam[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    double accF[2] = {0, 0}, accU[2][2] = {{0, 0}, {0, 0}};
    double num0 = 0.0, num1 = 0.0;

    Pt mid, far, pn;
    Edge eold = {0, 0, 0, 0, 0, 0, 0, 0};
    Raw nxt, now;
    fetch_point(Lv, s1, s2, j0 + dir, nxt);
    make_point(Lv, j0, true, ct, nxt, Lv.sid[(size_t)j0 * n], c, Vt, iVt, mid);
    far = mid;
    // visit v: the next point, the edge to it, the row of the point before (from v = 2). The
    // first interior visit has no row and runs before the loop, followed by stores to the
    // slot the first row then overwrites: every loop visit, and the loop's entry, then have a
    // point's loads followed by a row's stores, and no wait for those loads covers the stores.
    auto visit = [&](const int v, const bool has_row) {
        const int j = j0 + dir * v;
        now = nxt;
        fetch_point(Lv, s1, s2, v + 1 < nvis ? j + dir : j, nxt);
        make_point(Lv, j, false, ct, now, now.sid, c, Vt, iVt, pn);
        // the edge between mid and the new point, left to right
        const Pt el = dir > 0 ? mid : pn, er = dir > 0 ? pn : mid;
        const Edge en = edge_flux<BK>(Lv, el, er, Vt, iVt, bern);
        if constexpr (K > 1) {
            // the ion numbers over the half's intervals: left 0..m-1, right m..NP
            if (dir > 0 ? j <= m : j >= m) {
                const double hk = el.h * L;
                num0 += 0.5 * (mid.an + pn.an) * (mid.am * pn.am) * hk;
                num1 += 0.5 * (mid.ca + pn.ca) * (mid.cm * pn.cm) * hk;
            }
        }
        if (has_row) {
            // the row of mid (point i), its neighbours far and the new point
            const int i = j - dir;
            const Pt lft = dir > 0 ? far : pn, rgt = dir > 0 ? pn : far;
            const Edge eL = dir > 0 ? eold : en, eR = dir > 0 ? en : eold;
            double A[9], B[9], C[9], r[3 * K];
            row_blocks<K>(Lv, lft, mid, rgt.V, eL, eR, i == NP, L2, iVt, c.share, c.g_frac, A,
                          B, C, r, c.trapq != 0);
            if constexpr (sizeof(T) == 4) {
#pragma unroll
                for (int row = 0; row < 3; ++row) {
                    double mx = 0.0;
#pragma unroll
                    for (int col = 0; col < 3; ++col) {
                        mx = fmax(mx, fabs(A[3 * row + col]));
                        mx = fmax(mx, fabs(B[3 * row + col]));
                        mx = fmax(mx, fabs(C[3 * row + col]));
                    }
                    if (mx > 0 && mx < INFINITY) {
                        int e;
                        frexp(mx, &e);
                        const double sc = ldexp(1.0, -e);
#pragma unroll
                        for (int col = 0; col < 3; ++col) {
                            A[3 * row + col] *= sc; B[3 * row + col] *= sc; C[3 * row + col] *= sc;
                        }
#pragma unroll
                        for (int q = 0; q < K; ++q) r[row * K + q] *= sc;
                    }
                }
            }
            // N: the coupling to the eliminated side, F: to the side still ahead
            T N[9], F[9], Bt[9], rt[3 * K];
#pragma unroll
            for (int q = 0; q < 9; ++q) {
                N[q] = (T)(dir > 0 ? A[q] : C[q]);
                F[q] = (T)(dir > 0 ? C[q] : A[q]);
                Bt[q] = (T)B[q];
            }
#pragma unroll
            for (int q = 0; q < 3 * K; ++q) rt[q] = (T)r[q];
            T M[9], rh[3 * K];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    M[3 * p + q] = Bt[3 * p + q] - (N[3 * p + 0] * cp[0 + q] + N[3 * p + 1] * cp[3 + q] + N[3 * p + 2] * cp[6 + q]);
#pragma unroll
                for (int q = 0; q < K; ++q)
                    rh[p * K + q] = rt[p * K + q] - (N[3 * p + 0] * dp[0 * K + q] + N[3 * p + 1] * dp[1 * K + q] + N[3 * p + 2] * dp[2 * K + q]);
            }
            T Mi[9];
            inverse3(M, Mi);
            T cn_[9], dn_[3 * K];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    cn_[3 * p + q] = Mi[3 * p + 0] * F[0 + q] + Mi[3 * p + 1] * F[3 + q] + Mi[3 * p + 2] * F[6 + q];
#pragma unroll
                for (int q = 0; q < K; ++q)
                    dn_[p * K + q] = Mi[3 * p + 0] * rh[0 * K + q] + Mi[3 * p + 1] * rh[1 * K + q] + Mi[3 * p + 2] * rh[2 * K + q];
            }
            DDH_GLOBAL S *f = fac + (size_t)(i - 1) * NFAC * n;
#pragma unroll
            for (int q = 0; q < 9; ++q) f[q * n] = (S)cn_[q];
#pragma unroll
            for (int q = 0; q < 3 * K; ++q) f[(9 + q) * n] = (S)dn_[q];
            if constexpr (K > 1) {
                // gamma_i = W_i - p'^T gamma_prev, p' the previous row's factor
                const double hLm = lft.h * L, hLp = mid.h * L;
                const double mida = mid.am, midc = mid.cm;
                const double w0 = 0.5 * (mida * lft.am * hLm + rgt.am * mida * hLp) * mid.an * iVt * itot0;
                const double w1 = -0.5 * (midc * lft.cm * hLm + rgt.cm * midc * hLp) * mid.ca * iVt * itot1;
                double ng[3][2];
#pragma unroll
                for (int row = 0; row < 3; ++row) {
                    const double c0 = (double)cp[0 + row], c1 = (double)cp[3 + row], c2 = (double)cp[6 + row];
                    ng[row][0] = (row == 0 ? w0 : 0.0) - (c0 * gam[0][0] + c1 * gam[1][0] + c2 * gam[2][0]);
                    ng[row][1] = (row == 0 ? w1 : 0.0) - (c0 * gam[0][1] + c1 * gam[1][1] + c2 * gam[2][1]);
                }
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double sF = 0.0, s1_ = 0.0, s2_ = 0.0;
#pragma unroll
                    for (int row = 0; row < 3; ++row) {
                        gam[row][k] = ng[row][k];
                        sF += ng[row][k] * (double)dn_[row * K + 0];
                        s1_ += ng[row][k] * (double)dn_[row * K + 1];
                        s2_ += ng[row][k] * (double)dn_[row * K + 2];
                    }
                    accF[k] += sF; accU[k][0] += s1_; accU[k][1] += s2_;
                }
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) cp[q] = cn_[q];
#pragma unroll
            for (int q = 0; q < 3 * K; ++q) dp[q] = dn_[q];
        }
        far = mid;
        mid = pn;
        eold = en;
    };
    visit(1, false);
    {
        DDH_GLOBAL S *f = fac + (size_t)(half ? NP - 1 : 0) * NFAC * n;
#pragma unroll
        for (int q = 0; q < 9 + 3 * K; ++q) f[q * n] = S(0);
    }
    for (int v = 2; v < nvis; ++v) visit(v, true);
#pragma unroll
    for (int q = 0; q < 9; ++q) out.cf[q] = (double)cp[q];
#pragma unroll
    for (int row = 0; row < 3; ++row)
#pragma unroll
        for (int q = 0; q < 3; ++q) out.df[3 * row + q] = q < K ? (double)dp[row * K + q] : 0.0;
#pragma unroll
    for (int row = 0; row < 3; ++row) { out.gam[2 * row] = gam[row][0]; out.gam[2 * row + 1] = gam[row][1]; }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        out.acc[3 * k] = accF[k];
        out.acc[3 * k + 1] = accU[k][0];
        out.acc[3 * k + 2] = accU[k][1];
    }
    out.num[0] = num0;
    out.num[1] = num1;
}

// The middle: x_m, x_{m+1} and the ion step z from the two halves' records (the same
// arithmetic on both lanes). x_m = f'_m - e'_m x_{m+1}, x_{m+1} = d'_{m+1} - c'_{m+1} x_m for
// each column; the border's W^T X (src/jaxdd/batch.py:595-603) from both recursions.
DDH_FN void pair_middle(const Args &a, int lane, const SweepCtl &c, int K, const PairElim &Lh,
                        const PairElim &Rh, double z[2], double xm[3], double xm1[3]) {
    const double *E = Lh.cf, *Cq = Rh.cf;
    double Mm[9], Mi[9];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            Mm[3 * p + q] = (p == q ? 1.0 : 0.0) - (E[3 * p] * Cq[q] + E[3 * p + 1] * Cq[3 + q] + E[3 * p + 2] * Cq[6 + q]);
    inverse3<double>(Mm, Mi);
    double X[3][3], X1[3][3];  // [row][column]
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        double rhs[3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
            rhs[p] = Lh.df[3 * p + q] - (E[3 * p] * Rh.df[q] + E[3 * p + 1] * Rh.df[3 + q] + E[3 * p + 2] * Rh.df[6 + q]);
#pragma unroll
        for (int p = 0; p < 3; ++p) X[p][q] = Mi[3 * p] * rhs[0] + Mi[3 * p + 1] * rhs[1] + Mi[3 * p + 2] * rhs[2];
#pragma unroll
        for (int p = 0; p < 3; ++p)
            X1[p][q] = Rh.df[3 * p + q] - (Cq[3 * p] * X[0][q] + Cq[3 * p + 1] * X[1][q] + Cq[3 * p + 2] * X[2][q]);
    }
    z[0] = z[1] = 0.0;
    if (K > 1) {
        // W^T X = the halves' sums - gamma_m^T e'_m X_{m+1} - gamma_{m+1}^T c'_{m+1} X_m
        double WX[2][3];
#pragma unroll
        for (int k = 0; k < 2; ++k)
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double tl = 0.0, tr = 0.0;
#pragma unroll
                for (int p = 0; p < 3; ++p) {
                    const double ex = E[3 * p] * X1[0][q] + E[3 * p + 1] * X1[1][q] + E[3 * p + 2] * X1[2][q];
                    const double cx = Cq[3 * p] * X[0][q] + Cq[3 * p + 1] * X[1][q] + Cq[3 * p + 2] * X[2][q];
                    tl += Lh.gam[2 * p + k] * ex;
                    tr += Rh.gam[2 * p + k] * cx;
                }
                WX[k][q] = (Lh.acc[3 * k + q] + Rh.acc[3 * k + q]) - tl - tr;
            }
        const double tot0 = c.mobile ? SC(SC_AT) * c.share : 0.0;
        const double tot1 = c.mobile ? SC(SC_CT) * c.share : 0.0;
        const double itot0 = tot0 > 0 ? 1.0 / tot0 : 0.0, itot1 = tot1 > 0 ? 1.0 / tot1 : 0.0;
        const double num0 = Lh.num[0] + Rh.num[0], num1 = Lh.num[1] + Rh.num[1];
        const double g0 = tot0 > 0 ? num0 * itot0 - 1.0 : 0.0;
        const double g1 = tot1 > 0 ? num1 * itot1 - 1.0 : 0.0;
        const double D0 = tot0 > 0 ? num0 * itot0 : 1.0;
        const double D1 = tot1 > 0 ? num1 * itot1 : 1.0;
        z[0] = -g0 - WX[0][0];
        z[1] = -g1 - WX[1][0];
        solve2(D0 - WX[0][1], -WX[0][2], -WX[1][1], D1 - WX[1][2], z[0], z[1]);
    }
#pragma unroll
    for (int p = 0; p < 3; ++p) {
        xm[p] = X[p][0] - (X[p][1] * z[0] + X[p][2] * z[1]);
        xm1[p] = X1[p][0] - (X1[p][1] * z[0] + X1[p][2] * z[1]);
    }
}

// The back-substitution of one half, outwards from the middle: the step (x of the middle
// point given), its caps, the new state, the step size, the current's parts.
template <typename T, typename S, int K>
DDH_PAIRFN void backsub_half(const Args &a, int lane, int half, const SweepCtl &cref,
                             const double *zin, const double *xfirst, PairBack &out) {
    const SweepCtl c = cref;
    const int NP = DDH_UNIFORM(a.NP);
    const int m = NP / 2;
    const int dir = half ? 1 : -1;
    const int jf = half ? m + 1 : m;
    const int cnt = half ? NP - m : m;
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, nfl = SC(SC_NFLOOR), lnfloor = log(nfl);
    const double capv = CAP_V * Vt;
    const Lane Lv = lane_view(a, lane, c, PAIR_LS);
    const size_t n = Lv.n;
    DDH_GLOBAL const S *fac = (DDH_GLOBAL const S *)(factors<S>(a) + lane);
    const double z0 = zin[0], z1 = zin[1];
    struct Next {
        S f[9 + 3 * K];
        double y0, y1, y2, G, x;
        int sid;
    };
    auto fetch = [&](int j, Next &r) {
        DDH_GLOBAL const S *f = fac + (size_t)(j - 1) * NFAC * n;
#pragma unroll
        for (int q = 0; q < 9 + 3 * K; ++q) r.f[q] = f[q * n];
        gdouble *y = Lv.yw + (size_t)(j - 1) * 3 * n;
        r.y0 = y[0]; r.y1 = y[n]; r.y2 = y[2 * n];
        r.G = Lv.pt[((size_t)j * PT_N + PT_G) * n];
        r.x = Lv.pt[((size_t)j * PT_N + PT_X) * n];
        r.sid = Lv.sid[(size_t)j * n];
    };
    double xs0 = xfirst[0], xs1 = xfirst[1], xs2 = xfirst[2];
    double size = 0.0;
    int finite = 1;
    // the current's running parts
    double Sn = 0.0, Dn = 0.0, Sp = 0.0, Dp = 0.0, iu = 0.0;
    double pwn = 0.0, pwp = 0.0, pU = 0.0, px = 0.0;  // the previous point's
    Next nxt, now;
    fetch(jf, nxt);
    // visit v; the first (the point next to the middle, its x given) runs before the loop, so
    // that the loop is entered, as it repeats, after a visit's stores
    auto visit = [&](const int v) {
        const int j = jf + dir * v;
        now = nxt;
        fetch(v + 1 < cnt ? j + dir : j, nxt);
        const int sid = now.sid;
        const S *f = now.f;
        double x0 = xs0, x1 = xs1, x2 = xs2;
        if (v > 0) {
            double dF[3];
#pragma unroll
            for (int row = 0; row < 3; ++row) {
                double t = (double)f[9 + row * K + 0];
                if constexpr (K > 1) t -= (double)f[9 + row * K + 1] * z0 + (double)f[9 + row * K + 2] * z1;
                dF[row] = t;
            }
            x0 = dF[0] - ((double)f[0] * xs0 + (double)f[1] * xs1 + (double)f[2] * xs2);
            x1 = dF[1] - ((double)f[3] * xs0 + (double)f[4] * xs1 + (double)f[5] * xs2);
            x2 = dF[2] - ((double)f[6] * xs0 + (double)f[7] * xs1 + (double)f[8] * xs2);
        }
        xs0 = x0; xs1 = x1; xs2 = x2;
        const CurSeg g = cur_seg(Lv, sid);
        gdouble *y = Lv.yw + (size_t)(j - 1) * 3 * n;
        double y0 = now.y0, y1 = now.y1, y2 = now.y2;
        const double st0 = clipd(x0, -capv, capv);
        double st1 = clipd(x1, -CAP_N, CAP_N), st2 = clipd(x2, -CAP_N, CAP_N);
        if (y1 < lnfloor) {  // density_steps (src/jaxdd/solver.py:74-88)
            const double in_n = x1 > -1 ? log1p(x1) : -CAP_N;
            st1 = clipd(in_n, -FLOOR_DROP_CAP, CAP_N);
        }
        if (y2 < lnfloor) {
            const double in_n = x2 > -1 ? log1p(x2) : -CAP_N;
            st2 = clipd(in_n, -FLOOR_DROP_CAP, CAP_N);
        }
        y0 += st0; y1 += st1; y2 += st2;
        y[0] = y0; y[n] = y1; y[2 * n] = y2;
        if (!(isfinite(y0) && isfinite(y1) && isfinite(y2))) finite = 0;
        const double nn = exp(y1), pp = exp(y2);
        size = nmax(size, fabs(st0) * iVt);
        size = nmax(size, fabs(st1) * (nn / (nn + nfl)));
        size = nmax(size, fabs(st2) * (pp / (pp + nfl)));
        const double R = recombination_value(nn, pp, g.kd, g.ni2, g.nt, g.cn, g.cp, g.nt0, g.pt0);
        const double U = R - now.G * c.g_frac;
        const double wn = exp(-(y0 + g.offn) * iVt) * g.imun;
        const double wp = exp((y0 + g.offp) * iVt) * g.imup;
        const double xj = now.x;
        if (v == 0) {
            out.wn = wn; out.wp = wp; out.U = U; out.x = xj;
        } else if (dir > 0) {
            // right: intervals left to right from point m+1, iu from zero there
            const double dx = xj - px;
            const double iun = iu + 0.5 * dx * (pU + U);
            Sn += 0.5 * dx * (pwn + wn);
            Dn += 0.5 * dx * (pwn * iu + wn * iun);
            Sp += 0.5 * dx * (pwp + wp);
            Dp += 0.5 * dx * (pwp * iu + wp * iun);
            iu = iun;
        } else {
            // left: intervals right to left from point m; sum_k u_k T_k with T_k the sum of the
            // point weights w_j (j >= k, over intervals 1..m)
            const double dx = px - xj;
            const double u = 0.5 * dx * (U + pU);
            Dn += u * (Sn + 0.5 * dx * pwn);
            Dp += u * (Sp + 0.5 * dx * pwp);
            Sn += 0.5 * dx * (wn + pwn);
            Sp += 0.5 * dx * (wp + pwp);
            iu += u;
        }
        pwn = wn; pwp = wp; pU = U; px = xj;
    };
    visit(0);
    for (int v = 1; v < cnt; ++v) visit(v);
    if (half) {
        const int sidN = Lv.sid[(size_t)(NP + 1) * n];
        const double VN = contact_V(a, lane, c.Vint, true);
        out.endn = exp(SC(SC_LNR) - (VN + segv(Lv, sidN, SEG_OFFN)) * iVt);
        out.endp = exp(SC(SC_LPR) + (VN + segv(Lv, sidN, SEG_OFFP)) * iVt);
    } else {
        // the left contact: the last interval (0, 1), U = 0 at the contact
        const int sid0 = Lv.sid[0];
        const double V0 = contact_V(a, lane, c.Vint, false);
        const double Vg0n = V0 + segv(Lv, sid0, SEG_OFFN), Vg0p = V0 + segv(Lv, sid0, SEG_OFFP);
        const double wn = exp(-Vg0n * iVt) * segv(Lv, sid0, SEG_IMUN);
        const double wp = exp(Vg0p * iVt) * segv(Lv, sid0, SEG_IMUP);
        const double dx = px - Lv.pt[PT_X * n];
        const double u = 0.5 * dx * (0.0 + pU);
        Dn += u * (Sn + 0.5 * dx * pwn);
        Dp += u * (Sp + 0.5 * dx * pwp);
        Sn += 0.5 * dx * (wn + pwn);
        Sp += 0.5 * dx * (wp + pwp);
        iu += u;
        out.endn = exp(SC(SC_LNL) - Vg0n * iVt);
        out.endp = exp(SC(SC_LPL) + Vg0p * iVt);
    }
    out.size = size;
    out.finite = finite;
    out.Sn = Sn; out.Dn = Dn; out.Sp = Sp; out.Dp = Dp; out.iu = iu;
}

// The pair's step size, finiteness and current from both halves (the same order on both
// lanes): the interval between the middle points joins the left sums to the right's, whose
// iu starts from the left part's total (src/jaxdd/batch.py:321-373).
DDH_FN void pair_combine(const Args &a, int lane, const PairBack &Lh, const PairBack &Rh,
                         double &size, double &J, int &finite) {
    const double Vt = SC(SC_VT);
    const double dx = Rh.x - Lh.x;
    const double iu1 = Lh.iu + 0.5 * dx * (Lh.U + Rh.U);
    const double Sn = (Lh.Sn + 0.5 * dx * (Lh.wn + Rh.wn)) + Rh.Sn;
    const double Sp = (Lh.Sp + 0.5 * dx * (Lh.wp + Rh.wp)) + Rh.Sp;
    const double Dn = (Lh.Dn + 0.5 * dx * (Lh.wn * Lh.iu + Rh.wn * iu1)) + (iu1 * Rh.Sn + Rh.Dn);
    const double Dp = (Lh.Dp + 0.5 * dx * (Lh.wp * Lh.iu + Rh.wp * iu1)) + (iu1 * Rh.Sp + Rh.Dp);
    const double Kn = (Vt * (-Rh.endn + Lh.endn) + Dn) / Sn;
    const double Kp = (Vt * (Rh.endp - Lh.endp) - Dp) / Sp;
    J = -SC(SC_ELEC) * Q_E * (Kn + Kp);
    size = nmax(Lh.size, Rh.size);
    finite = Lh.finite != 0.0 && Rh.finite != 0.0;
}

// One Newton step of a pair, in three phases around the two exchanges; `other_*` are the
// partner lane's records of the phase before.
struct PairStep {
    SweepCtl c;
    PairElim e;
    PairBack b;
    double z[2];
};

template <typename T, int BK>
DDH_FN void pair_elim(const Args &a, int lane, int half, bool ions, PairStep &st) {
    if constexpr (sizeof(T) == 4) {
        if (ions) elim_half<float, float, 3, BK>(a, lane, half, st.c, st.e);
        else elim_half<float, float, 1, BK>(a, lane, half, st.c, st.e);
    } else if (a.mixed) {
        if (ions) elim_half<double, float, 3, BK>(a, lane, half, st.c, st.e);
        else elim_half<double, float, 1, BK>(a, lane, half, st.c, st.e);
    } else {
        if (ions) elim_half<double, double, 3, BK>(a, lane, half, st.c, st.e);
        else elim_half<double, double, 1, BK>(a, lane, half, st.c, st.e);
    }
}

template <typename T>
DDH_FN void pair_phase1(const Args &a, int lane, int half, bool ions, lcdouble *lseg,
                        const LaneState &S, PairStep &st) {
    st.c = sweep_ctl(a, lane, S, lseg);
    if (DDH_UNIFORM(a.bern)) pair_elim<T, 1>(a, lane, half, ions, st);
    else pair_elim<T, 0>(a, lane, half, ions, st);
}

template <typename T>
DDH_FN void pair_phase2(const Args &a, int lane, int half, bool ions, PairStep &st,
                        const PairElim &other) {
    double xm[3], xm1[3];
    const PairElim &Lh = half ? other : st.e, &Rh = half ? st.e : other;
    pair_middle(a, lane, st.c, ions ? 3 : 1, Lh, Rh, st.z, xm, xm1);
    st.c.fresh = 0;
    const double *x0 = half ? xm1 : xm;
    if constexpr (sizeof(T) == 4) {
        if (ions) backsub_half<float, float, 3>(a, lane, half, st.c, st.z, x0, st.b);
        else backsub_half<float, float, 1>(a, lane, half, st.c, st.z, x0, st.b);
    } else if (a.mixed) {
        if (ions) backsub_half<double, float, 3>(a, lane, half, st.c, st.z, x0, st.b);
        else backsub_half<double, float, 1>(a, lane, half, st.c, st.z, x0, st.b);
    } else {
        if (ions) backsub_half<double, double, 3>(a, lane, half, st.c, st.z, x0, st.b);
        else backsub_half<double, double, 1>(a, lane, half, st.c, st.z, x0, st.b);
    }
}

DDH_FN void pair_phase3(const Args &a, int lane, int half, LaneState &S, PairStep &st,
                        const PairBack &other) {
    double size, J;
    int finite;
    pair_combine(a, lane, half ? other : st.b, half ? st.b : other, size, J, finite);
    finish_step(a, lane, S, st.z, size, J, finite);
}

}  // namespace ddh
