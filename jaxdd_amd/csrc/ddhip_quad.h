// Four lanes per device, for batches of devices without mobile ions that are too small to fill
// the GPU at two lanes per device: each Newton step's block-tridiagonal solve split into four
// chunks of rows at m1 = NP/4, m2 = NP/2, m3 = 3NP/4.
//
//   lane 0: rows 1..m1,      eliminated from the left contact towards m1
//   lane 1: rows m2..m1+1,   from the middle towards m1, with x_{m2+1} (xi) kept as an unknown
//   lane 2: rows m2+1..m3,   from the middle towards m3, with x_{m2} (eta) kept as an unknown
//   lane 3: rows NP..m3+1,   from the right contact towards m3
//
// An unknown kept is three more right-hand-side columns: x_i = d'_i [1, -s] - c'_i x_next.
// Lanes 0 and 1 meet at (m1, m1+1) and lanes 2 and 3 at (m3, m3+1) as the two lanes of a pair
// do. Lanes 1 and 2 also accumulate, during their elimination, the prefix product of their
// factors (x of their first row as an affine function of the point beyond their last row; in
// the wavefront's local data share, not in registers), so that xi and eta follow from one 3x3
// solve at the middle without another pass over the rows. Each
// lane then back-substitutes its chunk outwards from its meeting point; the current's sums of
// the four chunks join over the three intervals between them, pairwise and then as the two
// lanes of a pair join theirs. Every joint value is computed from the same records in the same
// order on the lanes that use it, so a device's four state machines stay identical. The rows,
// edges, caps and steps are the one-lane kernel's arithmetic; results agree with it to
// round-off.
#pragma once

#include "ddhip_pair.h"

namespace ddh {

// Factors per grid point of the four-lane solve: c' (3x3), the residual column of d' (3) and
// its three columns of the kept unknown (3x3; zero in the chunks of lanes 0 and 3).
constexpr int NFAC_Q = 21;

// The part of one chunk's elimination the other lane of its meeting needs: the factors of its
// last row, d' as [row][column] over the columns (residual, s0, s1, s2).
struct QuadEdge {
    double cf[9];
    double df[12];
};

// One chunk's elimination: its meeting record, and (lanes 1 and 2) the prefix of its rows'
// factors: x_first = acc [1, -s] + R x_beyond.
struct QuadElim {
    QuadEdge m;
    double acc[12];
    double R[9];
};

// One chunk's back-substitution: its step size and finiteness, its current's sums (the forward
// form with the integral of U from zero at the chunk's left end), the point values at both of
// its ends (first: where its back-substitution started; last: where it ended), its contact's
// terms (lanes 0 and 3).
struct QuadBack {
    double size, finite;
    double Sn, Dn, Sp, Dp, iu;
    double wn0, wp0, U0, x0;
    double wn1, wp1, U1, x1;
    double endn, endp;
};

// Lane q's chunk: its first point j0 (a contact, or the point beyond the middle that is kept
// as an unknown), direction of elimination, number of rows and whether it keeps an unknown.
struct QuadChunk {
    int j0, dir, rows, spike;
};

DDH_FN QuadChunk quad_chunk(int NP, int q) {
    const int m1 = NP / 4, m2 = NP / 2, m3 = (3 * NP) / 4;
    QuadChunk c;
    c.spike = (q == 1 || q == 2);
    if (q == 0) { c.j0 = 0; c.dir = 1; c.rows = m1; }
    else if (q == 1) { c.j0 = m2 + 1; c.dir = -1; c.rows = m2 - m1; }
    else if (q == 2) { c.j0 = m2; c.dir = 1; c.rows = m3 - m2; }
    else { c.j0 = NP + 1; c.dir = -1; c.rows = NP - m3; }
    return c;
}

// The smallest grid the four-lane split is defined on (every chunk at least four rows).
constexpr int QUAD_MIN_NP = 16;

#ifdef DDH_CPU_EMULATION
#define DDH_QUADFN inline
#else
#define DDH_QUADFN __device__ __attribute__((noinline))
#endif

// The wavefront's local data share: the 16 devices' segment tables (row stride QUAD_LS), then
// the prefixes of lanes 1 and 2 of each device (QUAD_PRE rows of QUAD_PRE_LS columns): 25 KB
// for seven segments, so that four workgroups (every SIMD of a CU) fit in one CU's LDS. Lane
// q of the device in slot `slot` keeps its prefix at column 2 slot + (q - 1), element e at
// pre[e * QUAD_PRE_LS]. The prefix is read and written once a row, off the row's critical
// path, and in the local data share its waits do not drain the row's memory loads in flight.
constexpr int QUAD_LS = 16;
constexpr int QUAD_PRE = 21;
constexpr int QUAD_PRE_LS = 32;
DDH_MFN size_t quad_lds_doubles(int nseg) {
    return (size_t)nseg * SEG_N * QUAD_LS + (size_t)QUAD_PRE * QUAD_PRE_LS;
}
DDH_FN int quad_pre_column(int nseg, int slot, int q) {
    return nseg * SEG_N * QUAD_LS + 2 * slot + ((q - 1) & 1);
}

// The elimination of one chunk (src/jaxdd/solver.py:1021-1187 rows, :1534-1560 elimination),
// for the Bernoulli function of kind BK (0 exact, 1 SIMsalabim's; the launch's `bern`).
template <typename T, typename S, int BK>
DDH_QUADFN void elim_quarter(const Args &a, int lane, int q, const SweepCtl &cref, ldouble *pre,
                             QuadElim &out) {
    const SweepCtl c = cref;
    const int NP = DDH_UNIFORM(a.NP), bern = BK;
    const QuadChunk ch = quad_chunk(NP, q);
    const int dir = ch.dir, j0 = ch.j0, nvis = ch.rows + 2;
    const bool spike = ch.spike != 0;
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, L = SC(SC_L), L2 = L * L;
    const Lane Lv = lane_view(a, lane, c, QUAD_LS);
    const size_t n = Lv.n;
    DDH_GLOBAL S *fac = (DDH_GLOBAL S *)(factors<S>(a) + lane);
    gdouble *s1 = c.fresh ? Lv.ys1 : Lv.yw;
    gdouble *s2 = (c.fresh && c.ratio != 0.0) ? Lv.ys2 : s1;
    const Contact ct = contact_point(a, lane, Lv, c.Vint, q == 3);

    // x_prev = dp [1, -s] - cp x_this; a chunk that keeps an unknown starts from x_prev = s
    T cp[9], dp[12];
#pragma unroll
    for (int k = 0; k < 9; ++k) cp[k] = T(0);
#pragma unroll
    for (int row = 0; row < 3; ++row)
#pragma unroll
        for (int col = 0; col < 4; ++col)
            dp[4 * row + col] = (spike && col == row + 1) ? T(-1) : T(0);
    if (spike) {  // R = I, acc = 0
#pragma unroll
        for (int k = 0; k < QUAD_PRE; ++k) pre[k * QUAD_PRE_LS] = (k < 9 && k % 4 == 0) ? 1.0 : 0.0;
    }

    Pt mid, far, pn;
    Edge eold = {0, 0, 0, 0, 0, 0, 0, 0};
    Raw nxt, now;
    if (!spike) {
        fetch_point(Lv, s1, s2, j0 + dir, nxt);
        make_point(Lv, j0, true, ct, nxt, Lv.sid[(size_t)j0 * n], c, Vt, iVt, mid);
    } else {
        fetch_point(Lv, s1, s2, j0, now);
        fetch_point(Lv, s1, s2, j0 + dir, nxt);
        make_point(Lv, j0, false, ct, now, now.sid, c, Vt, iVt, mid);
    }
    far = mid;
    // visit v: the next point, the edge to it, the row of the point before (from v = 2); the
    // first visit and the stores its row then overwrites run before the loop, as in a pair
    auto visit = [&](const int v, const bool has_row) {
        const int j = j0 + dir * v;
        now = nxt;
        fetch_point(Lv, s1, s2, v + 1 < nvis ? j + dir : j, nxt);
        make_point(Lv, j, false, ct, now, now.sid, c, Vt, iVt, pn);
        const Pt el = dir > 0 ? mid : pn, er = dir > 0 ? pn : mid;
        const Edge en = edge_flux<BK>(Lv, el, er, Vt, iVt, bern);
        if (has_row) {
            const int i = j - dir;
            const Pt lft = dir > 0 ? far : pn, rgt = dir > 0 ? pn : far;
            const Edge eL = dir > 0 ? eold : en, eR = dir > 0 ? en : eold;
            double A[9], B[9], C[9], r[3];
            row_blocks<1>(Lv, lft, mid, rgt.V, eL, eR, i == NP, L2, iVt, c.share, c.g_frac, A, B,
                          C, r, c.trapq != 0);
            if constexpr (sizeof(T) == 4) {
                // row equilibration by powers of two (exact) before the rounding to FP32
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
                        r[row] *= sc;
                    }
                }
            }
            // N: the coupling to the eliminated side, F: to the side still ahead
            T N[9], F[9], Bt[9];
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                N[k] = (T)(dir > 0 ? A[k] : C[k]);
                F[k] = (T)(dir > 0 ? C[k] : A[k]);
                Bt[k] = (T)B[k];
            }
            T M[9], rh[12];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    M[3 * p + k] = Bt[3 * p + k] - (N[3 * p + 0] * cp[0 + k] + N[3 * p + 1] * cp[3 + k] + N[3 * p + 2] * cp[6 + k]);
#pragma unroll
                for (int col = 0; col < 4; ++col)
                    rh[4 * p + col] = (col == 0 ? (T)r[p] : T(0)) - (N[3 * p + 0] * dp[0 + col] + N[3 * p + 1] * dp[4 + col] + N[3 * p + 2] * dp[8 + col]);
            }
            T Mi[9];
            inverse3(M, Mi);
            T cn_[9], dn_[12];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int k = 0; k < 3; ++k)
                    cn_[3 * p + k] = Mi[3 * p + 0] * F[0 + k] + Mi[3 * p + 1] * F[3 + k] + Mi[3 * p + 2] * F[6 + k];
#pragma unroll
                for (int col = 0; col < 4; ++col)
                    dn_[4 * p + col] = Mi[3 * p + 0] * rh[0 + col] + Mi[3 * p + 1] * rh[4 + col] + Mi[3 * p + 2] * rh[8 + col];
            }
            DDH_GLOBAL S *f = fac + (size_t)(i - 1) * NFAC_Q * n;
#pragma unroll
            for (int k = 0; k < 9; ++k) f[k * n] = (S)cn_[k];
#pragma unroll
            for (int row = 0; row < 3; ++row) f[(9 + row) * n] = (S)dn_[4 * row];
            // (lanes 0 and 3 store their zero columns too: memory operations under a lane
            // condition would leave the compiler's waits for the next point's loads covering
            // this row's stores)
#pragma unroll
            for (int row = 0; row < 3; ++row)
#pragma unroll
                for (int k = 0; k < 3; ++k) f[(12 + 3 * row + k) * n] = (S)dn_[4 * row + 1 + k];
            if (spike) {
                // the prefix: acc += R d'_i, R = -R c'_i
                double R[9];
#pragma unroll
                for (int k = 0; k < 9; ++k) R[k] = pre[k * QUAD_PRE_LS];
#pragma unroll
                for (int p = 0; p < 3; ++p) {
#pragma unroll
                    for (int col = 0; col < 4; ++col)
                        pre[(9 + 4 * p + col) * QUAD_PRE_LS] += R[3 * p] * (double)dn_[col] + R[3 * p + 1] * (double)dn_[4 + col] + R[3 * p + 2] * (double)dn_[8 + col];
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        pre[(3 * p + k) * QUAD_PRE_LS] = -(R[3 * p] * (double)cn_[k] + R[3 * p + 1] * (double)cn_[3 + k] + R[3 * p + 2] * (double)cn_[6 + k]);
                }
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) cp[k] = cn_[k];
#pragma unroll
            for (int k = 0; k < 12; ++k) dp[k] = dn_[k];
        }
        far = mid;
        mid = pn;
        eold = en;
    };
    visit(1, false);
    {
        DDH_GLOBAL S *f = fac + (size_t)(j0 + dir - 1) * NFAC_Q * n;
#pragma unroll
        for (int k = 0; k < NFAC_Q; ++k) f[k * n] = S(0);
    }
    for (int v = 2; v < nvis; ++v) visit(v, true);
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        out.m.cf[k] = (double)cp[k];
        out.R[k] = spike ? pre[k * QUAD_PRE_LS] : (k % 4 == 0 ? 1.0 : 0.0);
    }
#pragma unroll
    for (int k = 0; k < 12; ++k) {
        out.m.df[k] = (double)dp[k];
        out.acc[k] = spike ? pre[(9 + k) * QUAD_PRE_LS] : 0.0;
    }
}

// The two points where the chunks of lanes 2k and 2k + 1 meet, each as a 3x4 map of
// [1, -s]: x_a = D_a - E x_b (lane 2k's last row) and x_b = D_b - Cq x_a (lane 2k + 1's).
DDH_FN void quad_meet(const QuadEdge &ev, const QuadEdge &od, double *Xa, double *Xb) {
    const double *E = ev.cf, *Da = ev.df, *Cq = od.cf, *Db = od.df;
    double Mm[9], Mi[9];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Mm[3 * p + k] = (p == k ? 1.0 : 0.0) - (E[3 * p] * Cq[k] + E[3 * p + 1] * Cq[3 + k] + E[3 * p + 2] * Cq[6 + k]);
    inverse3<double>(Mm, Mi);
#pragma unroll
    for (int col = 0; col < 4; ++col) {
        double rhs[3];
#pragma unroll
        for (int p = 0; p < 3; ++p)
            rhs[p] = Da[4 * p + col] - (E[3 * p] * Db[col] + E[3 * p + 1] * Db[4 + col] + E[3 * p + 2] * Db[8 + col]);
#pragma unroll
        for (int p = 0; p < 3; ++p) Xa[4 * p + col] = Mi[3 * p] * rhs[0] + Mi[3 * p + 1] * rhs[1] + Mi[3 * p + 2] * rhs[2];
#pragma unroll
        for (int p = 0; p < 3; ++p)
            Xb[4 * p + col] = Db[4 * p + col] - (Cq[3 * p] * Xa[col] + Cq[3 * p + 1] * Xa[4 + col] + Cq[3 * p + 2] * Xa[8 + col]);
    }
}

// Lane q's contribution to the middle: lane 1 gives eta = x_{m2} = acc1 + R1 x_{m1} as a map of
// [1, -xi] (P), lane 2 gives xi = x_{m2+1} = acc2 + R2 x_{m3+1} as a map of [1, -eta] (Q).
DDH_FN void quad_middle_part(const QuadElim &e, int q, const double *Xa, const double *Xb,
                             double *W) {
    const double *X = q == 1 ? Xa : Xb;
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int col = 0; col < 4; ++col)
            W[4 * p + col] = (q == 1 || q == 2)
                                 ? e.acc[4 * p + col] + (e.R[3 * p] * X[col] + e.R[3 * p + 1] * X[4 + col] + e.R[3 * p + 2] * X[8 + col])
                                 : 0.0;
}

// The unknowns kept from P and Q (the same arithmetic on every lane):
// (I - P_xi Q_eta) eta = P_0 - P_xi Q_0, xi = Q_0 - Q_eta eta; then lane q's first x of its
// back-substitution (from its meeting's Xa or Xb) and the value of its chunk's unknown.
DDH_FN void quad_center(const double *P, const double *Q, int q, const double *Xa,
                        const double *Xb, double xstart[3], double s[3]) {
    double Mm[9], Mi[9], rhs[3], eta[3], xi[3];
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
        for (int k = 0; k < 3; ++k)
            Mm[3 * p + k] = (p == k ? 1.0 : 0.0) -
                            (P[4 * p + 1] * Q[1 + k] + P[4 * p + 2] * Q[5 + k] + P[4 * p + 3] * Q[9 + k]);
#pragma unroll
    for (int p = 0; p < 3; ++p)
        rhs[p] = P[4 * p] - (P[4 * p + 1] * Q[0] + P[4 * p + 2] * Q[4] + P[4 * p + 3] * Q[8]);
    inverse3<double>(Mm, Mi);
#pragma unroll
    for (int p = 0; p < 3; ++p) eta[p] = Mi[3 * p] * rhs[0] + Mi[3 * p + 1] * rhs[1] + Mi[3 * p + 2] * rhs[2];
#pragma unroll
    for (int p = 0; p < 3; ++p) xi[p] = Q[4 * p] - (Q[4 * p + 1] * eta[0] + Q[4 * p + 2] * eta[1] + Q[4 * p + 3] * eta[2]);
    // lanes 0 and 2 start at their meeting's a point (m1, m3), lanes 1 and 3 at its b point
    const double *X = (q & 1) ? Xb : Xa;
    const double *sv = q < 2 ? xi : eta;
#pragma unroll
    for (int p = 0; p < 3; ++p)
        xstart[p] = X[4 * p] - (X[4 * p + 1] * sv[0] + X[4 * p + 2] * sv[1] + X[4 * p + 3] * sv[2]);
#pragma unroll
    for (int p = 0; p < 3; ++p) s[p] = (q == 1 || q == 2) ? sv[p] : 0.0;
}

// The back-substitution of one chunk, outwards from its meeting point: the step (the first
// point's x given), its caps, the new state, the step size, the current's sums.
template <typename T, typename S>
DDH_QUADFN void backsub_quarter(const Args &a, int lane, int q, const SweepCtl &cref,
                                const double *xfirst, const double *sv, QuadBack &out) {
    const SweepCtl c = cref;
    const int NP = DDH_UNIFORM(a.NP);
    const QuadChunk ch = quad_chunk(NP, q);
    // against the elimination: from the chunk's last row
    const int dir = -ch.dir;
    const int jf = ch.j0 + ch.dir * ch.rows;
    const int cnt = ch.rows;
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, nfl = SC(SC_NFLOOR), lnfloor = log(nfl);
    const double capv = CAP_V * Vt;
    const Lane Lv = lane_view(a, lane, c, QUAD_LS);
    const size_t n = Lv.n;
    DDH_GLOBAL const S *fac = (DDH_GLOBAL const S *)(factors<S>(a) + lane);
    const double s0 = sv[0], s1 = sv[1], s2 = sv[2];
    struct Next {
        S f[NFAC_Q];
        double y0, y1, y2, G, x;
        int sid;
    };
    auto fetch = [&](int j, Next &r) {
        DDH_GLOBAL const S *f = fac + (size_t)(j - 1) * NFAC_Q * n;
#pragma unroll
        for (int k = 0; k < NFAC_Q; ++k) r.f[k] = f[k * n];
        gdouble *y = Lv.yw + (size_t)(j - 1) * 3 * n;
        r.y0 = y[0]; r.y1 = y[n]; r.y2 = y[2 * n];
        r.G = Lv.pt[((size_t)j * PT_N + PT_G) * n];
        r.x = Lv.pt[((size_t)j * PT_N + PT_X) * n];
        r.sid = Lv.sid[(size_t)j * n];
    };
    double xs0 = xfirst[0], xs1 = xfirst[1], xs2 = xfirst[2];
    double size = 0.0;
    int finite = 1;
    double Sn = 0.0, Dn = 0.0, Sp = 0.0, Dp = 0.0, iu = 0.0;
    double pwn = 0.0, pwp = 0.0, pU = 0.0, px = 0.0;
    Next nxt, now;
    fetch(jf, nxt);
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
            for (int row = 0; row < 3; ++row)
                dF[row] = (double)f[9 + row] - ((double)f[12 + 3 * row] * s0 + (double)f[13 + 3 * row] * s1 + (double)f[14 + 3 * row] * s2);
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
            out.wn0 = wn; out.wp0 = wp; out.U0 = U; out.x0 = xj;
        } else if (dir > 0) {
            // left to right: the integral of U from zero at the chunk's left end
            const double dx = xj - px;
            const double iun = iu + 0.5 * dx * (pU + U);
            Sn += 0.5 * dx * (pwn + wn);
            Dn += 0.5 * dx * (pwn * iu + wn * iun);
            Sp += 0.5 * dx * (pwp + wp);
            Dp += 0.5 * dx * (pwp * iu + wp * iun);
            iu = iun;
        } else {
            // right to left: sum_k u_k T_k, T_k the intervals' point weights right of interval k
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
    out.wn1 = pwn; out.wp1 = pwp; out.U1 = pU; out.x1 = px;
    out.endn = out.endp = 0.0;
    if (q == 3) {
        const int sidN = Lv.sid[(size_t)(NP + 1) * n];
        const double VN = contact_V(a, lane, c.Vint, true);
        out.endn = exp(SC(SC_LNR) - (VN + segv(Lv, sidN, SEG_OFFN)) * iVt);
        out.endp = exp(SC(SC_LPR) + (VN + segv(Lv, sidN, SEG_OFFP)) * iVt);
    } else if (q == 0) {
        // the left contact: the interval (0, 1), U = 0 at the contact
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

// Two adjacent chunks' back-substitutions as one half of the grid (a pair lane's record): the
// left half from chunks 0 and 1 (points 0..m2; its end point m2), the right half from chunks 2
// and 3 (points m2+1..NP; its end point m2+1), joined over (m1, m1+1) or (m3, m3+1)
// (src/jaxdd/batch.py:321-373).
DDH_FN void quad_half(const QuadBack &ev, const QuadBack &od, int right, PairBack &h) {
    double IU = right ? 0.0 : ev.iu;
    double Sn = right ? 0.0 : ev.Sn, Dn = right ? 0.0 : ev.Dn;
    double Sp = right ? 0.0 : ev.Sp, Dp = right ? 0.0 : ev.Dp;
    auto chunk = [&](const QuadBack &c) {
        Dn += IU * c.Sn + c.Dn; Sn += c.Sn;
        Dp += IU * c.Sp + c.Dp; Sp += c.Sp;
        IU += c.iu;
    };
    if (right) chunk(ev);  // chunk 2, from m2+1
    // the interval between the chunks' meeting points: chunk 0's m1 and chunk 1's m1+1, or
    // chunk 2's m3 and chunk 3's m3+1 (both their back-substitutions' first points)
    const double dx = od.x0 - ev.x0;
    const double IUb = IU + 0.5 * dx * (ev.U0 + od.U0);
    Sn += 0.5 * dx * (ev.wn0 + od.wn0);
    Dn += 0.5 * dx * (ev.wn0 * IU + od.wn0 * IUb);
    Sp += 0.5 * dx * (ev.wp0 + od.wp0);
    Dp += 0.5 * dx * (ev.wp0 * IU + od.wp0 * IUb);
    IU = IUb;
    chunk(od);
    h.size = nmax(ev.size, od.size);
    h.finite = (ev.finite != 0.0 && od.finite != 0.0) ? 1.0 : 0.0;
    h.Sn = Sn; h.Dn = Dn; h.Sp = Sp; h.Dp = Dp; h.iu = IU;
    // the half's point next to the middle: chunk 1's last (m2), chunk 2's last (m2+1)
    const QuadBack &e = right ? ev : od;
    h.wn = e.wn1; h.wp = e.wp1; h.U = e.U1; h.x = e.x1;
    const QuadBack &k = right ? od : ev;
    h.endn = k.endn; h.endp = k.endp;
}

// One Newton step of a device's four lanes, in three phases around the exchanges of records.
struct QuadStep {
    SweepCtl c;
    QuadElim e;
    QuadBack b;
};

template <typename T>
DDH_FN void quad_phase1(const Args &a, int lane, int q, lcdouble *lseg, ldouble *pre,
                        const LaneState &S, QuadStep &st) {
    st.c = sweep_ctl(a, lane, S, lseg);
    const bool simss = DDH_UNIFORM(a.bern) != 0;
    if constexpr (sizeof(T) == 4) {
        if (simss) elim_quarter<float, float, 1>(a, lane, q, st.c, pre, st.e);
        else elim_quarter<float, float, 0>(a, lane, q, st.c, pre, st.e);
    } else if (a.mixed) {
        if (simss) elim_quarter<double, float, 1>(a, lane, q, st.c, pre, st.e);
        else elim_quarter<double, float, 0>(a, lane, q, st.c, pre, st.e);
    } else {
        if (simss) elim_quarter<double, double, 1>(a, lane, q, st.c, pre, st.e);
        else elim_quarter<double, double, 0>(a, lane, q, st.c, pre, st.e);
    }
}

template <typename T>
DDH_FN void quad_phase2(const Args &a, int lane, int q, QuadStep &st, const double *xs,
                        const double *sv) {
    st.c.fresh = 0;
    if constexpr (sizeof(T) == 4) backsub_quarter<float, float>(a, lane, q, st.c, xs, sv, st.b);
    else if (a.mixed) backsub_quarter<double, float>(a, lane, q, st.c, xs, sv, st.b);
    else backsub_quarter<double, double>(a, lane, q, st.c, xs, sv, st.b);
}

DDH_FN void quad_phase3(const Args &a, int lane, LaneState &S, const PairBack &Lh,
                        const PairBack &Rh) {
    double size, J;
    int finite;
    pair_combine(a, lane, Lh, Rh, size, J, finite);
    const double z[2] = {0.0, 0.0};
    finish_step(a, lane, S, z, size, J, finite);
}

// A device the four-lane kernel does not solve (mobile ions): status 2, no converged voltage.
DDH_FN void quad_refuse(const Args &a, int lane) {
    for (int v = 0; v < a.nV; ++v) {
        a.J[(size_t)lane * a.nV + v] = NAN;
        a.ok[(size_t)lane * a.nV + v] = 0;
        a.its[(size_t)lane * a.nV + v] = 0;
    }
    a.status[lane] = 2;
}

DDH_FN bool quad_refused(const Args &a, int lane) { return SC(SC_AT) > 0 || SC(SC_CT) > 0; }

}  // namespace ddh
