// The adjoint of the steady-state J-V solve: dJ(V)/dtheta at every converged voltage by the
// implicit function theorem (src/jaxdd/batch.py:650-737, `_converged_jvp` / `_tangent`).
//
// At a converged state y*, F(y*, theta) = 0, so dy*/dtheta = -T^-1 dF/dtheta with T = dF/dy
// (the Newton matrix, F the row-scaled residual), and
//     dJ/dtheta = dJ/dtheta|_y - lambda^T dF/dtheta|_y,    T^T lambda = dJ/dy.
// One wavefront per device, one lane per voltage; at each converged voltage:
//   forward pass (left to right)   the integral current's sums at y* (src/jaxdd/batch.py:302-403)
//   elimination (right to left)    T's blocks at y* (the Newton sweep's own row assembly),
//                                  T = U'L' eliminated as the Newton sweep does; dJ/dy by the
//                                  current's reverse mode; w = L'^-T dJ/dy
//   substitution (left to right)   lambda = U'^-T w
// then, per parameter direction k, one pass (left to right) evaluating the residual rows and
// the current in dual numbers along the packed coefficients' tangents dc/dtheta_k, which
// gives dF/dtheta_k|_y (dotted with lambda) and dJ/dtheta_k|_y.
//
// The parameters reach the kernel only as those tangents (the device builder's forward mode),
// so any parameter of the packed arrays can be differentiated. With mobile ions the system is
// bordered by each species' normalisation and number constraint, and the transposed solve
// gives the constraints' multipliers mu beside lambda (adj_lambda): dJ/dtheta then also takes
// -mu . dg/dtheta. Ions frozen at a pre-bias make the scan's fixed charge the pre-bias
// solution's ions: a second pass per voltage solves the same bordered transpose at that
// solution (adj_lambda's PRE) and adds its part; without the pre-bias solutions (pre NULL)
// such a device is refused (status 1, gradients NaN).
#pragma once

#include "ddhip_core.h"

namespace ddh {

// A value and its derivative along one direction.
struct D1 {
    double v, d;
};
DDH_FN D1 dc(double v) { return {v, 0.0}; }
DDH_FN D1 operator+(D1 a, D1 b) { return {a.v + b.v, a.d + b.d}; }
DDH_FN D1 operator-(D1 a, D1 b) { return {a.v - b.v, a.d - b.d}; }
DDH_FN D1 operator-(D1 a) { return {-a.v, -a.d}; }
DDH_FN D1 operator*(D1 a, D1 b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
DDH_FN D1 operator*(double a, D1 b) { return {a * b.v, a * b.d}; }
DDH_FN D1 operator*(D1 a, double b) { return {a.v * b, a.d * b}; }
DDH_FN D1 operator+(D1 a, double b) { return {a.v + b, a.d}; }
DDH_FN D1 operator+(double a, D1 b) { return {a + b.v, b.d}; }
DDH_FN D1 operator-(double a, D1 b) { return {a - b.v, -b.d}; }
DDH_FN D1 operator-(D1 a, double b) { return {a.v - b, a.d}; }
DDH_FN D1 operator/(D1 a, D1 b) {
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
DDH_FN D1 dexp(D1 a) {
    const double e = exp(a.v);
    return {e, e * a.d};
}

struct AdjArgs {
    int n, NP, nV, nseg, ntheta, bern;
    const double *pt;
    const uint8_t *sid;
    const double *seg;
    const double *sc;
    const double *V;        // [nV] (ascending, as the solve's)
    const double *states;   // [nV][NP][3][n] the solve's final states
    const double *pre;      // [NP][3][n] the pre-bias solutions (devices whose ions are frozen
                            // at them; NULL: such devices refused)
    const uint8_t *ok;      // [n][nV]
    const double *dpt;      // [ntheta][NP + 2][PT_N][n]  tangents of the packed arrays
    const double *dseg;     // [ntheta][nseg][SEG_N][n]
    const double *dsc;      // [ntheta][SC_N][n]
    double *grad;           // [n][nV][ntheta]
    int32_t *status;        // [n]
    // Each launched lane's own rows, lane stride ls (the launch's lane count):
    double *fac;            // [NP][ADJ_F][ls] Mi, C, w per row
    double *st;             // [NP][3][ls] the state's V, n, p
    double *lam;            // [NP][3][ls]
    double *iu;             // [NP + 1][ls] the current's running integral of U
    double *ts;             // [NP + 2][2][ls] the current's suffix sums of the point weights
    double *ion;            // [4][ls] the ions' normalisations l_a, l_c and their multipliers
    double *nu;             // [NP][ls] dJ/d(fixed charge) of a scan with frozen ions
    size_t ls;
    int passes;             // 1: one pass per direction even where a reverse pass would do
};
// Per row: Mi (9), C (9), w (3), and with mobile ions each constraint's gamma (2 x 3).
constexpr int ADJ_F = 27;
// Directions contracted in one reverse pass (more go through one pass per direction), and the
// points whose tangents the wavefront stages in LDS at a time for the contraction.
constexpr int ADJ_MAX_K = 16;
constexpr int ADJ_CHUNK = 4;
// The points' segment ids in the table: bytes, (NP + 2 + 7) / 8 doubles' room.
DDH_MFN int adj_sid_words(int NP) { return (NP + 2 + 7) / 8; }
// Workgroups of the adjoint launch (one wavefront each, looping over the devices): the
// per-lane workspace is sized by them, not by the batch.
constexpr int ADJ_SLOTS = 1024;

// A point of a dual pass: its potential and densities (tangents only at the contacts, whose
// values are parameters), its mobile ions' densities (through V_T: the normalisations are
// unknowns, held) and its spacing.
struct PD {
    D1 V, n, p, h, an, ca;
    int sid;
};

// The device's segment field with its tangent along the current direction, from the
// wavefront's table in LDS (one device per wavefront: every lane reads the same entry).
struct SegD {
    lcdouble *val, *tan;
    DDH_FN D1 operator()(int id, int f) const {
        const int o = id * SEG_N + f;
        return {val[o], tan[o]};
    }
};

// The wavefront's LDS table. K = 0 (one pass per direction): the device's segment fields and
// their tangents along the current direction, its scalars and their tangents, its points'
// (h, x, G, fixed ions) and their tangents, and its points' segment ids. K > 0 (the reverse
// pass): the values, the segment ids, then the segment fields' and the scalars' tangents of all
// K directions ([k][field]; the points' tangents are read from global memory).
struct AdjTab {
    int S, sct, ptv, ptt, sid, segT, scT, size;
    DDH_FN AdjTab(int nseg, int NP, int K = 0) {
        S = nseg * SEG_N;
        if (K == 0) {
            sct = 2 * S;
            ptv = sct + 2 * SC_N;
            ptt = ptv + (NP + 2) * PT_N;
            sid = ptt + (NP + 2) * PT_N;
            segT = scT = size = sid + adj_sid_words(NP);
        } else {
            sct = S;
            ptv = sct + SC_N;
            ptt = ptv;
            sid = ptv + (NP + 2) * PT_N;
            segT = sid + adj_sid_words(NP);
            scT = segT + K * S;
            ptt = scT + K * SC_N;  // a chunk of ADJ_CHUNK points' tangents, [k][point][field]
            size = ptt + K * ADJ_CHUNK * PT_N;
        }
    }
};
DDH_MFN int adj_table_size(int nseg, int NP, int K = 0) {
    return K == 0 ? 2 * nseg * SEG_N + 2 * SC_N + 2 * (NP + 2) * PT_N + adj_sid_words(NP)
                  : nseg * SEG_N + SC_N + (NP + 2) * PT_N + adj_sid_words(NP) +
                        K * (nseg * SEG_N + SC_N + ADJ_CHUNK * PT_N);
}
// The directions of one reverse pass: all of them up to ADJ_MAX_K, else none (one pass each).
DDH_MFN int adj_k(int ntheta, int passes) { return (passes || ntheta > ADJ_MAX_K) ? 0 : ntheta; }

DDH_FN D1 recomb_d1(D1 n, D1 p, const SegD &sd, int id) {
    const D1 np_ = n * p;
    D1 R = sd(id, SEG_KD) * (np_ - sd(id, SEG_NI2));
    const D1 Nt = sd(id, SEG_NT);
    if (Nt.v > 0) {
        const D1 Cn = sd(id, SEG_CN), Cp = sd(id, SEG_CP);
        const D1 base = Nt * Cn * Cp;
        const D1 num = base * (np_ - sd(id, SEG_NT0) * sd(id, SEG_PT0));
        const D1 den = Cn * (n + sd(id, SEG_NT0)) + Cp * (p + sd(id, SEG_PT0));
        R = R + num / den;
    }
    return R;
}

// The charged bulk traps' charge along the direction (trap_charge's value, solver.py:762-773).
DDH_FN D1 trap_charge_d1(D1 n, D1 p, const SegD &sd, int id) {
    const double tt = sd(id, SEG_TT).v;
    const D1 Nt = sd(id, SEG_NT);
    if (tt == 0.0 || !(Nt.v > 0)) return dc(0.0);
    const D1 Cn = sd(id, SEG_CN), Cp = sd(id, SEG_CP), pt0 = sd(id, SEG_PT0);
    const D1 f = (Cn * n + Cp * pt0) / (Cn * (n + sd(id, SEG_NT0)) + Cp * (p + pt0));
    return Nt * (dc(tt > 0 ? 1.0 : 0.0) - f);
}

// The edge fluxes Fn, Fp over (l, r) along the direction (src/jaxdd/solver.py:1069-1108).
DDH_FN void edge_d1(const PD &l, const PD &r, const SegD &sd, D1 Vt, D1 iVt, int bern,
                    D1 &Fn, D1 &Fp) {
    const D1 lcn = Vt * sd(l.sid, SEG_MUN), lcp = Vt * sd(l.sid, SEG_MUP);
    const D1 xn = ((r.V + sd(r.sid, SEG_OFFN)) - (l.V + sd(l.sid, SEG_OFFN))) * iVt;
    const D1 xp = ((r.V + sd(r.sid, SEG_OFFP)) - (l.V + sd(l.sid, SEG_OFFP))) * iVt;
    double Bf, Bb, dBf, dBb;
    bernoulli_pair(xn.v, bern, Bf, Bb, dBf, dBb);
    const D1 Bnf = {Bf, dBf * xn.d}, Bnb = {Bb, -dBb * xn.d};
    bernoulli_pair(xp.v, bern, Bf, Bb, dBf, dBb);
    const D1 Bpf = {Bf, dBf * xp.d}, Bpb = {Bb, -dBb * xp.d};
    Fn = lcn * (Bnf * r.n - Bnb * l.n);
    Fp = lcp * (Bpf * l.p - Bpb * r.p);
}

// Row m's scaled residual F along the direction (row_blocks' F).
DDH_FN void row_d1(const PD &l, const PD &m, const PD &r, D1 G, D1 fix, D1 Ld, D1 gfr,
                   const SegD &sd, D1 Vt, D1 iVt, int bern, D1 F[3]) {
    const D1 hm = l.h, hp = m.h, hs = hp + hm, L2 = Ld * Ld;
    const D1 D = dc(1.0) / (hp * hm * hs * L2);
    const D1 fac_ = 0.5 * L2 * hp * hm * hs;
    const D1 meps = sd(m.sid, SEG_EPS);
    const D1 fu = 2.0 * meps * hm * D;
    const D1 fl0 = 2.0 * meps * hp * D;
    const D1 deps = meps - sd(l.sid, SEG_EPS);
    const D1 jump = deps.v != 0.0 ? deps / (hm * hm * L2) : dc(0.0);
    const D1 fac_m = jump - (fu + fl0);
    const D1 fac_l = fl0 - jump;
    const D1 sP = sd(m.sid, SEG_SP);
    const D1 sC = sd(m.sid, SEG_SCS) * (hm * hs * L2 * D);
    const D1 gen = G * gfr;
    const D1 R = recomb_d1(m.n, m.p, sd, m.sid);
    D1 FnL, FpL, FnR, FpR;
    edge_d1(l, m, sd, Vt, iVt, bern, FnL, FpL);
    edge_d1(m, r, sd, Vt, iVt, bern, FnR, FpR);
    const D1 elec = hm * FnR - hp * FnL + fac_ * (gen - R);
    const D1 hole = hp * FpL - hm * FpR + fac_ * (gen - R);
    const D1 pois = (fac_l * l.V + fac_m * m.V + fu * r.V -
                     Q_E * (m.n - m.p + sd(m.sid, SEG_DOP) + fix + m.an - m.ca)) +
                    Q_E * trap_charge_d1(m.n, m.p, sd, m.sid);
    F[0] = pois * sP;
    F[1] = elec * sC;
    F[2] = hole * sC;
}

DDH_FN void inverse3d(const double *m, double *out) { inverse3<double>(m, out); }


// Copies device d's values into the LDS table (lane `lane` of `nl`: the wavefront's lanes
// together; the caller waits for all of them before the table is read).
DDH_FN void adj_stage_values(const AdjArgs &a, int d, ldouble *tab, int lane, int nl) {
    const size_t n = (size_t)a.n;
    const AdjTab t(a.nseg, a.NP, adj_k(a.ntheta, a.passes));
    for (int q = lane; q < t.S; q += nl) tab[q] = a.seg[(size_t)q * n + d];
    for (int q = lane; q < SC_N; q += nl) tab[t.sct + q] = a.sc[(size_t)q * n + d];
    for (int q = lane; q < (a.NP + 2) * PT_N; q += nl) tab[t.ptv + q] = a.pt[(size_t)q * n + d];
    DDH_LDS uint8_t *sb = (DDH_LDS uint8_t *)(tab + t.sid);
    for (int q = lane; q < a.NP + 2; q += nl) sb[q] = a.sid[(size_t)q * n + d];
    if (t.segT < t.size) {  // the reverse pass: every direction's segment and scalar tangents
        const int K = adj_k(a.ntheta, a.passes);
        for (int q = lane; q < K * t.S; q += nl) {
            const int k = q / t.S, f = q - k * t.S;
            tab[t.segT + q] = a.dseg[((size_t)k * t.S + f) * n + d];
        }
        for (int q = lane; q < K * SC_N; q += nl) {
            const int k = q / SC_N, f = q - k * SC_N;
            tab[t.scT + q] = a.dsc[((size_t)k * SC_N + f) * n + d];
        }
    }
}

// Copies direction k's tangents of device d into the LDS table (as adj_stage_values).
DDH_FN void adj_stage_direction(const AdjArgs &a, int d, int k, ldouble *tab, int lane, int nl) {
    const size_t n = (size_t)a.n;
    const AdjTab t(a.nseg, a.NP);
    const double *dsg = a.dseg + (size_t)k * t.S * n + d;
    const double *dsc = a.dsc + (size_t)k * SC_N * n + d;
    const double *dpt = a.dpt + (size_t)k * (a.NP + 2) * PT_N * n + d;
    for (int q = lane; q < t.S; q += nl) tab[t.S + q] = dsg[(size_t)q * n];
    for (int q = lane; q < SC_N; q += nl) tab[t.sct + SC_N + q] = dsc[(size_t)q * n];
    for (int q = lane; q < (a.NP + 2) * PT_N; q += nl) tab[t.ptt + q] = dpt[(size_t)q * n];
}

#define TSC(f) tab[t.sct + (f)]
#define TPT(j, f) tab[t.ptv + (j) * PT_N + (f)]
#define TSID(j) ((int)((DDH_LDS const uint8_t *)(tab + t.sid))[j])
#define TSEG(id, f) tab[(id) * SEG_N + (f)]
#define FAC(i, q) a.fac[((size_t)((i) - 1) * ADJ_F + (q)) * ls + sl]
#define STV(j, f) a.st[((size_t)((j) - 1) * 3 + (f)) * ls + sl]
#define LAM(i, f) a.lam[((size_t)((i) - 1) * 3 + (f)) * ls + sl]
#define IU(j) a.iu[(size_t)(j) * ls + sl]
#define TS(j, f) a.ts[((size_t)(j) * 2 + (f)) * ls + sl]
#define YS(j, f) a.states[(((size_t)v * NP + ((j) - 1)) * 3 + (f)) * n + d]

#define ION(f) a.ion[(size_t)(f) * ls + sl]
#define NU(j) a.nu[(size_t)((j) - 1) * ls + sl]
#define PS(j, f) a.pre[(((size_t)(j) - 1) * 3 + (f)) * n + d]

// Whether device d's mobile ions are frozen at its pre-bias (src/jaxdd/batch.py:859-872,
// 955-956): the scan then has no mobile ions, and its fixed charge is the pre-bias solution's
// ions, a function of that solution and so of the parameters through it.
DDH_FN bool adj_frozen(const AdjArgs &a, int d) {
    const size_t n = (size_t)a.n;
    const bool mobile = a.sc[(size_t)SC_AT * n + d] > 0 || a.sc[(size_t)SC_CT * n + d] > 0;
    return mobile && a.sc[(size_t)SC_PRE * n + d] > 0 && a.sc[(size_t)SC_FIX * n + d] > 0;
}

// Whether device d's gradients are refused: its ions frozen, and no pre-bias solution given.
DDH_FN bool adj_refused(const AdjArgs &a, int d) { return adj_frozen(a, d) && !a.pre; }

// The closed-form normalisations of device d's pre-bias solution (as adj_lambda's of a scan
// state): the logarithms its frozen ions' densities take.
DDH_FN void adj_pre_logs(const AdjArgs &a, int d, lcdouble *tab, double lg[2]) {
    const int NP = a.NP;
    const size_t n = (size_t)a.n;
    const AdjTab t(a.nseg, a.NP, adj_k(a.ntheta, a.passes));
    const double Vt = TSC(SC_VT), iVt = 1.0 / Vt, L = TSC(SC_L), elec = TSC(SC_ELEC);
    const double Vpre = TSC(SC_VPRE);
    const double V0c = TSC(SC_V0) + 0.5 * elec * Vpre, VLc = TSC(SC_VL) - 0.5 * elec * Vpre;
    double num0 = 0.0, num1 = 0.0, Vl = V0c;
    for (int j = 1; j <= NP + 1; ++j) {
        const double Vr = j <= NP ? PS(j, 0) : VLc;
        const int il = TSID(j - 1), ir = TSID(j);
        const double aml = TSEG(il, SEG_AM), amr = TSEG(ir, SEG_AM);
        const double cml = TSEG(il, SEG_CM), cmr = TSEG(ir, SEG_CM);
        const double hL = TPT(j - 1, PT_H) * L;
        num0 += 0.5 * (aml * exp(Vl * iVt) + amr * exp(Vr * iVt)) * (aml * amr) * hL;
        num1 += 0.5 * (cml * exp(-Vl * iVt) + cmr * exp(-Vr * iVt)) * (cml * cmr) * hL;
        Vl = Vr;
    }
    lg[0] = TSC(SC_AT) > 0 ? log(TSC(SC_AT) / num0) : 0.0;
    lg[1] = TSC(SC_CT) > 0 ? log(TSC(SC_CT) / num1) : 0.0;
}

// Point j's frozen ions (the anions' density less the cations'), from the pre-bias solution's
// potential there and its normalisations lg (the solver's freeze: make_point).
DDH_FN double adj_frozen_ions(const AdjArgs &a, int d, lcdouble *tab, int j, const double lg[2],
                              double &an, double &ca) {
    const size_t n = (size_t)a.n;
    const AdjTab t(a.nseg, a.NP, adj_k(a.ntheta, a.passes));
    const double iVt = 1.0 / TSC(SC_VT), V = PS(j, 0);
    const int id = TSID(j);
    const double am = TSEG(id, SEG_AM), cm = TSEG(id, SEG_CM);
    an = am != 0.0 ? am * exp(lg[0] + V * iVt) : 0.0;
    ca = cm != 0.0 ? cm * exp(lg[1] - V * iVt) : 0.0;
    return an - ca;
}

// Device d at converged voltage v, by the launched lane sl: lambda and, with mobile ions, the
// multipliers mu of their number constraints, and the state's V, n, p, into the lane's
// workspace. With mobile ions the system is bordered (src/jaxdd/batch.py:595-618): unknowns y
// and each species' normalisation l, Jacobian [T U; W^T D] (U = dF/dl, W = dg/dy, D = dg/dl
// = 1 at a converged state), and its transpose [T^T W; U^T D] (lambda, mu) = (dJ/dy, 0) is
//     mu = (D - U^T T^-T W)^-1 (-U^T T^-T dJ/dy),    lambda = T^-T (dJ/dy - W mu),
// where U^T T^-T v = v^T X with X = T^-1 U. The one right-to-left elimination gives all of it:
// X's d', the gammas of W (gamma_i = W_i - c'_{i+1}^T gamma_{i+1}, as w of dJ/dy), and by
// sum_i v_i . x_i = sum_i gamma(v)_i . d'_i the contractions (dJ/dy)^T X and W^T X; lambda's
// substitution then takes w - mu_a gamma_a - mu_c gamma_c.
// PRE (a device whose ions are frozen at its pre-bias, after the scan's pass at v): the same
// transposed bordered system at the pre-bias solution, with the right-hand side the scan's
// dJ/d(fixed charge) b_j = q sP_j lambda_j,V (kept in NU) carried through the frozen ions'
// dependence on that solution: b_j (a_j + c_j) / V_T in the potential's rows, and
// (sum b_j a_j, -sum b_j c_j) in the border's, so that the pre-bias solution's part of
// dJ/dtheta is -lambda_pre . dF_pre/dtheta - mu_pre . dg_pre/dtheta (no current's part).
DDH_FN void adj_lambda(const AdjArgs &a, int d, int v, size_t sl, lcdouble *tab,
                       bool pre = false) {
    const AdjTab t(a.nseg, a.NP, adj_k(a.ntheta, a.passes));
    const int NP = a.NP;
    const size_t n = (size_t)a.n, ls = a.ls;
    const double Vt = TSC(SC_VT), iVt = 1.0 / Vt, L = TSC(SC_L), L2 = L * L;
    const double gfr = TSC(SC_GFRAC), elec = TSC(SC_ELEC), cJ = -elec * Q_E;
    const int sid0 = TSID(0), sidN = TSID(NP + 1);
    const double Vint = pre ? TSC(SC_VPRE) : a.V[v];
    const double V0c = TSC(SC_V0) + 0.5 * elec * Vint, VLc = TSC(SC_VL) - 0.5 * elec * Vint;
    const double tot0 = TSC(SC_AT), tot1 = TSC(SC_CT);
    // (all of these the same for every lane of the wavefront: one device)
    const bool frozen = adj_frozen(a, d);
    const bool ions = (tot0 > 0 || tot1 > 0) && (pre || !frozen);  // the pass's mobile ions
    auto ystate = [&](int j, int f) -> double { return pre ? PS(j, f) : YS(j, f); };
    double lpre[2] = {0.0, 0.0};  // a frozen scan's fixed ions: the pre-bias's normalisations
    if (frozen && !pre) {
        adj_pre_logs(a, d, tab, lpre);
        ION(0) = lpre[0];
        ION(1) = lpre[1];
    }
    if (pre) {  // the scan's dJ/d(fixed charge), before its lambda is overwritten
        for (int j = 1; j <= NP; ++j) NU(j) = Q_E * TSEG(TSID(j), SEG_SP) * LAM(j, 0);
    }
    // the current's sums (the Newton sweep's order), the state's densities kept for the passes;
    // with ions each species' number at l = 0 (the constraint's trapezoid, solver.py:1241-1265)
    const double Vg0n = V0c + TSEG(sid0, SEG_OFFN), Vg0p = V0c + TSEG(sid0, SEG_OFFP);
    double wn_prev = exp(-Vg0n * iVt) * TSEG(sid0, SEG_IMUN);
    double wp_prev = exp(Vg0p * iVt) * TSEG(sid0, SEG_IMUP);
    const double endn0 = exp(TSC(SC_LNL) - Vg0n * iVt), endp0 = exp(TSC(SC_LPL) + Vg0p * iVt);
    double U_prev = 0.0, iu_prev = 0.0, x_prev = TPT(0, PT_X);
    double Sn = 0.0, Dn = 0.0, Sp = 0.0, Dp = 0.0;
    double num0 = 0.0, num1 = 0.0, V_prev = V0c;
    auto count = [&](int jl, int jr, double Vl, double Vr) {  // the edge (jl, jr)'s numbers
        const double aml = TSEG(TSID(jl), SEG_AM), amr = TSEG(TSID(jr), SEG_AM);
        const double cml = TSEG(TSID(jl), SEG_CM), cmr = TSEG(TSID(jr), SEG_CM);
        const double hL = TPT(jl, PT_H) * L;
        num0 += 0.5 * (aml * exp(Vl * iVt) + amr * exp(Vr * iVt)) * (aml * amr) * hL;
        num1 += 0.5 * (cml * exp(-Vl * iVt) + cmr * exp(-Vr * iVt)) * (cml * cmr) * hL;
    };
    if (!pre) IU(0) = 0.0;
    for (int j = 1; j <= NP; ++j) {
        const int id = TSID(j);
        const double y0 = ystate(j, 0), nn = exp(ystate(j, 1)), pp = exp(ystate(j, 2));
        STV(j, 0) = y0; STV(j, 1) = nn; STV(j, 2) = pp;
        if (ions) count(j - 1, j, V_prev, y0);
        V_prev = y0;
        if (pre) continue;  // no current at the pre-bias
        const double R = recombination_value(nn, pp, TSEG(id, SEG_KD), TSEG(id, SEG_NI2),
                                             TSEG(id, SEG_NT), TSEG(id, SEG_CN), TSEG(id, SEG_CP),
                                             TSEG(id, SEG_NT0), TSEG(id, SEG_PT0));
        const double U = R - TPT(j, PT_G) * gfr;
        const double wn = exp(-(y0 + TSEG(id, SEG_OFFN)) * iVt) * TSEG(id, SEG_IMUN);
        const double wp = exp((y0 + TSEG(id, SEG_OFFP)) * iVt) * TSEG(id, SEG_IMUP);
        const double xj = TPT(j, PT_X), dx = xj - x_prev;
        const double iu = iu_prev + 0.5 * dx * (U_prev + U);
        Sn += 0.5 * dx * (wn_prev + wn);
        Dn += 0.5 * dx * (wn_prev * iu_prev + wn * iu);
        Sp += 0.5 * dx * (wp_prev + wp);
        Dp += 0.5 * dx * (wp_prev * iu_prev + wp * iu);
        IU(j) = iu;
        wn_prev = wn; wp_prev = wp; U_prev = U; iu_prev = iu; x_prev = xj;
    }
    // the normalisations of the converged state in closed form (its numbers equal the totals:
    // solver.py:1695-1712's refit); a species without ions is held at l = 0
    double lg0 = 0.0, lg1 = 0.0;
    if (ions) {
        count(NP, NP + 1, V_prev, VLc);
        lg0 = tot0 > 0 ? log(tot0 / num0) : 0.0;
        lg1 = tot1 > 0 ? log(tot1 / num1) : 0.0;
        ION(0) = lg0;
        ION(1) = lg1;
    }
    // J = cJ (Kn + Kp): its partials in the four sums (the sums kept for the reverse pass in
    // the suffix sums' unused ends); none at the pre-bias
    double bSn = 0.0, bDn = 0.0, bSp = 0.0, bDp = 0.0;
    if (!pre) {
        const double endnN = exp(TSC(SC_LNR) - (VLc + TSEG(sidN, SEG_OFFN)) * iVt);
        const double endpN = exp(TSC(SC_LPR) + (VLc + TSEG(sidN, SEG_OFFP)) * iVt);
        const double Kn = (Vt * (-endnN + endn0) + Dn) / Sn;
        const double Kp = (Vt * (endpN - endp0) - Dp) / Sp;
        bSn = -cJ * Kn / Sn; bDn = cJ / Sn; bSp = -cJ * Kp / Sp; bDp = -cJ / Sp;
        TS(0, 0) = Sn; TS(0, 1) = Sp; TS(NP + 1, 0) = Dn; TS(NP + 1, 1) = Dp;
    }

    // elimination, right to left, with dJ/dy and w = L'^-T dJ/dy
    auto point = [&](int j, Pt &q) {
        q.sid = TSID(j);
        q.offn = TSEG(q.sid, SEG_OFFN); q.offp = TSEG(q.sid, SEG_OFFP);
        q.mun = TSEG(q.sid, SEG_MUN); q.mup = TSEG(q.sid, SEG_MUP);
        q.h = TPT(j, PT_H);
        if (j == 0 || j == NP + 1) {
            q.V = j == 0 ? V0c : VLc;
            q.n = exp(j == 0 ? TSC(SC_LNL) : TSC(SC_LNR));
            q.p = exp(j == 0 ? TSC(SC_LPL) : TSC(SC_LPR));
            q.G = q.fix = q.R = q.dRn = q.dRp = 0.0;
        } else {
            q.V = STV(j, 0);
            q.n = STV(j, 1);
            q.p = STV(j, 2);
            q.G = TPT(j, PT_G);
            q.fix = TPT(j, PT_FIX);
            if (frozen && !pre) {
                double fa, fc;
                q.fix += adj_frozen_ions(a, d, tab, j, lpre, fa, fc);
            }
            q.R = recombination(q.n, q.p, TSEG(q.sid, SEG_KD), TSEG(q.sid, SEG_NI2),
                                TSEG(q.sid, SEG_NT), TSEG(q.sid, SEG_CN), TSEG(q.sid, SEG_CP),
                                TSEG(q.sid, SEG_NT0), TSEG(q.sid, SEG_PT0), q.dRn, q.dRp);
        }
        q.am = ions ? TSEG(q.sid, SEG_AM) : 0.0;
        q.cm = ions ? TSEG(q.sid, SEG_CM) : 0.0;
        q.an = q.am != 0.0 ? q.am * exp(lg0 + q.V * iVt) : 0.0;
        q.ca = q.cm != 0.0 ? q.cm * exp(lg1 - q.V * iVt) : 0.0;
    };
    Pt right, mid, cur;
    point(NP + 1, right);
    point(NP, mid);
    Edge eR = edge_flux_pt(mid, right, Vt, iVt, a.bern);
    double cpn[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, wn_[3] = {0, 0, 0};
    // the border: X's d' ([row][species]) and the constraints' gammas ([species][row]) of row
    // i + 1, and the contractions (dJ/dy)^T X, W^T X ([constraint][species])
    double dun[3][2] = {{0, 0}, {0, 0}, {0, 0}}, gwn[2][3] = {{0, 0, 0}, {0, 0, 0}};
    double cX[2] = {0, 0}, WX[2][2] = {{0, 0}, {0, 0}};
    double rl[2] = {0, 0};  // the border's right-hand side (the pre-bias pass's)
    const double itot0 = tot0 > 0 ? 1.0 / tot0 : 0.0, itot1 = tot1 > 0 ? 1.0 / tot1 : 0.0;
    double Tn1 = 0.0, Tp1 = 0.0;  // the current's suffix sums from point i + 1
    double dx_next = 0.0;         // no interval after NP in the sums
    for (int i = NP; i >= 1; --i) {
        point(i - 1, cur);
        const Edge eL = edge_flux_pt(cur, mid, Vt, iVt, a.bern);
        double A[9], B[9], C[9], r[3];
        RowSeg g;
        g.meps = TSEG(mid.sid, SEG_EPS);
        g.ceps = TSEG(cur.sid, SEG_EPS);
        g.sP = TSEG(mid.sid, SEG_SP); g.scs = TSEG(mid.sid, SEG_SCS); g.dop = TSEG(mid.sid, SEG_DOP);
        if (TSEG(mid.sid, SEG_TT) != 0.0 && TSEG(mid.sid, SEG_NT) > 0.0)  // charged bulk traps
            g.qt = trap_charge(TSEG(mid.sid, SEG_TT), mid.n, mid.p, TSEG(mid.sid, SEG_NT),
                               TSEG(mid.sid, SEG_CN), TSEG(mid.sid, SEG_CP),
                               TSEG(mid.sid, SEG_NT0), TSEG(mid.sid, SEG_PT0), g.dqn, g.dqp);
        row_blocks_s<1>(g, cur, mid, right.V, eL, eR, i == NP, L2, iVt, 1.0, gfr, A, B, C, r);
        double M[9], Mi[9];
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                M[3 * p + q] = B[3 * p + q] - (C[3 * p] * cpn[q] + C[3 * p + 1] * cpn[3 + q] +
                                               C[3 * p + 2] * cpn[6 + q]);
        inverse3d(M, Mi);
        // the right-hand side: dJ/dy_i by the current's reverse mode; at the pre-bias, the
        // scan's dJ/d(fixed charge) through the frozen ions
        const double xi = TPT(i, PT_X), dx = xi - TPT(i - 1, PT_X);
        double gv[3], Tn = 0.0, Tp = 0.0;
        if (pre) {
            const double b = NU(i);
            gv[0] = b * (mid.an + mid.ca) * iVt;
            gv[1] = gv[2] = 0.0;
            rl[0] += b * mid.an;
            rl[1] -= b * mid.ca;
        } else {
            const double ai = 0.5 * (dx + dx_next);
            const double wn = exp(-(mid.V + mid.offn) * iVt) * TSEG(mid.sid, SEG_IMUN);
            const double wp = exp((mid.V + mid.offp) * iVt) * TSEG(mid.sid, SEG_IMUP);
            Tn = ai * wn + Tn1;
            Tp = ai * wp + Tp1;
            TS(i, 0) = Tn;
            TS(i, 1) = Tp;
            const double bU = bDn * (0.5 * dx * Tn + 0.5 * dx_next * Tn1) +
                              bDp * (0.5 * dx * Tp + 0.5 * dx_next * Tp1);
            const double iu = IU(i);
            gv[0] = (bSn + bDn * iu) * ai * (-wn * iVt) + (bSp + bDp * iu) * ai * (wp * iVt);
            gv[1] = bU * mid.dRn;
            gv[2] = bU * mid.dRp;
        }
        // w_i = g_i - c'_{i+1}^T w_{i+1}
        double w[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) w[q] = gv[q] - (cpn[q] * wn_[0] + cpn[3 + q] * wn_[1] + cpn[6 + q] * wn_[2]);
#pragma unroll
        for (int q = 0; q < 9; ++q) FAC(i, q) = Mi[q];
#pragma unroll
        for (int q = 0; q < 9; ++q) FAC(i, 9 + q) = C[q];
#pragma unroll
        for (int q = 0; q < 3; ++q) FAC(i, 18 + q) = w[q];
        if (ions) {
            // U_i: dF/dl in the potential's row (row_blocks_s' border columns); W_i: dg/dV_i
            // (the elimination sweep's border), both in the potential's row only
            const double U0[2] = {-Q_E * mid.an * g.sP, Q_E * mid.ca * g.sP};
            const double hLm = cur.h * L, hLp = mid.h * L;
            const double W0[2] = {
                0.5 * (mid.am * cur.am * hLm + right.am * mid.am * hLp) * mid.an * iVt * itot0,
                -0.5 * (mid.cm * cur.cm * hLm + right.cm * mid.cm * hLp) * mid.ca * iVt * itot1};
            double rh[3][2], du[3][2], gw[2][3];
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    rh[p][m] = (p == 0 ? U0[m] : 0.0) -
                               (C[3 * p] * dun[0][m] + C[3 * p + 1] * dun[1][m] + C[3 * p + 2] * dun[2][m]);
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int m = 0; m < 2; ++m)
                    du[p][m] = Mi[3 * p] * rh[0][m] + Mi[3 * p + 1] * rh[1][m] + Mi[3 * p + 2] * rh[2][m];
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    gw[k][q] = (q == 0 ? W0[k] : 0.0) -
                               (cpn[q] * gwn[k][0] + cpn[3 + q] * gwn[k][1] + cpn[6 + q] * gwn[k][2]);
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                cX[m] += w[0] * du[0][m] + w[1] * du[1][m] + w[2] * du[2][m];
#pragma unroll
                for (int k = 0; k < 2; ++k)
                    WX[k][m] += gw[k][0] * du[0][m] + gw[k][1] * du[1][m] + gw[k][2] * du[2][m];
            }
#pragma unroll
            for (int k = 0; k < 2; ++k)
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    FAC(i, 21 + 3 * k + q) = gw[k][q];
                    gwn[k][q] = gw[k][q];
                }
#pragma unroll
            for (int p = 0; p < 3; ++p)
#pragma unroll
                for (int m = 0; m < 2; ++m) dun[p][m] = du[p][m];
        }
        // c'_i = Mi A_i
#pragma unroll
        for (int p = 0; p < 3; ++p)
#pragma unroll
            for (int q = 0; q < 3; ++q)
                cpn[3 * p + q] = Mi[3 * p] * A[q] + Mi[3 * p + 1] * A[3 + q] + Mi[3 * p + 2] * A[6 + q];
#pragma unroll
        for (int q = 0; q < 3; ++q) wn_[q] = w[q];
        Tn1 = Tn; Tp1 = Tp; dx_next = dx;
        right = mid; mid = cur; eR = eL;
    }
    // the multipliers: (D - W^T X)^T mu = r_l - (dJ/dy)^T X, D = 1 (a held species' row too:
    // its W is zero, and so its constraint's derivatives)
    double mu0 = 0.0, mu1 = 0.0;
    if (ions) {
        mu0 = rl[0] - cX[0];
        mu1 = rl[1] - cX[1];
        solve2(1.0 - WX[0][0], -WX[1][0], -WX[0][1], 1.0 - WX[1][1], mu0, mu1);
        ION(2) = mu0;
        ION(3) = mu1;
    }
    // lambda_i = Mi_i^T (w_i - mu . gamma_i - C_{i-1}^T lambda_{i-1}), left to right
    {
        double lp[3] = {0, 0, 0}, Cp_[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int i = 1; i <= NP; ++i) {
            double t[3];
#pragma unroll
            for (int q = 0; q < 3; ++q)
                t[q] = FAC(i, 18 + q) - (Cp_[q] * lp[0] + Cp_[3 + q] * lp[1] + Cp_[6 + q] * lp[2]);
            if (ions) {
#pragma unroll
                for (int q = 0; q < 3; ++q) t[q] -= mu0 * FAC(i, 21 + q) + mu1 * FAC(i, 24 + q);
            }
            double Mi[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) Mi[q] = FAC(i, q);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                lp[q] = Mi[q] * t[0] + Mi[3 + q] * t[1] + Mi[6 + q] * t[2];
                LAM(i, q) = lp[q];
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) Cp_[q] = FAC(i, 9 + q);
        }
    }
}

// dJ/dtheta_k = dJ/dtheta_k|_y - lambda . dF/dtheta_k|_y - mu . dg/dtheta_k|_y at device d's
// converged voltage v (adj_lambda's workspace of lane sl), with direction k's tangents staged in
// the LDS table. PRE: the pre-bias solution's part, -lambda . dF/dtheta_k - mu . dg/dtheta_k
// at it (adj_lambda's PRE pass), the pre-bias voltage's tangent included.
DDH_FN double adj_direction(const AdjArgs &a, int d, int v, size_t sl, lcdouble *tab,
                            bool pre = false) {
    const int NP = a.NP;
    const size_t n = (size_t)a.n, ls = a.ls;
    const AdjTab t(a.nseg, a.NP);
    const int S = t.S;
    lcdouble *tsc = tab + t.sct;
    const double elec = tsc[SC_ELEC];
    const double cJ = -elec * Q_E;
    const int sid0 = TSID(0), sidN = TSID(NP + 1);
    const bool frozen = adj_frozen(a, d);
    const bool ions = (tsc[SC_AT] > 0 || tsc[SC_CT] > 0) && (pre || !frozen);
    const double lg0 = ions ? ION(0) : 0.0, lg1 = ions ? ION(1) : 0.0;
    const double lpre[2] = {frozen ? ION(0) : 0.0, frozen ? ION(1) : 0.0};
    {
        SegD sd;
        sd.val = tab;
        sd.tan = tab + S;
        auto ptd = [&](int j, int f) -> D1 {
            return {tab[t.ptv + j * PT_N + f], tab[t.ptt + j * PT_N + f]};
        };
        auto scd = [&](int f) -> D1 { return {tsc[f], tsc[SC_N + f]}; };
        const D1 Ld = scd(SC_L), Vt = scd(SC_VT), iVt = dc(1.0) / Vt, gfr = scd(SC_GFRAC);
        const D1 Vint = pre ? scd(SC_VPRE) : dc(a.V[v]);
        const D1 V0d = scd(SC_V0) + 0.5 * elec * Vint, VLd = scd(SC_VL) - 0.5 * elec * Vint;
        // a frozen scan's fixed charge at point j: the pre-bias solution's ions (its potential
        // and normalisations held), through V_T
        auto fixd = [&](int j) -> D1 {
            D1 f = ptd(j, PT_FIX);
            if (frozen && !pre) {
                const int id = TSID(j);
                const double am = TSEG(id, SEG_AM), cm = TSEG(id, SEG_CM), Vp = PS(j, 0);
                if (am != 0.0) f = f + am * dexp(lpre[0] + Vp * iVt);
                if (cm != 0.0) f = f - cm * dexp(lpre[1] - Vp * iVt);
            }
            return f;
        };
        auto pointd = [&](int j, PD &q) {
            q.sid = TSID(j);
            q.h = ptd(j, PT_H);
            if (j == 0 || j == NP + 1) {
                q.V = j == 0 ? V0d : VLd;
                q.n = dexp(scd(j == 0 ? SC_LNL : SC_LNR));
                q.p = dexp(scd(j == 0 ? SC_LPL : SC_LPR));
            } else {
                q.V = dc(STV(j, 0));
                q.n = dc(STV(j, 1));
                q.p = dc(STV(j, 2));
            }
            const double am = ions ? TSEG(q.sid, SEG_AM) : 0.0, cm = ions ? TSEG(q.sid, SEG_CM) : 0.0;
            q.an = am != 0.0 ? am * dexp(lg0 + q.V * iVt) : dc(0.0);
            q.ca = cm != 0.0 ? cm * dexp(lg1 - q.V * iVt) : dc(0.0);
        };
        // each species' number along the direction, the edge (l, r) weighted by h_l L
        D1 num0 = dc(0.0), num1 = dc(0.0);
        auto count = [&](const PD &l, const PD &r) {
            const D1 hL = l.h * Ld;
            num0 = num0 + 0.5 * (l.an + r.an) * (TSEG(l.sid, SEG_AM) * TSEG(r.sid, SEG_AM)) * hL;
            num1 = num1 + 0.5 * (l.ca + r.ca) * (TSEG(l.sid, SEG_CM) * TSEG(r.sid, SEG_CM)) * hL;
        };
        // the current along the direction
        const D1 Vg0nd = V0d + sd(sid0, SEG_OFFN), Vg0pd = V0d + sd(sid0, SEG_OFFP);
        D1 wnp = dexp(-Vg0nd * iVt) * sd(sid0, SEG_IMUN);
        D1 wpp = dexp(Vg0pd * iVt) * sd(sid0, SEG_IMUP);
        const D1 endn0d = dexp(scd(SC_LNL) - Vg0nd * iVt), endp0d = dexp(scd(SC_LPL) + Vg0pd * iVt);
        D1 Up = dc(0.0), iup = dc(0.0), xp = ptd(0, PT_X);
        D1 Snd = dc(0.0), Dnd = dc(0.0), Spd = dc(0.0), Dpd = dc(0.0);
        double acc = 0.0;  // lambda . dF/dtheta_k
        PD l, m, r;
        pointd(0, l);
        pointd(1, m);
        for (int j = 1; j <= NP; ++j) {
            pointd(j + 1, r);
            if (ions) count(l, m);
            // row j
            D1 F[3];
            row_d1(l, m, r, ptd(j, PT_G), fixd(j), Ld, gfr, sd, Vt, iVt, a.bern, F);
            acc += LAM(j, 0) * F[0].d + LAM(j, 1) * F[1].d + LAM(j, 2) * F[2].d;
            // the current's terms at point j
            const D1 R = recomb_d1(m.n, m.p, sd, m.sid);
            const D1 U = R - ptd(j, PT_G) * gfr;
            const D1 wn = dexp(-(m.V + sd(m.sid, SEG_OFFN)) * iVt) * sd(m.sid, SEG_IMUN);
            const D1 wp = dexp((m.V + sd(m.sid, SEG_OFFP)) * iVt) * sd(m.sid, SEG_IMUP);
            const D1 xj = ptd(j, PT_X), dx = xj - xp;
            const D1 iu = iup + 0.5 * dx * (Up + U);
            Snd = Snd + 0.5 * dx * (wnp + wn);
            Dnd = Dnd + 0.5 * dx * (wnp * iup + wn * iu);
            Spd = Spd + 0.5 * dx * (wpp + wp);
            Dpd = Dpd + 0.5 * dx * (wpp * iup + wp * iu);
            wnp = wn; wpp = wp; Up = U; iup = iu; xp = xj;
            l = m; m = r;
        }
        if (ions) {  // mu . dg/dtheta_k, g_s = number_s / total_s - 1 (a held species: none)
            count(l, m);
            if (tsc[SC_AT] > 0) acc += ION(2) * (num0 / scd(SC_AT)).d;
            if (tsc[SC_CT] > 0) acc += ION(3) * (num1 / scd(SC_CT)).d;
        }
        if (pre) return -acc;  // no current at the pre-bias
        const D1 endnNd = dexp(scd(SC_LNR) - (VLd + sd(sidN, SEG_OFFN)) * iVt);
        const D1 endpNd = dexp(scd(SC_LPR) + (VLd + sd(sidN, SEG_OFFP)) * iVt);
        const D1 Knd = (Vt * (endn0d - endnNd) + Dnd) / Snd;
        const D1 Kpd = (Vt * (endpNd - endp0d) - Dpd) / Spd;
        const D1 J = cJ * (Knd + Kpd);
        return J.d - acc;
    }
}
// dJ/dtheta_k for every direction k < K at device d's converged voltage v, in one reverse pass
// (lambda, the state and the current's sums and suffix sums from adj_lambda): the partials of
// J - lambda . F in every segment field, scalar and point field that the rows and the current
// read, the state held. The segment fields' and scalars' are contracted at once with the K
// directions' tangents (from the LDS table) into gk; the points' are kept in the lane's
// workspace for adj_contract_points. The same derivative as adj_direction's K passes: the rows
// of row_d1 and the current of adj_lambda, reversed.
DDH_FN void adj_contract(const AdjArgs &a, int d, int v, size_t sl, lcdouble *tab, double *gk,
                         bool pre = false) {
    const int NP = a.NP, K = adj_k(a.ntheta, a.passes);
    const size_t n = (size_t)a.n, ls = a.ls;
    const AdjTab t(a.nseg, NP, K);
    const int S = t.S;
    const double Vt = TSC(SC_VT), iVt = 1.0 / Vt, L = TSC(SC_L), L2 = L * L;
    const double gfr = TSC(SC_GFRAC), elec = TSC(SC_ELEC), cJ = -elec * Q_E;
    const int sid0 = TSID(0), sidN = TSID(NP + 1);
    const double Vint = pre ? TSC(SC_VPRE) : a.V[v];
    const double V0c = TSC(SC_V0) + 0.5 * elec * Vint, VLc = TSC(SC_VL) - 0.5 * elec * Vint;
    double bL = 0.0, bVt = 0.0, bgfr = 0.0;  // the partials in L, V_T and the generation's factor
    const bool frozen = adj_frozen(a, d);
    auto seg_now = [&](int id, int f, double b) {
        lcdouble *tg = tab + t.segT + id * SEG_N + f;
#pragma unroll
        for (int k = 0; k < ADJ_MAX_K; ++k)
            if (k < K) gk[k] += b * tg[k * S];
    };
    // a segment's partials are summed over its run of points (a layer's interior) and
    // contracted once the run ends
    double acc[SEG_N];
#pragma unroll
    for (int f = 0; f < SEG_N; ++f) acc[f] = 0.0;
    int run = -1;
    auto flush = [&]() {
        if (run >= 0) {
#pragma unroll
            for (int f = 0; f < SEG_N; ++f) {
                seg_now(run, f, acc[f]);
                acc[f] = 0.0;
            }
        }
    };
    auto seg_add = [&](int id, int f, double b) {
        if (id == run) acc[f] += b;
        else seg_now(id, f, b);
    };
    auto sc_one = [&](int f, double b) {
        lcdouble *tg = tab + t.scT + f;
#pragma unroll
        for (int k = 0; k < ADJ_MAX_K; ++k)
            if (k < K) gk[k] += b * tg[k * SC_N];
    };
    // at the pre-bias the contacts' potentials also move with the pre-bias voltage
    auto sc_add = [&](int f, double b) {
        sc_one(f, b);
        if (pre && (f == SC_V0 || f == SC_VL)) sc_one(SC_VPRE, (f == SC_V0 ? 0.5 : -0.5) * elec * b);
    };
    // the points' partials: into the lane's workspace (the elimination's factors, free once
    // lambda is known), [point 0..NP][field], contracted by adj_contract_points
    auto bp = [&](int j, int f) -> double & { return a.fac[((size_t)j * PT_N + f) * ls + sl]; };
    double pendH = 0.0, pendX = 0.0;  // point j - 1's h and x partials, from rows j - 1, j
    bp(0, PT_G) = 0.0;
    bp(0, PT_FIX) = 0.0;
    // J = cJ (Kn + Kp) and its partials in the sums and the contacts' terms (none at the
    // pre-bias: every partial of the current below carries one of the four seeds)
    const double Vg0n = V0c + TSEG(sid0, SEG_OFFN), Vg0p = V0c + TSEG(sid0, SEG_OFFP);
    double bSn = 0.0, bDn = 0.0, bSp = 0.0, bDp = 0.0;
    if (!pre) {
        const double Sn = TS(0, 0), Sp = TS(0, 1), Dn = TS(NP + 1, 0), Dp = TS(NP + 1, 1);
        const double endn0 = exp(TSC(SC_LNL) - Vg0n * iVt), endp0 = exp(TSC(SC_LPL) + Vg0p * iVt);
        const double endnN = exp(TSC(SC_LNR) - (VLc + TSEG(sidN, SEG_OFFN)) * iVt);
        const double endpN = exp(TSC(SC_LPR) + (VLc + TSEG(sidN, SEG_OFFP)) * iVt);
        const double Kn = (Vt * (-endnN + endn0) + Dn) / Sn;
        const double Kp = (Vt * (endpN - endp0) - Dp) / Sp;
        bSn = -cJ * Kn / Sn; bDn = cJ / Sn; bSp = -cJ * Kp / Sp; bDp = -cJ / Sp;
        const double VgNn = VLc + TSEG(sidN, SEG_OFFN), VgNp = VLc + TSEG(sidN, SEG_OFFP);
        const double iVt2 = iVt * iVt;
        bVt += cJ * ((endn0 - endnN) / Sn + (endpN - endp0) / Sp);  // Kn, Kp's factor V_T
        double e = cJ * Vt / Sn * endn0;  // endn0 = exp(LNL - (V0 + offn) / V_T)
        sc_add(SC_LNL, e); sc_add(SC_V0, -e * iVt); seg_add(sid0, SEG_OFFN, -e * iVt);
        bVt += e * Vg0n * iVt2;
        e = -cJ * Vt / Sp * endp0;        // endp0 = exp(LPL + (V0 + offp) / V_T)
        sc_add(SC_LPL, e); sc_add(SC_V0, e * iVt); seg_add(sid0, SEG_OFFP, e * iVt);
        bVt -= e * Vg0p * iVt2;
        e = -cJ * Vt / Sn * endnN;
        sc_add(SC_LNR, e); sc_add(SC_VL, -e * iVt); seg_add(sidN, SEG_OFFN, -e * iVt);
        bVt += e * VgNn * iVt2;
        e = cJ * Vt / Sp * endpN;
        sc_add(SC_LPR, e); sc_add(SC_VL, e * iVt); seg_add(sidN, SEG_OFFP, e * iVt);
        bVt -= e * VgNp * iVt2;
    }
    // the edge fluxes reversed: partials of bFn Fn + bFp Fp over (l, r) (edge_d1)
    auto edge_rev = [&](int sl_, int sr_, double Vl, double nl, double pl, double Vr, double nr,
                        double pr, bool lc, bool rc, double bFn, double bFp) {
        const double mun = TSEG(sl_, SEG_MUN), mup = TSEG(sl_, SEG_MUP);
        const double lcn = Vt * mun, lcp = Vt * mup;
        const double xn = ((Vr + TSEG(sr_, SEG_OFFN)) - (Vl + TSEG(sl_, SEG_OFFN))) * iVt;
        const double xp = ((Vr + TSEG(sr_, SEG_OFFP)) - (Vl + TSEG(sl_, SEG_OFFP))) * iVt;
        double Bnf, Bnb, dBnf, dBnb, Bpf, Bpb, dBpf, dBpb;
        bernoulli_pair(xn, a.bern, Bnf, Bnb, dBnf, dBnb);
        bernoulli_pair(xp, a.bern, Bpf, Bpb, dBpf, dBpb);
        seg_add(sl_, SEG_MUN, bFn * (Bnf * nr - Bnb * nl) * Vt);
        seg_add(sl_, SEG_MUP, bFp * (Bpf * pl - Bpb * pr) * Vt);
        const double bxn = bFn * lcn * nr * dBnf + bFn * lcn * nl * dBnb;
        const double bxp = bFp * lcp * pl * dBpf + bFp * lcp * pr * dBpb;
        // V_T: the factor V_T mu and the arguments' 1 / V_T
        bVt += bFn * mun * (Bnf * nr - Bnb * nl) + bFp * mup * (Bpf * pl - Bpb * pr) -
               (bxn * xn + bxp * xp) * iVt;
        seg_add(sr_, SEG_OFFN, bxn * iVt); seg_add(sl_, SEG_OFFN, -bxn * iVt);
        seg_add(sr_, SEG_OFFP, bxp * iVt); seg_add(sl_, SEG_OFFP, -bxp * iVt);
        if (lc) {  // the left contact: V0 and its densities
            sc_add(SC_V0, -(bxn + bxp) * iVt);
            sc_add(SC_LNL, -bFn * lcn * Bnb * nl);
            sc_add(SC_LPL, bFp * lcp * Bpf * pl);
        }
        if (rc) {
            sc_add(SC_VL, (bxn + bxp) * iVt);
            sc_add(SC_LNR, bFn * lcn * Bnf * nr);
            sc_add(SC_LPR, -bFp * lcp * Bpb * pr);
        }
    };
    // the current's weights at the left contact (iu = 0 there)
    const double x0 = TPT(0, PT_X), x1 = TPT(1, PT_X);
    {
        const double en = exp(-Vg0n * iVt), ep = exp(Vg0p * iVt);
        const double bwn = bSn * 0.5 * (x1 - x0), bwp = bSp * 0.5 * (x1 - x0);
        const double wn = en * TSEG(sid0, SEG_IMUN), wp = ep * TSEG(sid0, SEG_IMUP);
        seg_add(sid0, SEG_IMUN, bwn * en); seg_add(sid0, SEG_IMUP, bwp * ep);
        sc_add(SC_V0, -bwn * wn * iVt + bwp * wp * iVt);
        seg_add(sid0, SEG_OFFN, -bwn * wn * iVt); seg_add(sid0, SEG_OFFP, bwp * wp * iVt);
        bVt += (bwn * wn * Vg0n - bwp * wp * Vg0p) * iVt * iVt;
    }
    // mobile ions: the rows' densities at the held normalisations, and the constraints'
    // partials, -mu . dg/dtheta with g_s = number_s / total_s - 1 (a held species: none)
    const double tot0 = TSC(SC_AT), tot1 = TSC(SC_CT);
    const bool ions = (tot0 > 0 || tot1 > 0) && (pre || !frozen);  // the pass's mobile ions
    const double lg0 = ions ? ION(0) : 0.0, lg1 = ions ? ION(1) : 0.0;
    const double lpre[2] = {frozen ? ION(0) : 0.0, frozen ? ION(1) : 0.0};
    const double ma = ions && tot0 > 0 ? -ION(2) / tot0 : 0.0;  // -mu_a / total_a
    const double mc = ions && tot1 > 0 ? -ION(3) / tot1 : 0.0;
    double nsum0 = 0.0, nsum1 = 0.0;
    auto ion_an = [&](int id, double V) {
        const double m = TSEG(id, SEG_AM);
        return ions && m != 0.0 ? m * exp(lg0 + V * iVt) : 0.0;
    };
    auto ion_ca = [&](int id, double V) {
        const double m = TSEG(id, SEG_CM);
        return ions && m != 0.0 ? m * exp(lg1 - V * iVt) : 0.0;
    };
    // the edge (jl, jl + 1) of the numbers: its partials in L, V_T and the contacts'
    // potentials; returns the one in h_jl
    auto ion_edge = [&](int jl, double Vl, double Vr) -> double {
        const int il = TSID(jl), ir = TSID(jl + 1);
        const double mma = TSEG(il, SEG_AM) * TSEG(ir, SEG_AM);
        const double mmc = TSEG(il, SEG_CM) * TSEG(ir, SEG_CM);
        const double anl = ion_an(il, Vl), anr = ion_an(ir, Vr);
        const double cal = ion_ca(il, Vl), car = ion_ca(ir, Vr);
        const double h = TPT(jl, PT_H), hL = h * L;
        const double na = 0.5 * (anl + anr) * mma, nc = 0.5 * (cal + car) * mmc;
        nsum0 += na * hL;
        nsum1 += nc * hL;
        bL += (ma * na + mc * nc) * h;
        // d an / d V_T = -an V / V_T^2, d ca / d V_T = ca V / V_T^2
        bVt += 0.5 * hL * iVt * iVt *
               (ma * mma * (-anl * Vl - anr * Vr) + mc * mmc * (cal * Vl + car * Vr));
        if (jl == 0) sc_add(SC_V0, 0.5 * hL * iVt * (ma * mma * anl - mc * mmc * cal));
        if (jl == NP) sc_add(SC_VL, 0.5 * hL * iVt * (ma * mma * anr - mc * mmc * car));
        return (ma * na + mc * nc) * L;
    };
    double wn_prev = exp(-Vg0n * iVt) * TSEG(sid0, SEG_IMUN);
    double wp_prev = exp(Vg0p * iVt) * TSEG(sid0, SEG_IMUP);
    double U_prev = 0.0, iu_prev = 0.0, x_prev = x0;
    for (int j = 1; j <= NP; ++j) {
        const bool lc = j == 1, rc = j == NP;
        const int sl_ = TSID(j - 1), sm = TSID(j), sr_ = TSID(j + 1);
        if (sm != run) {
            flush();
            run = sm;
        }
        const double Vl = lc ? V0c : STV(j - 1, 0), nl = lc ? exp(TSC(SC_LNL)) : STV(j - 1, 1);
        const double pl = lc ? exp(TSC(SC_LPL)) : STV(j - 1, 2);
        const double Vm = STV(j, 0), nm = STV(j, 1), pm = STV(j, 2);
        const double Vr = rc ? VLc : STV(j + 1, 0), nr = rc ? exp(TSC(SC_LNR)) : STV(j + 1, 1);
        const double pr = rc ? exp(TSC(SC_LPR)) : STV(j + 1, 2);
        const double hm = TPT(j - 1, PT_H), hp = TPT(j, PT_H), hs = hp + hm;
        // the row's values (row_d1)
        const double fac_ = 0.5 * L2 * hp * hm * hs;
        const double meps = TSEG(sm, SEG_EPS), deps = meps - TSEG(sl_, SEG_EPS);
        const double fu = 2.0 * meps / (hp * hs * L2), fl0 = 2.0 * meps / (hm * hs * L2);
        const double jump = deps != 0.0 ? deps / (hm * hm * L2) : 0.0;
        const double fac_m = jump - (fu + fl0), fac_l = fl0 - jump;
        const double sP = TSEG(sm, SEG_SP), scs = TSEG(sm, SEG_SCS), sC = scs / hp;
        const double G = TPT(j, PT_G), gen = G * gfr;
        const double kd = TSEG(sm, SEG_KD), ni2 = TSEG(sm, SEG_NI2), Nt = TSEG(sm, SEG_NT);
        const double Cn = TSEG(sm, SEG_CN), Cp = TSEG(sm, SEG_CP);
        const double nt0 = TSEG(sm, SEG_NT0), pt0 = TSEG(sm, SEG_PT0);
        const double R = recombination_value(nm, pm, kd, ni2, Nt, Cn, Cp, nt0, pt0);
        double FnL, FpL, FnR, FpR;
        {
            double Bf, Bb, dBf, dBb;
            const double xnL = ((Vm + TSEG(sm, SEG_OFFN)) - (Vl + TSEG(sl_, SEG_OFFN))) * iVt;
            bernoulli_pair(xnL, a.bern, Bf, Bb, dBf, dBb);
            FnL = Vt * TSEG(sl_, SEG_MUN) * (Bf * nm - Bb * nl);
            const double xpL = ((Vm + TSEG(sm, SEG_OFFP)) - (Vl + TSEG(sl_, SEG_OFFP))) * iVt;
            bernoulli_pair(xpL, a.bern, Bf, Bb, dBf, dBb);
            FpL = Vt * TSEG(sl_, SEG_MUP) * (Bf * pl - Bb * pm);
            const double xnR = ((Vr + TSEG(sr_, SEG_OFFN)) - (Vm + TSEG(sm, SEG_OFFN))) * iVt;
            bernoulli_pair(xnR, a.bern, Bf, Bb, dBf, dBb);
            FnR = Vt * TSEG(sm, SEG_MUN) * (Bf * nr - Bb * nm);
            const double xpR = ((Vr + TSEG(sr_, SEG_OFFP)) - (Vm + TSEG(sm, SEG_OFFP))) * iVt;
            bernoulli_pair(xpR, a.bern, Bf, Bb, dBf, dBb);
            FpR = Vt * TSEG(sm, SEG_MUP) * (Bf * pm - Bb * pr);
        }
        const double elec_ = hm * FnR - hp * FnL + fac_ * (gen - R);
        const double hole = hp * FpL - hm * FpR + fac_ * (gen - R);
        const double anm = ion_an(sm, Vm), cam = ion_ca(sm, Vm);
        double fixm = TPT(j, PT_FIX), fan = 0.0, fca = 0.0;  // a frozen scan's: with the ions
        if (frozen && !pre) fixm += adj_frozen_ions(a, d, tab, j, lpre, fan, fca);
        const double ttq = TSEG(sm, SEG_TT), Ntq = TSEG(sm, SEG_NT);
        const bool charged = ttq != 0.0 && Ntq > 0.0;  // charged bulk traps (solver.py:1136)
        double qt = 0.0;
        if (charged) {
            double dqn_, dqp_;
            qt = trap_charge(ttq, nm, pm, Ntq, TSEG(sm, SEG_CN), TSEG(sm, SEG_CP),
                             TSEG(sm, SEG_NT0), TSEG(sm, SEG_PT0), dqn_, dqp_);
        }
        const double pois = (fac_l * Vl + fac_m * Vm + fu * Vr -
                             Q_E * (nm - pm + TSEG(sm, SEG_DOP) + fixm + anm - cam)) +
                            Q_E * qt;
        // the row reversed, seeded with -lambda_j
        const double b0 = -LAM(j, 0), b1 = -LAM(j, 1), b2 = -LAM(j, 2);
        const double bpois = b0 * sP;
        if (charged) {  // Nt (cwe - f), f = N / D: its partials in the level's fields
            const double Cn = TSEG(sm, SEG_CN), Cp = TSEG(sm, SEG_CP);
            const double nt0 = TSEG(sm, SEG_NT0), pt0 = TSEG(sm, SEG_PT0);
            const double N = Cn * nm + Cp * pt0, D = Cn * (nm + nt0) + Cp * (pm + pt0);
            const double iD2 = 1.0 / (D * D), bq = Q_E * bpois;
            seg_add(sm, SEG_NT, bq * ((ttq > 0 ? 1.0 : 0.0) - N / D));
            seg_add(sm, SEG_CN, -bq * Ntq * (nm * D - N * (nm + nt0)) * iD2);
            seg_add(sm, SEG_CP, -bq * Ntq * (pt0 * D - N * (pm + pt0)) * iD2);
            seg_add(sm, SEG_NT0, bq * Ntq * N * Cn * iD2);
            seg_add(sm, SEG_PT0, -bq * Ntq * Cp * (D - N) * iD2);
        }
        seg_add(sm, SEG_SP, b0 * pois);
        const double belec = b1 * sC, bhole = b2 * sC, bsC = b1 * elec_ + b2 * hole;
        seg_add(sm, SEG_SCS, bsC / hp);
        double bhp = -bsC * scs / (hp * hp), bhm = 0.0;
        const double bfac_l = bpois * Vl, bfac_m = bpois * Vm;
        double bfu = bpois * Vr;
        if (lc) sc_add(SC_V0, bpois * fac_l);
        if (rc) sc_add(SC_VL, bpois * fu);
        seg_add(sm, SEG_DOP, -Q_E * bpois);
        bp(j, PT_FIX) = -Q_E * bpois;
        bVt += bpois * Q_E * (anm + cam) * Vm * iVt * iVt;
        // the frozen ions through V_T (the pre-bias solution held; through it: the PRE pass)
        if (frozen && !pre) bVt += -Q_E * bpois * (-(fan + fca) * PS(j, 0) * iVt * iVt);
        const double bjump = bfac_m - bfac_l;
        bfu -= bfac_m;
        const double bfl0 = bfac_l - bfac_m;
        if (deps != 0.0) {
            const double bdeps = bjump / (hm * hm * L2);
            seg_add(sm, SEG_EPS, bdeps);
            seg_add(sl_, SEG_EPS, -bdeps);
            bhm += -2.0 * bjump * jump / hm;
            bL += -2.0 * bjump * jump / L;
        }
        seg_add(sm, SEG_EPS, bfu * 2.0 / (hp * hs * L2) + bfl0 * 2.0 / (hm * hs * L2));
        bhp += -bfu * fu * (1.0 / hp + 1.0 / hs) - bfl0 * fl0 / hs;
        bhm += -bfu * fu / hs - bfl0 * fl0 * (1.0 / hm + 1.0 / hs);
        bL += -2.0 * (bfu * fu + bfl0 * fl0) / L;
        bhm += belec * FnR - bhole * FpR;
        bhp += -belec * FnL + bhole * FpL;
        const double bfac = (belec + bhole) * (gen - R);
        bhp += bfac * 0.5 * L2 * hm * (hs + hp);
        bhm += bfac * 0.5 * L2 * hp * (hs + hm);
        bL += bfac * L * hp * hm * hs;
        double bR = -(belec + bhole) * fac_;
        double bG = (belec + bhole) * fac_ * gfr;
        edge_rev(sl_, sm, Vl, nl, pl, Vm, nm, pm, lc, false, -belec * hp, bhole * hp);
        edge_rev(sm, sr_, Vm, nm, pm, Vr, nr, pr, false, rc, belec * hm, -bhole * hm);
        // the current at point j (adj_lambda's sums)
        const double xj = TPT(j, PT_X), dx = xj - x_prev;
        const double dxn = j < NP ? TPT(j + 1, PT_X) - xj : 0.0;
        const double aj = 0.5 * (dx + dxn);
        const double en = exp(-(Vm + TSEG(sm, SEG_OFFN)) * iVt), ep = exp((Vm + TSEG(sm, SEG_OFFP)) * iVt);
        const double wn = en * TSEG(sm, SEG_IMUN), wp = ep * TSEG(sm, SEG_IMUP);
        const double U = R - G * gfr, iu = pre ? 0.0 : IU(j);
        const double Tn = pre ? 0.0 : TS(j, 0), Tp = pre ? 0.0 : TS(j, 1);
        const double Tn1 = (!pre && j < NP) ? TS(j + 1, 0) : 0.0;
        const double Tp1 = (!pre && j < NP) ? TS(j + 1, 1) : 0.0;
        const double bwn = (bSn + bDn * iu) * aj, bwp = (bSp + bDp * iu) * aj;
        seg_add(sm, SEG_IMUN, bwn * en); seg_add(sm, SEG_IMUP, bwp * ep);
        seg_add(sm, SEG_OFFN, -bwn * wn * iVt); seg_add(sm, SEG_OFFP, bwp * wp * iVt);
        bVt += (bwn * wn * (Vm + TSEG(sm, SEG_OFFN)) - bwp * wp * (Vm + TSEG(sm, SEG_OFFP))) * iVt * iVt;
        const double bU = bDn * (0.5 * dx * Tn + 0.5 * dxn * Tn1) + bDp * (0.5 * dx * Tp + 0.5 * dxn * Tp1);
        bR += bU;
        bG -= bU * gfr;
        bgfr += ((belec + bhole) * fac_ - bU) * G;
        // interval j (x_{j-1}, x_j)
        const double bdx = bSn * 0.5 * (wn_prev + wn) + bDn * (0.5 * (wn_prev * iu_prev + wn * iu) + 0.5 * (U_prev + U) * Tn) +
                           bSp * 0.5 * (wp_prev + wp) + bDp * (0.5 * (wp_prev * iu_prev + wp * iu) + 0.5 * (U_prev + U) * Tp);
        bp(j, PT_G) = bG;
        bp(j - 1, PT_H) = pendH + bhm + (ions ? ion_edge(j - 1, Vl, Vm) : 0.0);
        bp(j - 1, PT_X) = pendX - bdx;
        pendH = bhp;
        pendX = bdx;
        // the recombination's parameters (recomb_d1), the state held
        const double np_ = nm * pm;
        seg_add(sm, SEG_KD, bR * (np_ - ni2));
        seg_add(sm, SEG_NI2, -bR * kd);
        if (Nt > 0) {
            const double X = np_ - nt0 * pt0, den = Cn * (nm + nt0) + Cp * (pm + pt0);
            const double T = Nt * Cn * Cp * X / den;
            seg_add(sm, SEG_NT, bR * Cn * Cp * X / den);
            seg_add(sm, SEG_CN, bR * (Nt * Cp * X / den - T * (nm + nt0) / den));
            seg_add(sm, SEG_CP, bR * (Nt * Cn * X / den - T * (pm + pt0) / den));
            seg_add(sm, SEG_NT0, bR * (-Nt * Cn * Cp * pt0 / den - T * Cn / den));
            seg_add(sm, SEG_PT0, bR * (-Nt * Cn * Cp * nt0 / den - T * Cp / den));
        }
        wn_prev = wn; wp_prev = wp; U_prev = U; iu_prev = iu; x_prev = xj;
    }
    bp(NP, PT_H) = pendH + (ions ? ion_edge(NP, STV(NP, 0), VLc) : 0.0);
    bp(NP, PT_X) = pendX;
    flush();
    if (ions) {
        if (tot0 > 0) sc_add(SC_AT, ION(2) * nsum0 / (tot0 * tot0));
        if (tot1 > 0) sc_add(SC_CT, ION(3) * nsum1 / (tot1 * tot1));
    }
    sc_add(SC_L, bL);
    sc_add(SC_VT, bVt);
    sc_add(SC_GFRAC, bgfr);
}

// The points' part of the contraction, for points c0 .. c0 + ADJ_CHUNK - 1 (at most NP): their
// partials from the lane's workspace against the chunk's tangents the wavefront staged in LDS.
DDH_FN void adj_contract_points(const AdjArgs &a, int c0, size_t sl, lcdouble *tab, double *gk) {
    const int NP = a.NP, K = adj_k(a.ntheta, a.passes);
    const size_t ls = a.ls;
    const AdjTab t(a.nseg, NP, K);
    for (int jj = 0; jj < ADJ_CHUNK && c0 + jj <= NP; ++jj) {
#pragma unroll
        for (int f = 0; f < PT_N; ++f) {
            const double b = a.fac[((size_t)(c0 + jj) * PT_N + f) * ls + sl];
            lcdouble *tg = tab + t.ptt + jj * PT_N + f;
#pragma unroll
            for (int k = 0; k < ADJ_MAX_K; ++k)
                if (k < K) gk[k] += b * tg[k * ADJ_CHUNK * PT_N];
        }
    }
}

// Stages the tangents of points c0 .. c0 + ADJ_CHUNK - 1 of every direction (as
// adj_stage_values: lane `lane` of `nl`).
DDH_FN void adj_stage_points(const AdjArgs &a, int d, int c0, ldouble *tab, int lane, int nl) {
    const int NP = a.NP, K = adj_k(a.ntheta, a.passes);
    const size_t n = (size_t)a.n;
    const AdjTab t(a.nseg, NP, K);
    for (int q = lane; q < K * ADJ_CHUNK * PT_N; q += nl) {
        const int k = q / (ADJ_CHUNK * PT_N), r = q - k * ADJ_CHUNK * PT_N;
        const int j = c0 + r / PT_N, f = r % PT_N;
        tab[t.ptt + q] = j <= NP + 1 ? a.dpt[(((size_t)k * (NP + 2) + j) * PT_N + f) * n + d] : 0.0;
    }
}
#undef TS

#undef TSC
#undef TPT
#undef TSID
#undef TSEG
#undef FAC
#undef STV
#undef LAM
#undef IU
#undef YS
#undef ION


}  // namespace ddh
