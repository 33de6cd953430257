// Per-lane logic of the batched J-V solve, shared by the HIP kernels (ddhip_solve.hip) and a
// test-only CPU build of the same source (tests/kernel_emu), which replays each wavefront's
// lanes in turn so that the kernels' algorithm can be debugged and checked on a host.
//
// One lane per device runs that device's whole sequence -- Poisson start, the start solves
// with the ions' stages, the pre-bias and its freezing, and the voltage scan -- so no lane
// waits for another device's Newton iterations (the reference's vmap runs every voltage until
// its slowest lane has converged, src/jaxdd/batch.py:740-848, docs/batching.md).
//
// Each Newton step is two sweeps over the grid:
//   elimination (right to left): the state's Scharfetter-Gummel residual and 3x3 Jacobian
//     blocks in FP64 (src/jaxdd/solver.py:1021-1187), eliminated block by block as they are
//     made, the factors (c', d') written in T (double, or float on row-equilibrated blocks);
//     the ion border's Schur complement (src/jaxdd/batch.py:595-603) closes at the sweep's
//     end through an adjoint recursion gamma_i = W_i - c'_{i+1}^T gamma_{i+1}, so that
//     W^T x = sum_i gamma_i^T d'_i needs no second pass;
//   back-substitution (left to right): the step, its caps (solver.py:59-88), the new state,
//     the step size, and the integral current of the new state in the reference's prefix
//     order (src/jaxdd/batch.py:302-403).
#pragma once

#include <stdint.h>

#ifdef DDH_CPU_EMULATION
#include <cmath>
#include <cstring>
#define DDH_FN inline
#define DDH_CONST static const
#define DDH_GLOBAL
#define DDH_LDS
#define DDH_UNIFORM(x) (x)
using std::isfinite;
#else
#include <hip/hip_runtime.h>
#define DDH_FN __device__ inline
#define DDH_CONST __constant__
// Pointers into device memory: global address space, so that a sweep compiled as a function
// of its own loads and stores with global (not flat) instructions.
#define DDH_GLOBAL __attribute__((address_space(1)))
// The wavefront's copy of its lanes' segment tables: local data share (LDS).
#define DDH_LDS __attribute__((address_space(3)))
// A value every lane of the wavefront holds, moved to a scalar register: the sweeps are
// functions of their own, whose arguments arrive in vector registers, and a grid index or
// row stride kept there would make every address a vector computation.
#define DDH_UNIFORM(x) __builtin_amdgcn_readfirstlane(x)
#endif

#include "ddhip_layout.h"
#include "ddhip_math.h"

namespace ddh {

// The charged bulk traps' charge density (m^-3) at a point and its derivatives in ln n and
// ln p (src/jaxdd/solver.py:762-773 `_trap_charge`, one level; device.py:940-948): Nt (cwe -
// f), cwe 1 for donors (tt 1) and 0 for acceptors (tt -1), f = (Cn n + Cp pt0) / (Cn (n + nt0)
// + Cp (p + pt0)) the level's steady occupation.
DDH_FN double trap_charge(double tt, double n, double p, double Nt, double Cn, double Cp,
                          double nt0, double pt0, double &dqn, double &dqp) {
    const double num = Cn * n + Cp * pt0, den = Cn * (n + nt0) + Cp * (p + pt0);
    const double f = num / den, cwe = tt > 0 ? 1.0 : 0.0;
    const double iden2 = 1.0 / (den * den);
    dqn = -Nt * n * Cn * (Cn * nt0 + Cp * p) * iden2;
    dqp = Nt * p * Cp * num * iden2;
    return Nt * (cwe - f);
}

// ---------------------------------------------------------------------------------------
// Kernel arguments and per-lane state
// ---------------------------------------------------------------------------------------
enum Phase { PH_F = 0, PH_STAGE, PH_PRE, PH_SCAN, PH_DONE, PH_NEUTRAL };

struct LaneState {
    int phase, stage, j, below, it, maxit, err;
    int b_pois, b_f, b_c, b_work, b_last, b_before, b_start, b_src1, b_src2;
    int fresh, freeze_now, mobile_now, mobile0, ok_f, fin_f, ok_c, pre_ok, fin_work;
    int noise_state;  // jv mode: 0 no noise floor yet, 1 pending, 2 known
    int need64;       // mixed precision: this solve has switched to FP64 factors
    int part, parts;  // this lane's share of its device's state copies (lanes per device)
    int charged;      // the device holds charged bulk traps
    int neutral;      // this solve leaves their charge out (the start's first solve)
    double Vint, share, ratio, first, V_last, V_before, before_share, noise_floor;
    double lw[2], lpois[2], lf[2], lc[2], llast[2], lbefore[2], lstart[2];
    double hist[4], Jh[4];
    double z[2];
};

struct Args {
    int n, NP, nV, nseg;
    const double *pt;
    const uint8_t *sid;
    const double *seg;
    const double *sc;
    const double *V;
    double *Y;        // [NBUF][NP][3][n]
    // Factors, one region per element type: lane-innermost arrays of different element sizes
    // must not share memory (lane L's 4-byte element would overlay lane L/2's 8-byte one, and
    // wavefronts run their phases and precisions concurrently).
    double *fac64;    // [NP][NFAC][n] FP64 Newton factors
    float *fac32;     // [NP][NFAC][n] FP32 Newton factors
    double *pfac;     // [NP][4][n] Poisson's scalar factors
    double *fix;      // [NP][n] net fixed ions in use (frozen after the pre-bias)
    LaneState *st;    // [n]
    double *states;   // [nV][NP][3][n] each voltage's final state (NULL: not kept)
    double *pre;      // [NP][3][n] the pre-bias solution of a device with a pre-bias (NULL:
                      // not kept)
    double *J;        // [n][nV]
    uint8_t *ok;
    int32_t *its;
    int32_t *status;
    double tol;
    int max_iter, start_iter, strict, bern, jv;
    int mixed;  // FP32 factors for a scan solve's first MIXED_STEPS steps
};

// Mixed precision: the first Newton step of each voltage of the scan (the correction of the
// extrapolated start) factorises in FP32 on row-equilibrated blocks; every other step (the
// steps that decide convergence, src/jaxdd/batch.py:793-818: the quadratic and settled stops
// assume true Newton; and the start solves) factorises in FP64. Every step stores its factors
// c', d' in FP32, and every back-substitution runs in FP64 from them: an inexact Newton whose
// FP64 residual still drives each solve to the tolerance. On this GPU an FP32 operation costs
// what an FP64 one does (no packed math), so FP32 arithmetic saves no time; the stored
// factors are half the sweeps' memory traffic, and the sweeps wait on memory. More FP32
// factorisations (measured: up to 4 steps, or until a step falls below 1e-2) cost Newton steps
// where the Jacobian is ill-conditioned (high forward bias).
constexpr int MIXED_STEPS = 1;

#define SC(f) a.sc[(size_t)(f) * a.n + lane]
#define PT(j, f) a.pt[((size_t)(j) * PT_N + (f)) * a.n + lane]
#define SID(j) a.sid[(size_t)(j) * a.n + lane]
#define YB(b, j, f) a.Y[(((size_t)(b) * a.NP + ((j) - 1)) * 3 + (f)) * a.n + lane]
#define FIX(j) a.fix[(size_t)((j) - 1) * a.n + lane]

struct Seg {
    int id;
    double v[SEG_N];
};

DDH_FN void seg_at(const Args &a, int lane, int j, Seg &s) {
    int id = SID(j);
    if (id != s.id) {
        s.id = id;
#pragma unroll
        for (int f = 0; f < SEG_N; ++f) s.v[f] = a.seg[((size_t)id * SEG_N + f) * a.n + lane];
    }
}

// ---------------------------------------------------------------------------------------
// Poisson's start: src/jaxdd/batch.py:486-503 with src/jaxdd/solver.py:1657-1800
// (quasi-Fermi levels held, the ions re-fitted each step, the bordered step damped as a
// whole) and the linear start src/jaxdd/batch.py:276-299. Scalar tridiagonal per lane.
// ---------------------------------------------------------------------------------------
struct IonSum {  // trapezoidal number of each species over its region (solver.py:1241-1252)
    double num[2];
};

DDH_FN double contact_V(const Args &a, int lane, double Vint, bool right) {
    return right ? SC(SC_VL) - 0.5 * SC(SC_ELEC) * Vint : SC(SC_V0) + 0.5 * SC(SC_ELEC) * Vint;
}

// Closed-form Boltzmann normalisations from the number integrals at logs = 0
// (solver.py:1695-1712).
DDH_FN void refit_logs(const double num0[2], const double tot[2], double out[2]) {
    for (int k = 0; k < 2; ++k) out[k] = tot[k] > 0 ? log(tot[k] / num0[k]) : 0.0;
}

DDH_FN void poisson_lane(const Args &a, int lane, LaneState &S) {
    const int NP = a.NP;
    const double Vt = SC(SC_VT), L = SC(SC_L), L2 = L * L;
    const double tot[2] = {SC(SC_AT), SC(SC_CT)};
    const bool mobile = tot[0] > 0 || tot[1] > 0;
    // the pre-bias, else the first voltage (the lowest in batch mode, where V is sorted)
    const double first = SC(SC_PRE) > 0 ? SC(SC_VPRE) : a.V[0];
    const double Vl = contact_V(a, lane, first, false), Vr = contact_V(a, lane, first, true);
    const double elec = SC(SC_ELEC);
    double *cfac = a.pfac;  // [NP][4][n]: c', d' (3 right-hand sides)
#define PF(j, f) cfac[((size_t)((j) - 1) * 4 + (f)) * a.n + lane]

    // the linear start and the ions' numbers at logs = 0 (for the first refit)
    Seg sg; sg.id = -1;
    double num0[2] = {0.0, 0.0};
    double prevV = Vl, prevam = 0, prevcm = 0, prevh = 0;
    {
        seg_at(a, lane, 0, sg);
        prevam = sg.v[SEG_AM]; prevcm = sg.v[SEG_CM]; prevh = PT(0, PT_H);
    }
    for (int j = 1; j <= NP + 1; ++j) {
        seg_at(a, lane, j, sg);
        double V;
        if (j <= NP) {
            double x = PT(j, PT_X);
            V = SC(SC_V0) + 0.5 * elec * first + x * (SC(SC_VL) - SC(SC_V0) - elec * first) / L;
            double EF = SC(SC_WL) - x * (SC(SC_WL) - SC(SC_WR)) / L;
            double ln_n = log(sg.v[SEG_NC]) - (EF - sg.v[SEG_EC]) / Vt;
            YB(0, j, 0) = V;
            YB(0, j, 1) = ln_n;
            YB(0, j, 2) = 2 * log(sg.v[SEG_NI]) - ln_n;
            FIX(j) = PT(j, PT_FIX);
        } else {
            V = Vr;
        }
        if (mobile) {
            double am = sg.v[SEG_AM], cm = sg.v[SEG_CM];
            double hL = prevh * L;
            num0[0] += 0.5 * (am * exp(V / Vt) + prevam * exp(prevV / Vt)) * (am * prevam) * hL;
            num0[1] += 0.5 * (cm * exp(-V / Vt) + prevcm * exp(-prevV / Vt)) * (cm * prevcm) * hL;
            prevam = am; prevcm = cm;
        }
        prevV = V;
        prevh = PT(j, PT_H);
    }
    double logs[2] = {0.0, 0.0};
    if (mobile) refit_logs(num0, tot, logs);

    double size = INFINITY;
    int it = 0;
    while (size > POISSON_STOP && it < a.start_iter) {
        // elimination, right to left: m_i = mid_i - up_i c'_{i+1}, c'_i = lo_i/m_i, d'_i =
        // (r_i - up_i d'_{i+1})/m_i for r = (-F, U_anion, U_cation); gamma for the border.
        double cprev = 0.0, dprev[3] = {0.0, 0.0, 0.0}, gam[2] = {0.0, 0.0};
        double accF[2] = {0.0, 0.0}, accU[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        double numk[2] = {0.0, 0.0};
        Seg s2; s2.id = -1;
        // the right neighbour's values, rolled along: row i is assembled at its own point,
        // its left neighbour's potential read ahead
        seg_at(a, lane, NP + 1, s2);
        double Vp = Vr, amp = s2.v[SEG_AM], cmp = s2.v[SEG_CM];
        double ap = amp * exp(logs[0] + Vp / Vt), cp_ = cmp * exp(logs[1] - Vp / Vt);
        for (int i = NP; i >= 1; --i) {
            seg_at(a, lane, i, s2);
            double V = YB(0, i, 0), ln_n = YB(0, i, 1), ln_p = YB(0, i, 2);
            double n = exp(ln_n), p = exp(ln_p);
            double h = PT(i, PT_H), hm = PT(i - 1, PT_H);
            double Vm1 = (i == 1) ? Vl : YB(0, i - 1, 0);
            double eps = s2.v[SEG_EPS];
            const int idm = SID(i - 1);  // the left neighbour's segment
            double epsm = (idm == s2.id) ? eps : a.seg[((size_t)idm * SEG_N + SEG_EPS) * a.n + lane];
            double hs = h + hm;
            double fac_m = (eps - epsm) / (hm * hm * L2) - 2 * eps * (1 / (h * hs * L2) + 1 / (hm * hs * L2));
            double fac_l = 2 * eps / (hm * hs * L2) - (eps - epsm) / (hm * hm * L2);
            double fac_u = 2 * eps / (h * hs * L2);
            double sP = 1.0 / (Q_E * s2.v[SEG_NC]);
            double am = s2.v[SEG_AM], cm = s2.v[SEG_CM];
            double an = am * exp(logs[0] + V / Vt), ca = cm * exp(logs[1] - V / Vt);
            double F = fac_l * Vm1 + fac_m * V + fac_u * Vp -
                       Q_E * (n - p + s2.v[SEG_DOP] + FIX(i) + an - ca);
            double dtq = 0.0;  // the charged traps' charge, the densities moving with V
            if (s2.v[SEG_TT] != 0.0 && s2.v[SEG_NT] > 0.0) {  // (solver.py:1672-1673)
                double dqn, dqp;
                F = F + Q_E * trap_charge(s2.v[SEG_TT], n, p, s2.v[SEG_NT], s2.v[SEG_CN],
                                          s2.v[SEG_CP], s2.v[SEG_NT0], s2.v[SEG_PT0], dqn, dqp);
                dtq = Q_E * (dqn - dqp) / Vt;
            }
            double lo = (i == 1) ? 0.0 : fac_l * sP;
            double up = (i == NP) ? 0.0 : fac_u * sP;
            double mid = (fac_m - Q_E * (n / Vt + p / Vt + an / Vt + ca / Vt) + dtq) * sP;
            double r0 = -F * sP, r1 = -Q_E * an * sP * am, r2 = Q_E * ca * sP * cm;
            double m = mid - up * cprev;
            double rm = 1.0 / m;
            double c = lo * rm;
            double d0 = (r0 - up * dprev[0]) * rm;
            double d1 = (r1 - up * dprev[1]) * rm;
            double d2 = (r2 - up * dprev[2]) * rm;
            PF(i, 0) = c; PF(i, 1) = d0; PF(i, 2) = d1; PF(i, 3) = d2;
            // the constraints' gradient in V_i and the numbers (interval i, i+1)
            double hL = h * L, hLm = hm * L;
            double amm = (idm == s2.id) ? am : a.seg[((size_t)idm * SEG_N + SEG_AM) * a.n + lane];
            double cmm = (idm == s2.id) ? cm : a.seg[((size_t)idm * SEG_N + SEG_CM) * a.n + lane];
            double w0 = 0.0, w1 = 0.0;
            if (mobile) {
                double wsa = 0.5 * (am * amm * hLm + amp * am * hL);
                double wsc = 0.5 * (cm * cmm * hLm + cmp * cm * hL);
                w0 = tot[0] > 0 ? wsa * an * (1.0 / Vt) / tot[0] : 0.0;
                w1 = tot[1] > 0 ? wsc * ca * (-1.0 / Vt) / tot[1] : 0.0;
                numk[0] += 0.5 * (ap + an) * (amp * am) * hL;
                numk[1] += 0.5 * (cp_ + ca) * (cmp * cm) * hL;
            }
            gam[0] = w0 - cprev * gam[0];
            gam[1] = w1 - cprev * gam[1];
            accF[0] += gam[0] * d0; accF[1] += gam[1] * d0;
            accU[0][0] += gam[0] * d1; accU[0][1] += gam[0] * d2;
            accU[1][0] += gam[1] * d1; accU[1][1] += gam[1] * d2;
            cprev = c; dprev[0] = d0; dprev[1] = d1; dprev[2] = d2;
            Vp = V; amp = am; cmp = cm; ap = an; cp_ = ca;
        }
        double z0 = 0.0, z1 = 0.0;
        if (mobile) {
            // the interval (0, 1) of the numbers
            seg_at(a, lane, 0, s2);
            double am0 = s2.v[SEG_AM], cm0 = s2.v[SEG_CM], hL0 = PT(0, PT_H) * L;
            numk[0] += 0.5 * (ap + am0 * exp(logs[0] + Vl / Vt)) * (amp * am0) * hL0;
            numk[1] += 0.5 * (cp_ + cm0 * exp(logs[1] - Vl / Vt)) * (cmp * cm0) * hL0;
            double g[2], D[2];
            for (int k = 0; k < 2; ++k) {
                bool has = tot[k] > 0;
                g[k] = has ? numk[k] / tot[k] - 1.0 : 0.0;
                D[k] = has ? numk[k] / tot[k] : 1.0;
            }
            z0 = -g[0] - accF[0];
            z1 = -g[1] - accF[1];
            solve2(D[0] - accU[0][0], -accU[0][1], -accU[1][0], D[1] - accU[1][1], z0, z1);
        }
        // back-substitution, left to right; with mobile ions first the step's largest entry
        double scale = 1.0;
        if (mobile) {
            double xprev = 0.0, mx = 0.0;
            for (int i = 1; i <= NP; ++i) {
                double x = PF(i, 1) - (PF(i, 2) * z0 + PF(i, 3) * z1) - PF(i, 0) * xprev;
                mx = nmax(mx, fabs(x));
                xprev = x;
            }
            double mm = mx / Vt;
            scale = mm > 0 ? fmin(log1p(mm), CAP_V) / mm : 1.0;
            if (mm != mm) scale = mm;
        }
        double xprev = 0.0, sz = 0.0, cap = CAP_V * Vt;
        double nnum[2] = {0.0, 0.0};
        Seg s4; s4.id = -1;
        seg_at(a, lane, 0, s4);
        double pV = Vl, pam = s4.v[SEG_AM], pcm = s4.v[SEG_CM], ph = PT(0, PT_H);
        for (int i = 1; i <= NP + 1; ++i) {
            double V;
            seg_at(a, lane, i, s4);
            if (i <= NP) {
                double x = PF(i, 1) - (PF(i, 2) * z0 + PF(i, 3) * z1) - PF(i, 0) * xprev;
                xprev = x;
                double dV = mobile ? x * scale : clipd(x, -cap, cap);
                double shift = dV / Vt;
                V = YB(0, i, 0) + dV;
                YB(0, i, 0) = V;
                YB(0, i, 1) += shift;
                YB(0, i, 2) -= shift;
                sz = nmax(sz, fabs(shift));
            } else {
                V = Vr;
            }
            if (mobile) {
                double am = s4.v[SEG_AM], cm = s4.v[SEG_CM], hL = ph * L;
                nnum[0] += 0.5 * (am * exp(V / Vt) + pam * exp(pV / Vt)) * (am * pam) * hL;
                nnum[1] += 0.5 * (cm * exp(-V / Vt) + pcm * exp(-pV / Vt)) * (cm * pcm) * hL;
                pam = am; pcm = cm;
            }
            pV = V;
            ph = PT(i, PT_H);
        }
        if (mobile) refit_logs(nnum, tot, logs);
        size = sz;
        ++it;
    }
#undef PF
    S.lpois[0] = logs[0];
    S.lpois[1] = logs[1];
    S.first = first;
    S.b_pois = 0;
    S.mobile0 = mobile ? 1 : 0;
    S.charged = 0;
    S.neutral = 0;
    for (int q = 0; q < a.nseg; ++q)
        if (a.seg[((size_t)q * SEG_N + SEG_TT) * a.n + lane] != 0.0) S.charged = 1;
    a.status[lane] = 0;
}

// ---------------------------------------------------------------------------------------
// The Newton sweeps.
// ---------------------------------------------------------------------------------------
// The values of one grid point a sweep carries from visit to visit, with the material
// parameters its edges and the ion border read, copied from the lane's LDS segment table
// when the point is made (one batch of LDS reads, one wait); a row reads the rest of its
// point's fields in a second batch.
struct Pt {
    double V, n, p;        // the state
    double h;
    double R, dRn, dRp;    // recombination and its derivatives in ln n, ln p
    double G, fix;
    double an, ca;         // mobile ion densities (0 where none or frozen)
    double offn, offp, mun, mup;  // the segment's fields of the same names
    double am, cm;         // the ions' region masks (0 where none or frozen)
    int sid;               // the point's segment
};

struct Edge {  // Scharfetter-Gummel fluxes over one interval and their derivatives
    double Fn, Fp, gn, gp, an0, an1, ap0, ap1;
};

// The segment table of a wavefront's lanes in LDS, copied once per launch: field f of
// segment s of the lane at lseg[(s * SEG_N + f) * ls], ls the wavefront's device slots (64
// for one lane per device, 16 for four). A segment change then reloads from
// LDS, whose wait counter is not the one global stores count on, so the reload does not wait
// for the sweep's outstanding factor stores.
constexpr int WAVE = 64;
typedef DDH_LDS double ldouble;
typedef DDH_LDS const double lcdouble;

DDH_FN void fill_segments(const Args &a, int lane, ldouble *lseg, int ls = WAVE) {
    for (int q = 0; q < a.nseg * SEG_N; ++q) lseg[(size_t)q * ls] = a.seg[(size_t)q * a.n + lane];
}

struct SweepCtl {  // what one lane's sweeps need of its state
    double Vint, share, ratio, g_frac;
    double logs[2];     // normalisations at the state being assembled
    double lstart[2];   // the pre-bias's (for the freeze)
    int fresh, freeze_now, mobile, src1, src2, work;
    int trapq;          // the charged bulk traps' charge in the Poisson rows
    lcdouble *lseg;     // the lane's segment table in LDS
};

// Per-lane views of the arrays a sweep walks: each advances by its row stride per point.
typedef DDH_GLOBAL const double gcdouble;
typedef DDH_GLOBAL double gdouble;
typedef DDH_GLOBAL const uint8_t gcuint8;

struct Lane {
    gcdouble *pt;   // + j*PT_N*n + f*n
    gcuint8 *sid;   // + j*n
    lcdouble *seg;  // + (id*SEG_N + f)*ls (LDS)
    gdouble *yw, *ys1, *ys2;  // state buffers: + (j-1)*3n + f*n
    gdouble *fix;   // + (j-1)*n
    size_t n;
    int ls;         // the LDS table's row stride: the wavefront's device slots (a constant of
                    // each sweep, so that field offsets are immediates)
};

DDH_FN Lane lane_view(const Args &a, int lane, const SweepCtl &c, int ls = WAVE) {
    Lane L;
    L.ls = ls;
    const size_t n = (size_t)DDH_UNIFORM(a.n), ny = (size_t)DDH_UNIFORM(a.NP) * 3 * n;
    L.n = n;
    L.pt = (gcdouble *)(a.pt + lane);
    L.sid = (gcuint8 *)(a.sid + lane);
    L.seg = c.lseg;
    L.yw = (gdouble *)(a.Y + (size_t)c.work * ny + lane);
    L.ys1 = (gdouble *)(a.Y + (size_t)c.src1 * ny + lane);
    L.ys2 = (gdouble *)(a.Y + (size_t)c.src2 * ny + lane);
    L.fix = (gdouble *)(a.fix + lane);
    return L;
}

// Field f of this lane's segment id (LDS).
DDH_FN double segv(const Lane &L, int id, int f) { return L.seg[(id * SEG_N + f) * L.ls]; }

// The segment fields a point carries (q.sid set); the ion masks only while ions move.
DDH_FN void seg_fields(const Lane &L, bool mobile, Pt &q) {
    const int w = L.ls;
    lcdouble *s = L.seg + q.sid * SEG_N * w;
    q.offn = s[SEG_OFFN * w]; q.offp = s[SEG_OFFP * w];
    q.mun = s[SEG_MUN * w]; q.mup = s[SEG_MUP * w];
    const double am = s[SEG_AM * w], cm = s[SEG_CM * w];
    q.am = mobile ? am : 0.0;
    q.cm = mobile ? cm : 0.0;
}

// The segment fields the integral current reads at a point, in one batch of LDS reads.
struct CurSeg {
    double kd, ni2, nt, cn, cp, nt0, pt0, offn, offp, imun, imup;
};

DDH_FN CurSeg cur_seg(const Lane &L, int sid) {
    const int w = L.ls;
    lcdouble *s = L.seg + sid * SEG_N * w;
    CurSeg g;
    g.kd = s[SEG_KD * w]; g.ni2 = s[SEG_NI2 * w]; g.nt = s[SEG_NT * w];
    g.cn = s[SEG_CN * w]; g.cp = s[SEG_CP * w]; g.nt0 = s[SEG_NT0 * w];
    g.pt0 = s[SEG_PT0 * w]; g.offn = s[SEG_OFFN * w]; g.offp = s[SEG_OFFP * w];
    g.imun = s[SEG_IMUN * w]; g.imup = s[SEG_IMUP * w];
    return g;
}

// The fluxes over the interval from l to r (src/jaxdd/solver.py:1069-1108); the interval's
// mobilities are those of l's segment (the interval after the point). KIND: the Bernoulli
// function's kind fixed at compile time (0 exact, 1 SIMsalabim's; a sweep instantiated for
// one kind carries no code or registers of the other), or -1 for `bern` at run time.
template <int KIND = -1>
DDH_FN Edge edge_flux_pt(const Pt &l, const Pt &r, double Vt, double iVt, int bern) {
    if (KIND >= 0) bern = KIND;
    Edge e;
    const double lcn = Vt * l.mun, lcp = Vt * l.mup;
    double xn = ((r.V + r.offn) - (l.V + l.offn)) * iVt;
    double xp = ((r.V + r.offp) - (l.V + l.offp)) * iVt;
    double Bnf, Bnb, dBnf, dBnb, Bpf, Bpb, dBpf, dBpb;
    bernoulli_pair(xn, bern, Bnf, Bnb, dBnf, dBnb);
    bernoulli_pair(xp, bern, Bpf, Bpb, dBpf, dBpb);
    e.an0 = -lcn * Bnb * l.n;
    e.an1 = lcn * Bnf * r.n;
    e.Fn = e.an1 + e.an0;
    e.gn = lcn * (dBnf * r.n + dBnb * l.n) * iVt;
    e.ap0 = lcp * Bpf * l.p;
    e.ap1 = -lcp * Bpb * r.p;
    e.Fp = e.ap0 + e.ap1;
    e.gp = lcp * (dBpf * l.p + dBpb * r.p) * iVt;
    return e;
}

template <int KIND = -1>
DDH_FN Edge edge_flux(const Lane &L, const Pt &l, const Pt &r, double Vt, double iVt, int bern) {
    (void)L;
    return edge_flux_pt<KIND>(l, r, Vt, iVt, bern);
}

// An interior point's loads. A sweep issues them one point ahead of their use, so that their
// latency passes while the point before is computed: with one wavefront per SIMD nothing
// else would hide it.
struct Raw {
    double h, G, fix, y0, y1, y2, o0, o1, o2;  // o: the older state of the start's extrapolation
    int sid;
};

DDH_FN void fetch_point(const Lane &L, gdouble *s1, gdouble *s2, int j, Raw &r) {
    const size_t n = L.n, o = (size_t)(j - 1) * 3 * n;
    r.sid = L.sid[(size_t)j * n];
    r.h = L.pt[((size_t)j * PT_N + PT_H) * n];
    r.G = L.pt[((size_t)j * PT_N + PT_G) * n];
    r.fix = L.fix[(size_t)(j - 1) * n];
    r.y0 = s1[o]; r.y1 = s1[o + n]; r.y2 = s1[o + 2 * n];
    r.o0 = s2[o]; r.o1 = s2[o + n]; r.o2 = s2[o + 2 * n];
}

// A contact's values, read before a sweep's loop.
struct Contact {
    double V, n, p, h;
};

DDH_FN Contact contact_point(const Args &a, int lane, const Lane &L, double Vint, bool right) {
    Contact ct;
    const int j = right ? a.NP + 1 : 0;
    ct.h = L.pt[((size_t)j * PT_N + PT_H) * L.n];
    ct.V = contact_V(a, lane, Vint, right);
    ct.n = exp(right ? SC(SC_LNR) : SC(SC_LNL));
    ct.p = exp(right ? SC(SC_LPR) : SC(SC_LPL));
    return ct;
}

// Point j's values: interior points from their loads (forming a solve's start where `fresh`),
// the contacts from their boundary values.
DDH_FN void make_point(const Lane &L, int j, bool contact, const Contact &ct, const Raw &r,
                       int sid, const SweepCtl &c, double Vt, double iVt, Pt &q) {
    const size_t n = L.n;
    q.sid = sid;
    seg_fields(L, c.mobile, q);
    const int w = L.ls;
    lcdouble *s = L.seg + sid * SEG_N * w;
    const double kd = s[SEG_KD * w], ni2 = s[SEG_NI2 * w], nt = s[SEG_NT * w];
    const double cn = s[SEG_CN * w], cpp = s[SEG_CP * w], nt0 = s[SEG_NT0 * w];
    const double pt0 = s[SEG_PT0 * w];
    const double am = q.am, cm = q.cm;
    if (contact) {
        q.h = ct.h;
        q.V = ct.V;
        q.G = 0.0; q.fix = 0.0;
        q.n = ct.n;
        q.p = ct.p;
        q.R = q.dRn = q.dRp = 0.0;
    } else {
        q.h = r.h;
        double y0 = r.y0, y1 = r.y1, y2 = r.y2;
        q.fix = r.fix;
        if (c.fresh) {
            if (c.freeze_now) {
                // the pre-bias's mobile ions as fixed ones (src/jaxdd/batch.py:859-872)
                q.fix = L.pt[((size_t)j * PT_N + PT_FIX) * n] +
                        segv(L, sid, SEG_AM) * exp(c.lstart[0] + y0 * iVt) -
                        segv(L, sid, SEG_CM) * exp(c.lstart[1] - y0 * iVt);
                L.fix[(size_t)(j - 1) * n] = q.fix;
            }
            if (c.ratio != 0.0) {
                double capv = CAP_V * Vt;
                y0 += clipd(c.ratio * (y0 - r.o0), -capv, capv);
                y1 += clipd(c.ratio * (y1 - r.o1), -CAP_N, CAP_N);
                y2 += clipd(c.ratio * (y2 - r.o2), -CAP_N, CAP_N);
            }
            const size_t o = (size_t)(j - 1) * 3 * n;
            L.yw[o] = y0; L.yw[o + n] = y1; L.yw[o + 2 * n] = y2;
        }
        q.V = y0;
        q.G = r.G;
        q.n = exp(y1);
        q.p = exp(y2);
        q.R = recombination(q.n, q.p, kd, ni2, nt, cn, cpp, nt0, pt0, q.dRn, q.dRp);
    }
    q.an = am != 0.0 ? am * exp(c.logs[0] + q.V * iVt) : 0.0;
    q.ca = cm != 0.0 ? cm * exp(c.logs[1] - q.V * iVt) : 0.0;
}

template <typename T>
DDH_FN T *factors(const Args &a);
template <>
DDH_FN double *factors<double>(const Args &a) { return a.fac64; }
template <>
DDH_FN float *factors<float>(const Args &a) { return a.fac32; }

template <typename T>
DDH_FN void to_t(const double *in, T *out, int m) {
#pragma unroll
    for (int q = 0; q < m; ++q) out[q] = (T)in[q];
}

#ifdef DDH_CPU_EMULATION
#define DDH_SWEEP
#else
// Each sweep variant is its own function, so that its register allocation is its own and the
// state machine's values wait in memory, not in registers, while a sweep runs.
#define DDH_SWEEP __device__ __attribute__((noinline))
#endif

// Row i's Jacobian blocks of the scaled residual F (A: point i-1, B: point i, C: point i+1)
// and the Newton right-hand side -F, with the ion border's two columns where K = 3
// (src/jaxdd/solver.py:1021-1187, src/jaxdd/batch.py:518-549). cur, mid: points i-1, i; rV:
// point i+1's potential; eL, eR: the edges (i-1, i) and (i, i+1); last: i = NP, whose
// coupling to the contact is not an unknown.
// The segment fields a row reads: the permittivities of its point and of the point before,
// and its point's Poisson row scale, continuity row scale and doping.
struct RowSeg {
    double meps, ceps, sP, scs, dop;
    double qt = 0.0, dqn = 0.0, dqp = 0.0;  // the charged traps' charge, d/d ln n, d/d ln p
};

template <int K>
DDH_FN void row_blocks_s(const RowSeg &g, const Pt &cur, const Pt &mid, double rV,
                         const Edge &eL, const Edge &eR, bool last, double L2, double iVt,
                         double share, double g_frac, double *A, double *B, double *C,
                         double *r) {
    const double meps = g.meps, ceps = g.ceps, sP = g.sP, scs = g.scs, dop = g.dop;
    const double hm = cur.h, hp = mid.h, hs = hp + hm;
    const double D = 1.0 / (hp * hm * hs * L2);
    const double fac_ = 0.5 * L2 * hp * hm * hs;
    const double fu = 2 * meps * hm * D;   // 2 eps / (h_i (h_i + h_{i-1}) L^2)
    const double fl0 = 2 * meps * hp * D;  // 2 eps / (h_{i-1} (h_i + h_{i-1}) L^2)
    const double deps = meps - ceps;  // the jump of eps between layers
    // deps / (h_{i-1}^2 L^2) = deps (D h_i h_s)^2 L^2, D = 1 / (h_i h_{i-1} h_s L^2): no division
    const double Dhh = D * hp * hs;
    const double jump = deps != 0.0 ? deps * (Dhh * Dhh) * L2 : 0.0;
    const double fac_m = jump - (fu + fl0);
    const double fac_l = fl0 - jump;
    const double sC = scs * (hm * hs * L2 * D);  // (1/h_i) = h_{i-1} h_s L^2 D
    const double gen = mid.G * g_frac;
    const double elec = hm * eR.Fn - hp * eL.Fn + fac_ * (gen - mid.R);
    const double hole = hp * eL.Fp - hm * eR.Fp + fac_ * (gen - mid.R);
    const double pois = (fac_l * cur.V + fac_m * mid.V + fu * rV -
                         Q_E * (mid.n - mid.p + dop + share * mid.fix + mid.an - mid.ca)) +
                        Q_E * g.qt;  // solver.py:1136-1140
    A[0] = fac_l * sP; A[1] = 0.0; A[2] = 0.0;
    A[3] = hp * eL.gn * sC; A[4] = -hp * eL.an0 * sC; A[5] = 0.0;
    A[6] = -hp * eL.gp * sC; A[7] = 0.0; A[8] = hp * eL.ap0 * sC;
    B[0] = (fac_m - Q_E * (mid.an + mid.ca) * iVt) * sP;
    B[1] = (-Q_E * mid.n + Q_E * g.dqn) * sP;
    B[2] = (Q_E * mid.p + Q_E * g.dqp) * sP;
    B[3] = (-hp * eL.gn - hm * eR.gn) * sC;
    B[4] = (-hp * eL.an1 + hm * eR.an0 - fac_ * mid.dRn) * sC;
    B[5] = (-fac_ * mid.dRp) * sC;
    B[6] = (hp * eL.gp + hm * eR.gp) * sC;
    B[7] = (-fac_ * mid.dRn) * sC;
    B[8] = (hp * eL.ap1 - hm * eR.ap0 - fac_ * mid.dRp) * sC;
    if (last) {  // the coupling to the contact is not an unknown
#pragma unroll
        for (int q = 0; q < 9; ++q) C[q] = 0.0;
    } else {
        C[0] = fu * sP; C[1] = 0.0; C[2] = 0.0;
        C[3] = hm * eR.gn * sC; C[4] = hm * eR.an1 * sC; C[5] = 0.0;
        C[6] = -hm * eR.gp * sC; C[7] = 0.0; C[8] = -hm * eR.ap1 * sC;
    }
    r[0 * K] = -pois * sP; r[1 * K] = -elec * sC; r[2 * K] = -hole * sC;
    if constexpr (K > 1) {
        r[0 * K + 1] = -Q_E * mid.an * sP; r[0 * K + 2] = Q_E * mid.ca * sP;
        r[1 * K + 1] = 0.0; r[1 * K + 2] = 0.0; r[2 * K + 1] = 0.0; r[2 * K + 2] = 0.0;
    }
}

template <int K>
DDH_FN void row_blocks(const Lane &Lv, const Pt &cur, const Pt &mid, double rV, const Edge &eL,
                       const Edge &eR, bool last, double L2, double iVt, double share,
                       double g_frac, double *A, double *B, double *C, double *r,
                       bool trapq = false) {
    const int w = Lv.ls;
    lcdouble *sm = Lv.seg + mid.sid * SEG_N * w;
    RowSeg g;
    g.meps = sm[SEG_EPS * w];
    g.ceps = Lv.seg[(cur.sid * SEG_N + SEG_EPS) * w];
    g.sP = sm[SEG_SP * w]; g.scs = sm[SEG_SCS * w]; g.dop = sm[SEG_DOP * w];
    if (trapq) {  // the lane's device holds charged traps, and this solve counts their charge
        const double tt = sm[SEG_TT * w], Nt = sm[SEG_NT * w];
        if (tt != 0.0 && Nt > 0.0)
            g.qt = trap_charge(tt, mid.n, mid.p, Nt, sm[SEG_CN * w], sm[SEG_CP * w],
                               sm[SEG_NT0 * w], sm[SEG_PT0 * w], g.dqn, g.dqp);
    }
    row_blocks_s<K>(g, cur, mid, rV, eL, eR, last, L2, iVt, share, g_frac, A, B, C, r);
}

// Elimination sweep (right to left). Returns the ion step z. NOISE: the right-hand side is
// the residual's round-off instead of -F (src/jaxdd/solver.py:1728-1751, `_noise_step`).
template <typename T, typename S, int K, bool NOISE = false>
DDH_SWEEP void elim_sweep(const Args &a, int lane, const SweepCtl &cref, double *z) {
    // everything the loop reads of the arguments, in registers: the loop's stores to global
    // memory would otherwise keep these generic-pointer loads inside it
    const SweepCtl c = cref;
    const int NP = DDH_UNIFORM(a.NP), bern = DDH_UNIFORM(a.bern);
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, L = SC(SC_L), L2 = L * L;
    // the ions' totals (scaled in the start's stages); none once they are frozen (the slots
    // are then held, src/jaxdd/batch.py:859-872)
    const double tot0 = c.mobile ? SC(SC_AT) * c.share : 0.0;
    const double tot1 = c.mobile ? SC(SC_CT) * c.share : 0.0;
    const double itot0 = tot0 > 0 ? 1.0 / tot0 : 0.0, itot1 = tot1 > 0 ? 1.0 / tot1 : 0.0;
    const Lane Lv = lane_view(a, lane, c);
    const size_t n = Lv.n;
    DDH_GLOBAL S *fac = (DDH_GLOBAL S *)(factors<S>(a) + lane);
    // a solve's first step reads its start from the source buffers, later steps the iterate
    gdouble *s1 = c.fresh ? Lv.ys1 : Lv.yw;
    gdouble *s2 = (c.fresh && c.ratio != 0.0) ? Lv.ys2 : s1;
    const Contact left = contact_point(a, lane, Lv, c.Vint, false);
    const Contact right = contact_point(a, lane, Lv, c.Vint, true);
    const int sid0 = Lv.sid[0];
    Pt mid, cur;
    double rV = 0.0;  // the row's right neighbour: its potential and ion masks
    double ram = 0.0, rcm = 0.0;
    double rlnN = 0.0, rlnP = 0.0;  // and its log densities (the noise's right-hand side)
    Raw nxt;
    fetch_point(Lv, s1, s2, NP, nxt);
    make_point(Lv, NP + 1, true, right, nxt, Lv.sid[(size_t)(NP + 1) * n], c, Vt, iVt, mid);  // nxt unread
    Edge eR = {0, 0, 0, 0, 0, 0, 0, 0};
    T cp[9], dp[3 * K];
#pragma unroll
    for (int q = 0; q < 9; ++q) cp[q] = T(0);
#pragma unroll
    for (int q = 0; q < 3 * K; ++q) dp[q] = T(0);
    double gam[3][2] = {{0, 0}, {0, 0}, {0, 0}};
    double accF[2] = {0, 0}, accU[2][2] = {{0, 0}, {0, 0}};
    double num0 = 0.0, num1 = 0.0;
    // Point j: its values, the edge to its right neighbour `mid`, and mid's row where mid is an
    // interior point. Every visit of the loop issues the same memory operations (the next
    // point's loads, then the row's factor stores), so that the wait for the loads never
    // has to cover the stores.
    auto visit = [&](const int j, const bool has_row, const Raw &now, Raw &nxt) {
        fetch_point(Lv, s1, s2, j > 1 ? j - 1 : 1, nxt);
        make_point(Lv, j, j == 0, left, now, j > 0 ? now.sid : sid0, c, Vt, iVt, cur);
        const Edge eL = edge_flux(Lv, cur, mid, Vt, iVt, bern);
        if constexpr (K > 1) {
            const double hLj = cur.h * L;
            num0 += 0.5 * (mid.an + cur.an) * (mid.am * cur.am) * hLj;
            num1 += 0.5 * (mid.ca + cur.ca) * (mid.cm * cur.cm) * hLj;
        }
        if (has_row) {
            // row i = j + 1: left = cur, mid, right; edges eL (i-1, i), eR (i, i+1)
            const int i = j + 1;
            double A[9], B[9], C[9], r[3 * K];
            // (the noise's right-hand side takes the last row's coupling to the contact too)
            row_blocks<K>(Lv, cur, mid, rV, eL, eR, !NOISE && i == NP, L2, iVt, c.share, c.g_frac,
                          A, B, C, r, c.trapq != 0);
            if constexpr (NOISE) {
                // eps * sum |dF/dy| |y| over the row's three points, y = (V, ln n, ln p)
                const double yl[3] = {cur.V, log(cur.n), log(cur.p)};
                const double ym[3] = {mid.V, log(mid.n), log(mid.p)};
                const double yr[3] = {rV, rlnN, rlnP};
#pragma unroll
                for (int row = 0; row < 3; ++row) {
                    double acc = 0.0;
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        acc += fabs(A[3 * row + q]) * fabs(yl[q]);
                        acc += fabs(B[3 * row + q]) * fabs(ym[q]);
                        acc += fabs(C[3 * row + q]) * fabs(yr[q]);
                    }
                    r[row * K] = DBL_EPS * acc;
                }
            }
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
#pragma unroll
                        for (int q = 0; q < K; ++q) r[row * K + q] *= sc;
                    }
                }
            }
            T At[9], Bt[9], Ct[9], rt[3 * K];
            to_t(A, At, 9); to_t(B, Bt, 9); to_t(C, Ct, 9); to_t(r, rt, 3 * K);
            // M = B - C c'_prev ; rhs = r - C d'_prev
            T M[9], rh[3 * K];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    M[3 * p + q] = Bt[3 * p + q] - (Ct[3 * p + 0] * cp[0 + q] + Ct[3 * p + 1] * cp[3 + q] + Ct[3 * p + 2] * cp[6 + q]);
#pragma unroll
                for (int q = 0; q < K; ++q)
                    rh[p * K + q] = rt[p * K + q] - (Ct[3 * p + 0] * dp[0 * K + q] + Ct[3 * p + 1] * dp[1 * K + q] + Ct[3 * p + 2] * dp[2 * K + q]);
            }
            T Mi[9];
            inverse3(M, Mi);
            T cn_[9], dn_[3 * K];
#pragma unroll
            for (int p = 0; p < 3; ++p) {
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    cn_[3 * p + q] = Mi[3 * p + 0] * At[0 + q] + Mi[3 * p + 1] * At[3 + q] + Mi[3 * p + 2] * At[6 + q];
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
                // gamma_i = W_i - c'_{i+1}^T gamma_{i+1}; W_i only in the potential's row
                const double hLm = cur.h * L, hLp = mid.h * L;
                const double mida = mid.am, midc = mid.cm;
                const double w0 = 0.5 * (mida * cur.am * hLm + ram * mida * hLp) * mid.an * iVt * itot0;
                const double w1 = -0.5 * (midc * cur.cm * hLm + rcm * midc * hLp) * mid.ca * iVt * itot1;
                double ng[3][2];
#pragma unroll
                for (int row = 0; row < 3; ++row) {
                    const double c0 = (double)cp[0 + row], c1 = (double)cp[3 + row], c2 = (double)cp[6 + row];
                    ng[row][0] = (row == 0 ? w0 : 0.0) - (c0 * gam[0][0] + c1 * gam[1][0] + c2 * gam[2][0]);
                    ng[row][1] = (row == 0 ? w1 : 0.0) - (c0 * gam[0][1] + c1 * gam[1][1] + c2 * gam[2][1]);
                }
#pragma unroll
                for (int k = 0; k < 2; ++k) {
                    double sF = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
                    for (int row = 0; row < 3; ++row) {
                        gam[row][k] = ng[row][k];
                        sF += ng[row][k] * (double)dn_[row * K + 0];
                        s1 += ng[row][k] * (double)dn_[row * K + 1];
                        s2 += ng[row][k] * (double)dn_[row * K + 2];
                    }
                    accF[k] += sF; accU[k][0] += s1; accU[k][1] += s2;
                }
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) cp[q] = cn_[q];
#pragma unroll
            for (int q = 0; q < 3 * K; ++q) dp[q] = dn_[q];
        }
        rV = mid.V; ram = mid.am; rcm = mid.cm;
        if constexpr (NOISE) { rlnN = log(mid.n); rlnP = log(mid.p); }
        mid = cur;
        eR = eL;
    };
    Raw now = nxt;
    visit(NP, false, now, nxt);
    {   // the stores row NP will overwrite: the loop is then entered, as it repeats, with
        // the next point's loads followed by a row's stores
        DDH_GLOBAL S *f = fac + (size_t)(NP - 1) * NFAC * n;
#pragma unroll
        for (int q = 0; q < 9 + 3 * K; ++q) f[q * n] = S(0);
    }
    for (int j = NP - 1; j >= 0; --j) {
        now = nxt;
        visit(j, true, now, nxt);
    }
    z[0] = z[1] = 0.0;
    if constexpr (K > 1) {
        // the constraints (solver.py:1255-1265) and their Schur complement; a slot without
        // ions is held (an identity row)
        const double g0 = tot0 > 0 ? num0 * itot0 - 1.0 : 0.0;
        const double g1 = tot1 > 0 ? num1 * itot1 - 1.0 : 0.0;
        const double D0 = tot0 > 0 ? num0 * itot0 : 1.0;
        const double D1 = tot1 > 0 ? num1 * itot1 : 1.0;
        z[0] = -g0 - accF[0];
        z[1] = -g1 - accF[1];
        solve2(D0 - accU[0][0], -accU[0][1], -accU[1][0], D1 - accU[1][1], z[0], z[1]);
    }
}

// Back-substitution sweep (left to right): the step, the new state, its size and current.
// NOISE: only the step's size, measured as a Newton step's (the state is left as it is).
template <typename T, typename S, int K, bool NOISE = false>
DDH_SWEEP void backsub_sweep(const Args &a, int lane, const SweepCtl &cref, const double *z,
                             double *out /* size, J, finite */) {
    const SweepCtl c = cref;  // in registers (as in the elimination)
    const int NP = DDH_UNIFORM(a.NP);
    const double Vt = SC(SC_VT), iVt = 1.0 / Vt, nfl = SC(SC_NFLOOR), lnfloor = log(nfl);
    const double capv = CAP_V * Vt;
    const Lane Lv = lane_view(a, lane, c);
    const size_t n = Lv.n;
    DDH_GLOBAL const S *fac = (DDH_GLOBAL const S *)(factors<S>(a) + lane);
    const double z0 = z[0], z1 = z[1];
    // point j's loads, issued one point ahead of their use (as in the elimination)
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
    Next nxt;
    fetch(1, nxt);
    // the left contact
    const int sid0 = Lv.sid[0];
    const double V0 = contact_V(a, lane, c.Vint, false);
    const double Vg0n = V0 + segv(Lv, sid0, SEG_OFFN), Vg0p = V0 + segv(Lv, sid0, SEG_OFFP);
    double wn_prev = exp(-Vg0n * iVt) * segv(Lv, sid0, SEG_IMUN);
    double wp_prev = exp(Vg0p * iVt) * segv(Lv, sid0, SEG_IMUP);
    const double endn0 = exp(SC(SC_LNL) - Vg0n * iVt);
    const double endp0 = exp(SC(SC_LPL) + Vg0p * iVt);
    double U_prev = 0.0, iu_prev = 0.0, x_prev = Lv.pt[PT_X * n];
    double single_n = 0.0, double_n = 0.0, single_p = 0.0, double_p = 0.0;
    double xs0 = 0.0, xs1 = 0.0, xs2 = 0.0;
    double size = 0.0;
    int finite = 1;
    auto visit = [&](const int j, const Next &now, Next &nxt) {
        fetch(j < NP ? j + 1 : NP, nxt);  // the same loads every visit (as in the elimination)
        const int sid = now.sid;
        const S *f = now.f;
        double dF[3];
#pragma unroll
        for (int row = 0; row < 3; ++row) {
            double v = (double)f[9 + row * K + 0];
            if constexpr (K > 1) v -= (double)f[9 + row * K + 1] * z0 + (double)f[9 + row * K + 2] * z1;
            dF[row] = v;
        }
        const double x0 = dF[0] - ((double)f[0] * xs0 + (double)f[1] * xs1 + (double)f[2] * xs2);
        const double x1 = dF[1] - ((double)f[3] * xs0 + (double)f[4] * xs1 + (double)f[5] * xs2);
        const double x2 = dF[2] - ((double)f[6] * xs0 + (double)f[7] * xs1 + (double)f[8] * xs2);
        xs0 = x0; xs1 = x1; xs2 = x2;
        if constexpr (NOISE) {
            const double dn = exp(now.y1), dp = exp(now.y2);
            size = nmax(size, fabs(x0) / Vt);
            size = nmax(size, fabs(x1) * (dn / (dn + nfl)));
            size = nmax(size, fabs(x2) * (dp / (dp + nfl)));
            return;
        }
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
        // the integral current of the new state (src/jaxdd/batch.py:321-373)
        const double R = recombination_value(nn, pp, g.kd, g.ni2, g.nt, g.cn, g.cp, g.nt0, g.pt0);
        const double U = R - now.G * c.g_frac;
        const double wn = exp(-(y0 + g.offn) * iVt) * g.imun;
        const double wp = exp((y0 + g.offp) * iVt) * g.imup;
        const double xj = now.x;
        const double dx = xj - x_prev;
        const double iu = iu_prev + 0.5 * dx * (U_prev + U);
        single_n += 0.5 * dx * (wn_prev + wn);
        double_n += 0.5 * dx * (wn_prev * iu_prev + wn * iu);
        single_p += 0.5 * dx * (wp_prev + wp);
        double_p += 0.5 * dx * (wp_prev * iu_prev + wp * iu);
        wn_prev = wn; wp_prev = wp; U_prev = U; iu_prev = iu; x_prev = xj;
    };
    // two load buffers taking turns, and the first visit before the loop, so that the loop
    // is entered as it repeats: a buffer's loads, then a visit's stores (as in the elimination)
    Next other;
    visit(1, nxt, other);
    int j = 2;
    for (; j + 1 <= NP; j += 2) {
        visit(j, other, nxt);
        visit(j + 1, nxt, other);
    }
    if (j == NP) visit(NP, other, nxt);
    const int sidN = Lv.sid[(size_t)(NP + 1) * n];
    const double VN = contact_V(a, lane, c.Vint, true);
    const double endnN = exp(SC(SC_LNR) - (VN + segv(Lv, sidN, SEG_OFFN)) * iVt);
    const double endpN = exp(SC(SC_LPR) + (VN + segv(Lv, sidN, SEG_OFFP)) * iVt);
    // K = (V_T (sn dens_N ev_N - sn dens_0 ev_0) - sn double) / single, sn = -1 electrons
    const double Kn = (Vt * (-endnN + endn0) + double_n) / single_n;
    const double Kp = (Vt * (endpN - endp0) - double_p) / single_p;
    // The cumulative recombination terms of the two carriers cancel exactly in the sum (the
    // same U with opposite signs), so the device average is q (K_n + K_p), sign turned with
    // the anode on the left.
    out[0] = size;
    out[1] = NOISE ? 0.0 : -SC(SC_ELEC) * Q_E * (Kn + Kp);
    out[2] = finite;
}

// ---------------------------------------------------------------------------------------
// The per-lane state machine: src/jaxdd/batch.py:875-1006 (_start, _solve_one) and
// src/jaxdd/solver.py:2105-2245 (_scan of solve_jv).
// ---------------------------------------------------------------------------------------
DDH_CONST double ION_STAGES[4] = {1.0 / 64, 1.0 / 16, 1.0 / 4, 1.0};

DDH_FN void begin_solve(const Args &a, LaneState &S, double Vint, double share,
                                   int maxit, int src1, int src2, double ratio, const double lg[2],
                                   int work) {
    S.Vint = Vint; S.share = share; S.maxit = maxit;
    S.b_src1 = src1; S.b_src2 = src2; S.ratio = ratio; S.b_work = work;
    S.lw[0] = lg[0]; S.lw[1] = lg[1];
    // FP32 factors only in the scan: the start solves (far from their solutions, with the ions'
    // stages) converge in fewer FP64 steps, and their slowest devices set the kernel's time
    S.it = 0; S.fresh = 1; S.need64 = S.phase != PH_SCAN;
    for (int k = 0; k < 4; ++k) { S.hist[k] = INFINITY; S.Jh[k] = INFINITY; }
    S.noise_state = 0;
}

DDH_FN int free_buffer(const LaneState &S, int a0, int a1, int a2) {
    for (int b = 0; b < NBUF; ++b)
        if (b != a0 && b != a1 && b != a2) return b;
    return NBUF - 1;
}

DDH_FN void begin_scan_voltage(const Args &a, int lane, LaneState &S) {
    const int nV = a.nV;
    int j = S.j;
    int idx;
    double ratio = 0.0;
    if (a.jv) {
        idx = j;
    } else {
        idx = j < nV - S.below ? S.below + j : nV - 1 - j;
        if (j == nV - S.below) {  // the turn: from the pre-bias solution again
            S.b_last = S.b_before = S.b_start;
            S.llast[0] = S.lbefore[0] = S.lstart[0];
            S.llast[1] = S.lbefore[1] = S.lstart[1];
            S.V_last = S.V_before = S.first;
        }
        double Vi = a.V[idx];
        ratio = S.V_last != S.V_before ? (Vi - S.V_last) / (S.V_last - S.V_before) : 0.0;
    }
    double Vi = a.V[idx];
    double lg[2];
    for (int k = 0; k < 2; ++k)
        lg[k] = S.llast[k] + clipd(ratio * (S.llast[k] - S.lbefore[k]), -ION_LOG_CAP, ION_LOG_CAP);
    int work = free_buffer(S, S.b_last, S.b_before, S.b_start);
    begin_solve(a, S, Vi, 1.0, a.max_iter, S.b_last, S.b_before, ratio, lg, work);
}

// Copies state buffer `buf` of the device to dst [NP][3][n]: each of the device's lanes its
// share (the points its own sweeps wrote), eight loads in flight at a time.
DDH_FN void keep_state(const Args &a, int lane, const LaneState &S, int buf, double *dst0) {
    const size_t n = (size_t)a.n, m = (size_t)a.NP * 3;
    const double *src = a.Y + (size_t)buf * m * n + lane;
    double *dst = dst0 + lane;
    const size_t hi = m * (size_t)(S.part + 1) / (size_t)S.parts;
    size_t q = m * (size_t)S.part / (size_t)S.parts;
    for (; q + 8 <= hi; q += 8) {
        double t[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) t[k] = src[(q + k) * n];
#pragma unroll
        for (int k = 0; k < 8; ++k) dst[(q + k) * n] = t[k];
    }
    for (; q < hi; ++q) dst[q * n] = src[q * n];
}

DDH_FN void scan_setup(const Args &a, int lane, LaneState &S, int y_buf,
                                  const double ly[2]) {
    // the pre-bias solution (the adjoint's second linearisation point where it freezes ions)
    if (a.pre && !a.jv && SC(SC_PRE) > 0) keep_state(a, lane, S, y_buf, a.pre);
    S.b_start = y_buf;
    S.lstart[0] = ly[0]; S.lstart[1] = ly[1];
    bool frozen = SC(SC_PRE) > 0 && SC(SC_FIX) > 0;
    S.freeze_now = frozen ? 1 : 0;
    S.mobile_now = frozen ? 0 : S.mobile0;
    if (!a.jv) S.pre_ok = S.pre_ok || !frozen;
    int below = 0;
    if (!a.jv)
        while (below < a.nV && a.V[below] < S.first) ++below;
    S.below = below;
    S.j = 0;
    S.b_last = S.b_before = y_buf;
    S.llast[0] = S.lbefore[0] = ly[0];
    S.llast[1] = S.lbefore[1] = ly[1];
    S.V_last = S.V_before = S.first;
    S.phase = PH_SCAN;
    begin_scan_voltage(a, lane, S);
}

// Where the start stands after the F solve or the stages: the pre-bias solve, or straight to
// the scan where the start is already the solution.
DDH_FN void to_pre(const Args &a, int lane, LaneState &S, int y_buf, const double ly[2],
                              int solved) {
    if (solved) {
        S.pre_ok = 1;
        scan_setup(a, lane, S, y_buf, ly);
        return;
    }
    S.phase = PH_PRE;
    int work = free_buffer(S, y_buf, -1, -1);
    begin_solve(a, S, S.first, 1.0, a.jv ? a.max_iter : a.start_iter, y_buf, y_buf, 0.0, ly, work);
}

DDH_FN void begin_stage(const Args &a, LaneState &S) {
    double share = ION_STAGES[S.stage];
    double sh = log(share / S.before_share);
    double lg[2] = {S.lc[0] + sh, S.lc[1] + sh};
    S.lc[0] = lg[0]; S.lc[1] = lg[1];
    int work = (S.b_c == 2) ? 3 : 2;
    S.phase = PH_STAGE;
    begin_solve(a, S, S.first, share, a.jv ? a.max_iter : a.start_iter, S.b_c, S.b_c, 0.0, lg, work);
}

// The lane's first solve after Poisson's start. With charged bulk traps, first the solve with
// them neutral, whose solution (where it converges) stands in for Poisson's start
// (src/jaxdd/batch.py:898-900, solver.py:2192-2200): from Poisson's, Newton finds no way to the
// charged solution.
DDH_FN void start_lane(const Args &a, int lane, LaneState &S, bool neutral_done = false) {
    S.err = 0;
    S.mobile_now = S.mobile0;
    S.freeze_now = 0;
    S.pre_ok = 1;
    if (S.charged && !neutral_done) {
        S.phase = PH_NEUTRAL;
        S.neutral = 1;
        begin_solve(a, S, S.first, 1.0, a.jv ? a.max_iter : a.start_iter, S.b_pois, S.b_pois,
                    0.0, S.lpois, 4);
        return;
    }
    bool pre = SC(SC_PRE) > 0;
    if (a.jv && !S.mobile0) {
        // solve_jv: without mobile ions no solve at the first voltage (solver.py:2211-2226)
        to_pre(a, lane, S, S.b_pois, S.lpois, pre ? 0 : 1);
        return;
    }
    S.phase = PH_F;
    begin_solve(a, S, S.first, 1.0, a.jv ? a.max_iter : a.start_iter, S.b_pois, S.b_pois, 0.0,
                S.lpois, 1);
}

// One finished solve of the lane: the next solve, or the end.
DDH_FN void end_solve(const Args &a, int lane, LaneState &S, int ok, double J) {
    const int fin = S.fin_work && isfinite(S.lw[0]) && isfinite(S.lw[1]);
    switch (S.phase) {
    case PH_NEUTRAL: {
        S.neutral = 0;
        if (ok) {  // buffer 4: no later solve of the start writes it
            S.b_pois = S.b_work;
            S.lpois[0] = S.lw[0]; S.lpois[1] = S.lw[1];
        }
        start_lane(a, lane, S, true);
        return;
    }
    case PH_F: {
        S.ok_f = ok;
        S.fin_f = fin;
        S.b_f = S.b_work;
        S.lf[0] = S.lw[0]; S.lf[1] = S.lw[1];
        if (a.jv) {
            if (ok) { to_pre(a, lane, S, S.b_f, S.lf, SC(SC_PRE) > 0 ? 0 : 1); return; }
            S.stage = 0; S.b_c = S.b_pois; S.lc[0] = S.lpois[0]; S.lc[1] = S.lpois[1];
            S.before_share = 1.0;
            begin_stage(a, S);
            return;
        }
        bool done = ok || !S.mobile0;
        if (done) {
            int yb = fin ? S.b_f : S.b_pois;
            const double *ly = fin ? S.lf : S.lpois;
            to_pre(a, lane, S, yb, ly, ok);
            return;
        }
        S.stage = 0; S.b_c = S.b_pois; S.lc[0] = S.lpois[0]; S.lc[1] = S.lpois[1];
        S.before_share = 1.0;
        begin_stage(a, S);
        return;
    }
    case PH_STAGE: {
        if (a.jv) {
            // solve_jv: each stage from the last, stopping at the first failure
            S.b_c = S.b_work; S.lc[0] = S.lw[0]; S.lc[1] = S.lw[1];
            S.before_share = ION_STAGES[S.stage];
            S.ok_c = ok;
            S.stage++;
            if (ok && S.stage < 4) { begin_stage(a, S); return; }
            if (ok) { to_pre(a, lane, S, S.b_c, S.lc, SC(SC_PRE) > 0 ? 0 : 1); return; }
            to_pre(a, lane, S, S.b_pois, S.lpois, SC(SC_PRE) > 0 ? 0 : 1);
            return;
        }
        if (fin) { S.b_c = S.b_work; S.lc[0] = S.lw[0]; S.lc[1] = S.lw[1]; }
        S.ok_c = ok;
        S.before_share = ION_STAGES[S.stage];
        S.stage++;
        if (S.stage < 4) { begin_stage(a, S); return; }
        int staged = S.ok_c;
        int yb; const double *ly;
        if (staged) { yb = S.b_c; ly = S.lc; }
        else if (S.fin_f) { yb = S.b_f; ly = S.lf; }
        else { yb = S.b_pois; ly = S.lpois; }
        to_pre(a, lane, S, yb, ly, staged);
        return;
    }
    case PH_PRE: {
        if (a.jv) {
            if (!ok) {  // the reference raises RuntimeError at a failed pre-bias
                S.err = 1;
                a.status[lane] = 1;
                S.phase = PH_DONE;
                return;
            }
            S.pre_ok = 1;
        } else {
            S.pre_ok = ok;
        }
        scan_setup(a, lane, S, S.b_work, S.lw);
        return;
    }
    case PH_SCAN: {
        int idx = a.jv ? S.j : (S.j < a.nV - S.below ? S.below + S.j : a.nV - 1 - S.j);
        size_t o = (size_t)lane * a.nV + idx;
        a.J[o] = J;
        a.ok[o] = (uint8_t)((ok && S.pre_ok) ? 1 : 0);
        a.its[o] = S.it;
        if (a.states)  // the state the voltage ended with (the adjoint's linearisation point)
            keep_state(a, lane, S, S.b_work, a.states + (size_t)idx * a.NP * 3 * a.n);
        if (ok && fin) {  // only a converged, finite solution is carried on
            S.b_before = S.b_last; S.b_last = S.b_work;
            S.lbefore[0] = S.llast[0]; S.lbefore[1] = S.llast[1];
            S.llast[0] = S.lw[0]; S.llast[1] = S.lw[1];
            S.V_before = S.V_last; S.V_last = S.Vint;
        }
        S.freeze_now = 0;
        S.j++;
        if (S.j >= a.nV) { S.phase = PH_DONE; return; }
        begin_scan_voltage(a, lane, S);
        return;
    }
    default:
        S.phase = PH_DONE;
    }
}

// After a Newton step: the histories and the stop rules (src/jaxdd/batch.py:793-839; with
// jv, src/jaxdd/solver.py:2150-2166).
DDH_FN void after_step(const Args &a, int lane, LaneState &S, double size, double J) {
    for (int k = 0; k < 3; ++k) { S.hist[k] = S.hist[k + 1]; S.Jh[k] = S.Jh[k + 1]; }
    S.hist[3] = size;
    S.Jh[3] = J;
    S.it++;
    bool anynan = false;
    double hmax = 0.0, hmin_tail = INFINITY;
    for (int k = 0; k < 4; ++k) { anynan |= S.hist[k] != S.hist[k]; hmax = fmax(hmax, S.hist[k]); }
    for (int k = 1; k < 4; ++k) hmin_tail = fmin(hmin_tail, S.hist[k]);
    bool stalled = S.it > STALL_STEPS && !anynan && hmax < ROUND_OFF_FLOOR && hmin_tail >= 0.1 * S.hist[0];
    bool done;
    if (a.jv || a.strict) {
        done = size < a.tol || stalled;
    } else {
        double Jn = S.Jh[3], scale = SC(SC_JS) + fabs(Jn), jdev = 0.0;
        bool jnan = false;
        for (int k = 0; k < 4; ++k) {
            double v = fabs(S.Jh[k] - Jn);
            jnan |= v != v;
            jdev = fmax(jdev, v);
        }
        bool settled = S.it > STALL_STEPS && size < CURRENT_STEP_CAP && !jnan && jdev < CURRENT_TOL * scale;
        bool quadratic = S.it > 1 && size < QUADRATIC_TOL && size < QUADRATIC_RATIO * S.hist[2];
        done = size < a.tol || stalled || settled || quadratic;
    }
    if (a.jv && !done && S.it > STALL_STEPS && !anynan && hmax < NOISE_FLOOR_CAP) {
        // stalled above the round-off floor: stop below the problem's own noise floor
        // (src/jaxdd/solver.py:2162-2166), computed once per solve, on the lane's next turn
        if (S.noise_state == 2) {
            done = hmax < S.noise_floor;
        } else {
            S.noise_state = 1;
            S.fresh = 0;
            return;
        }
    }
    if (done) { end_solve(a, lane, S, 1, J); return; }
    if (S.it >= S.maxit || size != size) { end_solve(a, lane, S, 0, J); return; }
    S.fresh = 0;
    if (a.mixed && S.it >= MIXED_STEPS) S.need64 = 1;
}

// One Newton step of one lane (both sweeps and the state machine). `ions`: some lane of
// the wavefront has mobile ions (the sweeps then carry the border's right-hand sides).
// The sweeps' view of the lane's state.
DDH_FN SweepCtl sweep_ctl(const Args &a, int lane, const LaneState &S, lcdouble *lseg) {
    SweepCtl c;
    c.lseg = lseg;
    c.Vint = S.Vint; c.share = S.share; c.ratio = S.ratio; c.g_frac = SC(SC_GFRAC);
    c.logs[0] = S.lw[0]; c.logs[1] = S.lw[1];
    c.lstart[0] = S.lstart[0]; c.lstart[1] = S.lstart[1];
    c.fresh = S.fresh; c.freeze_now = S.freeze_now && S.fresh; c.mobile = S.mobile_now;
    c.src1 = S.b_src1; c.src2 = S.b_src2; c.work = S.b_work;
    c.trapq = S.charged && !S.neutral;
    return c;
}

// After the sweeps: the ion step's cap, the step size with it, the stop rules.
DDH_FN void finish_step(const Args &a, int lane, LaneState &S, const double z[2], double size,
                        double J, int finite) {
    const double dl0 = clipd(z[0], -ION_LOG_CAP, ION_LOG_CAP);
    const double dl1 = clipd(z[1], -ION_LOG_CAP, ION_LOG_CAP);
    S.lw[0] += dl0;
    S.lw[1] += dl1;
    S.fin_work = finite;
    after_step(a, lane, S, nmax(size, nmax(fabs(dl0), fabs(dl1))), J);
}

// One Newton step's two sweeps: T the factorisation's arithmetic, S the stored factors' type.
template <typename T, typename S>
DDH_FN void step_sweeps(const Args &a, int lane, bool ions, SweepCtl &c, double *z, double *out) {
    if (ions) elim_sweep<T, S, 3>(a, lane, c, z);
    else elim_sweep<T, S, 1>(a, lane, c, z);
    c.fresh = 0;
    if (ions) backsub_sweep<T, S, 3>(a, lane, c, z, out);
    else backsub_sweep<T, S, 1>(a, lane, c, z, out);
}

template <typename T>
DDH_FN void newton_lane_step(const Args &a, int lane, bool ions, lcdouble *lseg, LaneState &S) {
    SweepCtl c = sweep_ctl(a, lane, S, lseg);
    double z[2], out[3];
    if (S.noise_state == 1) {
        // the noise floor of the state the last step left (src/jaxdd/solver.py:1728-1751):
        // one FP64 solve without the ions' border, then the stop the last step deferred
        elim_sweep<double, double, 1, true>(a, lane, c, z);
        backsub_sweep<double, double, 1, true>(a, lane, c, z, out);
        S.noise_floor = out[0];
        S.noise_state = 2;
        double hmax = 0.0;
        for (int k = 0; k < 4; ++k) hmax = fmax(hmax, S.hist[k]);
        if (hmax < S.noise_floor) end_solve(a, lane, S, 1, S.Jh[3]);
        else if (S.it >= S.maxit) end_solve(a, lane, S, 0, S.Jh[3]);
        return;
    }
    if constexpr (sizeof(T) == 4) {
        step_sweeps<float, float>(a, lane, ions, c, z, out);
    } else {
        if (a.mixed) step_sweeps<double, float>(a, lane, ions, c, z, out);
        else step_sweeps<double, double>(a, lane, ions, c, z, out);
    }
    finish_step(a, lane, S, z, out[0], out[1], out[2] != 0.0);
}

DDH_FN bool lane_active(const LaneState &S) { return S.phase != PH_DONE; }
DDH_FN bool lane_mobile(const LaneState &S) { return S.mobile_now != 0; }
DDH_FN bool lane_needs64(const LaneState &S) { return S.need64 != 0; }

}  // namespace ddh
