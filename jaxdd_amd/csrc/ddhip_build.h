// The batch's arrays built on the GPU from each device's structure and per-layer numbers
// (reference: src/jaxdd/device.py:877-1063 `device_arrays` with the grid placed by `_spacing`
// :225-233, and src/jaxdd/batch.py:151-223 `prepare`'s packing), with their derivatives along
// given directions in those numbers (forward mode): the tangents the adjoint contracts, in place
// of central differences of the host builder.
//
// The path's subset of the physics: Boltzmann carriers, each layer's uniform generation
// (`G_ehp`, the reference's generation mode "none": every interior point of the layer) or a
// generation profile (a table of G over the position, stretched to the stack), ohmic
// contacts at the given work functions, bulk traps of one level, mobile ions' masks, totals and
// fixed densities, direct recombination by each layer's k_direct or by Langevin's rate. What is discrete (the grid's layout, each point's layer and segment, which
// layers generate light, hold traps or ions) is the structure's and has no derivative; the
// thicknesses move the points by scaling each layer's unscaled spacing, with the mean of the
// three spacings around each interface (SIMsalabim's rule).
//
// One lane per (device, direction): the lane computes every value with its derivative along
// its direction (a dual number), writes the derivative to its direction's slice of the
// tangents and, on direction 0, the values. The points are visited in order, each emitted two
// points after it is computed: the interface rule changes the spacing of the two points before
// an interface's first point.
#pragma once

#ifdef DDH_CPU_EMULATION
#include <cmath>
#include <cstddef>
#include <cstdint>
#define DDH_BFN inline
#else
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#define DDH_BFN __device__ inline
#endif

#include "ddhip_layout.h"

namespace ddh {

constexpr int BUILD_MAX_LAYERS = 8;
// SIMsalabim 5.36's constants (src/jaxdd/constants.py:8-12), bit-equal to the host builder's
constexpr double BUILD_Q = 1.6022e-19;
constexpr double BUILD_KB = 1.3807e-23;
constexpr double BUILD_EPS0 = 8.8542e-12;
constexpr double BUILD_DENSITY_FLOOR = 1e-12;  // src/jaxdd/solver.py:54

// Per-layer numbers [nl][BL_N][n] (the reference's `LAYER_PARAMETERS` of the path).
enum BuildLayer {
    BL_L, BL_EPSR, BL_EC, BL_EV, BL_NC, BL_ND, BL_NA, BL_MUN, BL_MUP, BL_NUN, BL_NUP, BL_KD,
    BL_NT, BL_CN, BL_CP, BL_ET, BL_G, BL_NANION, BL_NCATION, BL_MUA, BL_MUC, BL_PRELANGEVIN,
    BL_TRAPTYPE,  // the bulk traps' charge type (1 donor, -1 acceptor, 0 neutral): no derivative
    BL_N
};
// Per-device numbers [BS_N][n].
// BS_GENPROF: 1 for a device whose generation is its profile (`gen`), 0 for each layer's G_ehp.
enum BuildScalar {
    BS_T, BS_WL, BS_WR, BS_GFRAC, BS_VPRE, BS_GSH, BS_PRE, BS_FIX, BS_ELEC, BS_GENPROF, BS_N
};
// Per-layer flags [nl][n] (bits): bulk traps, mobile and fixed anions and cations, Langevin's
// rate for direct recombination, and that rate at the mobilities of the interval after each
// point (SIMsalabim's) rather than the layer's.
enum BuildFlag {
    BF_TRAPS = 1, BF_AMOBILE = 2, BF_AFIXED = 4, BF_CMOBILE = 8, BF_CFIXED = 16,
    BF_LANGEVIN = 32, BF_LANGEVIN_INTERVAL = 64, BF_LAYERGEN = 128
};

// The packed arrays' fields (batch.py SEG_FIELDS, SC_FIELDS, the points' PT fields).
enum BuildSeg {
    BG_EPS, BG_OFFN, BG_OFFP, BG_NC, BG_DOP, BG_KD, BG_NI2, BG_NT, BG_CN, BG_CP, BG_NT0,
    BG_PT0, BG_AM, BG_CM, BG_MUN, BG_MUP, BG_EC, BG_NI, BG_SP, BG_SCS, BG_IMUN, BG_IMUP,
    BG_TT, BG_N  // BG_TT: the traps' charge type (held: no tangent)
};
static_assert((int)BG_N == (int)SEG_N, "the builder writes every packed segment field");
enum BuildSc {
    BC_L, BC_VT, BC_GFRAC, BC_V0, BC_VL, BC_WL, BC_WR, BC_LNL, BC_LNR, BC_LPL, BC_LPR, BC_ELEC,
    BC_AT, BC_CT, BC_VPRE, BC_JS, BC_GSH, BC_NFLOOR, BC_PRE, BC_FIX, BC_N
};
enum BuildPt { BP_H, BP_X, BP_G, BP_FIX, BP_N };

struct BuildArgs {
    int n, NP, nl, nseg, K;
    const uint8_t *lid;       // [NP + 2][n] each point's layer
    const uint8_t *sid;       // [NP + 2][n] each point's segment
    const uint8_t *flags;     // [nl][n]
    const int32_t *i0;        // [nl][n] each layer's first point
    const double *unit;       // [NP + 2][n] the unscaled spacing (the interval after the point)
    const double *unit_sum;   // [nl][n] its sum over each layer
    const double *layers;     // [nl][BL_N][n]
    const double *scalars;    // [BS_N][n]
    const double *gen;        // [2][gen_m][n] each device's generation profile: positions, G
    int gen_m;                // its nodes (0: no device takes a profile)
    const double *dlayers;    // [K][nl][BL_N][n] the directions (K = 0: none)
    const double *dscalars;   // [K][BS_N][n]
    double *pt, *seg, *sc, *imu;     // [NP + 2][4][n], [nseg][22][n], [20][n], [nseg][2][n]
    double *dpt, *dseg, *dsc;        // [K][...] the tangents (K > 0)
};

// A number and its derivative along the lane's direction.
struct BD {
    double v, d;
};
DDH_BFN BD bd(double v, double d = 0.0) { BD r; r.v = v; r.d = d; return r; }
DDH_BFN BD operator+(BD a, BD b) { return bd(a.v + b.v, a.d + b.d); }
DDH_BFN BD operator-(BD a, BD b) { return bd(a.v - b.v, a.d - b.d); }
DDH_BFN BD operator-(BD a) { return bd(-a.v, -a.d); }
DDH_BFN BD operator*(BD a, BD b) { return bd(a.v * b.v, a.d * b.v + a.v * b.d); }
DDH_BFN BD operator*(BD a, double c) { return bd(a.v * c, a.d * c); }
DDH_BFN BD operator*(double c, BD a) { return bd(a.v * c, a.d * c); }
DDH_BFN BD operator/(BD a, BD b) {
    const double q = a.v / b.v;
    return bd(q, (a.d - q * b.d) / b.v);
}
DDH_BFN BD operator/(BD a, double c) { return bd(a.v / c, a.d / c); }
DDH_BFN BD operator/(double c, BD b) {
    const double q = c / b.v;
    return bd(q, -q * b.d / b.v);
}
DDH_BFN BD dexp1(BD a) { const double e = exp(a.v); return bd(e, e * a.d); }
DDH_BFN BD dlog1(BD a) { return bd(log(a.v), a.d / a.v); }
DDH_BFN BD dmax1(BD a, BD b) { return b.v > a.v ? b : a; }
// max(a, b) as `jnp.maximum` differentiates it: at a tie the mean of the two derivatives
// (what central differences see at the kink).
DDH_BFN BD dmax_mean(BD a, BD b) {
    if (a.v == b.v) return bd(a.v, 0.5 * (a.d + b.d));
    return b.v > a.v ? b : a;
}

// A device's view for one lane.
struct BuildLane {
    int d, k;  // the device, the direction (-1: none)
    size_t n;
};

DDH_BFN BD blayer(const BuildArgs &a, const BuildLane &L, int l, int f) {
    const size_t i = ((size_t)l * BL_N + f) * L.n + L.d;
    const double dv = L.k >= 0 ? a.dlayers[(size_t)L.k * a.nl * BL_N * L.n + i] : 0.0;
    return bd(a.layers[i], dv);
}

DDH_BFN BD bscalar(const BuildArgs &a, const BuildLane &L, int f) {
    const size_t i = (size_t)f * L.n + L.d;
    const double dv = L.k >= 0 ? a.dscalars[(size_t)L.k * BS_N * L.n + i] : 0.0;
    return bd(a.scalars[i], dv);
}

// The per-layer quantities the points take (reference: src/jaxdd/device.py:890 the intrinsic
// density, :933-941 the trap level's reference densities, :976 and :1034 the band offsets
// from the first layer's with the density of states' part).
struct BuildLayerVals {
    BD offn, offp, ni, nt0, pt0;
};

DDH_BFN BuildLayerVals blayer_vals(const BuildArgs &a, const BuildLane &L, int l, BD Vt) {
    BuildLayerVals r;
    const BD Ec = blayer(a, L, l, BL_EC), Ev = blayer(a, L, l, BL_EV), Nc = blayer(a, L, l, BL_NC);
    const BD Ec0 = blayer(a, L, 0, BL_EC), Ev0 = blayer(a, L, 0, BL_EV), Nc0 = blayer(a, L, 0, BL_NC);
    const BD Et = blayer(a, L, l, BL_ET);
    const BD dos = Vt * dlog1(Nc / Nc0);
    r.offn = (Ec - Ec0) + dos;
    r.offp = (Ev - Ev0) - dos;
    r.ni = Nc * dexp1((Ev - Ec) * (-0.5) / Vt);
    r.nt0 = Nc * dexp1((Ec - Et) / Vt);
    r.pt0 = Nc * dexp1((Et - Ev) / Vt);
    return r;
}

// What the lane writes: values on direction 0 (or without directions), its derivative always.
DDH_BFN void bput(const BuildLane &L, double *val, double *tan, size_t off, size_t slice, BD x) {
    if (L.k <= 0) val[off] = x.v;
    if (L.k >= 0) tan[(size_t)L.k * slice + off] = x.d;
}

// One segment's 22 fields at its first point j of layer l (batch.py `_point_fields`): the
// interval after a layer's last point (but the device's) takes the interface's mobility
// L h nu / V_T (device.py:894-898).
DDH_BFN void bsegment(const BuildArgs &a, const BuildLane &L, int j, int l, BD total, BD hj,
                      BD Vt, bool crossing, BD *f) {
    const int P = a.NP + 2;
    const bool interior = j > 0 && j < P - 1;
    const uint8_t fl = a.flags[(size_t)l * L.n + L.d];
    const BuildLayerVals lv = blayer_vals(a, L, l, Vt);
    const BD zero = bd(0.0);
    const BD Nc = blayer(a, L, l, BL_NC);
    f[BG_EPS] = blayer(a, L, l, BL_EPSR) * BUILD_EPS0;
    f[BG_OFFN] = lv.offn;
    f[BG_OFFP] = lv.offp;
    f[BG_NC] = Nc;
    f[BG_DOP] = blayer(a, L, l, BL_NA) - blayer(a, L, l, BL_ND);
    const BD mun = crossing ? total * hj * blayer(a, L, l, BL_NUN) / Vt : blayer(a, L, l, BL_MUN);
    const BD mup = crossing ? total * hj * blayer(a, L, l, BL_NUP) / Vt : blayer(a, L, l, BL_MUP);
    BD kd = blayer(a, L, l, BL_KD);
    if (fl & BF_LANGEVIN) {  // preLangevin q (mu_n + mu_p) / eps (device.py:900-911)
        const BD rate = blayer(a, L, l, BL_PRELANGEVIN) * BUILD_Q / (blayer(a, L, l, BL_EPSR) * BUILD_EPS0);
        kd = (fl & BF_LANGEVIN_INTERVAL) ? rate * (mun + mup)
                                          : rate * (blayer(a, L, l, BL_MUN) + blayer(a, L, l, BL_MUP));
    }
    f[BG_KD] = interior ? kd : zero;
    f[BG_NI2] = lv.ni * lv.ni;
    const bool trap = interior && (fl & BF_TRAPS);
    f[BG_NT] = trap ? blayer(a, L, l, BL_NT) : zero;
    f[BG_CN] = trap ? blayer(a, L, l, BL_CN) : zero;
    f[BG_CP] = trap ? blayer(a, L, l, BL_CP) : zero;
    f[BG_NT0] = trap ? lv.nt0 : zero;
    f[BG_PT0] = trap ? lv.pt0 : zero;
    f[BG_AM] = bd((fl & BF_AMOBILE) ? 1.0 : 0.0);
    f[BG_CM] = bd((fl & BF_CMOBILE) ? 1.0 : 0.0);
    f[BG_MUN] = mun;
    f[BG_MUP] = mup;
    f[BG_EC] = blayer(a, L, l, BL_EC);
    f[BG_NI] = lv.ni;
    f[BG_SP] = 1.0 / (Nc * BUILD_Q);
    f[BG_SCS] = 1.0 / (dmax_mean(mun, mup) * Vt * Nc);
    f[BG_IMUN] = 1.0 / mun;
    f[BG_IMUP] = 1.0 / mup;
    f[BG_TT] = trap ? bd(blayer(a, L, l, BL_TRAPTYPE).v) : zero;
}

DDH_BFN void bput_segment(const BuildArgs &a, const BuildLane &L, int s, const BD *f, BD mua,
                          BD muc) {
    const size_t slice = (size_t)a.nseg * BG_N * L.n;
    for (int q = 0; q < BG_N; ++q)
        bput(L, a.seg, a.dseg, ((size_t)s * BG_N + q) * L.n + L.d, slice, f[q]);
    if (L.k <= 0) {
        a.imu[((size_t)s * 2 + 0) * L.n + L.d] = mua.v;
        a.imu[((size_t)s * 2 + 1) * L.n + L.d] = muc.v;
    }
}

// The points' window: each point's spacing before the interface rule, emitted two points later.
struct BuildPoint {
    BD h;
    int j, l;
};

struct BuildRun {
    BD x, total, Vt, hG, hsum;  // the running position, sums for the current scale
    BD gscale;                  // the profile's stretch to the stack: total / its last position
    int gk;                     // the profile's interval at the running position
    BD Ncmax;
    int seg_prev;
    BD last[BG_N];              // the device's last segment (its padding)
    BD last_mua, last_muc;
};

// The generation profile at the running position (device.py:804-814 `_generation`): the
// profile's positions a_k stretched to the stack; the interval k = #{a_1..a_{m-1} <= x}
// (walked forward as the points are emitted in order), its value
// G_k + (x - a_k)(G_{k+1} - G_k) / (a_{k+1} - a_k), extrapolated before a_1; none at or beyond
// the last position. A shorter table padded by repeating its last node takes the same values.
DDH_BFN BD bprofile(const BuildArgs &a, const BuildLane &L, BuildRun &R) {
    const int m = a.gen_m;
    const size_t n = L.n;
    const double *gx = a.gen + L.d, *gg = a.gen + (size_t)m * n + L.d;
    auto axis = [&](int k) { return gx[(size_t)k * n] * R.gscale; };
    const BD x = R.x;
    if (!(x.v < axis(m - 1).v)) return bd(0.0);
    while (R.gk + 1 < m - 1 && axis(R.gk + 1).v <= x.v) ++R.gk;
    const int k = R.gk;
    const BD ak = axis(k), ak1 = axis(k + 1);
    const double Gk = gg[(size_t)k * n], Gk1 = gg[(size_t)(k + 1) * n];
    return bd(Gk) + (x - ak) * bd(Gk1 - Gk) / (ak1 - ak);
}

DDH_BFN void bemit(const BuildArgs &a, const BuildLane &L, const BuildPoint &p, BuildRun &R) {
    const int P = a.NP + 2, j = p.j, l = p.l;
    const size_t pslice = (size_t)P * BP_N * L.n;
    const uint8_t fl = a.flags[(size_t)l * L.n + L.d];
    const bool interior = j > 0 && j < P - 1;
    bput(L, a.pt, a.dpt, ((size_t)j * BP_N + BP_H) * L.n + L.d, pslice, p.h);
    bput(L, a.pt, a.dpt, ((size_t)j * BP_N + BP_X) * L.n + L.d, pslice, R.x);
    BD G = bd(0.0);
    if (interior) {
        if (a.scalars[(size_t)BS_GENPROF * L.n + L.d] > 0) {
            // a profile asked for without one handed over: NaN, not a read past `gen`
            if (a.gen_m < 2) G = bd(__builtin_nan(""));
            else if (fl & BF_LAYERGEN) G = bprofile(a, L, R);
        } else {
            G = blayer(a, L, l, BL_G);
        }
    }
    bput(L, a.pt, a.dpt, ((size_t)j * BP_N + BP_G) * L.n + L.d, pslice, G);
    const BD af = (fl & BF_AFIXED) ? blayer(a, L, l, BL_NANION) : bd(0.0);
    const BD cf = (fl & BF_CFIXED) ? blayer(a, L, l, BL_NCATION) : bd(0.0);
    bput(L, a.pt, a.dpt, ((size_t)j * BP_N + BP_FIX) * L.n + L.d, pslice, af - cf);
    if (j < P - 1) {  // the current scale's sums over the intervals
        R.hG = R.hG + p.h * G;
        R.hsum = R.hsum + p.h;
    }
    const BD Nc = blayer(a, L, l, BL_NC);
    R.Ncmax = dmax1(R.Ncmax, Nc);
    const int s = a.sid[(size_t)j * L.n + L.d];
    if (s != R.seg_prev) {
        // the interval after point j crosses an interface when j is a layer's last point
        // (not the device's): the next point starts another layer
        const bool crossing = j < P - 1 && a.lid[(size_t)(j + 1) * L.n + L.d] != l;
        bsegment(a, L, j, l, R.total, p.h, R.Vt, crossing, R.last);
        R.last_mua = blayer(a, L, l, BL_MUA);
        R.last_muc = blayer(a, L, l, BL_MUC);
        bput_segment(a, L, s, R.last, R.last_mua, R.last_muc);
        R.seg_prev = s;
    }
    R.x = R.x + R.total * p.h;
}

// One lane: device d along direction k (-1: values only).
DDH_BFN void build_lane(const BuildArgs &a, int d, int k) {
    BuildLane L;
    L.d = d; L.k = k; L.n = (size_t)a.n;
    const int P = a.NP + 2, nl = a.nl;
    BuildRun R;
    R.Vt = bscalar(a, L, BS_T) * BUILD_KB / BUILD_Q;
    R.total = bd(0.0);
    for (int l = 0; l < nl; ++l) R.total = R.total + blayer(a, L, l, BL_L);
    R.x = bd(0.0); R.hG = bd(0.0); R.hsum = bd(0.0);
    R.gk = 0;
    R.gscale = a.gen_m > 0 ? R.total / a.gen[(size_t)(a.gen_m - 1) * L.n + d] : bd(0.0);
    R.Ncmax = blayer(a, L, 0, BL_NC);
    R.seg_prev = -1;
    // each layer's scale of its unscaled spacing: L_l / (total unit_sum_l)
    BD scale[BUILD_MAX_LAYERS];
    for (int l = 0; l < nl; ++l)
        scale[l] = blayer(a, L, l, BL_L) / (R.total * a.unit_sum[(size_t)l * L.n + d]);
    BuildPoint w0 = {}, w1 = {};  // the two points not yet emitted (w0 the older)
    int held = 0;
    int next_iface = 1;  // the next layer whose first point is an interface's
    for (int j = 0; j < P; ++j) {
        const int l = a.lid[(size_t)j * L.n + d];
        BuildPoint c;
        c.j = j; c.l = l;
        c.h = scale[l] * a.unit[(size_t)j * L.n + d];
        if (next_iface < nl && j == a.i0[(size_t)next_iface * L.n + d]) {
            // SIMsalabim's rule: the three spacings around the interface take their mean
            const BD m = (c.h + w1.h + w0.h) / 3.0;
            w0.h = m; w1.h = m; c.h = m;
            ++next_iface;
        }
        if (held == 0) {
            w0 = c;
            held = 1;
        } else if (held == 1) {
            w1 = c;
            held = 2;
        } else {
            bemit(a, L, w0, R);
            w0 = w1;
            w1 = c;
        }
    }
    bemit(a, L, w0, R);
    if (held == 2) bemit(a, L, w1, R);
    // the segments beyond the device's own count repeat its last (never read)
    for (int s = R.seg_prev + 1; s < a.nseg; ++s) bput_segment(a, L, s, R.last, R.last_mua, R.last_muc);
    // the scalars (batch.py `_scalars`)
    const size_t cslice = (size_t)BC_N * L.n;
    auto put = [&](int f, BD v) { bput(L, a.sc, a.dsc, (size_t)f * L.n + d, cslice, v); };
    const BD WL = bscalar(a, L, BS_WL), WR = bscalar(a, L, BS_WR), Vt = R.Vt;
    const BD Ec0 = blayer(a, L, 0, BL_EC), Ev0 = blayer(a, L, 0, BL_EV), Nc0 = blayer(a, L, 0, BL_NC);
    const BD EcN = blayer(a, L, nl - 1, BL_EC), EvN = blayer(a, L, nl - 1, BL_EV);
    const BD NcN = blayer(a, L, nl - 1, BL_NC);
    put(BC_L, R.total);
    put(BC_VT, Vt);
    put(BC_GFRAC, bscalar(a, L, BS_GFRAC));
    put(BC_V0, (WR - WL) * 0.5);
    put(BC_VL, (WL - WR) * 0.5);
    put(BC_WL, WL);
    put(BC_WR, WR);
    put(BC_LNL, dlog1(Nc0) + (Ec0 - WL) / Vt);
    put(BC_LNR, dlog1(NcN) + (EcN - WR) / Vt);
    put(BC_LPL, dlog1(Nc0) + (WL - Ev0) / Vt);
    put(BC_LPR, dlog1(NcN) + (WR - EvN) / Vt);
    put(BC_ELEC, bscalar(a, L, BS_ELEC));
    BD at = bd(0.0), ct = bd(0.0);
    for (int l = 0; l < nl; ++l) {
        const uint8_t fl = a.flags[(size_t)l * L.n + d];
        if (fl & BF_AMOBILE) at = at + blayer(a, L, l, BL_L) * blayer(a, L, l, BL_NANION);
        if (fl & BF_CMOBILE) ct = ct + blayer(a, L, l, BL_L) * blayer(a, L, l, BL_NCATION);
    }
    put(BC_AT, at);
    put(BC_CT, ct);
    put(BC_VPRE, bscalar(a, L, BS_VPRE));
    // the current scale (batch.py:96, 115): the generation's current, at least 1 A/m^2
    const BD js = bd(BUILD_Q) * R.total * R.hG / R.hsum;
    put(BC_JS, js.v >= 1.0 ? js : bd(1.0));
    put(BC_GSH, bscalar(a, L, BS_GSH));
    put(BC_NFLOOR, R.Ncmax * BUILD_DENSITY_FLOOR);
    put(BC_PRE, bscalar(a, L, BS_PRE));
    put(BC_FIX, bscalar(a, L, BS_FIX));
}

// ---- the graded grid's layout on the GPU (jaxdd_amd/grid.py `graded_layout`; reference:
// src/jaxdd/device.py:160-214 the graded rule, :114-136 `_layout`) -------------------------
//
// The point density 1/H + beta/(d + d0) in the distance d to either end of a layer (H the
// stack's thickness over NP) integrates in closed form to F(u); each layer gets its share of
// F over the stack (at least five points beyond its first, the largest layer giving back what
// the minimum added), and its points equidistribute F: u_i with F(u_i) = F(L_k) i/(m + 1),
// found by Newton steps kept inside a shrinking bracket (a bisection step where Newton would
// leave it), each point until its step no longer moves it. Three passes: the structure (one
// lane per device), every spacing (one lane per point), each layer's sum (one lane per device).

constexpr int GRADED_MIN_POINTS = 5;
constexpr int GRADED_NEWTON_STEPS = 80;

struct LayoutArgs {
    int n, NP, nl;
    double beta, d0;
    const double *L;     // [nl][n] the thicknesses
    uint8_t *lid, *sid;  // [NP + 2][n]
    int32_t *i0;         // [nl][n]
    double *unit;        // [NP + 2][n]
    double *unit_sum;    // [nl][n]
    uint8_t *iface;      // [NP + 1][n] 1 where the interval after the point crosses layers
    int32_t *nseg;       // [n] each device's segments; -1: the counts do not fill the grid
};

DDH_BFN void graded_integral(double u, double Lk, double H, double beta, double d0, double &F,
                             double &dF) {
    F = u / H + beta * (log1p(u / d0) - log1p((Lk - u) / d0) + log1p(Lk / d0));
    dF = 1.0 / H + beta / (d0 + u) + beta / (d0 + Lk - u);
}

// u in [0, Lk] with F(u) = t, from the linear guess (t / t_end) Lk
DDH_BFN double graded_invert(double t, double t_end, double Lk, double H, double beta, double d0) {
    double lo = 0.0, hi = Lk;
    double u = t_end > 0 ? t / t_end * Lk : 0.0;
    for (int it = 0; it < GRADED_NEWTON_STEPS; ++it) {
        double F, dF;
        graded_integral(u, Lk, H, beta, d0, F, dF);
        const bool below = F < t;
        if (below) lo = u;
        else hi = u;
        const double step = u - (F - t) / dF;
        const double next = (step > lo && step < hi) ? step : 0.5 * (lo + hi);
        if (next == u) break;
        u = next;
    }
    return u;
}

// A device's layer weights and counts: its thicknesses, H, each layer's share w_k of the
// density's integral and its first and last point.
struct LayoutDev {
    double L[BUILD_MAX_LAYERS], w[BUILD_MAX_LAYERS], H;
    int starts[BUILD_MAX_LAYERS], ends[BUILD_MAX_LAYERS];
    bool fits;
};

DDH_BFN void layout_device(const LayoutArgs &a, int d, LayoutDev &D) {
    const size_t n = (size_t)a.n;
    const int NP = a.NP, nl = a.nl, P = NP + 2;
    int counts[BUILD_MAX_LAYERS] = {};
    double Ls = 0.0;
    for (int k = 0; k < nl; ++k) {
        D.L[k] = a.L[(size_t)k * n + d];
        Ls += D.L[k];
    }
    D.H = Ls / NP;
    double wsum = 0.0;
    for (int k = 0; k < nl; ++k) {
        double dF;
        graded_integral(D.L[k], D.L[k], D.H, a.beta, a.d0, D.w[k], dF);
        wsum += D.w[k];
    }
    int total = 0, big = 0;
    for (int k = 0; k < nl; ++k) {
        const int c = (int)floor((NP + 2 - nl) * D.w[k] / wsum);
        counts[k] = c > GRADED_MIN_POINTS ? c : GRADED_MIN_POINTS;
        total += counts[k];
        if (counts[k] > counts[big]) big = k;
    }
    const int over = total + nl - P;  // points the minimum added beyond NP + 2
    if (over > 0) counts[big] -= over;
    int start = 0;
    for (int k = 0; k < nl; ++k) {
        D.starts[k] = start;
        D.ends[k] = k == nl - 1 ? P - 1 : start + counts[k];
        start = D.ends[k] + 1;
    }
    D.fits = D.ends[nl - 1] >= D.starts[nl - 1];
}

// Pass 1, one lane per device: each layer's first point, each point's layer, the segments
// (a new one wherever a packed field may change), the interfaces.
DDH_BFN void layout_structure(const LayoutArgs &a, int d) {
    const size_t n = (size_t)a.n;
    const int nl = a.nl, P = a.NP + 2;
    LayoutDev D;
    layout_device(a, d, D);
    for (int k = 0; k < nl; ++k) a.i0[(size_t)k * n + d] = D.starts[k];
    if (!D.fits) {
        a.nseg[d] = -1;
        return;
    }
    int s = 0, key_prev = -1, k = 0;
    for (int j = 0; j < P; ++j) {
        while (k < nl - 1 && j > D.ends[k]) ++k;
        a.lid[(size_t)j * n + d] = (uint8_t)k;
        const bool crossing = k < nl - 1 && j == D.ends[k];
        const int key = k * 8 + (j == 0 ? 1 : 0) + (j == P - 1 ? 2 : 0) + (crossing ? 4 : 0);
        if (j > 0 && key != key_prev) ++s;
        key_prev = key;
        a.sid[(size_t)j * n + d] = (uint8_t)s;
        if (j < P - 1) a.iface[(size_t)j * n + d] = crossing ? 1 : 0;
    }
    a.nseg[d] = s + 1;
}

// Pass 2, one lane per (device, point j): the spacing after point j, the difference of the
// two positions around it (the layer's first position 0, its last its thickness).
DDH_BFN void layout_point(const LayoutArgs &a, int d, int j) {
    if (a.nseg[d] < 0) return;
    const size_t n = (size_t)a.n;
    LayoutDev D;
    layout_device(a, d, D);
    const int k = a.lid[(size_t)j * n + d];
    const int m = D.ends[k] - D.starts[k], i = j - D.starts[k] + 1;
    const double t_end = D.w[k] * (m + 1) / (m + 1);
    auto position = [&](int q) {
        if (q == 0) return 0.0;
        if (q == m + 1) return D.L[k];
        return graded_invert(D.w[k] * q / (m + 1), t_end, D.L[k], D.H, a.beta, a.d0);
    };
    a.unit[(size_t)j * n + d] = position(i) - position(i - 1);
}

// Pass 3, one lane per device: each layer's spacing sum.
DDH_BFN void layout_sums(const LayoutArgs &a, int d) {
    if (a.nseg[d] < 0) return;
    const size_t n = (size_t)a.n;
    const int nl = a.nl, P = a.NP + 2;
    double sum[BUILD_MAX_LAYERS] = {};
    for (int j = 0; j < P; ++j) sum[a.lid[(size_t)j * n + d]] += a.unit[(size_t)j * n + d];
    for (int k = 0; k < nl; ++k) a.unit_sum[(size_t)k * n + d] = sum[k];
}

}  // namespace ddh
