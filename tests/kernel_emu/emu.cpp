// Test-only CPU build of the HIP kernels' per-lane logic (jaxdd_amd/csrc/ddhip_core.h).
//
// It replays the 64-lane wavefronts of the persistent solve kernel lane by lane, with the
// same "any lane active" / "any lane with mobile ions" / "any lane needs FP64" wave votes and
// the wavefronts interleaved step by step, so that the kernels' algorithm can be debugged and
// compared with the oracle on a host without a GPU. Nothing in the product loads it.
#define DDH_CPU_EMULATION 1
#include "ddhip_core.h"
#include "ddhip_adjoint.h"
#include "ddhip_assemble.h"
#include "ddhip_build.h"
#include "ddhip_impedance.h"
#include "ddhip_pair.h"
#include "ddhip_quad.h"
#include "ddhip_transient.h"

#include <cstdlib>
#include <vector>

using namespace ddh;

// The wavefronts of the GPU run concurrently, each at its own phase (Poisson's start or a
// Newton step) and precision: here they take turns, one unit of work each per round, so that
// memory one wavefront's phase shares with another's would show here too.
static void run(Args &a) {
    const int nw = (a.n + 63) / 64;
    std::vector<char> started(nw, 0), done(nw, 0);
    // each wavefront's local data share (the lanes' segment tables)
    std::vector<double> lds_all((size_t)nw * a.nseg * SEG_N * WAVE);
    int remaining = nw;
    while (remaining > 0) {
        for (int w = 0; w < nw; ++w) {
            if (done[w]) continue;
            int w0 = w * 64, w1 = w0 + 64 < a.n ? w0 + 64 : a.n;
            if (!started[w]) {
                for (int lane = w0; lane < w1; ++lane) {
                    fill_segments(a, lane, &lds_all[(size_t)w * a.nseg * SEG_N * WAVE] + (lane - w0));
                    poisson_lane(a, lane, a.st[lane]);
                    a.st[lane].part = 0; a.st[lane].parts = 1;
                    start_lane(a, lane, a.st[lane]);
                }
                started[w] = 1;
                continue;
            }
            double *lds = &lds_all[(size_t)w * a.nseg * SEG_N * WAVE];
            bool any = false, ions = false, f64 = !a.mixed;
            for (int lane = w0; lane < w1; ++lane) {
                bool act = lane_active(a.st[lane]);
                any |= act;
                ions |= act && lane_mobile(a.st[lane]);
                f64 |= act && lane_needs64(a.st[lane]);
            }
            if (!any) { done[w] = 1; --remaining; continue; }
            for (int lane = w0; lane < w1; ++lane) {
                if (!lane_active(a.st[lane])) continue;
                if (f64) newton_lane_step<double>(a, lane, ions, lds + (lane - w0), a.st[lane]);
                else newton_lane_step<float>(a, lane, ions, lds + (lane - w0), a.st[lane]);
            }
        }
    }
}

// Two lanes per device (ddhip_pair.h): wavefronts of 32 devices; each pair's three phases
// run lane by lane, the records exchanged between the phases as the GPU's shuffles do.
static void run_pair(Args &a) {
    const int nw = (a.n + 31) / 32;
    std::vector<char> started(nw, 0), done(nw, 0);
    std::vector<double> lds_all((size_t)nw * a.nseg * SEG_N * PAIR_LS);
    int remaining = nw;
    while (remaining > 0) {
        for (int w = 0; w < nw; ++w) {
            if (done[w]) continue;
            const int d0 = w * 32, d1 = d0 + 32 < a.n ? d0 + 32 : a.n;
            double *lds = &lds_all[(size_t)w * a.nseg * SEG_N * PAIR_LS];
            if (!started[w]) {
                for (int d = d0; d < d1; ++d) {
                    fill_segments(a, d, lds + (d - d0), PAIR_LS);
                    poisson_lane(a, d, a.st[2 * d]);
                    a.st[2 * d + 1] = a.st[2 * d];
                    for (int h = 0; h < 2; ++h) {
                        a.st[2 * d + h].part = h; a.st[2 * d + h].parts = 2;
                        start_lane(a, d, a.st[2 * d + h]);
                    }
                }
                started[w] = 1;
                continue;
            }
            bool any = false, ions = false, f64 = !a.mixed;
            for (int q = 2 * d0; q < 2 * d1; ++q) {
                bool act = lane_active(a.st[q]);
                any |= act;
                ions |= act && lane_mobile(a.st[q]);
                f64 |= act && lane_needs64(a.st[q]);
            }
            if (!any) { done[w] = 1; --remaining; continue; }
            for (int d = d0; d < d1; ++d) {
                LaneState &S0 = a.st[2 * d], &S1 = a.st[2 * d + 1];
                if (!lane_active(S0)) continue;
                PairStep p0, p1;
                const double *l = lds + (d - d0);
                if (f64) {
                    pair_phase1<double>(a, d, 0, ions, l, S0, p0);
                    pair_phase1<double>(a, d, 1, ions, l, S1, p1);
                    pair_phase2<double>(a, d, 0, ions, p0, p1.e);
                    pair_phase2<double>(a, d, 1, ions, p1, p0.e);
                } else {
                    pair_phase1<float>(a, d, 0, ions, l, S0, p0);
                    pair_phase1<float>(a, d, 1, ions, l, S1, p1);
                    pair_phase2<float>(a, d, 0, ions, p0, p1.e);
                    pair_phase2<float>(a, d, 1, ions, p1, p0.e);
                }
                pair_phase3(a, d, 0, S0, p0, p1.b);
                pair_phase3(a, d, 1, S1, p1, p0.b);
            }
        }
    }
}

// Four lanes per device (ddhip_quad.h): wavefronts of 16 devices; each device's phases run
// lane by lane, the records exchanged between them as the GPU's shuffles do.
static void run_quad(Args &a) {
    const int nw = (a.n + 15) / 16;
    std::vector<char> started(nw, 0), done(nw, 0);
    const size_t per = quad_lds_doubles(a.nseg);
    std::vector<double> lds_all((size_t)nw * per);
    int remaining = nw;
    while (remaining > 0) {
        for (int w = 0; w < nw; ++w) {
            if (done[w]) continue;
            const int d0 = w * 16, d1 = d0 + 16 < a.n ? d0 + 16 : a.n;
            double *lds = &lds_all[(size_t)w * per];
            if (!started[w]) {
                for (int d = d0; d < d1; ++d) {
                    if (quad_refused(a, d)) {
                        quad_refuse(a, d);
                        for (int q = 0; q < 4; ++q) a.st[4 * d + q].phase = PH_DONE;
                        continue;
                    }
                    fill_segments(a, d, lds + (d - d0), QUAD_LS);
                    poisson_lane(a, d, a.st[4 * d]);
                    for (int q = 1; q < 4; ++q) a.st[4 * d + q] = a.st[4 * d];
                    for (int q = 0; q < 4; ++q) {
                        a.st[4 * d + q].part = q; a.st[4 * d + q].parts = 4;
                        start_lane(a, d, a.st[4 * d + q]);
                    }
                }
                started[w] = 1;
                continue;
            }
            bool any = false, f64 = !a.mixed;
            for (int k = 4 * d0; k < 4 * d1; ++k) {
                bool act = lane_active(a.st[k]);
                any |= act;
                f64 |= act && lane_needs64(a.st[k]);
            }
            if (!any) { done[w] = 1; --remaining; continue; }
            for (int d = d0; d < d1; ++d) {
                LaneState *S = &a.st[4 * d];
                if (!lane_active(S[0])) continue;
                QuadStep st[4];
                const double *l = lds + (d - d0);
                for (int q = 0; q < 4; ++q) {
                    double *pre = lds + quad_pre_column(a.nseg, d - d0, q);
                    if (f64) quad_phase1<double>(a, d, q, l, pre, S[q], st[q]);
                    else quad_phase1<float>(a, d, q, l, pre, S[q], st[q]);
                }
                double Xa[4][12], Xb[4][12], W[4][12];
                for (int q = 0; q < 4; ++q) {
                    quad_meet(st[q & ~1].e.m, st[q | 1].e.m, Xa[q], Xb[q]);
                    quad_middle_part(st[q].e, q, Xa[q], Xb[q], W[q]);
                }
                for (int q = 0; q < 4; ++q) {
                    double xs[3], sv[3];
                    quad_center(W[1], W[2], q, Xa[q], Xb[q], xs, sv);
                    if (f64) quad_phase2<double>(a, d, q, st[q], xs, sv);
                    else quad_phase2<float>(a, d, q, st[q], xs, sv);
                }
                PairBack h[4];
                for (int q = 0; q < 4; ++q) quad_half(st[q & ~1].b, st[q | 1].b, q >> 1, h[q]);
                for (int q = 0; q < 4; ++q) quad_phase3(a, d, S[q], h[q & 1], h[2 | (q & 1)]);
            }
        }
    }
}

static int g_lanes = 1;  // the next solve's lanes per device (jv mode: 1)
extern "C" void emu_set_lanes(int lanes) { g_lanes = lanes; }

static int g_poison = 0;  // fill the workspace with this byte before a solve (0: zeros)
extern "C" void emu_set_poison(int byte) { g_poison = byte; }

static double *g_states = nullptr;  // where the next solve keeps each voltage's final state
static double *g_pre = nullptr;     // and the pre-bias solutions
extern "C" void emu_keep_states(double *states, double *pre) { g_states = states; g_pre = pre; }

extern "C" int emu_solve_batch(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                               const double *seg, const double *sc, const double *V, int nV,
                               double tol, int max_iter, int start_iter, int strict, int bern,
                               int precision, int jv, double *J, uint8_t *ok, int32_t *its,
                               int32_t *status) {
    std::vector<double> Y((size_t)NBUF * NP * 3 * n);
    const size_t nf = NFAC_Q > NFAC ? NFAC_Q : NFAC;
    std::vector<double> fac64((size_t)NP * nf * n), pfac((size_t)NP * 4 * n);
    std::vector<float> fac32((size_t)NP * nf * n);
    std::vector<double> fix((size_t)NP * n);
    std::vector<LaneState> st(4 * (size_t)n);
    if (g_poison) {
        memset(Y.data(), g_poison, Y.size() * sizeof(double));
        memset(fac64.data(), g_poison, fac64.size() * sizeof(double));
        memset(fac32.data(), g_poison, fac32.size() * sizeof(float));
        memset(pfac.data(), g_poison, pfac.size() * sizeof(double));
        memset(fix.data(), g_poison, fix.size() * sizeof(double));
        memset((void *)st.data(), g_poison, st.size() * sizeof(LaneState));
    }
    Args a;
    a.n = n; a.NP = NP; a.nV = nV; a.nseg = nseg;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.V = V;
    a.Y = Y.data(); a.fac64 = fac64.data(); a.fac32 = fac32.data(); a.pfac = pfac.data();
    a.fix = fix.data(); a.st = st.data(); a.states = g_states; a.pre = g_pre;
    a.J = J; a.ok = ok; a.its = its; a.status = status;
    a.tol = tol; a.max_iter = max_iter; a.start_iter = start_iter; a.strict = strict;
    a.bern = bern; a.jv = jv; a.mixed = precision ? 1 : 0;
    if (g_lanes == 4 && !jv) run_quad(a);
    else if (g_lanes == 2 && !jv) run_pair(a);
    else run(a);
    return 0;
}

// The adjoint kernel's wavefronts in turn: one per device, a lane per voltage.
static int g_adj_passes = 0;  // 1: the adjoint's directions one pass each
extern "C" void emu_set_adjoint_passes(int passes) { g_adj_passes = passes; }

extern "C" int emu_adjoint(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                           const double *seg, const double *sc, const double *V, int nV,
                           const double *states, const double *pre, const uint8_t *ok, int ntheta,
                           const double *dpt, const double *dseg, const double *dsc, int bern,
                           double *grad, int32_t *status) {
    const size_t ls = 64;  // one workgroup: the devices in turn, a lane per voltage
    std::vector<double> fac((size_t)NP * ADJ_F * ls), st((size_t)NP * 3 * ls),
        lam((size_t)NP * 3 * ls), iu((size_t)(NP + 1) * ls), ts((size_t)(NP + 2) * 2 * ls),
        ion((size_t)4 * ls), nu((size_t)NP * ls);
    AdjArgs a;
    a.n = n; a.NP = NP; a.nV = nV; a.nseg = nseg; a.ntheta = ntheta; a.bern = bern;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.V = V; a.states = states; a.ok = ok;
    a.pre = pre;
    a.dpt = dpt; a.dseg = dseg; a.dsc = dsc; a.grad = grad; a.status = status;
    a.fac = fac.data(); a.st = st.data(); a.lam = lam.data(); a.iu = iu.data(); a.ls = ls;
    a.ts = ts.data();
    a.ion = ion.data();
    a.nu = nu.data();
    a.passes = g_adj_passes;
    const int Kc = adj_k(ntheta, a.passes);
    std::vector<double> tab(adj_table_size(nseg, NP, Kc));
    for (int d = 0; d < n; ++d) {
        const bool refused = adj_refused(a, d);
        const bool frozen = adj_frozen(a, d);
        if (refused) status[d] = 1;
        adj_stage_values(a, d, tab.data(), 0, 1);
        for (int v0 = 0; v0 < nV; v0 += 64) {
            const int v1 = v0 + 64 < nV ? v0 + 64 : nV;
            std::vector<char> act(64, 0);
            for (int v = v0; v < v1; ++v) {
                act[v - v0] = !refused && ok[(size_t)d * nV + v];
                double *out = grad + ((size_t)d * nV + v) * ntheta;
                if (!act[v - v0])
                    for (int k = 0; k < ntheta; ++k) out[k] = refused ? NAN : 0.0;
                else
                    adj_lambda(a, d, v, (size_t)(v - v0), tab.data());
            }
            if (Kc > 0) {  // every direction in one reverse pass
                std::vector<double> g((size_t)64 * ADJ_MAX_K, 0.0);
                for (int pass = 0; pass < (frozen ? 2 : 1); ++pass) {
                    for (int v = v0; v < v1; ++v) {
                        if (!act[v - v0]) continue;
                        if (pass) adj_lambda(a, d, v, (size_t)(v - v0), tab.data(), true);
                        adj_contract(a, d, v, (size_t)(v - v0), tab.data(), &g[(size_t)(v - v0) * ADJ_MAX_K], pass == 1);
                    }
                    for (int c0 = 0; c0 <= NP; c0 += ADJ_CHUNK) {
                        adj_stage_points(a, d, c0, tab.data(), 0, 1);
                        for (int v = v0; v < v1; ++v)
                            if (act[v - v0])
                                adj_contract_points(a, c0, (size_t)(v - v0), tab.data(), &g[(size_t)(v - v0) * ADJ_MAX_K]);
                    }
                }
                for (int v = v0; v < v1; ++v)
                    if (act[v - v0])
                        for (int k = 0; k < ntheta; ++k)
                            grad[((size_t)d * nV + v) * ntheta + k] = g[(size_t)(v - v0) * ADJ_MAX_K + k];
                continue;
            }
            for (int pass = 0; pass < (frozen ? 2 : 1); ++pass) {
                if (pass)
                    for (int v = v0; v < v1; ++v)
                        if (act[v - v0]) adj_lambda(a, d, v, (size_t)(v - v0), tab.data(), true);
                for (int k = 0; k < ntheta; ++k) {
                    adj_stage_direction(a, d, k, tab.data(), 0, 1);
                    for (int v = v0; v < v1; ++v)
                        if (act[v - v0]) {
                            double &o = grad[((size_t)d * nV + v) * ntheta + k];
                            o = (pass ? o : 0.0) + adj_direction(a, d, v, (size_t)(v - v0), tab.data(), pass == 1);
                        }
                }
            }
        }
    }
    return 0;
}

// The assembly operator's lanes in turn.
extern "C" int emu_assemble(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                            const double *seg, const double *sc, const double *y,
                            const double *Vint, const double *logs, int bern, double *F,
                            double *A, double *B, double *C, double *U, double *J) {
    AsmArgs a = {};
    a.n = n; a.NP = NP; a.nseg = nseg; a.bern = bern;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.y = y; a.Vint = Vint; a.logs = logs;
    a.F = F; a.A = A; a.B = B; a.C = C; a.U = U; a.J = J;
    for (int d = 0; d < n; ++d) assemble_lane<false>(a, d);
    return 0;
}

// The impedance operator: the per-device table, then every (device, frequency) lane.
extern "C" int emu_impedance(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                             const double *seg, const double *sc, const double *y,
                             const double *Vint, const uint8_t *iface, const double *laps, int nf,
                             double dt, int bern, double *tab, double *dsc, double *Y,
                             int32_t *status, int ions, const double *imu) {
    ImpArgs a = {};
    a.n = n; a.NP = NP; a.nseg = nseg; a.bern = bern; a.nf = nf;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.y = y; a.Vint = Vint; a.iface = iface;
    a.laps = laps; a.dt = dt; a.tab = tab; a.dsc = dsc; a.Y = Y; a.status = status;
    a.imu = imu; a.row = ions ? (int)IMI_ROW : (int)IM_ROW;
    // the table's wavefront per device: its phases in turn, the scans across its lanes serial
    for (int d = 0; d < n; ++d) {
        const ImpDev k = imp_device(a, d);
        ImpIon is;
        if (ions) {
            is = imp_ion_setup(a, d, k);
            if (!is.ok) {
                status[d] = 2;
                continue;
            }
        } else if (imp_mobile_ions(a, d)) {
            status[d] = 1;
            continue;
        }
        status[d] = 0;
        for (int r = 0; r < NP; ++r) {
            imp_assemble_row(a, d, k, r);
            if (ions) imp_ion_row(a, d, k, is, r);
        }
        const int seg = (NP + 63) / 64;
        double tot[64][4], run[4] = {0.0, k.cn0, k.cp0, k.h0}, Dn = 0.0, Dp = 0.0;
        for (int lane = 0; lane < 64; ++lane) {
            const int r0 = lane * seg < NP ? lane * seg : NP;
            const int r1 = r0 + seg < NP ? r0 + seg : NP;
            imp_segment_totals(a, d, r0, r1, tot[lane]);
        }
        for (int lane = 0; lane < 64; ++lane) {
            const int r0 = lane * seg < NP ? lane * seg : NP;
            const int r1 = r0 + seg < NP ? r0 + seg : NP;
            double off[4] = {run[0], run[1], run[2], run[3]}, dn, dp;
            imp_segment_prefix(a, d, r0, r1, off, dn, dp);
            Dn += dn;
            Dp += dp;
            for (int q = 0; q < 4; ++q) run[q] += tot[lane][q];
        }
        const ImpTot t = imp_totals(a, d, k, run[1], run[2], Dn, Dp, run[3], true);
        for (int r = 0; r < NP; ++r) imp_coefficients(a, d, k, t, r);
        if (ions) imp_ion_current(a, d, k, is, t.H);
    }
    for (int d = 0; d < n; ++d)
        for (int f = 0; f < nf; ++f) {
            if (ions) impedance_ion_lane(a, d, f);
            else impedance_lane(a, d, f);
        }
    return 0;
}

// The 5x5 assembly operator's lanes in turn.
extern "C" int emu_assemble_ions(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                                 const double *seg, const double *sc, const double *y,
                                 const double *Vint, const double *logs, int bern, double *F,
                                 double *A, double *B, double *C, double *U) {
    AsmArgs a = {};
    a.n = n; a.NP = NP; a.nseg = nseg; a.bern = bern;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.y5 = y; a.Vint = Vint; a.logs = logs;
    a.F5 = F; a.A5 = A; a.B5 = B; a.C5 = C; a.U5 = U;
    for (int d = 0; d < n; ++d) assemble_lane<true>(a, d);
    return 0;
}

// How many (workgroup, k) pairs of a launch of nblocks workgroups the XCD map gives each of n
// devices (hits[n]); returns the number of pairs that fell outside [0, n).
extern "C" int emu_xcd_hits(int n, int nblocks, int32_t *hits) {
    int outside = 0;
    for (int d = 0; d < n; ++d) hits[d] = 0;
    for (int b = 0; b < nblocks; ++b) {
        const XcdMap m = xcd_map(b, nblocks, n);
        for (int k = 0;; ++k) {
            const int d = xcd_device(m, k, n);
            if (d < 0) break;
            if (d >= n) { ++outside; break; }
            ++hits[d];
            if (k > n) { ++outside; break; }  // a map that never ends
        }
    }
    return outside;
}

// The transient (ddhip_transient.h): each lane's whole sequence of steps in turn (the kernel's
// lanes share nothing).
extern "C" int emu_transient(int n, int NP, int nseg, int nt, const double *pt, const uint8_t *sid,
                             const uint8_t *iface, const double *seg, const double *sc,
                             const double *imu, const double *t, const double *V, const double *G,
                             const double *y0, double tol, int max_iter, int bern, int zimt,
                             int failure_mode, int ions, double *J, double *Jext, uint8_t *ok,
                             int32_t *its, uint8_t *kept, int32_t *status) {
    TrArgs a = {};
    a.n = n; a.NP = NP; a.nseg = nseg; a.nt = nt;
    a.pt = pt; a.sid = sid; a.iface = iface; a.seg = seg; a.sc = sc; a.imu = imu;
    a.t = t; a.V = V; a.G = G; a.y0 = y0;
    std::vector<double> Y((size_t)3 * NP * TR_NC * n), fb((size_t)NP * n), fac((size_t)NP * TR_FAC * n);
    a.Y = Y.data(); a.fb = fb.data(); a.fac = fac.data();
    a.J = J; a.Jext = Jext; a.ok = ok; a.kept = kept; a.its = its; a.status = status;
    a.tol = tol; a.max_iter = max_iter; a.bern = bern; a.zimt = zimt; a.failure_mode = failure_mode;
    for (int lane = 0; lane < n; ++lane) {
        TrLane S;
        if (ions) {
            tr_init<5>(a, lane, S);
            while (!S.done) tr_newton<5>(a, lane, S);
        } else {
            tr_init<3>(a, lane, S);
            while (!S.done) tr_newton<3>(a, lane, S);
        }
    }
    return 0;
}

// The builder's lanes in turn (one per device and direction; k = -1 without directions).
extern "C" int emu_build(int n, int NP, int nl, int nseg, int K, const uint8_t *lid,
                         const uint8_t *sid, const uint8_t *flags, const int32_t *i0,
                         const double *unit, const double *unit_sum, const double *layers,
                         const double *scalars, const double *gen, int gen_m,
                         const double *dlayers, const double *dscalars,
                         double *pt, double *seg, double *sc, double *imu, double *dpt,
                         double *dseg, double *dsc) {
    BuildArgs a = {};
    a.n = n; a.NP = NP; a.nl = nl; a.nseg = nseg; a.K = K;
    a.lid = lid; a.sid = sid; a.flags = flags; a.i0 = i0; a.unit = unit; a.unit_sum = unit_sum;
    a.layers = layers; a.scalars = scalars; a.dlayers = dlayers; a.dscalars = dscalars;
    a.gen = gen; a.gen_m = gen_m;
    a.pt = pt; a.seg = seg; a.sc = sc; a.imu = imu; a.dpt = dpt; a.dseg = dseg; a.dsc = dsc;
    for (int k = (K > 0 ? 0 : -1); k < (K > 0 ? K : 0); ++k)
        for (int d = 0; d < n; ++d) build_lane(a, d, k);
    return 0;
}

// The layout's lanes in turn.
extern "C" int emu_graded_layout(int n, int NP, int nl, const double *L, double beta, double d0,
                                 uint8_t *lid, uint8_t *sid, int32_t *i0, double *unit,
                                 double *unit_sum, uint8_t *iface, int32_t *nseg) {
    LayoutArgs a = {};
    a.n = n; a.NP = NP; a.nl = nl; a.beta = beta; a.d0 = d0; a.L = L;
    a.lid = lid; a.sid = sid; a.i0 = i0; a.unit = unit; a.unit_sum = unit_sum; a.iface = iface;
    a.nseg = nseg;
    for (int d = 0; d < n; ++d) layout_structure(a, d);
    for (int j = 0; j < NP + 2; ++j)
        for (int d = 0; d < n; ++d) layout_point(a, d, j);
    for (int d = 0; d < n; ++d) layout_sums(a, d);
    return 0;
}
