// Test-only CPU build of the HIP kernels' per-lane logic (jaxdd_amd/csrc/ddhip_core.h).
//
// It replays each 64-lane wavefront of the persistent solve kernel lane by lane, with the
// same "any lane active" / "any lane with mobile ions" wave votes, so that the kernels'
// algorithm can be debugged and compared with the oracle on a host without a GPU. Nothing in
// the product loads it.
#define DDH_CPU_EMULATION 1
#include "ddhip_core.h"

#include <cstdlib>
#include <vector>

using namespace ddh;

static void run(Args &a) {
    for (int w0 = 0; w0 < a.n; w0 += 64) {
        int w1 = w0 + 64 < a.n ? w0 + 64 : a.n;
        for (int lane = w0; lane < w1; ++lane) {
            poisson_lane(a, lane);
            start_lane(a, lane, a.st[lane]);
        }
        for (;;) {
            bool any = false, ions = false, f64 = !a.mixed;
            for (int lane = w0; lane < w1; ++lane) {
                bool act = lane_active(a, lane);
                any |= act;
                ions |= act && lane_mobile(a, lane);
                f64 |= act && lane_needs64(a, lane);
            }
            if (!any) break;
            for (int lane = w0; lane < w1; ++lane) {
                if (!lane_active(a, lane)) continue;
                if (f64) newton_lane_step<double>(a, lane, ions);
                else newton_lane_step<float>(a, lane, ions);
            }
        }
    }
}

extern "C" int emu_solve_batch(int n, int NP, int nseg, const double *pt, const uint8_t *sid,
                               const double *seg, const double *sc, const double *V, int nV,
                               double tol, int max_iter, int start_iter, int strict, int bern,
                               int precision, int jv, double *J, uint8_t *ok, int32_t *its,
                               int32_t *status) {
    size_t tsize = 8;
    std::vector<double> Y((size_t)NBUF * NP * 3 * n);
    size_t facb = (size_t)NP * NFAC * n * tsize, poisb = (size_t)NP * 4 * n * 8;
    std::vector<char> fac(facb > poisb ? facb : poisb);
    std::vector<double> fix((size_t)NP * n);
    std::vector<LaneState> st(n);
    Args a;
    a.n = n; a.NP = NP; a.nV = nV; a.nseg = nseg;
    a.pt = pt; a.sid = sid; a.seg = seg; a.sc = sc; a.V = V;
    a.Y = Y.data(); a.fac = fac.data(); a.fix = fix.data(); a.st = st.data();
    a.J = J; a.ok = ok; a.its = its; a.status = status;
    a.tol = tol; a.max_iter = max_iter; a.start_iter = start_iter; a.strict = strict;
    a.bern = bern; a.jv = jv; a.mixed = precision ? 1 : 0;
    run(a);
    return 0;
}
