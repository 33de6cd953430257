// HIP kernels of the batched steady-state J-V solve (gfx950) and the C-ABI of libddhip.so.
// The per-lane logic is in ddhip_core.h; here the launches: the persistent solve kernel (one
// lane per device runs its Poisson start, start solves, pre-bias and voltage scan) and the
// block-tridiagonal sweep operator.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "ddhip_core.h"
#include "ddhip_adjoint.h"
#include "ddhip_assemble.h"
#include "ddhip_build.h"
#include "ddhip_impedance.h"
#include "ddhip_pair.h"
#include "ddhip_quad.h"
#include "ddhip_transient.h"

namespace ddh {

// The whole solve of every device in one launch: each lane runs Poisson's start, then its
// Newton state machine; each wavefront loops until its last lane has finished its scan.
__global__ void __launch_bounds__(64) solve_kernel(Args a) {
    extern __shared__ double lds[];  // a.nseg * SEG_N * WAVE doubles
    const int lane = blockIdx.x * 64 + threadIdx.x;
    const bool valid = lane < a.n;
    ldouble *lseg = (ldouble *)lds + threadIdx.x;
    LaneState &S = a.st[valid ? lane : 0];
    if (valid) {
        fill_segments(a, lane, lseg);
        poisson_lane(a, lane, S);
        S.part = 0; S.parts = 1;
        start_lane(a, lane, S);
    }
    while (true) {
        const bool act = valid && lane_active(S);
        if (!__any(act)) break;
        const bool ions = __any(act && lane_mobile(S));
        const bool f64 = !a.mixed || __any(act && lane_needs64(S));
        if (act) {
            if (f64) newton_lane_step<double>(a, lane, ions, lseg, S);
            else newton_lane_step<float>(a, lane, ions, lseg, S);
        }
    }
}

// A register copy of the record of lane (this lane ^ M): with M = 1 the partner lane of a pair
// (lanes 2d and 2d + 1 hold device d).
template <int M = 1, typename R>
__device__ inline R partner(const R &r) {
    static_assert(sizeof(R) % 4 == 0, "records of 32-bit words");
    R o;
    const int *src = reinterpret_cast<const int *>(&r);
    int *dst = reinterpret_cast<int *>(&o);
#pragma unroll
    for (int q = 0; q < (int)(sizeof(R) / 4); ++q) dst[q] = __shfl_xor(src[q], M);
    return o;
}

// A register copy of lane `from`'s record.
template <typename R>
__device__ inline R lane_record(const R &r, int from) {
    static_assert(sizeof(R) % 4 == 0, "records of 32-bit words");
    R o;
    const int *src = reinterpret_cast<const int *>(&r);
    int *dst = reinterpret_cast<int *>(&o);
#pragma unroll
    for (int q = 0; q < (int)(sizeof(R) / 4); ++q) dst[q] = __shfl(src[q], from);
    return o;
}

// Two lanes per device (ddhip_pair.h): 32 devices per wavefront. The left lane of each pair
// runs the Poisson start; both lanes then run the same state machine on their own state.
__global__ void __launch_bounds__(64) solve_pair_kernel(Args a) {
    extern __shared__ double lds[];  // a.nseg * SEG_N * PAIR_LS doubles (32 device slots)
    const int t = threadIdx.x, half = t & 1;
    const int lane = blockIdx.x * 32 + (t >> 1);  // the device
    const bool valid = lane < a.n;
    ldouble *lseg = (ldouble *)lds + (t >> 1);
    LaneState &S = a.st[valid ? 2 * (size_t)lane + half : 0];
    if (valid && half == 0) {
        fill_segments(a, lane, lseg, PAIR_LS);
        poisson_lane(a, lane, S);
    }
    __syncthreads();  // the segment table and the Poisson start, for the right lanes
    if (valid && half == 1) S = a.st[2 * (size_t)lane];
    if (valid) {
        S.part = half; S.parts = 2;
        start_lane(a, lane, S);
    }
    while (true) {
        const bool act = valid && lane_active(S);
        if (!__any(act)) break;
        const bool ions = __any(act && lane_mobile(S));
        const bool f64 = !a.mixed || __any(act && lane_needs64(S));
        if (act) {
            PairStep st;
            if (f64) pair_phase1<double>(a, lane, half, ions, lseg, S, st);
            else pair_phase1<float>(a, lane, half, ions, lseg, S, st);
            const PairElim oe = partner(st.e);
            if (f64) pair_phase2<double>(a, lane, half, ions, st, oe);
            else pair_phase2<float>(a, lane, half, ions, st, oe);
            const PairBack ob = partner(st.b);
            pair_phase3(a, lane, half, S, st, ob);
        }
    }
}

// Four lanes per device (ddhip_quad.h): 16 devices per wavefront, devices without mobile ions
// (a device with them is refused: status 2); lanes 4d..4d+3 hold device d. The local data share
// holds the 16 segment tables and lanes 1 and 2's prefixes (quad_lds_doubles), 25 KB for seven
// segments, so that four workgroups fit one CU. Lane 0 of each device runs the Poisson start;
// the four lanes then run the same state machine. A step's records move by shuffles: the
// meeting records within lanes (0, 1) and (2, 3), the middle's two maps from lanes 1 and 2 to
// all four, the chunks' current sums within the pairs and then the halves' across them.
struct QuadMid {
    double W[12];
};

__global__ void __launch_bounds__(64) solve_quad_kernel(Args a) {
    extern __shared__ double lds[];  // quad_lds_doubles(a.nseg) doubles
    const int t = threadIdx.x, q = t & 3, base = t & ~3, slot = t >> 2;
    const int lane = blockIdx.x * 16 + slot;  // the device
    const bool inb = lane < a.n;
    const bool refused = inb && quad_refused(a, lane);
    const bool valid = inb && !refused;
    ldouble *lseg = (ldouble *)lds + slot;
    ldouble *pre = (ldouble *)lds + quad_pre_column(a.nseg, slot, q);
    LaneState &S = a.st[valid ? 4 * (size_t)lane + q : 0];
    if (refused && q == 0) quad_refuse(a, lane);
    if (valid && q == 0) {
        fill_segments(a, lane, lseg, QUAD_LS);
        poisson_lane(a, lane, S);
    }
    __syncthreads();  // the segment table and the Poisson start, for lanes 1..3
    if (valid && q != 0) S = a.st[4 * (size_t)lane];
    if (valid) {
        S.part = q; S.parts = 4;
        start_lane(a, lane, S);
    }
    while (true) {
        const bool act = valid && lane_active(S);
        if (!__any(act)) break;
        const bool f64 = !a.mixed || __any(act && lane_needs64(S));
        if (act) {
            QuadStep st;
            if (f64) quad_phase1<double>(a, lane, q, lseg, pre, S, st);
            else quad_phase1<float>(a, lane, q, lseg, pre, S, st);
            const QuadEdge oe = partner<1>(st.e.m);
            double Xa[12], Xb[12];
            quad_meet((q & 1) ? oe : st.e.m, (q & 1) ? st.e.m : oe, Xa, Xb);
            QuadMid w;
            quad_middle_part(st.e, q, Xa, Xb, w.W);
            const QuadMid P = lane_record(w, base + 1), Q = lane_record(w, base + 2);
            double xs[3], sv[3];
            quad_center(P.W, Q.W, q, Xa, Xb, xs, sv);
            if (f64) quad_phase2<double>(a, lane, q, st, xs, sv);
            else quad_phase2<float>(a, lane, q, st, xs, sv);
            const QuadBack ob = partner<1>(st.b);
            PairBack h;
            quad_half((q & 1) ? ob : st.b, (q & 1) ? st.b : ob, q >> 1, h);
            const PairBack oh = partner<2>(h);
            quad_phase3(a, lane, S, (q >> 1) ? oh : h, (q >> 1) ? h : oh);
        }
    }
}

// The assembly operators: one lane per device; FIVE the 5x5 form with the ions as unknowns.
template <bool FIVE>
__global__ void __launch_bounds__(64) assemble_kernel(AsmArgs a) {
    const int d = blockIdx.x * 64 + threadIdx.x;
    if (d < a.n) assemble_lane<FIVE>(a, d);
}

// The impedance operator: the per-device tables (one wavefront per device), then one lane per
// (device, frequency), the frequencies of a device on neighbouring lanes so that their reads of
// its table coincide.
__device__ inline double wave_sum(double x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m);
    return x;
}

// exclusive prefix sum over the wavefront's lanes
__device__ inline double wave_exclusive_scan(double x, int lane) {
    double s = x;
#pragma unroll
    for (int k = 1; k < 64; k <<= 1) {
        const double y = __shfl_up(s, k);
        if (lane >= k) s += y;
    }
    return s - x;
}

// One device per workgroup, each XCD a contiguous run of devices (xcd_map, ddhip_layout.h), so
// that a line of the lane-innermost arrays it fetches serves the next devices it runs.
// IONS: the 5 x 5 path (the ions' logarithms as unknowns; their rows, the slots' constraint rows
// and the ions' current coefficients after the carriers'), else the 3 x 3 path, which marks a
// device with mobile ions (status 1). A device outside the 5 x 5 path gets status 2.
template <bool IONS>
__global__ void __launch_bounds__(64) impedance_prep_kernel(ImpArgs a) {
    const int d = xcd_device(xcd_map((int)blockIdx.x, (int)gridDim.x, a.n), 0, a.n);
    if (d < 0) return;
    const int lane = threadIdx.x, NP = a.NP;
    const ImpDev k = imp_device(a, d);
    ImpIon ions;
    if (IONS) {
        ions = imp_ion_setup(a, d, k);
        if (!ions.ok) {
            if (lane == 0) a.status[d] = 2;
            return;
        }
    } else if (imp_mobile_ions(a, d)) {
        if (lane == 0) a.status[d] = 1;
        return;
    }
    if (lane == 0) a.status[d] = 0;
    for (int r = lane; r < NP; r += 64) {
        imp_assemble_row(a, d, k, r);
        if (IONS) imp_ion_row(a, d, k, ions, r);
    }
    __syncthreads();
    const int seg = (NP + 63) / 64;
    const int r0 = lane * seg < NP ? lane * seg : NP;
    const int r1 = r0 + seg < NP ? r0 + seg : NP;
    double tot[4], off[4];
    imp_segment_totals(a, d, r0, r1, tot);
    const double base[4] = {0.0, k.cn0, k.cp0, k.h0};
    for (int q = 0; q < 4; ++q) off[q] = base[q] + wave_exclusive_scan(tot[q], lane);
    double Dn, Dp;
    imp_segment_prefix(a, d, r0, r1, off, Dn, Dp);
    Dn = wave_sum(Dn);
    Dp = wave_sum(Dp);
    const double Sn = k.cn0 + wave_sum(tot[1]), Sp = k.cp0 + wave_sum(tot[2]);
    const double H = k.h0 + wave_sum(tot[3]);
    const ImpTot t = imp_totals(a, d, k, Sn, Sp, Dn, Dp, H, lane == 0);
    __syncthreads();
    for (int r = lane; r < NP; r += 64) imp_coefficients(a, d, k, t, r);
    if (IONS) {
        __syncthreads();
        if (lane == 0) imp_ion_current(a, d, k, ions, t.H);
    }
}

template <bool IONS>
__global__ void __launch_bounds__(256) impedance_kernel(ImpArgs a) {
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long)a.n * a.nf) return;
    if (IONS) impedance_ion_lane(a, (int)(gid / a.nf), (int)(gid % a.nf));
    else impedance_lane(a, (int)(gid / a.nf), (int)(gid % a.nf));
}

// The batch's arrays and their tangents from each device's structure and numbers
// (ddhip_build.h): one lane per (device, direction), devices innermost so that a direction's
// slice of the lane-innermost arrays is written by neighbouring lanes.
__global__ void __launch_bounds__(64) build_kernel(BuildArgs a) {
    const long gid = (long)blockIdx.x * 64 + threadIdx.x;
    const long lanes = (long)a.n * (a.K > 0 ? a.K : 1);
    if (gid >= lanes) return;
    build_lane(a, (int)(gid % a.n), a.K > 0 ? (int)(gid / a.n) : -1);
}

// The graded grid's layout of every device (ddhip_build.h): the structure and the layers' sums
// one lane per device, the spacings one lane per (device, point), devices innermost.
__global__ void __launch_bounds__(64) layout_structure_kernel(LayoutArgs a) {
    const int d = blockIdx.x * 64 + threadIdx.x;
    if (d < a.n) layout_structure(a, d);
}

__global__ void __launch_bounds__(256) layout_point_kernel(LayoutArgs a) {
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid < (long)a.n * (a.NP + 2)) layout_point(a, (int)(gid % a.n), (int)(gid / a.n));
}

__global__ void __launch_bounds__(64) layout_sums_kernel(LayoutArgs a) {
    const int d = blockIdx.x * 64 + threadIdx.x;
    if (d < a.n) layout_sums(a, d);
}

// The transient (ddhip_transient.h): one lane per device through all of its time steps, one
// Newton step per turn of the loop; a lane whose step's solve ends begins its next step at
// once. NC = 5 with mobile ions (their densities as unknowns), 3 without.
template <int NC>
__global__ void __launch_bounds__(64) transient_kernel(TrArgs a) {
    const int lane = blockIdx.x * 64 + threadIdx.x;
    const bool valid = lane < a.n;
    TrLane S;
    S.done = 1;
    if (valid) tr_init<NC>(a, lane, S);
    while (true) {
        const bool act = valid && !S.done;
        if (!__any(act)) break;
        if (act) tr_newton<NC>(a, lane, S);
    }
}

// The gradients of every device's J-V curve: one wavefront per device (the workgroups loop
// over the devices), one lane per voltage; the device's tables and tangents are staged in LDS
// by the whole wavefront, read there by every lane. REV: every direction in one reverse pass
// over the parameters (up to ADJ_MAX_K directions), else one dual-number pass per direction.
template <bool REV>
__global__ void __launch_bounds__(64) adjoint_kernel(AdjArgs a) {
    extern __shared__ double lds[];  // adj_table_size(a.nseg, a.NP, K or 0) doubles
    ldouble *tab = (ldouble *)lds;
    const int lane = threadIdx.x;
    const size_t sl = (size_t)blockIdx.x * 64 + lane;
    const int nV = a.nV, K = a.ntheta;
    // the XCDs in turn (as the impedance tables): each runs a contiguous run of devices, so
    // that the lines of the lane-innermost arrays one wavefront fetches serve its neighbours
    const XcdMap map = xcd_map((int)blockIdx.x, (int)gridDim.x, a.n);
    for (int k = 0;; ++k) {
        const int d = xcd_device(map, k, a.n);
        if (d < 0) break;
        const bool refused = adj_refused(a, d);
        const bool frozen = adj_frozen(a, d);  // a second pass at the pre-bias solution
        if (lane == 0) a.status[d] = refused ? 1 : 0;
        __syncthreads();  // every lane is done with the previous device's table
        adj_stage_values(a, d, tab, lane, 64);
        __syncthreads();
        for (int v0 = 0; v0 < nV; v0 += 64) {
            const int v = v0 + lane;
            const bool act = v < nV && !refused && a.ok[(size_t)d * nV + v];
            double *out = a.grad + ((size_t)d * nV + (v < nV ? v : 0)) * K;
            if (v < nV && !act)
                for (int k = 0; k < K; ++k) out[k] = refused ? NAN : 0.0;
            if (act) adj_lambda(a, d, v, sl, tab);
            if (REV) {  // every direction in one reverse pass
                double g[ADJ_MAX_K];
#pragma unroll
                for (int k = 0; k < ADJ_MAX_K; ++k) g[k] = 0.0;
                for (int pass = 0; pass < (frozen ? 2 : 1); ++pass) {
                    if (act && pass) adj_lambda(a, d, v, sl, tab, true);
                    if (act) adj_contract(a, d, v, sl, tab, g, pass == 1);
                    for (int c0 = 0; c0 <= a.NP; c0 += ADJ_CHUNK) {
                        __syncthreads();  // every lane is done with the previous chunk
                        adj_stage_points(a, d, c0, tab, lane, 64);
                        __syncthreads();
                        if (act) adj_contract_points(a, c0, sl, tab, g);
                    }
                }
                if (act) {
#pragma unroll
                    for (int k = 0; k < ADJ_MAX_K; ++k)
                        if (k < K) out[k] = g[k];
                }
                continue;
            }
            for (int pass = 0; pass < (frozen ? 2 : 1); ++pass) {
                if (act && pass) adj_lambda(a, d, v, sl, tab, true);
                for (int k = 0; k < K; ++k) {
                    __syncthreads();  // every lane is done with the previous direction's tangents
                    adj_stage_direction(a, d, k, tab, lane, 64);
                    __syncthreads();
                    if (act) out[k] = (pass ? out[k] : 0.0) + adj_direction(a, d, v, sl, tab, pass == 1);
                }
            }
        }
    }
}

}  // namespace ddh

// ---------------------------------------------------------------------------------------
// The block-tridiagonal sweep operator (src/jaxdd/sweep.py:313-330): one lane per system.
// ---------------------------------------------------------------------------------------
namespace ddh {

template <typename T, int b>
__device__ void inverse_b(const T *m, T *out) {
    if (b == 3) { inverse3(m, out); return; }
    // Gauss-Jordan with partial pivoting (5x5 blocks)
    T t[b * b], inv[b * b];
#pragma unroll
    for (int q = 0; q < b * b; ++q) { t[q] = m[q]; inv[q] = (q / b == q % b) ? T(1) : T(0); }
    for (int col = 0; col < b; ++col) {
        int piv = col;
        for (int r = col + 1; r < b; ++r)
            if (fabs((double)t[r * b + col]) > fabs((double)t[piv * b + col])) piv = r;
        if (piv != col)
            for (int q = 0; q < b; ++q) {
                T s = t[col * b + q]; t[col * b + q] = t[piv * b + q]; t[piv * b + q] = s;
                s = inv[col * b + q]; inv[col * b + q] = inv[piv * b + q]; inv[piv * b + q] = s;
            }
        T rp = T(1) / t[col * b + col];
        for (int q = 0; q < b; ++q) { t[col * b + q] *= rp; inv[col * b + q] *= rp; }
        for (int r = 0; r < b; ++r) {
            if (r == col) continue;
            T f = t[r * b + col];
            for (int q = 0; q < b; ++q) { t[r * b + q] -= f * t[col * b + q]; inv[r * b + q] -= f * inv[col * b + q]; }
        }
    }
#pragma unroll
    for (int q = 0; q < b * b; ++q) out[q] = inv[q];
}

template <typename T, int b>
__global__ void __launch_bounds__(64) block_sweep_kernel(int n, int NP, int k, const T *A,
                                                         const T *B, const T *C, T *d, T *ws) {
    const int lane = blockIdx.x * 64 + threadIdx.x;
    if (lane >= n) return;
    const int bb = b * b, bk = b * k;
    // ws: c' [NP][bb][n]; d is overwritten by d' then x
    T cprev[b * b], dprev[b * 8];
#pragma unroll
    for (int q = 0; q < bb; ++q) cprev[q] = T(0);
    for (int q = 0; q < bk; ++q) dprev[q] = T(0);
    for (int i = 0; i < NP; ++i) {
        T a_[b * b], m[b * b], mi[b * b], rh[b * 8];
#pragma unroll
        for (int q = 0; q < bb; ++q) {
            a_[q] = A[((size_t)i * bb + q) * n + lane];
            m[q] = B[((size_t)i * bb + q) * n + lane];
        }
        for (int q = 0; q < bk; ++q) rh[q] = d[((size_t)i * bk + q) * n + lane];
        if (i > 0) {
#pragma unroll
            for (int p = 0; p < b; ++p) {
#pragma unroll
                for (int q = 0; q < b; ++q) {
                    T s = T(0);
#pragma unroll
                    for (int l = 0; l < b; ++l) s += a_[p * b + l] * cprev[l * b + q];
                    m[p * b + q] -= s;
                }
                for (int q = 0; q < k; ++q) {
                    T s = T(0);
#pragma unroll
                    for (int l = 0; l < b; ++l) s += a_[p * b + l] * dprev[l * k + q];
                    rh[p * k + q] -= s;
                }
            }
        }
        inverse_b<T, b>(m, mi);
        T c_[b * b];
#pragma unroll
        for (int q = 0; q < bb; ++q) c_[q] = C[((size_t)i * bb + q) * n + lane];
#pragma unroll
        for (int p = 0; p < b; ++p) {
#pragma unroll
            for (int q = 0; q < b; ++q) {
                T s = T(0);
#pragma unroll
                for (int l = 0; l < b; ++l) s += mi[p * b + l] * c_[l * b + q];
                cprev[p * b + q] = s;
                ws[((size_t)i * bb + p * b + q) * n + lane] = s;
            }
            for (int q = 0; q < k; ++q) {
                T s = T(0);
#pragma unroll
                for (int l = 0; l < b; ++l) s += mi[p * b + l] * rh[l * k + q];
                dprev[p * k + q] = s;
            }
        }
        for (int q = 0; q < bk; ++q) d[((size_t)i * bk + q) * n + lane] = dprev[q];
    }
    T xn[b * 8];
    for (int q = 0; q < bk; ++q) xn[q] = T(0);
    for (int i = NP - 1; i >= 0; --i) {
        T x[b * 8];
        for (int q = 0; q < bk; ++q) x[q] = d[((size_t)i * bk + q) * n + lane];
#pragma unroll
        for (int p = 0; p < b; ++p)
            for (int q = 0; q < k; ++q) {
                T s = T(0);
#pragma unroll
                for (int l = 0; l < b; ++l) s += ws[((size_t)i * bb + p * b + l) * n + lane] * xn[l * k + q];
                x[p * k + q] -= s;
            }
        for (int q = 0; q < bk; ++q) {
            d[((size_t)i * bk + q) * n + lane] = x[q];
            xn[q] = x[q];
        }
    }
}

}  // namespace ddh

// ---------------------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------------------
#include "ddhip.h"

static thread_local char g_err[512];

static int fail(int code, const char *msg) {
    strncpy(g_err, msg, sizeof(g_err) - 1);
    g_err[sizeof(g_err) - 1] = 0;
    return code;
}

static int hip_fail(hipError_t e, const char *where) {
    snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return DDH_E_HIP;
}

extern "C" int ddh_abi_version(void) { return DDH_ABI_VERSION; }
extern "C" const char *ddh_last_error(void) { return g_err; }

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct WsLayout {
    size_t y, fac64, fac32, pfac, fix, st, total;
};

// One region per array: the state buffers, the FP64 and FP32 Newton factors, Poisson's
// factors, the fixed ions in use and the lanes' state machines.
// Sized for every kernel: the four-lane solve's factors and lanes are the largest.
static WsLayout ws_layout(int32_t n, int32_t NP, int32_t precision, int32_t lanes = 4) {
    WsLayout w;
    const size_t nf = ddh::NFAC_Q > ddh::NFAC ? ddh::NFAC_Q : ddh::NFAC;
    size_t off = 0;
    w.y = off; off = align256(off + (size_t)ddh::NBUF * NP * 3 * n * sizeof(double));
    w.fac64 = off; off = align256(off + (size_t)NP * nf * n * sizeof(double));
    w.fac32 = off;
    if (precision == DDH_MIXED) off = align256(off + (size_t)NP * nf * n * sizeof(float));
    w.pfac = off; off = align256(off + (size_t)NP * 4 * n * sizeof(double));
    w.fix = off; off = align256(off + (size_t)NP * n * sizeof(double));
    w.st = off; off = align256(off + (size_t)lanes * n * sizeof(ddh::LaneState));
    w.total = off;
    return w;
}

extern "C" size_t ddh_workspace_bytes(int32_t n, int32_t NP, int32_t nV, int32_t precision) {
    (void)nV;
    if (n <= 0 || NP < 2) return 0;
    return ws_layout(n, NP, precision, 4).total;  // enough for every kernel
}

static int solve_batch(const ddh_batch *b, const double *V, int32_t nV, const ddh_opts *o,
                       double *J, uint8_t *ok, int32_t *its, int32_t *status, double *states,
                       double *pre,
                       void *workspace, size_t workspace_bytes, void *stream) {
    if (!b || !V || !o || !J || !ok || !its || !status || !workspace)
        return fail(DDH_E_ARG, "ddh_solve_batch: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || nV < 1 || b->nseg < 1 || b->nseg > DDH_MAX_SEGMENTS)
        return fail(DDH_E_SHAPE, "ddh_solve_batch: n > 0, 2 <= NP <= 65536, nV >= 1, 1 <= nseg <= 14");
    if (o->precision != DDH_FP64 && o->precision != DDH_MIXED)
        return fail(DDH_E_ARG, "ddh_solve_batch: precision must be DDH_FP64 or DDH_MIXED");
    if (o->bernoulli != 0 && o->bernoulli != 1)
        return fail(DDH_E_ARG, "ddh_solve_batch: bernoulli must be 0 (exact) or 1 (simss)");
    if (o->max_iter < 1 || o->start_iter < 1)
        return fail(DDH_E_ARG, "ddh_solve_batch: max_iter and start_iter must be >= 1");
    if (o->lanes != 0 && o->lanes != 1 && o->lanes != 2 && o->lanes != 4)
        return fail(DDH_E_ARG, "ddh_solve_batch: lanes must be 1, 2 or 4 (0: the default, 2)");
    const int lanes = o->jv_mode ? 1 : (o->lanes == 0 ? 2 : o->lanes);
    if (lanes == 4 && b->NP < ddh::QUAD_MIN_NP)
        return fail(DDH_E_SHAPE, "ddh_solve_batch: four lanes per device need NP >= 16");
    WsLayout w = ws_layout(b->n, b->NP, o->precision, lanes);
    if (workspace_bytes < w.total) return fail(DDH_E_WORKSPACE, "ddh_solve_batch: workspace too small");
    char *base = (char *)workspace;
    ddh::Args a;
    a.n = b->n; a.NP = b->NP; a.nV = nV; a.nseg = b->nseg;
    a.pt = b->pt; a.sid = b->sid; a.seg = b->seg; a.sc = b->sc; a.V = V;
    a.Y = (double *)(base + w.y);
    a.fac64 = (double *)(base + w.fac64);
    a.fac32 = (float *)(base + w.fac32);
    a.pfac = (double *)(base + w.pfac);
    a.fix = (double *)(base + w.fix);
    a.st = (ddh::LaneState *)(base + w.st);
    a.states = states;
    a.pre = pre;
    a.J = J; a.ok = ok; a.its = its; a.status = status;
    a.tol = o->tol; a.max_iter = o->max_iter; a.start_iter = o->start_iter;
    a.strict = o->strict; a.bern = o->bernoulli; a.jv = o->jv_mode;
    a.mixed = o->precision == DDH_MIXED ? 1 : 0;
    hipStream_t s = (hipStream_t)stream;
    const size_t lds = (size_t)b->nseg * ddh::SEG_N * ddh::WAVE * sizeof(double);
    if (lanes == 4) {
        dim3 grid((b->n + 15) / 16), block(64);
        const size_t qlds = ddh::quad_lds_doubles(b->nseg) * sizeof(double);
        (void)hipFuncSetAttribute((const void *)ddh::solve_quad_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)qlds);
        hipLaunchKernelGGL(ddh::solve_quad_kernel, grid, block, qlds, s, a);
    } else if (lanes == 2) {
        dim3 grid((b->n + 31) / 32), block(64);
        const size_t plds = (size_t)b->nseg * ddh::SEG_N * ddh::PAIR_LS * sizeof(double);
        (void)hipFuncSetAttribute((const void *)ddh::solve_pair_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)plds);
        hipLaunchKernelGGL(ddh::solve_pair_kernel, grid, block, plds, s, a);
    } else {
        if (lds > 160 * 1024)
            return fail(DDH_E_ARG, "ddh_solve_batch: one lane per device takes at most 13 "
                                   "segments (its segment tables fill the 160 KiB of LDS)");
        dim3 grid((b->n + 63) / 64), block(64);
        (void)hipFuncSetAttribute((const void *)ddh::solve_kernel,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        hipLaunchKernelGGL(ddh::solve_kernel, grid, block, lds, s, a);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "solve_kernel launch");
    return DDH_OK;
}

extern "C" int ddh_solve_batch(const ddh_batch *b, const double *V, int32_t nV,
                               const ddh_opts *o, double *J, uint8_t *ok, int32_t *its,
                               int32_t *status, void *workspace, size_t workspace_bytes,
                               void *stream) {
    return solve_batch(b, V, nV, o, J, ok, its, status, nullptr, nullptr, workspace,
                       workspace_bytes, stream);
}

extern "C" int ddh_solve_batch_states(const ddh_batch *b, const double *V, int32_t nV,
                                      const ddh_opts *o, double *J, uint8_t *ok, int32_t *its,
                                      int32_t *status, double *states, double *pre_states,
                                      void *workspace, size_t workspace_bytes, void *stream) {
    if (!states) return fail(DDH_E_ARG, "ddh_solve_batch_states: NULL states");
    return solve_batch(b, V, nV, o, J, ok, its, status, states, pre_states, workspace,
                       workspace_bytes, stream);
}

// ddh_solve_batch without an allocation: the workspace is this library's own device array of
// DDH_NOALLOC_SCRATCH_BYTES, one copy per GPU, held until the process ends. In host code the
// array's name is only a stand-in for its symbol; its device address comes from
// hipGetSymbolAddress, taken once per device with that device current, and only that address
// is handed to the kernels. Every call on one GPU shares its copy (the solve initialises every
// part of the workspace it reads, so nothing needs clearing between calls); calls that overlap
// on one GPU race.
static __device__ __align__(256) unsigned char scratch[DDH_NOALLOC_SCRATCH_BYTES];

// the scratch's device address on each GPU, by device ordinal
constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

static int scratch_address(void **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "ddh_solve_batch_noalloc: hipGetDevice");
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES)
        return fail(DDH_E_ARG, "ddh_solve_batch_noalloc: device ordinal beyond the library's table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "ddh_solve_batch_noalloc: hipGetSymbolAddress");
        if (!p) return fail(DDH_E_HIP, "ddh_solve_batch_noalloc: no device address for the scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return DDH_OK;
}

extern "C" size_t ddh_noalloc_scratch_bytes(void) { return (size_t)DDH_NOALLOC_SCRATCH_BYTES; }

extern "C" int ddh_solve_batch_noalloc(const ddh_batch *b, const double *V, int32_t nV,
                                       const ddh_opts *o, double *J, uint8_t *ok, int32_t *its,
                                       int32_t *status, void *stream) {
    if (!b || !o) return fail(DDH_E_ARG, "ddh_solve_batch_noalloc: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536)
        return fail(DDH_E_SHAPE, "ddh_solve_batch_noalloc: n > 0, 2 <= NP <= 65536");
    if (o->precision != DDH_FP64 && o->precision != DDH_MIXED)
        return fail(DDH_E_ARG, "ddh_solve_batch_noalloc: precision must be DDH_FP64 or DDH_MIXED");
    const int lanes = o->jv_mode ? 1 : (o->lanes == 0 ? 2 : o->lanes);
    const size_t need = ws_layout(b->n, b->NP, o->precision, lanes).total;
    if (need > (size_t)DDH_NOALLOC_SCRATCH_BYTES) {
        snprintf(g_err, sizeof(g_err),
                 "ddh_solve_batch_noalloc: %d devices of NP = %d need %zu workspace bytes, more "
                 "than the library's scratch of %zu (use ddh_solve_batch with a workspace)",
                 b->n, b->NP, need, (size_t)DDH_NOALLOC_SCRATCH_BYTES);
        return DDH_E_WORKSPACE;
    }
    void *ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != DDH_OK) return rc;
    return solve_batch(b, V, nV, o, J, ok, its, status, nullptr, nullptr, ws,
                       (size_t)DDH_NOALLOC_SCRATCH_BYTES, stream);
}

struct AdjLayout {
    size_t fac, st, lam, iu, ts, ion, nu, total, slots, ls;
};

// Per launched lane (ADJ_SLOTS workgroups of 64 at most): the factors, the state's V, n, p,
// lambda and the current's integral, each [rows][ls] lane-innermost.
static AdjLayout adj_layout(int32_t n, int32_t NP, int32_t nV) {
    (void)nV;
    AdjLayout w;
    w.slots = (size_t)(n < ddh::ADJ_SLOTS ? n : ddh::ADJ_SLOTS);
    w.ls = w.slots * 64;
    size_t off = 0;
    w.fac = off; off = align256(off + (size_t)NP * ddh::ADJ_F * w.ls * sizeof(double));
    w.st = off; off = align256(off + (size_t)NP * 3 * w.ls * sizeof(double));
    w.lam = off; off = align256(off + (size_t)NP * 3 * w.ls * sizeof(double));
    w.iu = off; off = align256(off + (size_t)(NP + 1) * w.ls * sizeof(double));
    w.ts = off; off = align256(off + (size_t)(NP + 2) * 2 * w.ls * sizeof(double));
    w.ion = off; off = align256(off + (size_t)4 * w.ls * sizeof(double));
    w.nu = off; off = align256(off + (size_t)NP * w.ls * sizeof(double));
    w.total = off;
    return w;
}

extern "C" size_t ddh_adjoint_workspace_bytes(int32_t n, int32_t NP, int32_t nV) {
    if (n <= 0 || NP < 2 || nV < 1) return 0;
    return adj_layout(n, NP, nV).total;
}

extern "C" int ddh_adjoint(const ddh_batch *b, const double *V, int32_t nV, const double *states,
                           const double *pre_states, const uint8_t *ok, int32_t ntheta,
                           const double *dpt,
                           const double *dseg, const double *dsc, int32_t bernoulli,
                           double *grad, int32_t *status, void *workspace,
                           size_t workspace_bytes, void *stream) {
    if (!b || !V || !states || !ok || !dpt || !dseg || !dsc || !grad || !status || !workspace)
        return fail(DDH_E_ARG, "ddh_adjoint: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || nV < 1 || ntheta < 1 || b->nseg < 1 ||
        b->nseg > DDH_ADJOINT_MAX_SEGMENTS)
        return fail(DDH_E_SHAPE, "ddh_adjoint: n > 0, 2 <= NP <= 65536, nV >= 1, ntheta >= 1, "
                                 "1 <= nseg <= 14");
    if (bernoulli != 0 && bernoulli != 1)
        return fail(DDH_E_ARG, "ddh_adjoint: bernoulli must be 0 (exact) or 1 (simss)");
    AdjLayout w = adj_layout(b->n, b->NP, nV);
    if (workspace_bytes < w.total) return fail(DDH_E_WORKSPACE, "ddh_adjoint: workspace too small");
    char *base = (char *)workspace;
    ddh::AdjArgs a;
    a.n = b->n; a.NP = b->NP; a.nV = nV; a.nseg = b->nseg; a.ntheta = ntheta; a.bern = bernoulli;
    a.pt = b->pt; a.sid = b->sid; a.seg = b->seg; a.sc = b->sc;
    a.V = V; a.states = states; a.pre = pre_states; a.ok = ok;
    a.dpt = dpt; a.dseg = dseg; a.dsc = dsc;
    a.grad = grad; a.status = status;
    a.fac = (double *)(base + w.fac);
    a.st = (double *)(base + w.st);
    a.lam = (double *)(base + w.lam);
    a.iu = (double *)(base + w.iu);
    a.ts = (double *)(base + w.ts);
    a.ion = (double *)(base + w.ion);
    a.nu = (double *)(base + w.nu);
    a.ls = w.ls;
    a.passes = 0;
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((unsigned)w.slots), block(64);
    const size_t lds =
        (size_t)ddh::adj_table_size(b->nseg, b->NP, ddh::adj_k(ntheta, 0)) * sizeof(double);
    if (lds > 160 * 1024) return fail(DDH_E_SHAPE, "ddh_adjoint: NP too large for the LDS table");
    const bool rev = ddh::adj_k(ntheta, 0) > 0;
    (void)hipFuncSetAttribute(rev ? (const void *)ddh::adjoint_kernel<true>
                                  : (const void *)ddh::adjoint_kernel<false>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (rev) hipLaunchKernelGGL(ddh::adjoint_kernel<true>, grid, block, lds, s, a);
    else hipLaunchKernelGGL(ddh::adjoint_kernel<false>, grid, block, lds, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adjoint_kernel launch");
    return DDH_OK;
}

namespace ddh {
// theta_bar[d][k] = sum_v dJbar[d][v] grad[d][v][k]: one thread per (device, direction)
__global__ void __launch_bounds__(256) vjp_kernel(int n, int nV, int K, const double *dJbar,
                                                  const double *grad, double *theta_bar) {
    const long gid = (long)blockIdx.x * 256 + threadIdx.x;
    if (gid >= (long)n * K) return;
    const int d = (int)(gid / K), k = (int)(gid % K);
    double s = 0.0;
    for (int v = 0; v < nV; ++v)
        s += dJbar[(size_t)d * nV + v] * grad[((size_t)d * nV + v) * K + k];
    theta_bar[gid] = s;
}
}  // namespace ddh

extern "C" size_t ddh_adjoint_vjp_workspace_bytes(int32_t n, int32_t NP, int32_t nV,
                                                  int32_t ntheta) {
    if (n <= 0 || NP < 2 || nV < 1 || ntheta < 1) return 0;
    return align256(ddh_adjoint_workspace_bytes(n, NP, nV)) +
           (size_t)n * nV * ntheta * sizeof(double);
}

extern "C" int ddh_adjoint_vjp(const ddh_batch *b, const double *V, int32_t nV,
                               const double *states, const double *pre_states,
                               const uint8_t *ok, int32_t ntheta,
                               const double *dpt, const double *dseg, const double *dsc,
                               int32_t bernoulli, const double *dJbar, double *theta_bar,
                               int32_t *status, void *workspace, size_t workspace_bytes,
                               void *stream) {
    if (!b || !dJbar || !theta_bar || !workspace)
        return fail(DDH_E_ARG, "ddh_adjoint_vjp: NULL argument");
    if (b->n <= 0 || nV < 1 || ntheta < 1)
        return fail(DDH_E_SHAPE, "ddh_adjoint_vjp: n > 0, nV >= 1, ntheta >= 1");
    const size_t need = ddh_adjoint_vjp_workspace_bytes(b->n, b->NP, nV, ntheta);
    if (need == 0 || workspace_bytes < need)
        return fail(DDH_E_WORKSPACE, "ddh_adjoint_vjp: workspace smaller than ddh_adjoint_vjp_workspace_bytes");
    const size_t abytes = align256(ddh_adjoint_workspace_bytes(b->n, b->NP, nV));
    double *grad = (double *)((char *)workspace + abytes);
    const int rc = ddh_adjoint(b, V, nV, states, pre_states, ok, ntheta, dpt, dseg, dsc,
                               bernoulli, grad, status, workspace, abytes, stream);
    if (rc != DDH_OK) return rc;
    const long threads = (long)b->n * ntheta;
    hipLaunchKernelGGL(ddh::vjp_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, b->n, nV, ntheta, dJbar, grad, theta_bar);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "vjp_kernel launch");
    return DDH_OK;
}

extern "C" int ddh_assemble(const ddh_batch *b, const double *y, const double *Vint,
                            const double *logs, int32_t bernoulli, double *F, double *A,
                            double *B, double *C, double *U, double *J, void *stream) {
    if (!b || !y || !Vint || !F || !A || !B || !C || !U || !J)
        return fail(DDH_E_ARG, "ddh_assemble: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || b->nseg < 1 || b->nseg > DDH_MAX_SEGMENTS)
        return fail(DDH_E_SHAPE, "ddh_assemble: n > 0, 2 <= NP <= 65536, 1 <= nseg <= 14");
    if (bernoulli != 0 && bernoulli != 1)
        return fail(DDH_E_ARG, "ddh_assemble: bernoulli must be 0 (exact) or 1 (simss)");
    ddh::AsmArgs a = {};
    a.n = b->n; a.NP = b->NP; a.nseg = b->nseg; a.bern = bernoulli;
    a.pt = b->pt; a.sid = b->sid; a.seg = b->seg; a.sc = b->sc;
    a.y = y; a.Vint = Vint; a.logs = logs;
    a.F = F; a.A = A; a.B = B; a.C = C; a.U = U; a.J = J;
    dim3 grid((b->n + 63) / 64), block(64);
    hipLaunchKernelGGL(ddh::assemble_kernel<false>, grid, block, 0, (hipStream_t)stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "assemble_kernel launch");
    return DDH_OK;
}

extern "C" int ddh_assemble_ions(const ddh_batch *b, const double *y, const double *Vint,
                                 const double *logs, int32_t bernoulli, double *F, double *A,
                                 double *B, double *C, double *U, void *stream) {
    if (!b || !y || !Vint || !logs || !F || !A || !B || !C || !U)
        return fail(DDH_E_ARG, "ddh_assemble_ions: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || b->nseg < 1 || b->nseg > DDH_MAX_SEGMENTS)
        return fail(DDH_E_SHAPE, "ddh_assemble_ions: n > 0, 2 <= NP <= 65536, 1 <= nseg <= 14");
    if (bernoulli != 0 && bernoulli != 1)
        return fail(DDH_E_ARG, "ddh_assemble_ions: bernoulli must be 0 (exact) or 1 (simss)");
    ddh::AsmArgs a = {};
    a.n = b->n; a.NP = b->NP; a.nseg = b->nseg; a.bern = bernoulli;
    a.pt = b->pt; a.sid = b->sid; a.seg = b->seg; a.sc = b->sc;
    a.y5 = y; a.Vint = Vint; a.logs = logs;
    a.F5 = F; a.A5 = A; a.B5 = B; a.C5 = C; a.U5 = U;
    dim3 grid((b->n + 63) / 64), block(64);
    hipLaunchKernelGGL(ddh::assemble_kernel<true>, grid, block, 0, (hipStream_t)stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "assemble_kernel launch");
    return DDH_OK;
}

static size_t imp_bytes(int32_t n, int32_t NP, int row) {
    if (n <= 0 || NP < 1) return 0;
    return ((size_t)n * NP * row + (size_t)n * ddh::IS_N) * sizeof(double);
}

extern "C" size_t ddh_impedance_workspace_bytes(int32_t n, int32_t NP) {
    return imp_bytes(n, NP, ddh::IM_ROW);
}

extern "C" size_t ddh_impedance_ions_workspace_bytes(int32_t n, int32_t NP) {
    return imp_bytes(n, NP, ddh::IMI_ROW);
}

static int impedance_launch(const char *name, bool ions, const ddh_batch *b, const double *ion_mu,
                            const double *y, const double *Vint, const uint8_t *iface,
                            const double *laps, int32_t nf, double dt, int32_t bernoulli,
                            double *Y, int32_t *status, void *workspace, size_t workspace_bytes,
                            void *stream) {
    char msg[160];
    if (!b || !y || !Vint || !iface || !laps || !Y || !status || !workspace || (ions && !ion_mu)) {
        snprintf(msg, sizeof msg, "%s: NULL argument", name);
        return fail(DDH_E_ARG, msg);
    }
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || b->nseg < 1 || b->nseg > DDH_MAX_SEGMENTS) {
        snprintf(msg, sizeof msg, "%s: n > 0, 2 <= NP <= 65536, 1 <= nseg <= 14", name);
        return fail(DDH_E_SHAPE, msg);
    }
    if (nf < 1 || (long)b->n * nf > (1L << 30)) {
        snprintf(msg, sizeof msg, "%s: nf >= 1 and n * nf <= 2^30", name);
        return fail(DDH_E_SHAPE, msg);
    }
    if (bernoulli != 0 && bernoulli != 1) {
        snprintf(msg, sizeof msg, "%s: bernoulli must be 0 (exact) or 1 (simss)", name);
        return fail(DDH_E_ARG, msg);
    }
    if (!(dt >= 0.0) || !isfinite(dt)) {
        snprintf(msg, sizeof msg, "%s: dt must be finite and >= 0 (0: continuous)", name);
        return fail(DDH_E_ARG, msg);
    }
    const int row = ions ? (int)ddh::IMI_ROW : (int)ddh::IM_ROW;
    if (workspace_bytes < imp_bytes(b->n, b->NP, row)) {
        snprintf(msg, sizeof msg, "%s: workspace smaller than its workspace_bytes", name);
        return fail(DDH_E_WORKSPACE, msg);
    }
    ddh::ImpArgs a = {};
    a.n = b->n; a.NP = b->NP; a.nseg = b->nseg; a.bern = bernoulli; a.nf = nf;
    a.pt = b->pt; a.sid = b->sid; a.seg = b->seg; a.sc = b->sc;
    a.y = y; a.Vint = Vint; a.iface = iface; a.laps = laps; a.dt = dt;
    a.tab = (double *)workspace;
    a.dsc = a.tab + (size_t)b->n * b->NP * row;
    a.Y = Y; a.status = status;
    a.imu = ion_mu; a.row = row;
    hipStream_t s = (hipStream_t)stream;
    const unsigned prep_groups = (unsigned)(((b->n + ddh::XCDS - 1) / ddh::XCDS) * ddh::XCDS);
    if (ions) hipLaunchKernelGGL(ddh::impedance_prep_kernel<true>, dim3(prep_groups), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(ddh::impedance_prep_kernel<false>, dim3(prep_groups), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "impedance_prep_kernel launch");
    const long lanes = (long)b->n * nf;
    const dim3 grid((unsigned)((lanes + 255) / 256));
    if (ions) hipLaunchKernelGGL(ddh::impedance_kernel<true>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(ddh::impedance_kernel<false>, grid, dim3(256), 0, s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "impedance_kernel launch");
    return DDH_OK;
}

extern "C" int ddh_impedance(const ddh_batch *b, const double *y, const double *Vint,
                             const uint8_t *iface, const double *laps, int32_t nf, double dt,
                             int32_t bernoulli, double *Y, int32_t *status, void *workspace,
                             size_t workspace_bytes, void *stream) {
    return impedance_launch("ddh_impedance", false, b, nullptr, y, Vint, iface, laps, nf, dt,
                            bernoulli, Y, status, workspace, workspace_bytes, stream);
}

extern "C" int ddh_impedance_ions(const ddh_batch *b, const double *ion_mu, const double *y,
                                  const double *Vint, const uint8_t *iface, const double *laps,
                                  int32_t nf, double dt, int32_t bernoulli, double *Y,
                                  int32_t *status, void *workspace, size_t workspace_bytes,
                                  void *stream) {
    return impedance_launch("ddh_impedance_ions", true, b, ion_mu, y, Vint, iface, laps, nf, dt,
                            bernoulli, Y, status, workspace, workspace_bytes, stream);
}

extern "C" int ddh_graded_layout(int32_t n, int32_t NP, int32_t nl, const double *L,
                                 double beta, double d0, uint8_t *lid, uint8_t *sid, int32_t *i0,
                                 double *unit, double *unit_sum, uint8_t *iface, int32_t *nseg,
                                 void *stream) {
    if (!L || !lid || !sid || !i0 || !unit || !unit_sum || !iface || !nseg)
        return fail(DDH_E_ARG, "ddh_graded_layout: NULL argument");
    if (n <= 0 || NP < 2 || NP > 65536 || nl < 1 || nl > ddh::BUILD_MAX_LAYERS ||
        NP + 2 < nl * (ddh::GRADED_MIN_POINTS + 1))
        return fail(DDH_E_SHAPE, "ddh_graded_layout: n > 0, 2 <= NP <= 65536, 1 <= nl <= 8, "
                                 "six points per layer");
    if (!(beta >= 0.0) || !(d0 > 0.0))
        return fail(DDH_E_ARG, "ddh_graded_layout: beta >= 0 and d0 > 0");
    ddh::LayoutArgs a = {};
    a.n = n; a.NP = NP; a.nl = nl; a.beta = beta; a.d0 = d0; a.L = L;
    a.lid = lid; a.sid = sid; a.i0 = i0; a.unit = unit; a.unit_sum = unit_sum; a.iface = iface;
    a.nseg = nseg;
    hipStream_t s = (hipStream_t)stream;
    const dim3 dev_grid((unsigned)((n + 63) / 64));
    hipLaunchKernelGGL(ddh::layout_structure_kernel, dev_grid, dim3(64), 0, s, a);
    const long points = (long)n * (NP + 2);
    hipLaunchKernelGGL(ddh::layout_point_kernel, dim3((unsigned)((points + 255) / 256)), dim3(256),
                       0, s, a);
    hipLaunchKernelGGL(ddh::layout_sums_kernel, dev_grid, dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "layout kernels launch");
    return DDH_OK;
}

extern "C" int ddh_build(const ddh_structure *st, const double *layers, const double *scalars,
                         const double *gen, int32_t gen_m, int32_t K, const double *dlayers,
                         const double *dscalars, double *pt, double *seg, double *sc,
                         double *ion_mu, double *dpt, double *dseg, double *dsc, void *stream) {
    if (!st || !layers || !scalars || !pt || !seg || !sc || !ion_mu || !st->lid || !st->sid ||
        !st->flags || !st->i0 || !st->unit || !st->unit_sum)
        return fail(DDH_E_ARG, "ddh_build: NULL argument");
    if (gen_m < 0 || gen_m == 1 || (gen_m > 0 && !gen))
        return fail(DDH_E_ARG, "ddh_build: gen_m 0 (no profile) or >= 2 nodes with gen");
    if (K < 0 || K > 1024 || (K > 0 && (!dlayers || !dscalars || !dpt || !dseg || !dsc)))
        return fail(DDH_E_ARG, "ddh_build: 0 <= K <= 1024, and the directions' arrays when K > 0");
    if (st->n <= 0 || st->NP < 2 || st->NP > 65536 || st->nl < 1 || st->nl > ddh::BUILD_MAX_LAYERS ||
        st->nseg < 1 || st->nseg > DDH_MAX_SEGMENTS || (long)st->n * (K > 0 ? K : 1) > (1L << 30))
        return fail(DDH_E_SHAPE, "ddh_build: n > 0, 2 <= NP <= 65536, 1 <= nl <= 8, 1 <= nseg <= 14");
    ddh::BuildArgs a = {};
    a.n = st->n; a.NP = st->NP; a.nl = st->nl; a.nseg = st->nseg; a.K = K;
    a.lid = st->lid; a.sid = st->sid; a.flags = st->flags; a.i0 = st->i0;
    a.unit = st->unit; a.unit_sum = st->unit_sum;
    a.layers = layers; a.scalars = scalars; a.dlayers = dlayers; a.dscalars = dscalars;
    a.gen = gen; a.gen_m = gen_m;
    a.pt = pt; a.seg = seg; a.sc = sc; a.imu = ion_mu; a.dpt = dpt; a.dseg = dseg; a.dsc = dsc;
    const long lanes = (long)st->n * (K > 0 ? K : 1);
    hipLaunchKernelGGL(ddh::build_kernel, dim3((unsigned)((lanes + 63) / 64)), dim3(64), 0,
                       (hipStream_t)stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "build_kernel launch");
    return DDH_OK;
}

extern "C" int ddh_block_sweep(int32_t n, int32_t NP, int32_t b, int32_t k, const void *A,
                               const void *B, const void *C, void *d, int32_t precision,
                               void *workspace, void *stream) {
    if (!A || !B || !C || !d || !workspace) return fail(DDH_E_ARG, "ddh_block_sweep: NULL argument");
    if (n <= 0 || NP < 1 || (b != 3 && b != 5) || k < 1 || k > 8)
        return fail(DDH_E_SHAPE, "ddh_block_sweep: n > 0, NP >= 1, b in {3, 5}, 1 <= k <= 8");
    hipStream_t s = (hipStream_t)stream;
    dim3 grid((n + 63) / 64), block(64);
    if (precision == DDH_FP64) {
        if (b == 3)
            hipLaunchKernelGGL((ddh::block_sweep_kernel<double, 3>), grid, block, 0, s, n, NP, k,
                               (const double *)A, (const double *)B, (const double *)C, (double *)d, (double *)workspace);
        else
            hipLaunchKernelGGL((ddh::block_sweep_kernel<double, 5>), grid, block, 0, s, n, NP, k,
                               (const double *)A, (const double *)B, (const double *)C, (double *)d, (double *)workspace);
    } else if (precision == DDH_MIXED) {
        if (b == 3)
            hipLaunchKernelGGL((ddh::block_sweep_kernel<float, 3>), grid, block, 0, s, n, NP, k,
                               (const float *)A, (const float *)B, (const float *)C, (float *)d, (float *)workspace);
        else
            hipLaunchKernelGGL((ddh::block_sweep_kernel<float, 5>), grid, block, 0, s, n, NP, k,
                               (const float *)A, (const float *)B, (const float *)C, (float *)d, (float *)workspace);
    } else {
        return fail(DDH_E_ARG, "ddh_block_sweep: precision must be DDH_FP64 or DDH_MIXED");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "block_sweep_kernel launch");
    return DDH_OK;
}


// ---- the transient (ddhip_transient.h) -----------------------------------------------------
struct TrLayout {
    size_t Y, fb, fac, total;
};

static TrLayout tr_layout(int32_t n, int32_t NP) {
    TrLayout w;
    size_t off = 0;
    w.Y = off; off = align256(off + (size_t)3 * NP * ddh::TR_NC * n * sizeof(double));
    w.fb = off; off = align256(off + (size_t)NP * n * sizeof(double));
    w.fac = off; off = align256(off + (size_t)NP * ddh::TR_FAC * n * sizeof(double));
    w.total = off;
    return w;
}

extern "C" size_t ddh_transient_workspace_bytes(int32_t n, int32_t NP) {
    if (n <= 0 || NP < 2) return 0;
    return tr_layout(n, NP).total;
}

extern "C" int ddh_transient(const ddh_batch *b, const double *ion_mu, const uint8_t *iface,
                             const double *t, const double *V, const double *G, int32_t nt,
                             const double *y0, const ddh_topts *o, int32_t ions, double *J,
                             double *J_ext, uint8_t *ok, int32_t *its, uint8_t *kept,
                             int32_t *status, void *workspace, size_t workspace_bytes,
                             void *stream) {
    if (!b || !ion_mu || !iface || !t || !V || !G || !y0 || !o || !J || !J_ext || !ok || !its ||
        !kept || !status || !workspace)
        return fail(DDH_E_ARG, "ddh_transient: NULL argument");
    if (b->n <= 0 || b->NP < 2 || b->NP > 65536 || nt < 1 || b->nseg < 1 || b->nseg > DDH_MAX_SEGMENTS)
        return fail(DDH_E_SHAPE, "ddh_transient: n > 0, 2 <= NP <= 65536, nt >= 1, 1 <= nseg <= 14");
    if (o->bernoulli != 0 && o->bernoulli != 1)
        return fail(DDH_E_ARG, "ddh_transient: bernoulli must be 0 (exact) or 1 (simss)");
    if (o->max_iter < 1) return fail(DDH_E_ARG, "ddh_transient: max_iter must be >= 1");
    if (o->failure_mode < 0 || o->failure_mode > 2)
        return fail(DDH_E_ARG, "ddh_transient: failure_mode must be 0, 1 or 2");
    const TrLayout w = tr_layout(b->n, b->NP);
    if (workspace_bytes < w.total) return fail(DDH_E_WORKSPACE, "ddh_transient: workspace too small");
    char *base = (char *)workspace;
    ddh::TrArgs a = {};
    a.n = b->n; a.NP = b->NP; a.nseg = b->nseg; a.nt = nt;
    a.pt = b->pt; a.sid = b->sid; a.iface = iface; a.seg = b->seg; a.sc = b->sc; a.imu = ion_mu;
    a.t = t; a.V = V; a.G = G; a.y0 = y0;
    a.Y = (double *)(base + w.Y);
    a.fb = (double *)(base + w.fb);
    a.fac = (double *)(base + w.fac);
    a.J = J; a.Jext = J_ext; a.ok = ok; a.kept = kept; a.its = its; a.status = status;
    a.tol = o->tol; a.max_iter = o->max_iter; a.bern = o->bernoulli; a.zimt = o->zimt_current;
    a.failure_mode = o->failure_mode;
    dim3 grid((b->n + 63) / 64), block(64);
    if (ions) hipLaunchKernelGGL(ddh::transient_kernel<5>, grid, block, 0, (hipStream_t)stream, a);
    else hipLaunchKernelGGL(ddh::transient_kernel<3>, grid, block, 0, (hipStream_t)stream, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "transient_kernel launch");
    return DDH_OK;
}
