// Packed batch layout shared by the HIP kernels and the host driver.
//
// Every array is lane-innermost ("batch-innermost"): element (row, field, lane) of an array
// with F fields lives at (row * F + field) * n + lane, n the number of lanes (devices). A
// wavefront of 64 devices reading one field of one grid point reads 64 consecutive values.
//
// A device is described by
//   pt   double [NP + 2][PT_N][n]  per grid point: spacing h (over L), position x, generation
//                                  G (at G_frac = 1), net fixed ions (anions - cations)
//   sid  uint8  [NP + 2][n]        the point's segment: a run of points sharing every material
//                                  parameter and the mobility of the interval after the point
//   seg  double [nseg][SEG_N][n]   each segment's parameters
//   sc   double [SC_N][n]          per-device scalars
// The reference keeps these as per-point arrays of its Device (src/jaxdd/device.py:251-382);
// layers are piecewise constant there, so a 3-layer device has 5 segments (the interface
// points carry the interface interval's mobility L*h*nu/V_T, device.py:894-898).
#pragma once

#include <stdint.h>

#ifdef DDH_CPU_EMULATION
#define DDH_MAP_FN inline
#else
#include <hip/hip_runtime.h>
#define DDH_MAP_FN __host__ __device__ inline
#endif

namespace ddh {

enum PtField { PT_H = 0, PT_X, PT_G, PT_FIX, PT_N };

enum SegField {
    SEG_EPS = 0,  // permittivity (F/m)
    SEG_OFFN,     // V_gn - V (E_c offset plus the density-of-states term)
    SEG_OFFP,     // V_gp - V
    SEG_NC,       // N_c
    SEG_DOP,      // N_A - N_D
    SEG_KD,       // direct recombination constant
    SEG_NI2,      // n_i^2
    SEG_NT,       // bulk trap density (0 without traps)
    SEG_CN,
    SEG_CP,
    SEG_NT0,
    SEG_PT0,
    SEG_AM,       // 1 where anions are mobile
    SEG_CM,       // 1 where cations are mobile
    SEG_MUN,      // electron mobility of the interval after the point
    SEG_MUP,
    SEG_EC,       // conduction band edge (eV), for the linear start
    SEG_NI,       // intrinsic density, for the linear start
    SEG_SP,       // Poisson's row scale 1/(q N_c)
    SEG_SCS,      // the continuity rows' scale times h: 1/(max(mu_n, mu_p) V_T N_c)
    SEG_IMUN,     // 1/mu_n, 1/mu_p of the interval after the point (the current's integrals)
    SEG_IMUP,
    SEG_TT,       // the bulk traps' charge type: 1 donor, -1 acceptor, 0 neutral or none
    SEG_N
};

enum ScalarField {
    SC_L = 0, SC_VT, SC_GFRAC, SC_V0, SC_VL, SC_WL, SC_WR,
    SC_LNL, SC_LNR, SC_LPL, SC_LPR,  // ln n and ln p at the left and right contacts
    SC_ELEC,                         // left_elec: -1 cathode left, +1 anode left
    SC_AT, SC_CT,                    // mobile anions and cations per area (0: none)
    SC_VPRE, SC_JS, SC_GSH, SC_NFLOOR,
    SC_PRE, SC_FIX,                  // pre_cond, fix_ions as 0/1
    SC_N
};

// Number of state buffers per lane (Newton iterate, continuation carries, start stages).
constexpr int NBUF = 6;
// Factors per grid point: c' (3x3) and d' (3 x 3 right-hand sides).
constexpr int NFAC = 18;

// Workgroups go to the 8 XCDs in turn (workgroup b to XCD b % 8), each XCD with its own L2,
// and a 128-byte line of the lane-innermost batch arrays holds 16 neighbouring devices' values.
// A launch that gives each workgroup one device at a time maps workgroup b's k-th device so
// that every XCD takes a contiguous run of the n devices: XCD x the devices
// [x * per, (x + 1) * per), per = ceil(n / G) over G = min(8, workgroups) XCDs, its c
// workgroups (the x-th of each round of G) the run's devices j, j + c, j + 2c, ... Each device
// goes to exactly one (workgroup, k); -1 past the run's end. With as many workgroups as
// devices rounded up to a multiple of 8 every workgroup takes at most one device.
constexpr int XCDS = 8;

struct XcdMap {
    int x, j, c, per;
};

DDH_MAP_FN XcdMap xcd_map(int block, int nblocks, int n) {
    const int G = nblocks < XCDS ? nblocks : XCDS;
    XcdMap m;
    m.x = block % G;
    m.j = block / G;
    m.c = (nblocks - m.x + G - 1) / G;  // workgroups of XCD x
    m.per = (n + G - 1) / G;
    return m;
}

// Workgroup's k-th device (k = 0, 1, ...), or -1 when it has no k-th device.
DDH_MAP_FN int xcd_device(const XcdMap &m, int k, int n) {
    const int i = m.j + k * m.c;
    if (i >= m.per) return -1;
    const int d = m.x * m.per + i;
    return d < n ? d : -1;
}

}  // namespace ddh
