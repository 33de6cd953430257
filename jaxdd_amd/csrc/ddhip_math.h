// Device-side arithmetic shared by the kernels: the Bernoulli function, recombination, small
// dense algebra, NaN-keeping clip/max. FP64 unless a template says otherwise.
#pragma once

#ifdef DDH_CPU_EMULATION
#include <cmath>
#define DDH_MFN inline
using std::expm1;
using std::log1p;
#else
#include <hip/hip_runtime.h>
#include <math.h>
#define DDH_MFN __device__ __host__ inline
#endif

namespace ddh {

constexpr double Q_E = 1.6022e-19;  // SIMsalabim's elementary charge (src/jaxdd/constants.py:8)

// Convergence constants, bit-equal to the reference's (src/jaxdd/solver.py:44-71,
// src/jaxdd/batch.py:79-87).
constexpr double ROUND_OFF_FLOOR = 1e-6;
constexpr int STALL_STEPS = 3;
constexpr double NOISE_FLOOR_CAP = 1e-3;
constexpr double CAP_V = 10.0;  // potential step cap in V_T
constexpr double CAP_N = 5.0;   // ln n, ln p step cap
constexpr double ION_LOG_CAP = 5.0;
constexpr double FLOOR_DROP_CAP = 50.0;
constexpr double SERIES_BELOW = 1e-2;
constexpr double CURRENT_TOL = 1e-5;
constexpr double CURRENT_STEP_CAP = 1e-3;
constexpr double QUADRATIC_TOL = 1e-5;
constexpr double QUADRATIC_RATIO = 1e-2;
constexpr double POISSON_STOP = 1e-8;
constexpr double DBL_EPS = 2.220446049250313e-16;  // 2^-52, the FP64 machine epsilon

// jnp.clip and jnp.maximum keep NaN; fmin/fmax would drop it.
DDH_MFN double clipd(double x, double lo, double hi) {
    return x < lo ? lo : (x > hi ? hi : x);
}
DDH_MFN double nmax(double a, double b) {
    return (a != a || b != b) ? (a + b) : (a > b ? a : b);
}

// B(x) = x/expm1(x) and B(-x), with derivatives B'(x) and B'(-x) (src/jaxdd/solver.py:95-121).
// kind 0: exact, the series 1 - x/2 + x^2/12 - x^4/720 below |x| = 1e-2; kind 1: SIMsalabim's
// Taylor orders 1/2/4 below |x| 0.05/0.45/1.5. Above the series, the function is evaluated
// once at y = |x| and the other sign follows from B(-y) = B(y) + y (a sum of positives, no
// cancellation), its derivative from B'(-y) = -B'(y) - 1.
DDH_MFN void bernoulli_pair(double x, int kind, double &Bf, double &Bb, double &dBf,
                                      double &dBb) {
    double ax = fabs(x);
    if (kind == 0 && ax < SERIES_BELOW) {
        // the reference's series, its divisions by constants as products with their
        // reciprocals (|x| < 1e-2: the terms they touch are below 1e-5 of B)
        constexpr double I12 = 1.0 / 12.0, I720 = 1.0 / 720.0, I6 = 1.0 / 6.0, I180 = 1.0 / 180.0;
        double x2 = x * x, x4 = x2 * x2;
        Bf = 1.0 - x * 0.5 + x2 * I12 - x4 * I720;
        Bb = 1.0 + x * 0.5 + x2 * I12 - x4 * I720;
        dBf = -0.5 + x * I6 - x * x2 * I180;
        dBb = -0.5 - x * I6 + x * x2 * I180;
        return;
    }
    if (kind == 1 && ax < 1.5) {
        if (ax < 0.05) {
            Bf = 1.0 - 0.5 * x; Bb = 1.0 + 0.5 * x;
            dBf = -0.5; dBb = -0.5;
        } else if (ax < 0.45) {
            Bf = 1.0 - x * (0.5 - x / 12.0); Bb = 1.0 + x * (0.5 + x / 12.0);
            dBf = -0.5 + x / 6.0; dBb = -0.5 - x / 6.0;
        } else {
            double x2 = x * x;
            Bf = 1.0 + x * (-0.5 + x * (1.0 / 12.0 - x2 / 720.0));
            Bb = 1.0 - x * (-0.5 - x * (1.0 / 12.0 - x2 / 720.0));
            dBf = -0.5 + x / 6.0 - x * x2 / 180.0;
            dBb = -0.5 - x / 6.0 + x * x2 / 180.0;
        }
        return;
    }
    double e = expm1(ax);
    double r = 1.0 / e;
    double By = ax * r;
    double dBy = r * (1.0 - ax * (1.0 + r));  // 1/expm1(y) - y*exp(y)/expm1(y)^2
    double Bm = By + ax;
    double dBm = -dBy - 1.0;
    if (x >= 0) { Bf = By; Bb = Bm; dBf = dBy; dBb = dBm; }
    else        { Bf = Bm; Bb = By; dBf = dBm; dBb = dBy; }
}

// Direct plus SRH recombination through one neutral level (src/jaxdd/solver.py:734-745) and
// its derivatives in ln n and ln p.
DDH_MFN double recombination(double n, double p, double kd, double ni2, double Nt,
                                       double Cn, double Cp, double nt0, double pt0,
                                       double &dln_n, double &dln_p) {
    double np_ = n * p;
    double R = kd * (np_ - ni2);
    double dn = kd * p, dp = kd * n;
    if (Nt > 0) {
        double base = Nt * Cn * Cp;
        double num = base * (np_ - nt0 * pt0);
        double den = Cn * (n + nt0) + Cp * (p + pt0);
        double rden = 1.0 / den;
        R += num * rden;
        dn += (base * p - num * rden * Cn) * rden;
        dp += (base * n - num * rden * Cp) * rden;
    }
    dln_n = dn * n;
    dln_p = dp * p;
    return R;
}

// Recombination alone (the current's integrand).
DDH_MFN double recombination_value(double n, double p, double kd, double ni2,
                                             double Nt, double Cn, double Cp, double nt0,
                                             double pt0) {
    double np_ = n * p;
    double R = kd * (np_ - ni2);
    if (Nt > 0) R += Nt * Cn * Cp * (np_ - nt0 * pt0) / (Cn * (n + nt0) + Cp * (p + pt0));
    return R;
}

// Inverse of a 3x3 block by its adjugate (src/jaxdd/solver.py:1481-1495).
template <typename T>
DDH_MFN void inverse3(const T *m, T *out) {
    T a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    T c00 = e * i - f * h, c01 = c * h - b * i, c02 = b * f - c * e;
    T c10 = f * g - d * i, c11 = a * i - c * g, c12 = c * d - a * f;
    T c20 = d * h - e * g, c21 = b * g - a * h, c22 = a * e - b * d;
    T det = a * c00 + b * c10 + c * c20;
    T r = T(1) / det;
    out[0] = c00 * r; out[1] = c01 * r; out[2] = c02 * r;
    out[3] = c10 * r; out[4] = c11 * r; out[5] = c12 * r;
    out[6] = c20 * r; out[7] = c21 * r; out[8] = c22 * r;
}

// 2x2 solve with partial pivoting (the K x K Schur complement of two ion slots).
DDH_MFN void solve2(double s00, double s01, double s10, double s11, double &z0,
                              double &z1) {
    if (fabs(s10) > fabs(s00)) {
        double t;
        t = s00; s00 = s10; s10 = t;
        t = s01; s01 = s11; s11 = t;
        t = z0; z0 = z1; z1 = t;
    }
    double l = s10 / s00;
    double u11 = s11 - l * s01;
    double r1 = z1 - l * z0;
    z1 = r1 / u11;
    z0 = (z0 - s01 * z1) / s00;
}

}  // namespace ddh
