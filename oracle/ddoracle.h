/* ddoracle: FP64 CPU restatement of jaxdd's steady-state J-V solve.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker the HIP product path is compared
 * against; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * Nothing in the product (jaxdd_amd/) links, imports or calls it.
 *
 * It restates, in plain C and in the reference's own order of operations:
 *   - bernoulli / bernoulli_simss           src/jaxdd/solver.py:95-121
 *   - _stencils (the subset on this path)   src/jaxdd/solver.py:210-249
 *   - _recombination                        src/jaxdd/solver.py:734-745
 *   - _assemble (by edges)                  src/jaxdd/solver.py:1021-1187
 *   - _block_thomas, _inverse3              src/jaxdd/solver.py:1481-1560
 *   - _ion_constraints, _ion_logs           src/jaxdd/solver.py:1241-1265, 1695-1725
 *   - _poisson_delta (bordered, damped)     src/jaxdd/solver.py:1757-1800
 *   - density_steps, step caps              src/jaxdd/solver.py:59-88
 *   - _current / _current_parts (integral)  src/jaxdd/batch.py:302-403
 *   - _newton_loop (4 stop rules)           src/jaxdd/batch.py:740-848
 *   - _start, _freeze, _solve_one           src/jaxdd/batch.py:859-1006
 *   - solve_jv / _scan / _noise_step        src/jaxdd/solver.py:1728-1751, 2041-2245
 */
#ifndef DDORACLE_H
#define DDORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One device, its arrays on all NP + 2 points (point-major, float64), as the reference's
 * Device holds them (src/jaxdd/device.py:251-382). mu_n/mu_p are interval mobilities
 * (mu[i] between x[i] and x[i+1]); iface has NP + 1 entries, 1 on the interval after the
 * last point of every layer but the last. */
typedef struct {
    int32_t NP;
    const double *x, *h, *eps, *offset_n, *offset_p, *N_c, *ni, *E_c, *N_D, *N_A;
    const double *mu_n, *mu_p, *k_direct, *Nt, *Cn, *Cp, *nt0, *pt0, *G;
    const double *a_mask, *c_mask, *a_fixed, *c_fixed;
    const uint8_t *iface;
    double L, Vt, g_frac, V0, VL, W_L, W_R;
    double ln_n_left, ln_n_right, ln_p_left, ln_p_right;
    double left_elec, a_total, c_total, V_pre, J_scale, G_shunt;
    int32_t pre_cond, fix_ions;
    const double *mu_a, *mu_c; /* the mobile ions' mobility per point (transients) */
    const double *tt; /* per point: the bulk traps' charge type, 1 donor, -1 acceptor, 0
                         neutral or none (NULL: none charged); the steady state only */
} ddo_device;

typedef struct {
    double tol;
    int32_t max_iter, start_iter, strict, bernoulli; /* bernoulli: 0 exact, 1 simss */
} ddo_opts;

int ddo_abi_version(void);

/* B(x) (kind 0 exact, 1 SimSS's); *dB receives B'(x) when not NULL. */
double ddo_bernoulli(double x, int kind, double *dB);

/* _solve_one for ascending voltages: J (without the shunt), converged, iterations. */
int ddo_solve_one(const ddo_device *dev, const double *V_sorted, int nV, const ddo_opts *o,
                  double *J, uint8_t *ok, int32_t *its);
/* The same, also each voltage's final iterate (nV x NP x 3) into states. */
int ddo_solve_one_states(const ddo_device *dev, const double *V_sorted, int nV,
                         const ddo_opts *o, double *J, uint8_t *ok, int32_t *its,
                         double *states);

/* solve_jv: voltages in the given order, J with the shunt. Returns -2 when the pre-bias does
 * not converge (the reference raises RuntimeError). */
int ddo_solve_jv(const ddo_device *dev, const double *V, int nV, double tol, int max_iter,
                 int bernoulli, double *J, uint8_t *ok, int32_t *its);

/* _assemble at a full state (NP + 2 x 3, row-major): F (NP x 3), A, B, C (NP x 3 x 3),
 * U (NP x 3 x 2). la/lc are the ions' normalisations (the point logs of one region each). */
int ddo_assemble(const ddo_device *dev, const double *full, double g_frac, double la, double lc,
                 int bernoulli, double *F, double *A, double *B, double *C, double *U);
int ddo_assemble_ions(const ddo_device *dev, const double *full5, double g_frac,
                      const double *logs, int bernoulli, double *F, double *A, double *B,
                      double *C, double *U);

/* _block_thomas: n systems in one call are not needed; one system of NP blocks b x b with k
 * right-hand sides, row-major (NP x b x b, NP x b x k). */
int ddo_block_thomas(int NP, int b, int k, const double *A, const double *B, const double *C,
                     const double *d, double *x);

/* _noise_step (solver.py:1728-1751) at the interior state y (NP x 3) with the ions' point
 * normalisations logs[2] and internal voltage Vint, the round-off unit eps (DBL_EPSILON in
 * solve_jv's stop). */
double ddo_noise_step(const ddo_device *dev, const double *y, const double *logs, double Vint,
                      double g_frac, int bernoulli, double eps);

/* Options of a transient (transient.py:537-637 `solve_transient`). */
typedef struct {
    double tol;
    int32_t max_iter, bernoulli, zimt_current, failure_mode;
} ddo_topts;

/* solve_transient for one device through nt steps (t, V internal = external, G the generation
 * fraction): per step J (internal current), J_ext (with the shunt), converged, iterations,
 * kept. Returns 0, -1 for a bad shape, -3 for a device outside the subset (a pre-bias, an ion
 * region reaching a contact). */
int ddo_transient(const ddo_device *dev, const double *t, const double *V, const double *G,
                  int nt, const ddo_topts *o, double *J, double *J_ext, uint8_t *ok,
                  int32_t *its, uint8_t *kept);

/* One time step's residual, blocks (nc = 3, or 5 with the ions as unknowns), slots' columns,
 * transient current and the ions' constraints g (2) with their rows W (NP x 2, the derivative in
 * each interior point's own ion logarithm) at the full state `full` ((NP + 2) x nc). fb NULL:
 * the bulk traps at their steady occupation (instantaneous) instead of over the step. J, g and
 * W may be NULL. */
int ddo_assemble_step(const ddo_device *dev, const double *full, int nc, double g_frac,
                      double dti, const double *full_prev, const double *fb,
                      const double *logs, int bernoulli, int zimt, double *F, double *A,
                      double *B, double *C, double *U, double *J, double *g, double *W);

/* The integral current of a full state (batch._current, no shunt). */
double ddo_current(const ddo_device *dev, const double *full, double g_frac);

#ifdef __cplusplus
}
#endif
#endif
