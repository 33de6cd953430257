/* ddhip: C-ABI of the MI355X (gfx950) batched steady-state drift-diffusion J-V solve.
 *
 * The reference (jaxdd 0.5.0) is pure Python on JAX and has no C/FFI boundary of its own; the
 * entry points below are what its Python API for this path binds to, each named with the
 * reference function it replaces. Every pointer argument is DEVICE memory unless stated;
 * sizes are plain integers; no torch or C++ types cross the boundary. The caller allocates
 * and frees every buffer, the workspace included (ddh_workspace_bytes); the library allocates
 * nothing persistent. Errors: 0 on success, a negative code otherwise, with a message from
 * ddh_last_error() (thread-local). No C++ exception crosses the ABI. Calls on distinct
 * streams are re-entrant; there is no global mutable state besides the error message.
 */
#ifndef DDHIP_H
#define DDHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DDH_ABI_VERSION 5

/* Segments (runs of points of one material) per device: the solve keeps each wavefront's
 * segment tables in the compute unit's 160 KB local data share. */
#define DDH_MAX_SEGMENTS 14
/* The adjoint reads the segment tables from global memory: the solve's limit. */
#define DDH_ADJOINT_MAX_SEGMENTS DDH_MAX_SEGMENTS

enum ddh_status {
    DDH_OK = 0,
    DDH_E_SHAPE = -1,   /* a size or shape the kernels do not take */
    DDH_E_HIP = -2,     /* a HIP runtime error (launch, sync) */
    DDH_E_WORKSPACE = -3,
    DDH_E_ARG = -4      /* a NULL pointer or an unknown option */
};

enum ddh_precision { DDH_FP64 = 0, DDH_MIXED = 1 };

/* A packed batch of n devices on NP interior points each, lane-innermost
 * (jaxdd_amd/csrc/ddhip_layout.h documents every field). Built by jaxdd_amd.batch.prepare,
 * the counterpart of src/jaxdd/batch.py:151-223 (`prepare`). A segment's 23rd field is its bulk
 * traps' charge type (1 donor, -1 acceptor, 0 neutral or none; device.py:940-948): their charge
 * enters the Poisson rows of ddh_solve_batch, ddh_solve_batch_states and ddh_assemble (the
 * start then solves first with them neutral, batch.py:898-900), ddh_assemble, ddh_adjoint and
 * ddh_adjoint_vjp (the charge's partials in the traps' fields). ddh_impedance,
 * ddh_impedance_ions and ddh_transient take no charged traps (the field 0). */
typedef struct {
    int32_t n;        /* devices */
    int32_t NP;       /* interior grid points */
    int32_t nseg;     /* segments per device (max over the batch), <= DDH_MAX_SEGMENTS */
    int32_t reserved;
    const double *pt;   /* [NP + 2][4][n] */
    const uint8_t *sid; /* [NP + 2][n] */
    const double *seg;  /* [nseg][23][n], the last field the bulk traps' charge type */
    const double *sc;   /* [20][n] */
} ddh_batch;

/* Options of src/jaxdd/batch.py:1136-1137 (`solve_batch`) and src/jaxdd/solver.py:2041-2043
 * (`solve_jv`). */
typedef struct {
    double tol;          /* Newton step tolerance (1e-10) */
    int32_t max_iter;    /* Newton steps per voltage (60 batch, 200 jv) */
    int32_t start_iter;  /* Poisson and start solves (max_iter) */
    int32_t strict;      /* 1: only the step below tol or the round-off floor stop Newton */
    int32_t bernoulli;   /* 0 exact B(x), 1 SIMsalabim's piecewise approximation */
    int32_t precision;   /* ddh_precision */
    int32_t jv_mode;     /* 0: solve_batch semantics; 1: solve_jv semantics (given order,
                            no extrapolation, tol/stall/noise-floor stops, pre-bias must hold) */
    int32_t lanes;       /* lanes per device in batch mode: 2 (0, the default) splits each
                            Newton solve at the grid's middle, 1 sweeps it from one end (the
                            reference's order of elimination), 4 splits it in quarters (for
                            batches too small to fill the GPU; devices without mobile ions
                            and NP >= 16 only: a device with ions gets status 2, NaN currents
                            and no converged voltage); jv mode always uses 1 */
} ddh_opts;

int ddh_abi_version(void);
const char *ddh_last_error(void);

/* Workspace bytes ddh_solve_batch needs for this shape and precision (any lanes). */
size_t ddh_workspace_bytes(int32_t n, int32_t NP, int32_t nV, int32_t precision);

/* The J-V curves of a batch: replaces src/jaxdd/batch.py:1033-1047 (`_solve_batch`, the
 * jit+vmap of `_solve_one` :937-1006) and, with opts.jv_mode = 1, src/jaxdd/solver.py:2105-2245
 * (`_scan` of `solve_jv`) for each device. V: nV voltages (device memory), ascending in batch
 * mode, the caller's order in jv mode. Outputs [n][nV]: J (A/m^2, without the shunt),
 * converged, iterations. status: [n] int32, 1 where a device's pre-bias failed in jv mode
 * (the reference raises RuntimeError there), 2 where opts.lanes = 4 met mobile ions. */
int ddh_solve_batch(const ddh_batch *b, const double *V, int32_t nV, const ddh_opts *opts,
                    double *J, uint8_t *ok, int32_t *its, int32_t *status, void *workspace,
                    size_t workspace_bytes, void *hip_stream);

/* ddh_solve_batch that also keeps each voltage's final state: states [nV][NP][3][n] doubles
 * (device memory), (V, ln n, ln p) of every interior point; voltages in the solve's order
 * (ascending in batch mode). pre_states: NULL, or [NP][3][n] doubles that receive the pre-bias
 * solution of each device with a pre-bias (batch mode; left as they are for the others). The
 * linearisation points of ddh_adjoint (the pre-bias's for devices whose ions it freezes). */
int ddh_solve_batch_states(const ddh_batch *b, const double *V, int32_t nV, const ddh_opts *opts,
                           double *J, uint8_t *ok, int32_t *its, int32_t *status, double *states,
                           double *pre_states, void *workspace, size_t workspace_bytes,
                           void *hip_stream);

/* The device memory the library holds for ddh_solve_batch_noalloc, per GPU it runs on: 256 MiB,
 * ddh_solve_batch's workspace for about 1,500 devices of NP = 400 (mixed precision). */
#define DDH_NOALLOC_SCRATCH_BYTES (256ull << 20)

/* DDH_NOALLOC_SCRATCH_BYTES, for callers that bind the library without its header. */
size_t ddh_noalloc_scratch_bytes(void);

/* ddh_solve_batch without allocating: no counterpart in the reference. For callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). The workspace is the library's own static device array of
 * DDH_NOALLOC_SCRATCH_BYTES on the current GPU (its address taken at this entry point's first
 * call on that GPU); the call allocates and frees nothing. Arguments and results as
 * ddh_solve_batch. A batch whose workspace (ddh_workspace_bytes for its lanes) exceeds the
 * scratch is refused with DDH_E_WORKSPACE. Every call on one GPU shares that GPU's scratch:
 * calls that overlap on one GPU (other streams, other host threads) race and must be
 * serialised by the caller. */
int ddh_solve_batch_noalloc(const ddh_batch *b, const double *V, int32_t nV, const ddh_opts *opts,
                            double *J, uint8_t *ok, int32_t *its, int32_t *status,
                            void *hip_stream);

/* Workspace bytes ddh_adjoint needs. */
size_t ddh_adjoint_workspace_bytes(int32_t n, int32_t NP, int32_t nV);

/* dJ(V)/dtheta at every converged voltage by the implicit function theorem: replaces
 * src/jaxdd/batch.py:606-737 (`_transposed`, `_converged_jvp`, `_tangent`: the transposed
 * bordered solve of reverse mode), free mobile ions included (each species' normalisation and
 * number constraint border the system; the normalisations of a converged state follow from
 * its potential in closed form), and ions frozen at a pre-bias (batch.py:937-960: the scan's
 * fixed charge is the pre-bias solution's ions, differentiated through that solution by a
 * second transposed bordered solve at it; pre_states from ddh_solve_batch_states). theta reaches the kernel as the tangents of the packed batch
 * arrays: dpt [ntheta][NP + 2][4][n], dseg [ntheta][nseg][23][n], dsc [ntheta][20][n]
 * (d array / d theta_k; every field the steady rows, the current and the ions' constraints
 * read, V_T and the generation's factor included). states and ok: from ddh_solve_batch_states
 * on the same batch and ascending V. grad: [n][nV][ntheta], of the current without the shunt
 * (the caller adds V d(1/R_shunt)/dtheta with the shunt's current); 0 where a voltage did not
 * converge (the reference's zero gradient there); NaN and status[d] = 1 where device d's
 * mobile ions are frozen at a pre-bias and pre_states is NULL. nseg <= DDH_ADJOINT_MAX_SEGMENTS; NP <= 2200 (each
 * device's arrays and tangents are staged in the local data share). Up to 16 directions take
 * one reverse pass over the parameters, more one forward pass each. bernoulli: as ddh_opts. */
int ddh_adjoint(const ddh_batch *b, const double *V, int32_t nV, const double *states,
                const double *pre_states, const uint8_t *ok, int32_t ntheta, const double *dpt,
                const double *dseg, const double *dsc, int32_t bernoulli, double *grad,
                int32_t *status, void *workspace, size_t workspace_bytes, void *hip_stream);

/* Workspace bytes ddh_adjoint_vjp needs. */
size_t ddh_adjoint_vjp_workspace_bytes(int32_t n, int32_t NP, int32_t nV, int32_t ntheta);

/* The vector-Jacobian product of the curves in the parameters, as the reference's `jax.grad`
 * of a loss of `solve_batch(device_tree(structure, params), V)` takes it (batch.py:142-148,
 * 650-737): theta_bar [n][ntheta], theta_bar[d][k] = sum_v dJbar[d][v] dJ(V_v)/dtheta_k, with
 * dJbar [n][nV] the loss's derivative in each curve's current. Arguments otherwise as
 * ddh_adjoint (its gradient kept in the workspace and contracted on the device); status and
 * the zero contribution of unconverged voltages as ddh_adjoint's (NaN for a device whose
 * mobile ions are frozen at a pre-bias, without pre_states). workspace: ddh_adjoint_vjp_workspace_bytes(n, NP, nV, ntheta) bytes. */
int ddh_adjoint_vjp(const ddh_batch *b, const double *V, int32_t nV, const double *states,
                    const double *pre_states, const uint8_t *ok, int32_t ntheta,
                    const double *dpt, const double *dseg, const double *dsc, int32_t bernoulli,
                    const double *dJbar, double *theta_bar, int32_t *status, void *workspace,
                    size_t workspace_bytes, void *hip_stream);

/* The batched block-tridiagonal sweep: replaces src/jaxdd/sweep.py:313-330 (`block_sweep`,
 * the custom_vmap operator over src/jaxdd/solver.py:1534-1560 `_block_thomas`). Solves
 * A[i] x[i-1] + B[i] x[i] + C[i] x[i+1] = d[i] for n systems of NP blocks b x b (b = 3 or 5)
 * with k <= 8 right-hand sides. Layout lane-innermost: A, B, C [NP][b*b][n]; d_inout
 * [NP][b*k][n], overwritten by x. precision: DDH_FP64 (double arrays) or DDH_MIXED (float
 * arrays). workspace: NP*(b*b + b*k)*n elements of the same type. */
int ddh_block_sweep(int32_t n, int32_t NP, int32_t b, int32_t k, const void *A, const void *B,
                    const void *C, void *d_inout, int32_t precision, void *workspace,
                    void *hip_stream);

/* The assembly operator at a given state: replaces src/jaxdd/solver.py:1021-1187
 * (`_assemble`, Boltzmann, no optional modules) with the ion slots' columns of
 * src/jaxdd/batch.py:518-549 (`_system`), and the integral current of src/jaxdd/batch.py:302-403
 * (`_current_parts`), for every device of the batch. y: [NP][3][n] the interior points'
 * (V, ln n, ln p); Vint: [n] the internal voltage (the contacts' potentials follow from it);
 * logs: [2][n] the anion and cation normalisations, or NULL for no mobile ions. Outputs
 * lane-innermost: F [NP][3][n] the row-scaled residual, A, B, C [NP][9][n] its Jacobian blocks
 * (row-major 3x3; the first row's A and the last row's C couple to the contacts, as the
 * reference returns them), U [NP][6][n] dF/d(logs) as [row][slot], J [n] the current without
 * the shunt. bernoulli: as ddh_opts. */
int ddh_assemble(const ddh_batch *b, const double *y, const double *Vint, const double *logs,
                 int32_t bernoulli, double *F, double *A, double *B, double *C, double *U,
                 double *J, void *hip_stream);

/* The 5x5 system with the mobile ions as unknowns, at a steady state: replaces
 * src/jaxdd/solver.py:1364-1473 (`_assemble_ions` with dti = 0, no sites, no gauge). y: the
 * interior state [NP][5][n] (V, ln n, ln p, ln a, ln c); Vint [n]; logs [2][n] the slots'
 * normalisations (one anion slot, one cation slot). Outputs (device memory): F [NP][5][n] the
 * scaled residual (the carriers' rows `_assemble`'s at each point's ion logarithms; each ion's
 * row ln a - la - V/V_T, ln c - lc + V/V_T in its region, its unknown outside), A, B, C
 * [NP][25][n] the 5x5 blocks row-major, U [NP][10][n] dF/d(logs) as [row][slot] (only the ions'
 * rows). bernoulli: as ddh_opts. */
int ddh_assemble_ions(const ddh_batch *b, const double *y, const double *Vint,
                      const double *logs, int32_t bernoulli, double *F, double *A, double *B,
                      double *C, double *U, void *hip_stream);

/* A batch's structures (what its numbers do not change), lane-innermost device arrays: each
 * point's layer lid [NP + 2][n] and segment sid [NP + 2][n] (a new segment wherever a field of
 * the packed arrays may change: a layer's first point, a contact, the last point of a layer
 * before an interface), each layer's flags [nl][n] (1 bulk traps, 2 mobile anions, 4 fixed
 * anions, 8 mobile cations, 16 fixed cations) and first point i0 [nl][n], each point's unscaled
 * spacing unit [NP + 2][n] (the interval after it) and its sum over each layer unit_sum
 * [nl][n]. nl layers (at most 8) and nseg segments (at most 14) for every device. */
typedef struct {
    int32_t n, NP, nl, nseg;
    const uint8_t *lid, *sid, *flags;
    const int32_t *i0;
    const double *unit, *unit_sum;
} ddh_structure;

/* The graded grid's layout of every device on the GPU: replaces src/jaxdd/device.py:160-214 (the
 * graded rule: the point density 1/H + beta/(d + d0) in the distance d to either end of a layer,
 * its integral shared among the layers, at least five points beyond each layer's first, each
 * layer's points equidistributing it) with :114-136 (`_layout`). L [nl][n] the thicknesses (m);
 * beta the points per e-fold, d0 (m) the distance below which it stops refining. Outputs the
 * ddh_structure arrays lid, sid, i0, unit, unit_sum (as ddh_build takes them), iface
 * [NP + 1][n] (1 where the interval after the point crosses from one layer to the next) and
 * nseg [n] each device's segment count (-1: its counts do not fill the grid). */
int ddh_graded_layout(int32_t n, int32_t NP, int32_t nl, const double *L, double beta, double d0,
                      uint8_t *lid, uint8_t *sid, int32_t *i0, double *unit, double *unit_sum,
                      uint8_t *iface, int32_t *nseg, void *hip_stream);

/* The batch's packed arrays from its structures and numbers on the GPU: replaces
 * src/jaxdd/device.py:877-1063 (`device_arrays`, the grid placed by `_spacing` :225-233) and
 * the packing of src/jaxdd/batch.py:151-223 (`prepare`), for the path's physics (Boltzmann
 * carriers, uniform generation per layer or a generation profile, ohmic contacts, one bulk trap
 * level, mobile and fixed ions' masks and totals, direct recombination by k_direct or Langevin's
 * rate, device.py:900-911).
 * layers [nl][23][n]: per layer L, eps_r, E_c, E_v, N_c, N_D, N_A, mu_n, mu_p, nu_int_n,
 * nu_int_p, k_direct, N_t_bulk, C_n_bulk, C_p_bulk, E_t_bulk, G_ehp, N_anion, N_cation, mu_anion,
 * mu_cation, preLangevin, bulkTrapType (the traps' charge type: 1 donor, -1 acceptor, 0
 * neutral; held, its direction ignored); structure->flags [nl][n] bits: 1 bulk traps, 2 mobile anions, 4 fixed
 * anions, 8 mobile cations, 16 fixed cations, 32 Langevin's rate preLangevin q (mu_n + mu_p) /
 * eps in place of k_direct, 64 that rate at the interval's mobilities (the interface's at a
 * layer's last point) instead of the layer's, 128 the layer generates (a profile's `layerGen`);
 * scalars [10][n]: T, W_L, W_R, G_frac, V_pre, 1/R_shunt (0: none), pre_cond, fix_ions,
 * left_elec, gen_profile (1: the device's generation is its profile, 0: each layer's G_ehp).
 * gen [2][gen_m][n]: each device's generation profile, positions (ascending, their last the
 * profile's length, stretched to the stack) then G (m^-3 s^-1) at G_frac = 1, a device's
 * shorter table padded by repeating its last node (device.py:804-814 `_generation`: NumPy's
 * interp, 0 at and beyond the last position, in generating layers only); gen_m 0 when no device
 * takes a profile, else >= 2. Outputs the ddh_batch arrays pt [NP + 2][4][n]
 * (h, x, G, fixed anions - cations), seg [nseg][23][n] (each segment's fields, a device's
 * segments beyond its own count repeating its last), sc [20][n], and ion_mu [nseg][2][n] (each
 * segment's anion and cation mobility). With K > 0 directions dlayers [K][nl][23][n] and
 * dscalars [K][10][n] (each a direction in the numbers), also their derivatives along each,
 * forward mode: dpt [K][NP + 2][4][n], dseg [K][nseg][23][n], dsc [K][20][n], the tangents
 * ddh_adjoint takes; the structure is held (the thicknesses scale each layer's spacing and
 * stretch the profile), as the reference differentiates `device_tree(structure, params)`
 * (batch.py:142-148). */
int ddh_build(const ddh_structure *structure, const double *layers, const double *scalars,
              const double *gen, int32_t gen_m, int32_t K, const double *dlayers,
              const double *dscalars, double *pt, double *seg, double *sc, double *ion_mu,
              double *dpt, double *dseg, double *dsc, void *hip_stream);

/* Device bytes of ddh_impedance's workspace (the per-device tables) for n devices of NP
 * interior points. */
size_t ddh_impedance_workspace_bytes(int32_t n, int32_t NP);

/* The admittance dJ/dV_int of every device at a steady state, per frequency: replaces
 * src/jaxdd/impedance.py:230-360 (`_admittance`) as `_batch_admittance` (:363-368) calls it for
 * devices without mobile ions (bulk traps one level, their occupations at their own rates; no
 * interface traps). y: [NP][3][n] the steady interior state (ddh_solve_batch_states at one
 * voltage); Vint [n] its internal voltage; iface [NP + 1][n] 1 where the interval after a point
 * crosses from one layer to the next; laps [nf][2] (re, im) of iw, or of implicit Euler's
 * (1 - e^(-iw dt))/dt; dt: implicit Euler's step for the traps' occupations (0: continuous, as
 * the reference's dt = None). Output Y [n][nf][2] (re, im) in SIMsalabim's sign;
 * status [n]: 0, or 1 for a device with mobile ions (Y NaN; the reference solves those with the
 * ions as unknowns). The external impedance is 1/(Y + G_shunt) + R_series (impedance.py:405).
 * workspace: ddh_impedance_workspace_bytes(n, NP) bytes of device memory. */
int ddh_impedance(const ddh_batch *b, const double *y, const double *Vint, const uint8_t *iface,
                  const double *laps, int32_t nf, double dt, int32_t bernoulli, double *Y,
                  int32_t *status, void *workspace, size_t workspace_bytes, void *hip_stream);

/* Device bytes of ddh_impedance_ions' workspace for n devices of NP interior points. */
size_t ddh_impedance_ions_workspace_bytes(int32_t n, int32_t NP);

/* ddh_impedance with the mobile ions' logarithms as unknowns of their own: replaces the branch
 * of src/jaxdd/impedance.py:242-262 (`_admittance` with `ions`: the 5x5 system of
 * `solver._assemble_ions` at a step of dti = 1 from the steady state with that step's time terms
 * taken back out, each species' number kept by its slot and its trapezoidal number as a bordered
 * constraint, the ions' currents in the linearised transient current), which the reference takes
 * for a whole batch when any device has mobile ions; a device without them gives the 3x3
 * result. The ions' densities are the steady state's closed-form Boltzmann profiles of the
 * potential (`solve_states`' anion and cation). ion_mu: [nseg][2][n] each segment's anion and
 * cation mobility. Arguments otherwise as ddh_impedance; status [n]: 0, or 2 for a device
 * outside this path (a species in several regions, or a region reaching a contact; Y NaN).
 * workspace: ddh_impedance_ions_workspace_bytes(n, NP) bytes of device memory. */
int ddh_impedance_ions(const ddh_batch *b, const double *ion_mu, const double *y,
                       const double *Vint, const uint8_t *iface, const double *laps, int32_t nf,
                       double dt, int32_t bernoulli, double *Y, int32_t *status, void *workspace,
                       size_t workspace_bytes, void *hip_stream);

/* Options of ddh_transient: src/jaxdd/transient.py:537-637 (`solve_transient`). */
typedef struct {
    double tol;            /* Newton step tolerance of each time step (1e-10) */
    int32_t max_iter;      /* Newton steps per time step (200) */
    int32_t bernoulli;     /* 0 exact B(x), 1 SIMsalabim's */
    int32_t zimt_current;  /* 1: ZimT's sums in the integral current (the interval after each
                              point, the left contact's change counted); 0: control volumes */
    int32_t failure_mode;  /* ZimT's failureMode: 0 stop at a failed step, 1 keep it, 2 skip it */
} ddh_topts;

/* Workspace bytes ddh_transient needs. */
size_t ddh_transient_workspace_bytes(int32_t n, int32_t NP);

/* Transients of a batch: replaces src/jaxdd/transient.py:467-480 (`_solve_batch`, the vmap of
 * `_solve_one` :270-464) for devices without an open circuit, a series resistance, interface
 * traps or ZimT's automatic stop (Track 0), each ion region inside the device. Every device
 * runs its own steps: t, V (internal voltage), G (generation fraction), [n][nt]; t[.][0] = 0
 * (a steady state) and a time 0 later is a steady state again. y0: [NP][3][n] each device's
 * steady state at its first voltage and generation fraction (ddh_solve_batch_states), from
 * which the first step starts. ion_mu: [nseg][2][n] the mobile anions' and cations' mobility of
 * each segment; iface: [NP + 1][n] 1 where the interval after a point crosses a layer
 * interface. ions: 1 (any device with mobile ions) solves 5 x 5 systems with the ions'
 * densities as unknowns, 0 the 3 x 3 ones. Outputs [n][nt]: J the internal current (A/m^2),
 * J_ext with the shunt, converged, Newton steps, kept (the steps ZimT keeps). status [n]: 0, or
 * 3 for a device outside this path (a pre-bias; ions reaching a contact), whose outputs are NaN
 * and 0. workspace: ddh_transient_workspace_bytes(n, NP) bytes of device memory. */
int ddh_transient(const ddh_batch *b, const double *ion_mu, const uint8_t *iface, const double *t,
                  const double *V, const double *G, int32_t nt, const double *y0,
                  const ddh_topts *opts, int32_t ions, double *J, double *J_ext, uint8_t *ok,
                  int32_t *its, uint8_t *kept, int32_t *status, void *workspace,
                  size_t workspace_bytes, void *hip_stream);

#ifdef __cplusplus
}
#endif
#endif
