/* vanguard.h -- C ABI of the MI355X-native explicit Godunov cell update.
 *
 * Drop-in replacement for the solve entry points of the reference's Solvers/
 * directory.  Plain C: pointers, sizes and PODs only.  All functions return a
 * status (VG_OK = 0, negative on error; vg_last_error() describes the last
 * failure on the calling thread).  No exception crosses this boundary.
 *
 * Host-pointer calls (prim_in / prim_out) are synchronous.  Primitive states are
 * arrays of structures in the reference's field order, row-major with x (the
 * column index j) contiguous -- bit-compatible with the reference's
 * std::vector<EulerStateVector> and std::vector<std::vector<EulerStateVector>>.
 *
 * Boundary conditions: VG_BC_TRANSMISSIVE reproduces the reference's ghost
 * cells (Solvers/solvers.cpp:7-32, 88-143); VG_BC_PERIODIC has no reference
 * counterpart.
 */
#ifndef VANGUARD_H
#define VANGUARD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    VG_OK = 0,
    VG_ERR_ARG = -1,        /* invalid argument (sizes, pointers, options) */
    VG_ERR_HIP = -2,        /* HIP runtime failure */
    VG_ERR_NONFINITE = -3,  /* the final state holds NaN/Inf or inadmissible cells (Euler: rho <= 0 or
                               p + p_inf <= 0; RGFM: in the material the level set makes real in
                               each cell, or a non-finite level set), or a time step could not
                               advance t; outputs are still written */
    VG_ERR_UNSUPPORTED = -4 /* option combination not implemented */
};

enum { VG_BC_TRANSMISSIVE = 0, VG_BC_PERIODIC = 1 };

/* VG_MODE_EXACT: the reference's operation order, bit-identical to its CPU
 * solver.  VG_MODE_FAST: algebraically reduced arithmetic (reciprocals reused,
 * no primitive/conserved round trips, multiply-adds fused), same scheme; not
 * bit-identical.  Over runs of hundreds of steps it agrees with EXACT to about
 * 1e-13 relative; over thousands, a last-bit difference can flip a limiter
 * branch or a level-set decision and grow (1e-3 relative and more seen on
 * batched MHD and RGFM runs), so use EXACT where the reference's results
 * matter. */
enum { VG_MODE_EXACT = 0, VG_MODE_FAST = 1 };

typedef struct vg_ctx vg_ctx;

/* EulerMaterialParameters (Euler/eulermaterialparameters.h:30-36), same field
 * order; EulerMaterialParameters(gamma) sets p_inf 0, cv 25, e_form 0,
 * T_ign 0.25, k_rate 250 (eulermaterialparameters.cpp:14-23). */
typedef struct {
    double gamma, p_inf, cv, e_form, t_ign, k_rate;
} vg_euler_params;

typedef struct {
    int32_t steps;          /* time steps taken by this call */
    int32_t total_steps;    /* steps since the state was loaded */
    double t;               /* simulated time reached */
    double dt_min, dt_max;  /* over this call's steps */
    double wall_ms;         /* host wall time of the call */
    double device_ms;       /* stream time of the step loop (HIP events) */
    int64_t nonfinite_cells;/* cells of the final state that are non-finite or inadmissible */
    int64_t first_nonfinite;/* flat index of the first such cell, -1 if none */
} vg_run_stats;

/* ---- context ------------------------------------------------------------ */

/* Creates a context on HIP device `device`: one stream, the scratch of the
 * solvers and the kernel-timing table.  One context per host thread. */
int vg_ctx_create(vg_ctx **ctx, int device);
void vg_ctx_destroy(vg_ctx *ctx);
/* Description of the last failure on this thread ("" if none). */
const char *vg_last_error(void);
/* Library version string. */
const char *vg_version(void);

/* Per-kernel timing: when enabled, every sweep launch is bracketed by HIP
 * events on the context's stream; totals are kept per kernel name. */
int vg_ctx_set_timing(vg_ctx *ctx, int enabled);
int vg_ctx_kernel_time(vg_ctx *ctx, const char *kernel, int64_t *launches, double *total_ms);
int vg_ctx_reset_timing(vg_ctx *ctx);

/* Step replay as HIP graphs (default on): runs capture a chunk of time steps
 * (all launches of each step, no memory operations) once per grid and replay
 * it until the clock stops.  Used only while per-kernel timing is off.  No
 * reference counterpart (the reference has no device). */
int vg_ctx_set_graphs(vg_ctx *ctx, int enabled);
/* Tile-shape autotuner (default on): the first run on a grid shape times the
 * candidate y-sweep shapes (threads per block x rows per strip) on that grid
 * and keeps the fastest for the context.  vg_ctx_tuned_shape returns the chosen
 * variant for a key "physics/order/mode/ny/nx" (forward Y sweeps, variants 0-7;
 * physics "euler", or "euler_planar" for grids swept with 4 fields)
 * or "euler_adjoint_x/ny/nx" / "euler_adjoint_y/ny/nx" (adjoint sweeps, variants
 * 0-5; "euler_planar_adjoint_..." for the 4-field adjoint sweeps; a 1D grid of n
 * cells has ny = 1, nx = n), -1 if not tuned
 * yet.  vg_ctx_set_tuned_shape pins a variant for a key (for example one saved
 * from an earlier process); -1 forgets it.  Shapes change speed, never results. */
int vg_ctx_set_autotune(vg_ctx *ctx, int enabled);
/* Constant-field specialisation (default on): an Euler grid whose out-of-plane
 * velocity is +0 and reaction progress exactly 1 in every cell (checked at the
 * start of each run) is swept with a 4-field kernel that carries rho, u, v, p
 * and leaves those two planes as they are -- the same bits as the 6-field
 * sweep, which keeps them at +0 and 1 exactly.  No reference counterpart. */
int vg_ctx_set_reduced_fields(vg_ctx *ctx, int enabled);
/* Streams a batch call may run its independent instances on at once (1..16,
 * default 4; 1 runs them one after another on the context's stream).  The 2D
 * RGFM batch (vg_rgfm_solve_batch_2d) hands instances, longest predicted run
 * first, to that many worker streams, each replaying its own captured steps,
 * so one instance's kernels fill the others' launch gaps.  The 1D batches
 * (vg_euler_solve_batch, vg_mhd_solve_batch) finish there the instances their
 * single launch stopped at its step cap or work budget.  Per-kernel timing
 * (vg_ctx_set_timing) runs a batch on one stream.  Results are the same bits
 * for every count.  No reference counterpart. */
int vg_ctx_set_batch_streams(vg_ctx *ctx, int streams);
int vg_ctx_tuned_shape(vg_ctx *ctx, const char *key, int *variant);
int vg_ctx_set_tuned_shape(vg_ctx *ctx, const char *key, int variant);
/* Frees the device workspace the context keeps between calls (the adjoint's
 * checkpoints and sweep inputs). */
int vg_ctx_release_workspace(vg_ctx *ctx);
/* Device memory the Euler adjoint may use for its recorded forward (bytes; 0,
 * the default: 85 % of the free device memory).  A smaller limit records fewer
 * steps' inner sweep inputs and recomputes the rest; results are the same
 * bits.  No reference counterpart. */
int vg_ctx_set_adjoint_memory(vg_ctx *ctx, int64_t bytes);

/* ---- Euler: whole-solve drop-ins ---------------------------------------- */

/* SecondOrderSolver::solve (Solvers/secondordersolver.h:55-56,
 * secondordersolver.cpp:288-314) when order == 2, FirstOrderSolver::solve
 * (firstordersolver.h:87-88, firstordersolver.cpp:435-461) when order == 1.
 * prim_in/prim_out: n cells x 6 doubles (may alias).  max_steps 0 runs while
 * t < t_final exactly as the reference; > 0 also stops after that many steps.
 * dt_log (may be NULL) receives up to dt_log_cap time steps.  The reference's
 * subcyclingIterations has no effect for Euler (an empty loop,
 * secondordersolver.cpp:302-305) and is therefore not an argument. */
int vg_euler_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                      double t_final, int max_steps, double bias, int limiter, int order, int bc,
                      const vg_euler_params *mp, int mode, double *dt_log, int dt_log_cap,
                      vg_run_stats *stats);

/* SecondOrderSolver::solve2D (secondordersolver.h:64-65, secondordersolver.cpp:392-424)
 * when order == 2, FirstOrderSolver::solve2D (firstordersolver.h:94-95,
 * firstordersolver.cpp:521-553) when order == 1: Strang split X(dt/2) Y(dt) X(dt/2).
 * prim_in/prim_out: ny rows x nx columns x 6 doubles. */
int vg_euler_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                      double cfl, double t_final, int max_steps, double bias, int limiter, int order,
                      int bc, const vg_euler_params *mp, int mode, double *dt_log, int dt_log_cap,
                      vg_run_stats *stats);

/* ---- Euler: device-resident state (stream-ordered) ----------------------- */

typedef struct vg_euler_state vg_euler_state;

/* A device-resident grid (dim 1: ny must be 1) held as structure-of-arrays
 * primitive planes, with its time-step clock.  Replaces the per-sweep calls
 * computeSLICTimeStep / computeX|YSLICTimeStep2D (secondordersolver.h:32-48)
 * and Solvers::computeStableTimeStep[2D] (solvers.h:31-45) for callers that keep
 * the grid on the device between steps. */
int vg_euler_state_create(vg_ctx *ctx, int dim, int ny, int nx, vg_euler_state **st);
void vg_euler_state_destroy(vg_euler_state *st);
/* Loads an array-of-structures primitive grid (host memory when on_device is 0,
 * device memory otherwise) and resets the clock to t = 0, step 0. */
int vg_euler_state_load(vg_euler_state *st, const double *prim, int on_device);
/* Writes the current grid as array-of-structures primitives. */
int vg_euler_state_store(vg_euler_state *st, double *prim, int on_device);
/* Device pointer of the current structure-of-arrays grid: 6 planes of
 * ny*nx doubles at a stride of *plane doubles.  The pointer, and writes through
 * it, are valid only until the next vg_euler_state_run, _sweep or _load on this
 * state (the grid alternates between two buffers): call this again before each
 * write.  Each call marks the planes as possibly modified, so the next run
 * checks again whether the 4-field sweeps apply (w = +0, lambda = 1). */
int vg_euler_state_planes(vg_euler_state *st, double **planes, int64_t *plane);
/* Advances up to max_steps steps (0 = until t >= t_final) continuing the clock. */
int vg_euler_state_run(vg_euler_state *st, int order, double dx, double cfl, double t_final, int max_steps,
                       double bias, int limiter, int bc, const vg_euler_params *mp, int mode,
                       double *dt_log, int dt_log_cap, vg_run_stats *stats);
/* Solvers::computeMaximumWaveSpeed[2D] (solvers.cpp:259-276, 318-339) over the
 * ghosted grid (2D transmissive: default-state corners included); NaN speeds
 * are skipped as in the reference.  Order-independent, hence deterministic. */
int vg_euler_state_max_wave_speed(vg_euler_state *st, int bc, const vg_euler_params *mp, double *smax);
/* One sweep of the device-resident grid with the caller's time step dt:
 * SecondOrderSolver::computeSLICTimeStep (1D, secondordersolver.h:32,
 * secondordersolver.cpp:80-98) or computeX|YSLICTimeStep2D (dir 0 / 1,
 * secondordersolver.h:39-48, secondordersolver.cpp:140-162, 214-236) when
 * order == 2; FirstOrderSolver::computeFORCETimeStep / computeX|YFORCETimeStep2D
 * (firstordersolver.h:63-81, firstordersolver.cpp:232-378) when order == 1.
 * Ghost cells come from `bc` (transmissive: the reference's insertBoundaryCells
 * mirror); the clock of vg_euler_state_run is not advanced. */
int vg_euler_state_sweep(vg_euler_state *st, int dir, int order, double dx, double dt, double bias, int limiter, int bc,
                         const vg_euler_params *mp, int mode);

/* The same per-sweep calls on host arrays, as the reference takes them:
 * `cells_with_boundary` is the ghosted snapshot the caller built (read as
 * given, corners included): n + 2G cells (1D) or (ny + 2G) x (nx + 2G) (2D),
 * G = 2 for SLIC (order 2), 1 for FORCE (order 1), 6 doubles per cell;
 * `cells` (n or ny x nx cells) receives the updated states, as the reference's
 * currentCells.  dir selects computeXSLICTimeStep2D (0) or computeYSLICTimeStep2D
 * (1).  Synchronous. */
int vg_euler_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                          double bias, int limiter, int order, const vg_euler_params *mp, int mode);
int vg_euler_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                          double dx, double dt, double bias, int limiter, int order, const vg_euler_params *mp,
                          int mode);

/* Batched 1D Euler solves in one persistent launch: n_inst independent
 * SecondOrderSolver::solve (order 2) / FirstOrderSolver::solve (order 1) calls
 * (secondordersolver.cpp:288-314, firstordersolver.cpp:435-461) with their own
 * sizes, spacings, material parameters and final times; transmissive boundaries.
 * offsets: n_inst + 1 cell offsets into prim_in/prim_out (6 doubles per cell).
 * steps/t_out (may be NULL) receive each instance's step count and final time. */
int vg_euler_solve_batch(vg_ctx *ctx, int n_inst, const int64_t *offsets, const double *prim_in, double *prim_out,
                         const double *dx, const vg_euler_params *mp, const double *t_final, double cfl,
                         int max_steps, double bias, int limiter, int order, int mode, int32_t *steps,
                         double *t_out, vg_run_stats *stats);

/* ---- Euler: adjoint (reverse-mode derivative) ---------------------------- */

/* Reverse-mode derivative of the SLIC solve (SecondOrderSolver::solve /
 * solve2D, EXACT arithmetic) with the time steps held at the values the forward
 * run chose.  No reference counterpart: the reference has no derivatives.
 * Runs the forward (prim_out receives its final state, dt_log its time steps),
 * then the adjoint backwards.  With max_steps > 0 the forward can write its
 * sweeps into recording buffers of the context's workspace: checkpoint_every 0
 * keeps every sweep input when they fit in free device memory (the reverse pass
 * recomputes nothing), else every step's start and, for as many of the last
 * steps as fit, their inner sweep inputs too (the reverse pass recomputes the
 * two inner sweeps of the other steps); checkpoint_every 1 keeps the step
 * starts only.  A grid swept with 4 fields records planes 0-4 only.  The
 * memory used is at most 85 % of the free device memory, or the context's
 * vg_ctx_set_adjoint_memory limit.  Otherwise a
 * checkpoint every `checkpoint_every` steps (0: ceil(sqrt(max_steps)), 16 for
 * runs to t_final) and recomputation per segment.  Every choice gives the same
 * bits.  seed: the
 * derivative of a scalar loss with respect to the final primitive state
 * (array of structures, 6 doubles per cell); grad receives its derivative with
 * respect to the initial primitive state.  The slope limiter is differentiated
 * piecewise: at a kink, the derivative of the branch the forward comparison
 * selected; the 1e-5 slope clamp counts as a constant. */
int vg_euler_adjoint_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad, int n,
                        double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
                        const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                        vg_run_stats *stats);
int vg_euler_adjoint_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad, int ny,
                        int nx, double dx, double cfl, double t_final, int max_steps, double bias, int limiter, int bc,
                        const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                        vg_run_stats *stats);
/* The same with prim_in, prim_out, seed and grad in device memory (same
 * array-of-structures layout); dt_log and stats stay on the host.  Checkpoints
 * and sweep inputs live in the context's workspace, kept between calls
 * (vg_ctx_release_workspace frees it); checkpoint_every chooses as above. */
int vg_euler_adjoint_1d_device(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                               int n, double dx, double cfl, double t_final, int max_steps, double bias, int limiter,
                               int bc, const vg_euler_params *mp, int checkpoint_every, double *dt_log, int dt_log_cap,
                               vg_run_stats *stats);
int vg_euler_adjoint_2d_device(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                               int ny, int nx, double dx, double cfl, double t_final, int max_steps, double bias,
                               int limiter, int bc, const vg_euler_params *mp, int checkpoint_every, double *dt_log,
                               int dt_log_cap, vg_run_stats *stats);

/* The SLIC solve with prescribed time steps dts[0..n_steps) (EXACT arithmetic):
 * the map the adjoint differentiates, for frozen-dt finite-difference checks. */
int vg_euler_solve_fixed_dt_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx,
                               const double *dts, int n_steps, double bias, int limiter, int bc,
                               const vg_euler_params *mp);
int vg_euler_solve_fixed_dt_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                               const double *dts, int n_steps, double bias, int limiter, int bc,
                               const vg_euler_params *mp);

/* ---- ideal MHD with GLM cleaning ----------------------------------------- */

/* MHDMaterialParameters (MHD/mhdmaterialparameters.h) field order: gamma, p_inf,
 * GLM hyperbolic speed c_h, parabolic damping c_p.  As in the reference the
 * second-order solvers set c_h to each step's maximum wave speed and
 * c_p = sqrt(0.18 c_h) (mhdsecondordersolver.cpp:330-332); the first-order
 * solvers use the c_h given here (MHDMaterialParameters(gamma) has c_h = 0). */
typedef struct {
    double gamma, p_inf, ch, cp;
} vg_mhd_params;

/* MHDSecondOrderSolver::solve (mhdsecondordersolver.h:58, mhdsecondordersolver.cpp:318-360)
 * when order == 2, with `subcycles` RK4 GLM-damping sub-cycles
 * (MHDForcingSolver::computeRungeKuttaTimeStep, mhdforcingsolver.cpp:7-44,
 * 341-356) before and after every SLIC step; MHDFirstOrderSolver::solve
 * (mhdfirstordersolver.cpp:457-477) when order == 1, which neither sub-cycles
 * nor sets c_h (the c_h, c_p given here are its GLM constants).  Primitive
 * order (rho, u, v, w, p, Bx, By, Bz, psi), 9 doubles per cell.  Without
 * sub-cycles it runs as one persistent workgroup. */
int vg_mhd_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_mhd_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats);

/* MHDSecondOrderSolver::solve2D (mhdsecondordersolver.h:70, mhdsecondordersolver.cpp:501-546,
 * sub-cycles by computeRungeKuttaTimeStep2D, mhdforcingsolver.cpp:134-194, 394-417)
 * when order == 2, MHDFirstOrderSolver::solve2D (mhdfirstordersolver.cpp:552-578)
 * when order == 1. */
int vg_mhd_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_mhd_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats);

/* n_inst independent MHDSecondOrderSolver::solve calls (transmissive
 * boundaries, zero sub-cycles) in one persistent launch; each instance has its
 * own size, spacing, parameters, final time, dt sequence and c_h.  Returns
 * VG_ERR_NONFINITE if an instance had to stop because its time step could not
 * advance t (where the reference would loop forever). */
int vg_mhd_solve_batch(vg_ctx *ctx, int n_inst, const int64_t *offsets, const double *prim_in, double *prim_out,
                       const double *dx, const vg_mhd_params *mp, const double *t_final, double cfl, int max_steps,
                       double bias, int limiter, int order, int mode, int32_t *steps, double *t_out,
                       vg_run_stats *stats);

/* vg_euler_time_step_2d without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * `cells` and `cells_with_boundary` are DEVICE arrays (same layouts as
 * vg_euler_time_step_2d); the intermediate planes and the clock live in a fixed
 * device array of the library (96 MiB per device, held until the process
 * ends), so (ny + 2G) * (nx + 2G) must not exceed 1048576 cells (VG_ERR_ARG
 * otherwise).  Calls that overlap on one device share that array and race.
 * Stream-ordered on the context's stream; returns after the work completes
 * unless the stream is being captured. */
int vg_euler_time_step_2d_noalloc(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx,
                                  int dir, double dx, double dt, double bias, int limiter, int order,
                                  const vg_euler_params *mp, int mode);

/* MHDSecondOrderSolver::computeSLICTimeStep / computeX|YSLICTimeStep2D
 * (mhdsecondordersolver.h:33-56, mhdsecondordersolver.cpp:91-108) when order
 * == 2 and MHDFirstOrderSolver's FORCE twins when order == 1, on host arrays as
 * vg_euler_time_step_*: 9 doubles per cell.  The GLM constants c_h, c_p are
 * used as given (the reference's solve sets them per step before calling these,
 * mhdsecondordersolver.cpp:330-332). */
int vg_mhd_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                        double bias, int limiter, int order, const vg_mhd_params *mp, int mode);
int vg_mhd_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                        double dx, double dt, double bias, int limiter, int order, const vg_mhd_params *mp, int mode);

/* ---- two-material Euler: diffuse-interface models ------------------------ */

/* The reference's EulerMultiphysicsStateVector (Euler/Multiphysics/
 * eulermultiphysicsstatevector.h:8-76): 8 primitives per cell, in the order of
 * its primitive vector (alpha1, u, v, w, rho1, p1, rho2, p2); and
 * EulerReducedStateVector (eulerreducedstatevector.h:8-71): 7 (alpha1, u, v, w,
 * rho1, rho2, p).  p1 / p2 are the two EulerMaterialParameters.  Both compute
 * in EXACT arithmetic, bit-identical to the reference. */

/* SecondOrderSolver::solve / solve2D for these states
 * (Solvers/secondordersolver.h:57-70, secondordersolver.cpp:316-517) when order
 * == 2, with MultiphysicsSolvers::reinitialiseVolumeFraction after every
 * reinit_every-th step (0: never); FirstOrderSolver::solve / solve2D
 * (firstordersolver.h:89-100, firstordersolver.cpp:463-640) when order == 1,
 * which never reinitialises.  The remaining arguments are vg_euler_solve_*'s. */
int vg_euler_multiphysics_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                                   double t_final, int max_steps, double bias, int limiter, int order, int bc,
                                   int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2,
                                   double *dt_log, int dt_log_cap, vg_run_stats *stats);
int vg_euler_multiphysics_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                                   double cfl, double t_final, int max_steps, double bias, int limiter, int order,
                                   int bc, int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2,
                                   double *dt_log, int dt_log_cap, vg_run_stats *stats);
int vg_euler_reduced_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                              double t_final, int max_steps, double bias, int limiter, int order, int bc,
                              int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2, double *dt_log,
                              int dt_log_cap, vg_run_stats *stats);
int vg_euler_reduced_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                              double cfl, double t_final, int max_steps, double bias, int limiter, int order, int bc,
                              int reinit_every, const vg_euler_params *p1, const vg_euler_params *p2, double *dt_log,
                              int dt_log_cap, vg_run_stats *stats);

/* computeSLICTimeStep / computeX|YSLICTimeStep2D for these states
 * (secondordersolver.h:34-36, 41-44, 49-52) when order == 2 and the FORCE twins
 * (firstordersolver.h:65-82) when order == 1, on host arrays with the caller's
 * ghosted snapshot, as vg_euler_time_step_*. */
int vg_euler_multiphysics_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n,
                                       double dx, double dt, double bias, int limiter, int order,
                                       const vg_euler_params *p1, const vg_euler_params *p2);
int vg_euler_multiphysics_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx,
                                       int dir, double dx, double dt, double bias, int limiter, int order,
                                       const vg_euler_params *p1, const vg_euler_params *p2);
int vg_euler_reduced_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx,
                                  double dt, double bias, int limiter, int order, const vg_euler_params *p1,
                                  const vg_euler_params *p2);
int vg_euler_reduced_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx,
                                  int dir, double dx, double dt, double bias, int limiter, int order,
                                  const vg_euler_params *p1, const vg_euler_params *p2);

/* MultiphysicsSolvers::reinitialiseVolumeFraction (multiphysicssolvers.h:12-17)
 * on a host array in place: 1D resets the four cells around the last falling
 * crossing of alpha1 = 0.5, 2D every crossing of every row, then of every
 * column.  Where the reference would index outside the grid (a crossing within
 * two cells of an edge, or no crossing in 1D) the outside cells are skipped. */
int vg_euler_multiphysics_reinitialise(vg_ctx *ctx, double *prim, int dim, int ny, int nx);
int vg_euler_reduced_reinitialise(vg_ctx *ctx, double *prim, int dim, int ny, int nx);

/* ---- two-material Euler: real ghost fluid method (RGFM) ----------------- */

enum {
    VG_RIEMANN_HLLC = 0,  /* HLLCSolver (RGFM/hllcsolver.cpp): RGFMSolver::solveHLLC[2D] */
    VG_RIEMANN_EXACT = 1  /* ExactSolver (RGFM/exactsolver.cpp): RGFMSolver::solveExact[2D] */
};

/* RGFMSolver::solveHLLC / solveExact (rgfmsolver.h:30-36, rgfmsolver.cpp:648-687, 745-784),
 * zero sub-cycles and no reinitialisation, as the reference runs them.
 * MultimaterialSystem (multimaterialsystem.h:8-39) is passed as its three
 * arrays: material 1 and material 2 primitive states over the whole domain
 * (n x 6 doubles each, EulerStateVector order) and the level set (n doubles).
 * The level set's sign changes mark the interface; the material flag starts at
 * material 1 on the left. */
int vg_rgfm_solve_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                     double *m2_out, double *phi_out, int n, double dx, double cfl, double t_final, int max_steps,
                     double bias, int limiter, int riemann, const vg_euler_params *p1, const vg_euler_params *p2,
                     int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats);

/* RGFMSolver::solveHLLC2D / solveExact2D (rgfmsolver.h:33-38, rgfmsolver.cpp:689-743, 786-840):
 * ny x nx grids, rows contiguous; each row and column starts in material 2 (the
 * reference's 2D flag starts "in the ghost region").  dy is the y spacing
 * (dy == dx is the reference; dy != dx uses dy in the y sweeps and the y
 * level-set pass and min(dx, dy) in the time step).  `circulation` (may be NULL)
 * receives sum over interior cells with phi < 0 of (dv/dx - du/dy) dx dy of
 * material 1's final velocity (central differences). */
int vg_rgfm_solve_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                     double *m2_out, double *phi_out, int ny, int nx, double dx, double dy, double cfl,
                     double t_final, int max_steps, double bias, int limiter, int riemann,
                     const vg_euler_params *p1, const vg_euler_params *p2, int mode, double *dt_log, int dt_log_cap,
                     vg_run_stats *stats, double *circulation);

/* n_inst independent vg_rgfm_solve_2d calls on grids of one shape (states
 * concatenated instance by instance), each with its own final time; per-instance
 * steps, final times and circulations out (each may be NULL). */
int vg_rgfm_solve_batch_2d(vg_ctx *ctx, int n_inst, const double *m1_in, const double *m2_in, const double *phi_in,
                           double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double dy,
                           double cfl, const double *t_final, int max_steps, double bias, int limiter, int riemann,
                           const vg_euler_params *p1, const vg_euler_params *p2, int mode, int32_t *steps,
                           double *t_out, double *circulation, vg_run_stats *stats);

/* Reverse-mode derivative of the HLLC RGFM solve (RGFMSolver::solveHLLC /
 * solveHLLC2D, EXACT arithmetic) with the time steps held at the values the
 * forward run chose.  No reference counterpart: the reference has no
 * derivatives.  Runs the forward (outputs as vg_rgfm_solve_*, dt_log its
 * steps), checkpointing both materials and the level set every
 * `checkpoint_every` steps (0: ceil(sqrt(max_steps)), 64 for runs to t_final),
 * then the adjoint backwards: per step the SLIC sweeps of both materials and
 * the ghost fills.  seed1/seed2: derivative of a scalar loss with respect to
 * the final states of materials 1 and 2 (array of structures, 6 doubles per
 * cell); grad1/grad2 receive its derivative with respect to their initial
 * states.  The level set only decides where the ghost fills write and carries
 * no derivative; the slope limiter is differentiated piecewise as in
 * vg_euler_adjoint_*, the HLLC interface state on the branch its comparisons
 * select. */
int vg_rgfm_adjoint_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, const double *seed1, const double *seed2, double *grad1,
                       double *grad2, int n, double dx, double cfl, double t_final, int max_steps, double bias,
                       int limiter, const vg_euler_params *p1, const vg_euler_params *p2, int checkpoint_every,
                       double *dt_log, int dt_log_cap, vg_run_stats *stats);
int vg_rgfm_adjoint_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, const double *seed1, const double *seed2, double *grad1,
                       double *grad2, int ny, int nx, double dx, double dy, double cfl, double t_final, int max_steps,
                       double bias, int limiter, const vg_euler_params *p1, const vg_euler_params *p2,
                       int checkpoint_every, double *dt_log, int dt_log_cap, vg_run_stats *stats);
/* The adjoints of a batch of independent 2D instances (arrays concatenated
 * instance by instance, as vg_rgfm_solve_batch_2d; each with its own final
 * time and seeds), run like vg_rgfm_adjoint_2d on the context's worker streams
 * (vg_ctx_set_batch_streams), longest predicted run first; each worker gets an
 * equal share of the free device memory for its checkpoints, with fewer workers
 * than streams when the longest predicted run would not fit in a share.  The
 * workers' device memory is released before the call returns.  Results are the
 * same bits as one vg_rgfm_adjoint_2d call per instance.  steps (may be NULL):
 * each instance's step count; status (may be NULL): each instance's status.
 * An instance without a derivative (an inadmissible forward, an overflowing
 * adjoint) gets VG_ERR_NONFINITE in its status while the others run on; the
 * call then returns VG_ERR_NONFINITE.  No reference counterpart. */
int vg_rgfm_adjoint_batch_2d(vg_ctx *ctx, int n_inst, const double *m1_in, const double *m2_in, const double *phi_in,
                             double *m1_out, double *m2_out, double *phi_out, const double *seed1,
                             const double *seed2, double *grad1, double *grad2, int ny, int nx, double dx, double dy,
                             double cfl, const double *t_final, int max_steps, double bias, int limiter,
                             const vg_euler_params *p1, const vg_euler_params *p2, int checkpoint_every,
                             int32_t *steps, int32_t *status, vg_run_stats *stats);
/* The HLLC RGFM solve with prescribed time steps dts[0..n_steps) (EXACT
 * arithmetic): the map the adjoint differentiates, for finite-difference checks. */
int vg_rgfm_solve_fixed_dt_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                              double *m1_out, double *m2_out, double *phi_out, int n, double dx, const double *dts,
                              int n_steps, double bias, int limiter, const vg_euler_params *p1,
                              const vg_euler_params *p2);
int vg_rgfm_solve_fixed_dt_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                              double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double dy,
                              const double *dts, int n_steps, double bias, int limiter, const vg_euler_params *p1,
                              const vg_euler_params *p2);

/* Self-check of the EXACT mode's division sequences; no counterpart in the
 * reference.  For n operand pairs (host arrays a, b) computes on the device the
 * shared-reciprocal quotient of the conversions (shared_q), the short division
 * sequence of the slope limiters (in_range_q) and the compiler's IEEE division
 * (ieee_q), so that a test can show them equal over the range the solvers use
 * them in (vanguard_amd/csrc/vg_quot.h).  n <= 2^26. */
int vg_division_selfcheck(vg_ctx *ctx, int64_t n, const double *a, const double *b, double *shared_q,
                          double *in_range_q, double *ieee_q);

/* ---- two-material ideal MHD: real ghost fluid method ------------------- */

/* MHDRGFMSolver::solve (mhdrgfmsolver.h:27-28, mhdrgfmsolver.cpp:543-581) and
 * solve2D (:29-30, :584-693).  MHDMultimaterialSystem is passed as its three
 * arrays: both materials' primitive states (rho, u, v, w, p, Bx, By, Bz, psi;
 * 9 doubles per cell, rows contiguous in 2D) and the level set.  Interface
 * states by MHDHLLCSolver (RGFM/mhdhllcsolver.cpp), ghost pressures floored at
 * 1e-8; dt = the smaller of the two materials' stable steps.  The 1D solve
 * uses the caller's c_h, c_p and ignores `subcycles`, as the reference does;
 * solve2D sets c_h = 4, c_p = sqrt(0.18 c_h) after the first step's ghost fill
 * and level set, and runs `subcycles` RK4 GLM sub-cycles before and after each
 * material's X(dt/2) Y(dt) X(dt/2) sweeps.  The reference's
 * reinitialisationFrequency is unused there and has no argument here.
 * Transmissive boundaries; lines of at most 16384 cells. */
int vg_mhd_rgfm_solve_1d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                         double *m1_out, double *m2_out, double *phi_out, int n, double dx, double cfl,
                         double t_final, int max_steps, double bias, int limiter, int subcycles,
                         const vg_mhd_params *p1, const vg_mhd_params *p2, int mode, double *dt_log, int dt_log_cap,
                         vg_run_stats *stats);
int vg_mhd_rgfm_solve_2d(vg_ctx *ctx, const double *m1_in, const double *m2_in, const double *phi_in,
                         double *m1_out, double *m2_out, double *phi_out, int ny, int nx, double dx, double cfl,
                         double t_final, int max_steps, double bias, int limiter, int subcycles,
                         const vg_mhd_params *p1, const vg_mhd_params *p2, int mode, double *dt_log, int dt_log_cap,
                         vg_run_stats *stats);

/* MHDRGFMSolver::applyRGFMBoundaryConditions (dim 1, mhdrgfmsolver.cpp:7-127:
 * psi left alone) and applyRGFMBoundaryConditions2D (dim 2, :129-378: rows,
 * then columns, psi written), in place on host arrays m1, m2 (ny x nx x 9; ny
 * ignored in 1D). */
int vg_mhd_rgfm_ghost_fill(vg_ctx *ctx, int dim, int ny, int nx, double *m1, double *m2, const double *phi,
                           const vg_mhd_params *p1, const vg_mhd_params *p2);

/* ---- Godunov-Peshkov-Romenski (HPR) continuum ---------------------------- */

enum {
    VG_HPR_EOS_STIFFENED_GAS = 0,       /* "StiffenedGas" */
    VG_HPR_EOS_SHOCK_MIE_GRUNEISEN = 1, /* "ShockMieGruneisen" */
    VG_HPR_EOS_GODUNOV_ROMENSKI = 2     /* "GodunovRomenski" */
};

/* HPRMaterialParameters (HPR/hprmaterialparameters.h:10-15), the full
 * constructor argument list in its order, the EOS name as an enum.  Every
 * path reads eos, is_thermal, is_plastic, reference_density,
 * reference_temperature, initial_temperature, specific_heat_capacity,
 * adiabatic_index, transverse_wave_speed, strain_dissipation_time,
 * thermal_impulse_relaxation_time and heat_wave_speed; the stiffened gas adds
 * stiffening_parameter; ShockMieGruneisen reference_pressure,
 * reference_sound_speed, reference_gruneisen_coefficient and
 * hugoniot_slope_coefficient; GodunovRomenski reference_sound_speed and
 * alpha_parameter (its Grueneisen coefficient is adiabatic_index); the plastic
 * relaxation (is_plastic != 0) elastic_plastic_transition_parameter (> 0) and
 * power_law_index (hprsourceterms.cpp:7-39).  The adjoints take the stiffened
 * gas without plastic relaxation only (VG_ERR_UNSUPPORTED otherwise).  As in the reference,
 * the two relaxation times are taken as given (the reference's
 * configureStrainDissipationTime / configureThermalImpulseRelaxationTime,
 * hprmaterialparameters.cpp:101-119, are helpers the caller applies). */
typedef struct {
    int32_t eos, is_thermal, is_plastic, reserved;
    double ideal_gas_constant, reference_density, reference_pressure, reference_temperature, initial_temperature,
        specific_heat_capacity, adiabatic_index, stiffening_parameter, reference_sound_speed,
        reference_gruneisen_coefficient, hugoniot_slope_coefficient, reference_internal_energy, transverse_wave_speed,
        strain_dissipation_time, thermal_impulse_relaxation_time, dynamic_viscosity_coefficient, prandtl_number,
        elastic_plastic_transition_parameter, power_law_index, heat_wave_speed, alpha_parameter, beta_parameter,
        gamma_parameter;
} vg_hpr_params;

/* HPRSecondOrderSolver::solve (hprsecondordersolver.h:56, hprsecondordersolver.cpp:314-352)
 * when order == 2, HPRFirstOrderSolver::solve when order == 1.  Each step:
 * `subcycles` RK4 relaxation sub-cycles over dt/2, the sweep, `subcycles` more
 * (HPRForcingSolver::computeRungeKuttaTimeStep, hprforcingsolver.cpp:341-356).
 * Primitive order (rho, p, A[3][3] row-major, u, v, w, J[3]), 17 doubles per
 * cell; J is carried unchanged when is_thermal is 0, as in the reference. */
int vg_hpr_solve_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats);

/* HPRSecondOrderSolver::solve2D (hprsecondordersolver.h:67, hprsecondordersolver.cpp:502-546):
 * source(dt/2), X(dt/2) Y(dt) X(dt/2), source(dt/2). */
int vg_hpr_solve_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx, double cfl,
                    double t_final, int max_steps, double bias, int limiter, int order, int bc, int subcycles,
                    const vg_hpr_params *mp, int mode, double *dt_log, int dt_log_cap, vg_run_stats *stats);

/* Device-resident HPR grid (as vg_euler_state). */
typedef struct vg_hpr_state vg_hpr_state;
int vg_hpr_state_create(vg_ctx *ctx, int dim, int ny, int nx, vg_hpr_state **st);
void vg_hpr_state_destroy(vg_hpr_state *st);
int vg_hpr_state_load(vg_hpr_state *st, const double *prim, int on_device);
int vg_hpr_state_store(vg_hpr_state *st, double *prim, int on_device);
int vg_hpr_state_planes(vg_hpr_state *st, double **planes, int64_t *plane);
int vg_hpr_state_run(vg_hpr_state *st, int order, double dx, double cfl, double t_final, int max_steps, double bias,
                     int limiter, int bc, int subcycles, const vg_hpr_params *mp, int mode, double *dt_log,
                     int dt_log_cap, vg_run_stats *stats);
/* Maximum wave speed (HPRSolvers::computeMaximumWaveSpeed[2D], hprsolvers.cpp:286-303,
 * 362-385) over the ghosted grid. */
int vg_hpr_state_max_wave_speed(vg_hpr_state *st, int bc, const vg_hpr_params *mp, double *smax);
/* One sweep with the caller's time step (HPRSecondOrderSolver::computeSLICTimeStep /
 * computeX|YSLICTimeStep2D, hprsecondordersolver.h:33-54, hprsecondordersolver.cpp:89-241,
 * when order == 2; HPRFirstOrderSolver's FORCE twins when order == 1), as
 * vg_euler_state_sweep; no relaxation source. */
int vg_hpr_state_sweep(vg_hpr_state *st, int dir, int order, double dx, double dt, double bias, int limiter, int bc,
                       const vg_hpr_params *mp);
/* The same on host arrays with the caller's ghosted snapshot, as
 * vg_euler_time_step_*: 17 doubles per cell. */
int vg_hpr_time_step_1d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int n, double dx, double dt,
                        double bias, int limiter, int order, const vg_hpr_params *mp);
int vg_hpr_time_step_2d(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny, int nx, int dir,
                        double dx, double dt, double bias, int limiter, int order, const vg_hpr_params *mp);

/* Reverse-mode derivative of the second-order HPR solve (as vg_hpr_solve_*
 * with order 2, EXACT arithmetic; stiffened gas, not plastic) with the time
 * steps held at the values the forward run chose.  No reference counterpart:
 * the reference has no derivatives.  Runs the forward (outputs as
 * vg_hpr_solve_*, dt_log its steps), checkpointing every `checkpoint_every`
 * steps (0: ceil(sqrt(max_steps)), 16 for runs to t_final), then the adjoint
 * of every operation backwards: the RK4 relaxation sub-cycles
 * (HPRForcingSolver::computeRungeKuttaTimeStep[2D], hprforcingsolver.cpp:341-356,
 * 394-417) and the SLIC sweeps (hprsecondordersolver.cpp:7-17, 89-241).  seed:
 * derivative of a scalar loss with respect to the final primitive state (17
 * doubles per cell); grad receives its derivative with respect to the initial
 * primitive state, grad_tau[0..1] (may be NULL) with respect to
 * strain_dissipation_time and thermal_impulse_relaxation_time.  The slope
 * limiter is differentiated piecewise as in vg_euler_adjoint_*.  Needs at
 * least 3 cells per direction. */
int vg_hpr_adjoint_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                      double *grad_tau, int n, double dx, double cfl, double t_final, int max_steps, double bias,
                      int limiter, int bc, int subcycles, const vg_hpr_params *mp, int checkpoint_every,
                      double *dt_log, int dt_log_cap, vg_run_stats *stats);
int vg_hpr_adjoint_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, const double *seed, double *grad,
                      double *grad_tau, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                      double bias, int limiter, int bc, int subcycles, const vg_hpr_params *mp, int checkpoint_every,
                      double *dt_log, int dt_log_cap, vg_run_stats *stats);
/* The second-order HPR solve with prescribed time steps dts[0..n_steps): the
 * map the adjoint differentiates, for finite-difference checks. */
int vg_hpr_solve_fixed_dt_1d(vg_ctx *ctx, const double *prim_in, double *prim_out, int n, double dx, const double *dts,
                             int n_steps, double bias, int limiter, int bc, int subcycles, const vg_hpr_params *mp);
int vg_hpr_solve_fixed_dt_2d(vg_ctx *ctx, const double *prim_in, double *prim_out, int ny, int nx, double dx,
                             const double *dts, int n_steps, double bias, int limiter, int bc, int subcycles,
                             const vg_hpr_params *mp);

#ifdef __cplusplus
}
#endif

#endif /* VANGUARD_H */
