/* CPU oracle: a plain-C restatement of the reference's explicit Godunov-type
 * cell update (Euler, MHD, HPR, RGFM), used ONLY as the checker by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg.  Nothing in the
 * product (vanguard_amd/, include/) links, loads or calls it.
 *
 * Every function restates the reference's arithmetic in the reference's
 * operation order (no FMA contraction, IEEE division and square root), so its
 * results are bit-identical to the compiled reference; each definition cites
 * the reference file:line it follows.  Pinning: see tests/test_oracle_*.py,
 * which compare these functions bit for bit with golden vectors written by the
 * compiled reference (tests/golden/) and with closed forms and invariants.
 *
 * Layout: primitive states are arrays of structures, nf doubles per cell, in
 * the reference's field order (Euler: rho, u, v, w, p, lambda), row-major with
 * x (column index j) contiguous, exactly as std::vector<std::vector<State>>.
 */
#ifndef VG_ORACLE_H
#define VG_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* EulerMaterialParameters (Euler/eulermaterialparameters.h:30-36) field order. */
typedef struct {
    double gamma, p_inf, cv, e_form, t_ign, k_rate;
} orc_euler_params;

/* Whole-solve restatement.  dim 1: nx cells (ny ignored, pass 1); dim 2: ny*nx.
 * order 1 = FORCE (FirstOrderSolver), 2 = SLIC (SecondOrderSolver).
 * bc 0 = transmissive (reference ghosts), 1 = periodic (not in the reference).
 * max_steps 0 = run while t < t_final (reference semantics).
 * dt_log (may be NULL) receives up to dt_cap time steps.  Returns steps taken. */
int orc_euler_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
                    int max_steps, double bias, int limiter, int bc, const orc_euler_params *mp,
                    const double *prim_in, double *prim_out, double *dt_log, int dt_cap);

/* Per-function entry points (same record layout as the reference harness's
 * euler_units mode: 92 doubles per record of five input states). */
void orc_euler_units(int n, double bias, int limiter, double dx, double ts, const orc_euler_params *mp,
                     const double *states, double *out);
void orc_limiter_units(int n, double bias, const double *r, double *out);

/* One sweep (for tests of single kernels): dir 0 = x, 1 = y. */
void orc_euler_sweep(int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
                     int limiter, int bc, const orc_euler_params *mp, const double *prim_in,
                     double *prim_out);
/* Maximum wave speed over the cells (no ghost corners). */
double orc_euler_max_wave_speed(int dim, int ny, int nx, const orc_euler_params *mp, const double *prim);

/* SLIC solve with prescribed time steps dts[0..n_steps), and (when tangent_in
 * is not NULL) its forward-mode derivative along tangent_in, by dual numbers
 * (euler_dual.cpp).  Limiter branches follow the primal comparisons; the 1e-5
 * slope clamp is a constant.  Same layout as orc_euler_solve. */
void orc_euler_fixed_dt_jvp(int dim, int ny, int nx, double dx, const double *dts, int n_steps, double bias,
                            int limiter, int bc, const orc_euler_params *mp, const double *prim_in,
                            const double *tangent_in, double *prim_out, double *tangent_out);
/* Hash of every branch decision (slope clamps, limiter cases, min selections)
 * of the last orc_euler_fixed_dt_jvp call: equal hashes, same smooth pieces. */
uint64_t orc_euler_branch_hash(void);

/* MHDMaterialParameters (MHD/mhdmaterialparameters.h) field order: adiabatic
 * index, stiffening parameter, GLM hyperbolic wave speed c_h, parabolic
 * damping c_p.  MHDMaterialParameters(gamma) sets the last three to 0. */
typedef struct {
    double gamma, p_inf, ch, cp;
} orc_mhd_params;

/* MHDSecondOrderSolver::solve/solve2D with `subcycles` RK4 GLM sub-cycles
 * (MHDForcingSolver) on each side of every step, and MHDFirstOrderSolver::solve/
 * solve2D (which neither sub-cycles nor sets c_h, so mp's c_h is used);
 * primitive order (rho, u, v, w, p, Bx, By, Bz, psi). */
int orc_mhd_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                  double bias, int limiter, int bc, int subcycles, const orc_mhd_params *mp, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap);
/* Per-function records of five input states (45 doubles) -> 85 doubles. */
void orc_mhd_sweep(int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias, int limiter, int bc,
                   const orc_mhd_params *mp, const double *prim_in, double *prim_out);
void orc_mhd_units(int n, double bias, int limiter, double dx, double ts, const orc_mhd_params *mp,
                   const double *states, double *out);

/* HPRMaterialParameters (HPR/hprmaterialparameters.h): the fields the solver
 * reads.  tau1/tau2 are the strain-dissipation and thermal relaxation times as
 * configureStrainDissipationTime / configureThermalImpulseRelaxationTime set
 * them (hprmaterialparameters.cpp:101-119).  eos: 0 StiffenedGas, 1
 * ShockMieGruneisen, 2 GodunovRomenski; p_ref, c0, gamma0, s_h, alpha_gr are
 * the reference pressure, reference sound speed, reference Grueneisen
 * coefficient, Hugoniot slope and alpha parameter those branches read; sigma_y
 * and n_pow the elastic-plastic transition parameter and power-law index of the
 * plastic relaxation (hprsourceterms.cpp:12-22). */
typedef struct {
    int32_t is_thermal, is_plastic;
    double rho0, t_ref, t0, cv, gamma, p_inf, cs, tau1, tau2, alpha;
    int32_t eos, reserved;
    double p_ref, c0, gamma0, s_h, alpha_gr, sigma_y, n_pow;
} orc_hpr_params;
/* HPRMieGruneisen records (rho, p, A: 11 doubles) -> e, p(rho, e), T, de/drho,
 * de/dp, dT/drho, dT/dp, theta1^-1 (8 doubles). */
void orc_hpr_eos_units(int n, const orc_hpr_params *mp, const double *states, double *out);

/* HPRSecondOrderSolver::solve / solve2D with `subcycles` RK4 source sub-cycles on
 * each side of the SLIC step.  Primitive order (rho, p, A row-major, u, v, w, J). */
int orc_hpr_solve(int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                  double bias, int limiter, int bc, int subcycles, const orc_hpr_params *mp, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap);
/* Per-function records of five states (85 doubles) -> 156 doubles. */
void orc_hpr_units(int n, double bias, int limiter, double dx, double ts, const orc_hpr_params *mp,
                   const double *states, double *out);
/* Dual-number restatement of the HPR second-order solve with prescribed time
 * steps (hpr_dual.cpp): primal and, with tangent_in, the derivative along
 * (tangent_in, dtau1, dtau2); sets the branch hash. */
void orc_hpr_fixed_dt_jvp(int dim, int ny, int nx, double dx, const double *dts, int n_steps, double bias, int limiter,
                          int bc, int subcycles, const orc_hpr_params *mp, const double *prim_in,
                          const double *tangent_in, double dtau1, double dtau2, double *prim_out, double *tangent_out);
/* Jacobian-vector products of p2c, c2p, Fx, Fy and S per record (5 x 17 doubles). */
void orc_hpr_jvp_units(int n, const orc_hpr_params *mp, double dtau1, double dtau2, const double *w,
                       const double *v, double *out);

/* RGFMSolver::solveHLLC / solveHLLC2D (riemann 0) and solveExact / solveExact2D
 * (riemann 1).  Two full-domain material grids (primitive AoS, 6 doubles per
 * cell) and the level set (1 double per cell).  dy is used by the Y sweep and
 * the Y level-set pass and min(dx, dy) by the time step; dy == dx is the
 * reference. */
int orc_rgfm_solve(int riemann, int dim, int ny, int nx, double dx, double dy, double cfl, double t_final,
                   int max_steps, double bias, int limiter, const orc_euler_params *p1, const orc_euler_params *p2,
                   const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out, double *m2_out,
                   double *phi_out, double *dt_log, int dt_cap);
/* applyRGFMBoundaryConditions{HLLC,Exact}[2D], in place. */
void orc_rgfm_ghost_fill(int riemann, int dim, int ny, int nx, const orc_euler_params *p1,
                         const orc_euler_params *p2, double *m1, double *m2, const double *phi);
/* updateLevelSetFunction (dim 1) / update2DLevelSetFunctionX (dir 0) / Y (dir 1). */
void orc_rgfm_level_set(int dim, int dir, int ny, int nx, double h, double dt, const double *m1, const double *m2,
                        const double *phi_in, double *phi_out);
/* HLLC records (12 doubles -> 24) and exact-solver records (12 -> 10: p*, u*,
 * rho*_a, rho*_b and the state sampled at x/t = s in direction dir). */
void orc_hllc_units(int n, const orc_euler_params *p1, const orc_euler_params *p2, const double *states,
                    double *out);
void orc_exact_units(int n, int dir, double s, const orc_euler_params *p1, const orc_euler_params *p2,
                     const double *states, double *out);
/* HLLC RGFM solve with prescribed time steps and (t1_in not NULL) the
 * forward-mode derivative of both final material states along (t1_in, t2_in),
 * by dual numbers (euler_dual.cpp); the level set is not differentiated.
 * orc_euler_branch_hash() afterwards hashes every branch decision taken. */
void orc_rgfm_fixed_dt_jvp(int dim, int ny, int nx, double dx, double dy, const double *dts, int n_steps, double bias,
                           int limiter, const orc_euler_params *p1, const orc_euler_params *p2, const double *m1_in,
                           const double *m2_in, const double *phi_in, const double *t1_in, const double *t2_in,
                           double *m1_out, double *m2_out, double *phi_out, double *t1_out, double *t2_out);
/* Interface circulation over phi < 0 from material 1's velocity. */
double orc_rgfm_circulation(int ny, int nx, double dx, double dy, const double *m1, const double *phi);

#ifdef __cplusplus
}
#endif
/* Two-material diffuse-interface Euler models (multiphysics_oracle.c).  kind 0:
 * EulerMultiphysicsStateVector, 8 primitives per cell (alpha1, u, v, w, rho1,
 * p1, rho2, p2); kind 1: EulerReducedStateVector, 7 (alpha1, u, v, w, rho1,
 * rho2, p).  order 2 reinitialises the volume fraction after every
 * reinit_every-th step (0: never); order 1 never does. */
int orc_mp_solve(int kind, int order, int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps,
                 double bias, int limiter, int bc, int reinit_every, const orc_euler_params *m1,
                 const orc_euler_params *m2, const double *prim_in, double *prim_out, double *dt_log, int dt_cap);
void orc_mp_sweep(int kind, int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
                  int limiter, int bc, const orc_euler_params *m1, const orc_euler_params *m2, const double *prim_in,
                  double *prim_out);
void orc_mp_reinit(int kind, int dim, int ny, int nx, double *prim);
void orc_mp_units(int kind, int n, const orc_euler_params *m1, const orc_euler_params *m2, double bias, int limiter,
                  double dx, double ts, const double *states, double *out);

/* Two-material ideal MHD with the real ghost fluid method (mhd_rgfm_oracle.c):
 * MHDRGFMSolver::solve (dim 1; the parameters' c_h as given, no sub-cycles)
 * and solve2D (dim 2; c_h = 4 and c_p = sqrt(0.18 c_h) from the first step,
 * `subcycles` RK4 GLM sub-cycles around each material's sweeps).  States are
 * 9 doubles per cell (rho, u, v, w, p, Bx, By, Bz, psi). */
int orc_mhd_rgfm_solve(int dim, int ny, int nx, double dx, double cfl, double t_final, int max_steps, double bias,
                       int limiter, int subcycles, const orc_mhd_params *p1, const orc_mhd_params *p2,
                       const double *m1_in, const double *m2_in, const double *phi_in, double *m1_out,
                       double *m2_out, double *phi_out, double *dt_log, int dt_cap);
void orc_mhd_rgfm_ghost_fill(int dim, int ny, int nx, const orc_mhd_params *p1, const orc_mhd_params *p2, double *m1,
                             double *m2, const double *phi);
void orc_mhd_hllc_solve(int d, const double *a, const double *b, const orc_mhd_params *m1,
                        const orc_mhd_params *m2, double *out);
void orc_mhd_hllc_units(int n, const orc_mhd_params *p1, const orc_mhd_params *p2, const double *states, double *out);

#endif
