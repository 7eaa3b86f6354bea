/* Generic SLIC / FORCE scheme of the oracle, over a physics table.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). */
#ifndef VG_SCHEME_ORACLE_H
#define VG_SCHEME_ORACLE_H

#define ORC_MAX_NF 17

typedef struct orc_phys orc_phys;
struct orc_phys {
    int nf;
    /* primitive <-> conserved, in the reference's operation order */
    void (*prim_to_cons)(const orc_phys *ph, const double *w, double *U);
    void (*cons_to_prim)(const orc_phys *ph, const double *U, double *w);
    /* flux of a conserved vector, dir 0 = x, 1 = y */
    void (*flux)(const orc_phys *ph, int dir, const double *U, double *F);
    /* per-cell wave speed entering the CFL maximum (dim 1 or 2) */
    double (*wave_speed)(const orc_phys *ph, const double *w, int dim);
    /* the state's default constructor (2D corner ghosts) */
    const double *default_state;
    /* called once per second-order step with the step's maximum wave speed
     * (MHD sets c_h; the first-order solvers never do) */
    void (*begin_step)(orc_phys *ph, double smax);
    /* source term S(U) of the RK4 sub-cycles (NULL: no source stage) */
    void (*source)(const orc_phys *ph, const double *U, double *S);
    /* after a sub-cycle, per cell: fix-ups of the new primitive state w_out
     * given the old one w_in (HPR without heat conduction keeps its own J) */
    void (*forcing_finish)(const orc_phys *ph, const double *w_in, double *w_out);
    void *params;
    /* primitive fields per cell when fewer than the conserved ones (0: nf);
     * the two-material Euler models carry 8 or 7 primitives for 9 or 8
     * conserved variables */
    int np;
    /* called after each step `it` (counted from 1) of a solve with the new
     * primitive grid (NULL: nothing); the two-material models reinitialise their
     * volume fraction here */
    void (*after_step)(const orc_phys *ph, int order, int dim, int ny, int nx, int it, double *prim);
};

/* Primitive fields per cell of a physics table. */
static inline int orc_np(const orc_phys *ph) { return ph->np ? ph->np : ph->nf; }

double orc_smin(double a, double b);
double orc_smax(double a, double b);
double orc_limiter_value(double r, double bias, int limiter);
double orc_r_coefficient(double r, double bias);
void orc_slope_vector(int nf, const double *UL, const double *UM, const double *UR, double bias, int limiter,
                      double *d);
void orc_evolve_half(const orc_phys *ph, int dir, const double *wl, const double *wm, const double *wr, double dx,
                     double ts, double bias, int limiter, int side, double *out);
void orc_force_flux(const orc_phys *ph, int dir, const double *wl, const double *wr, double dx, double ts,
                    double *F);
void orc_lax_friedrichs(int nf, const double *Ul, const double *Ur, const double *Fl, const double *Fr, double dx,
                        double ts, double *out);
void orc_richtmyer_state(int nf, const double *Ul, const double *Ur, const double *Fl, const double *Fr, double dx,
                         double ts, double *out);
void orc_slic_flux(const orc_phys *ph, int dir, const double *wa, const double *wb, const double *wc,
                   const double *wd, double dx, double ts, double bias, int limiter, double *F);
/* One sweep over a 1D (dim 1) or 2D grid of primitive states, AoS. */
void orc_sweep(const orc_phys *ph, int order, int dim, int dir, int ny, int nx, double dx, double ts, double bias,
               int limiter, int bc, const double *prim_in, double *prim_out);
/* Maximum over the ghosted grid, 2D transmissive corners included. */
double orc_ghosted_max(const orc_phys *ph, int dim, int ny, int nx, int bc, const double *prim);
double orc_stable_time_step(double S, double dx, double cfl, double t, double T, int it);
/* The solve loop; returns steps taken. */
int orc_solve(orc_phys *ph, int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
              int max_steps, double bias, int limiter, int bc, const double *prim_in, double *prim_out,
              double *dt_log, int dt_cap);
/* The same with `subcycles` RK4 source sub-cycles on each side of every
 * second-order step (first order never sub-cycles). */
int orc_solve_sub(orc_phys *ph, int order, int dim, int ny, int nx, double dx, double cfl, double t_final,
                  int max_steps, double bias, int limiter, int bc, int subcycles, const double *prim_in,
                  double *prim_out, double *dt_log, int dt_cap);
/* One RK4 source cell update from conserved neighbours (Ut/Ub NULL in 1D). */
void orc_forcing_cell(const orc_phys *ph, int dim, const double *Ul, const double *Um, const double *Ur,
                      const double *Ut, const double *Ub, double dx, double ts, double bias, int limiter,
                      double *Unew);
/* One sub-cycle over the grid (one-layer ghosts). */
void orc_forcing_step(const orc_phys *ph, int dim, int ny, int nx, int bc, double dx, double ts, double bias,
                      int limiter, const double *prim_in, double *prim_out);

/* Euler physics table and EOS sound speed (euler_oracle.c), for the RGFM oracle. */
#include "oracle.h"
orc_phys orc_euler_phys_table(const orc_euler_params *mp);
double orc_euler_sound_speed(double rho, double p, const orc_euler_params *mp);

#endif
