/* A plain C caller of the C ABI (include/vanguard.h), linked with
 * -lvanguard_hip: the whole-solve, per-sweep and device-resident entry points
 * on the 1D Sod problem of 1000 cells (minbee, CFL 0.9, t = 0.25; initial
 * condition as Euler/eulertests.cpp:14-24) and on a 2D grid read from a file.
 *
 *   capi_check <dir>
 *
 * Writes raw float64 files into <dir>: sod_solve.f64, sod_dt.f64 (the whole
 * solve and its time steps), sod_steps.f64 (the same run as a loop of
 * vg_euler_time_step_1d calls with those steps and mirror ghosts).  When
 * <dir>/disc_in.f64 (64 x 64 x 6) and <dir>/disc_dt.f64 (200 steps) exist:
 * disc_solve.f64 (vg_euler_solve_2d, superbee, CFL 0.8, 200 steps),
 * disc_steps.f64 (X(dt/2) Y(dt) X(dt/2) per step through vg_euler_time_step_2d)
 * and disc_state.f64 (the same sweeps through vg_euler_state_sweep).  Exits
 * non-zero on any error status. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vanguard.h"

#define NF 6

static int fail(const char *what, int rc)
{
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, vg_last_error());
    return 1;
}

static int write_f64(const char *dir, const char *name, const double *a, size_t n)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "wb");
    if (!f || fwrite(a, sizeof(double), n, f) != n) return 1;
    fclose(f);
    return 0;
}

static long read_f64(const char *dir, const char *name, double *a, size_t n)
{
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s", dir, name);
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    size_t got = fread(a, sizeof(double), n, f);
    fclose(f);
    return (long)got;
}

/* Solvers::insertBoundaryCells with two layers (mirror-ordered) along a line of
 * n cells with stride s (in cells) into out (n + 4 cells, contiguous). */
static void ghost_line(const double *c, size_t s, int n, double *out)
{
    memcpy(out + 0 * NF, c + 1 * s * NF, NF * sizeof(double));
    memcpy(out + 1 * NF, c + 0 * s * NF, NF * sizeof(double));
    for (int i = 0; i < n; i++) memcpy(out + (i + 2) * NF, c + i * s * NF, NF * sizeof(double));
    memcpy(out + (n + 2) * NF, c + (n - 1) * s * NF, NF * sizeof(double));
    memcpy(out + (n + 3) * NF, c + (n - 2) * s * NF, NF * sizeof(double));
}

/* Solvers::insertBoundaryCells2D with two layers: default-state corners. */
static void ghost_grid(const double *c, int ny, int nx, double *out)
{
    const double def[NF] = {1.0, 0.0, 0.0, 0.0, 1.0, 1.0};
    const int gx = nx + 4;
    for (int k = 0; k < (ny + 4) * gx; k++) memcpy(out + k * NF, def, sizeof(def));
    double *line = malloc(sizeof(double) * NF * (size_t)(nx + ny + 4));
    for (int i = 0; i < ny; i++) {
        ghost_line(c + (size_t)i * nx * NF, 1, nx, line);
        memcpy(out + ((size_t)(i + 2) * gx) * NF, line, sizeof(double) * NF * (nx + 4));
    }
    for (int j = 0; j < nx; j++) {
        ghost_line(c + (size_t)j * NF, (size_t)nx, ny, line);
        for (int r = 0; r < 2; r++) {
            memcpy(out + ((size_t)r * gx + j + 2) * NF, line + r * NF, NF * sizeof(double));
            memcpy(out + ((size_t)(ny + 2 + r) * gx + j + 2) * NF, line + (ny + 2 + r) * NF, NF * sizeof(double));
        }
    }
    free(line);
}

int main(int argc, char **argv)
{
    if (argc != 2) {
        fprintf(stderr, "usage: capi_check <dir>\n");
        return 2;
    }
    const char *dir = argv[1];
    vg_ctx *ctx = NULL;
    int rc = vg_ctx_create(&ctx, 0);
    if (rc) return fail("vg_ctx_create", rc);
    const vg_euler_params air = {1.4, 0.0, 25.0, 0.0, 0.25, 250.0};
    vg_run_stats st;

    /* 1D Sod, 1000 cells */
    const int n = 1000;
    const double dx = 1.0 / n;
    double *w = malloc(sizeof(double) * NF * n), *out = malloc(sizeof(double) * NF * n);
    double *wb = malloc(sizeof(double) * NF * (n + 4)), *dts = malloc(sizeof(double) * 4096);
    for (int i = 0; i < n; i++) {
        const int left = i <= 0.5 * n;
        const double s[NF] = {left ? 1.0 : 0.125, 0.0, 0.0, 0.0, left ? 1.0 : 0.1, 1.0};
        memcpy(w + i * NF, s, sizeof(s));
    }
    rc = vg_euler_solve_1d(ctx, w, out, n, dx, 0.9, 0.25, 0, 0.0, 2, 2, VG_BC_TRANSMISSIVE, &air, VG_MODE_EXACT, dts,
                           4096, &st);
    if (rc) return fail("vg_euler_solve_1d", rc);
    printf("sod: %d steps, t = %.17g\n", st.steps, st.t);
    if (write_f64(dir, "sod_solve.f64", out, (size_t)NF * n) || write_f64(dir, "sod_dt.f64", dts, st.steps))
        return fail("write", -1);
    for (int k = 0; k < st.steps; k++) {
        ghost_line(w, 1, n, wb);
        rc = vg_euler_time_step_1d(ctx, w, wb, n, dx, dts[k], 0.0, 2, 2, &air, VG_MODE_EXACT);
        if (rc) return fail("vg_euler_time_step_1d", rc);
    }
    if (write_f64(dir, "sod_steps.f64", w, (size_t)NF * n)) return fail("write", -1);

    /* 2D grid from the caller's file */
    const int ny = 64, nx = 64, nsteps = 200;
    const double dx2 = 1.0 / 64;
    double *g = malloc(sizeof(double) * NF * ny * nx), *go = malloc(sizeof(double) * NF * ny * nx);
    double *gb = malloc(sizeof(double) * NF * (ny + 4) * (nx + 4)), *gdt = malloc(sizeof(double) * nsteps);
    if (read_f64(dir, "disc_in.f64", g, (size_t)NF * ny * nx) == (long)(NF * ny * nx) &&
        read_f64(dir, "disc_dt.f64", gdt, nsteps) == nsteps) {
        rc = vg_euler_solve_2d(ctx, g, go, ny, nx, dx2, 0.8, INFINITY, nsteps, 0.0, 0, 2, VG_BC_TRANSMISSIVE, &air,
                               VG_MODE_EXACT, NULL, 0, &st);
        if (rc) return fail("vg_euler_solve_2d", rc);
        if (write_f64(dir, "disc_solve.f64", go, (size_t)NF * ny * nx)) return fail("write", -1);
        vg_euler_state *es = NULL;
        rc = vg_euler_state_create(ctx, 2, ny, nx, &es);
        if (!rc) rc = vg_euler_state_load(es, g, 0);
        if (rc) return fail("vg_euler_state", rc);
        for (int k = 0; k < nsteps; k++) {
            const double ts[3] = {0.5 * gdt[k], gdt[k], 0.5 * gdt[k]};
            for (int s = 0; s < 3; s++) {
                ghost_grid(g, ny, nx, gb);
                rc = vg_euler_time_step_2d(ctx, g, gb, ny, nx, s == 1, dx2, ts[s], 0.0, 0, 2, &air, VG_MODE_EXACT);
                if (rc) return fail("vg_euler_time_step_2d", rc);
                rc = vg_euler_state_sweep(es, s == 1, 2, dx2, ts[s], 0.0, 0, VG_BC_TRANSMISSIVE, &air, VG_MODE_EXACT);
                if (rc) return fail("vg_euler_state_sweep", rc);
            }
        }
        if (write_f64(dir, "disc_steps.f64", g, (size_t)NF * ny * nx)) return fail("write", -1);
        rc = vg_euler_state_store(es, go, 0);
        if (rc) return fail("vg_euler_state_store", rc);
        vg_euler_state_destroy(es);
        if (write_f64(dir, "disc_state.f64", go, (size_t)NF * ny * nx)) return fail("write", -1);
        printf("disc: %d steps through three entry points\n", nsteps);
    }
    vg_ctx_destroy(ctx);
    free(w); free(out); free(wb); free(dts); free(g); free(go); free(gb); free(gdt);
    return 0;
}
