/*
 * fen_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C, FP64 restatement of FEN's single-phase fractional-step
 * Navier-Stokes step (reference: fradevita/FEN).  It is the checker for the
 * HIP library and the CPU baseline timed beside it; the product never links
 * it.  Every routine cites the Fortran it restates (file:line in the
 * reference tree).  Expression trees follow Fortran's left-to-right
 * evaluation and explicit parentheses; build with -ffp-contract=off so that
 * no multiply-add is fused (the reference is built with gfortran without
 * -march, which cannot fuse either).
 *
 * The FFT is this file's own iterative radix-2 transform (FFTW is not part
 * of this build): FFTW's conventions are kept (unnormalised, forward sign
 * -1, r2c keeps bins 0..n/2, c2r ignores Im of bins 0 and n/2), its
 * rounding is not.  Sizes must be powers of two.
 */
#include "fen_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* -------------------------------------------------------------------------- */
/* Fields                                                                     */
/* -------------------------------------------------------------------------- */

typedef struct {
    double *f;       /* storage */
    int gl;          /* ghost level 0 or 1 */
    char loc;        /* 'c','x','y','z' (src/scalar.f90:46, src/vector.f90:58-62) */
    int type[6];     /* bc type per face: -1 halo, 0 periodic, 1 Dirichlet, 2 Neumann */
    double *plane[6];/* bc value planes (src/scalar.f90:89-106) */
} ofield;

typedef struct {
    int n, logn;
    double *wr, *wi; /* w_k = exp(-2 pi i k / n), k < n/2 */
    int *rev;
} offt;

struct orc_ctx {
    int dim, nx, ny, nz;
    double lx, ly, lz, x0[3], delta;
    int bc[6];
    int periodic[3];
    int solver;
    /* navier_stokes_mod state (src/navier_stokes.f90:17-45) */
    double density, viscosity, g[3], cfl;
    int constant_cfl;
    double dt_o, dt_visc, dt_conv, maxdiv, maxcfl;
    ofield p, phi, u, v, w;
    double *dv[3], *dvo[3], *S[3], *gp[3];
    /* poisson_mod state (src/poisson.f90:31-49) */
    double *mwx, *mwy, *mwz;
    double *ta, *tb, *tc;
    offt fx, fy, fz;
    double *sre, *sim;  /* half spectrum (nx/2+1)*ny*nz, kx fastest (real spectra use sre) */
    double *cosx, *cosy; /* cos(pi m / (2n)), m < 4n, for the DCTs along x and y */
    /* two-phase (MF) state: multiphase_mod and volume_of_fluid_mod */
    int mf, x_first;
    double rho0, rho1, mu0, mu1, sigma, beta, rhomin, irhomin;
    double dt_surf;      /* set by set_timestep when sigma > 0 (src/navier_stokes.f90:29,657) */
    int have_dt_surf;
    ofield vof, h, dd, curv, nrx, nry, lx2, ly2, rho, mu, p_hat, p_o;
};

#define GX(c) ((size_t)(c)->nx + 2)
#define GY(c) ((size_t)(c)->ny + 2)
#define GZ(c) ((size_t)(c)->nz + 2)
#define GI(c, i, j, k) ((size_t)(i) + GX(c) * ((size_t)(j) + GY(c) * (size_t)(k)))
#define II(c, i, j, k) ((size_t)((i) - 1) + (size_t)(c)->nx * ((size_t)((j) - 1) + (size_t)(c)->ny * (size_t)((k) - 1)))

static size_t ghosted_len(const orc_ctx *c) { return GX(c) * GY(c) * GZ(c); }
static size_t interior_len(const orc_ctx *c) { return (size_t)c->nx * c->ny * c->nz; }

static size_t plane_len(const orc_ctx *c, int face)
{
    if (face < 2) return GY(c) * GZ(c);
    if (face < 4) return GX(c) * GZ(c);
    return GX(c) * GY(c);
}

static void field_alloc(orc_ctx *c, ofield *f, char loc)
{
    f->gl = 1;
    f->loc = loc;
    f->f = (double *)calloc(ghosted_len(c), sizeof(double));   /* src/scalar.f90:84 zero */
    for (int s = 0; s < 6; s++) {
        f->type[s] = 0;                                         /* src/scalar.f90:109-116 */
        f->plane[s] = (double *)calloc(plane_len(c, s), sizeof(double));
    }
}

static void field_free(ofield *f)
{
    free(f->f);
    for (int s = 0; s < 6; s++) free(f->plane[s]);
}

/* -------------------------------------------------------------------------- */
/* Ghost fill: scalar%update_ghost_nodes, src/scalar.f90:223-396.             */
/* One rank: the 2decomp halo exchange (src/halo.f90:33) is fully overwritten */
/* by the plane fills below, which run in the order left, right, bottom,      */
/* top, front, back over full planes.                                         */
/* -------------------------------------------------------------------------- */
static void ghost_fill_array(const orc_ctx *c, double *f, char loc, const int *type, double *const *plane)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t gx = GX(c), gy = GY(c), gz = GZ(c);
    /* Left (src/scalar.f90:254-271) */
    for (size_t k = 0; k < gz; k++)
        for (size_t j = 0; j < gy; j++) {
            double *row = f + gx * (j + gy * k);
            const double bcv = plane[0][j + gy * k];
            if (type[0] == 0) row[0] = row[nx];
            else if (type[0] == 1) row[0] = (loc == 'x') ? bcv : 2.0 * bcv - row[1];
            else if (type[0] == 2) row[0] = row[1];
        }
    /* Right (src/scalar.f90:273-291) */
    for (size_t k = 0; k < gz; k++)
        for (size_t j = 0; j < gy; j++) {
            double *row = f + gx * (j + gy * k);
            const double bcv = plane[1][j + gy * k];
            if (type[1] == 0) row[nx + 1] = row[1];
            else if (type[1] == 1) {
                if (loc == 'x') { row[nx] = bcv; row[nx + 1] = bcv; }
                else row[nx + 1] = 2.0 * bcv - row[nx];
            } else if (type[1] == 2) row[nx + 1] = row[nx];
        }
    /* Bottom (src/scalar.f90:293-316); prow == 1 on one rank */
    for (size_t k = 0; k < gz; k++)
        for (size_t i = 0; i < gx; i++) {
            const double bcv = plane[2][i + gx * k];
            double *g0 = f + GI(c, i, 0, k), *g1 = f + GI(c, i, 1, k), *gn = f + GI(c, i, ny, k);
            if (type[2] == 0) *g0 = *gn;
            else if (type[2] == 1) *g0 = (loc == 'y') ? bcv : 2.0 * bcv - *g1;
            else if (type[2] == 2) *g0 = *g1;
        }
    /* Top (src/scalar.f90:318-342) */
    for (size_t k = 0; k < gz; k++)
        for (size_t i = 0; i < gx; i++) {
            const double bcv = plane[3][i + gx * k];
            double *gt = f + GI(c, i, ny + 1, k), *g1 = f + GI(c, i, 1, k), *gn = f + GI(c, i, ny, k);
            if (type[3] == 0) *gt = *g1;
            else if (type[3] == 1) {
                if (loc == 'y') { *gn = bcv; *gt = bcv; }
                else *gt = 2.0 * bcv - *gn;
            } else if (type[3] == 2) *gt = *gn;
        }
    if (c->dim != 3) return;   /* src/scalar.f90:344 #if DIM==3 */
    /* Front (src/scalar.f90:345-365) */
    for (size_t j = 0; j < gy; j++)
        for (size_t i = 0; i < gx; i++) {
            const double bcv = plane[4][i + gx * j];
            double *g0 = f + GI(c, i, j, 0), *g1 = f + GI(c, i, j, 1), *gn = f + GI(c, i, j, nz);
            if (type[4] == 0) *g0 = *gn;
            else if (type[4] == 1) *g0 = (loc == 'z') ? bcv : 2.0 * bcv - *g1;
            else if (type[4] == 2) *g0 = *g1;
        }
    /* Back (src/scalar.f90:367-388) */
    for (size_t j = 0; j < gy; j++)
        for (size_t i = 0; i < gx; i++) {
            const double bcv = plane[5][i + gx * j];
            double *gt = f + GI(c, i, j, nz + 1), *g1 = f + GI(c, i, j, 1), *gn = f + GI(c, i, j, nz);
            if (type[5] == 0) *gt = *g1;
            else if (type[5] == 1) {
                if (loc == 'z') { *gn = bcv; *gt = bcv; }
                else *gt = 2.0 * bcv - *gn;
            } else if (type[5] == 2) *gt = *gn;
        }
}

static void ghost_fill(const orc_ctx *c, ofield *f)
{
    ghost_fill_array(c, f->f, f->loc, f->type, f->plane);
}

/* -------------------------------------------------------------------------- */
/* FFT (radix-2, decimation in time)                                          */
/* -------------------------------------------------------------------------- */
static int ilog2(int n) { int l = 0; while ((1 << l) < n) l++; return l; }

static void fft_init(offt *p, int n)
{
    p->n = n;
    p->logn = ilog2(n);
    p->wr = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    p->wi = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    p->rev = (int *)malloc(sizeof(int) * (size_t)n);
    const long double tpi = 2.0L * acosl(-1.0L);
    for (int k = 0; k < n / 2 + 1; k++) {
        long double a = tpi * (long double)k / (long double)n;
        p->wr[k] = (double)cosl(a);
        p->wi[k] = (double)(-sinl(a));
    }
    for (int i = 0; i < n; i++) {
        int r = 0;
        for (int b = 0; b < p->logn; b++) if (i & (1 << b)) r |= 1 << (p->logn - 1 - b);
        p->rev[i] = r;
    }
}

static void fft_free(offt *p) { free(p->wr); free(p->wi); free(p->rev); }

/* In-place complex transform of one line; sign -1 = FFTW_FORWARD, +1 = FFTW_BACKWARD. */
static void fft_line(const offt *p, double *re, double *im, int sign)
{
    const int n = p->n;
    for (int i = 0; i < n; i++) {
        int r = p->rev[i];
        if (r > i) {
            double t = re[i]; re[i] = re[r]; re[r] = t;
            t = im[i]; im[i] = im[r]; im[r] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len >> 1, step = n / len;
        for (int b = 0; b < n; b += len)
            for (int j = 0; j < half; j++) {
                const double wr = p->wr[j * step];
                const double wi = (sign < 0) ? p->wi[j * step] : -p->wi[j * step];
                const int a = b + j, d = b + j + half;
                const double tr = wr * re[d] - wi * im[d];
                const double ti = wr * im[d] + wi * re[d];
                re[d] = re[a] - tr; im[d] = im[a] - ti;
                re[a] = re[a] + tr; im[a] = im[a] + ti;
            }
    }
}

/* r2c of n reals -> bins 0..n/2 (dfftw_plan_dft_r2c_1d). */
static void fft_r2c_line(const offt *p, const double *x, size_t xs, double *re, double *im, size_t os,
                         double *tr, double *ti)
{
    const int n = p->n;
    for (int i = 0; i < n; i++) { tr[i] = x[(size_t)i * xs]; ti[i] = 0.0; }
    fft_line(p, tr, ti, -1);
    for (int k = 0; k <= n / 2; k++) { re[(size_t)k * os] = tr[k]; im[(size_t)k * os] = ti[k]; }
}

/* c2r of bins 0..n/2 -> n reals (dfftw_plan_dft_c2r_1d): Hermitian extension,
 * Im of bins 0 and n/2 ignored. */
static void fft_c2r_line(const offt *p, const double *re, const double *im, size_t is, double *x, size_t xs,
                         double *tr, double *ti)
{
    const int n = p->n;
    for (int k = 0; k <= n / 2; k++) { tr[k] = re[(size_t)k * is]; ti[k] = im[(size_t)k * is]; }
    ti[0] = 0.0;
    if (n > 1) ti[n / 2] = 0.0;
    for (int k = 1; k < n / 2; k++) { tr[n - k] = tr[k]; ti[n - k] = -ti[k]; }
    fft_line(p, tr, ti, +1);
    for (int i = 0; i < n; i++) x[(size_t)i * xs] = tr[i];
}

void orc_fft_c2c(int n, double *re, double *im, int sign)
{
    offt p; fft_init(&p, n); fft_line(&p, re, im, sign); fft_free(&p);
}

void orc_fft_r2c(int n, const double *x, double *re, double *im)
{
    offt p; fft_init(&p, n);
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_r2c_line(&p, x, 1, re, im, 1, t, t + n);
    free(t); fft_free(&p);
}

void orc_fft_c2r(int n, const double *re, const double *im, double *x)
{
    offt p; fft_init(&p, n);
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    fft_c2r_line(&p, re, im, 1, x, 1, t, t + n);
    free(t); fft_free(&p);
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

/* -------------------------------------------------------------------------- */
/* Poisson fast direct solvers (src/poisson.f90)                              */
/* -------------------------------------------------------------------------- */

/* Modified wavenumber, src/poisson.f90:141,628,646,664 (factor 2 pi, periodic)
 * and :265,794,881,899 (factor pi, Neumann). */
static double *modified_wavenumbers(int n, double delta, double factor)
{
    double *m = (double *)malloc(sizeof(double) * (size_t)n);
    const double pi = acos(-1.0);                     /* src/global.f90:14 */
    for (int i = 1; i <= n; i++)
        m[i - 1] = 2.0 * (cos(factor * pi * ((double)i - 1.0) / (double)(float)n) - 1.0) / (delta * delta);
    return m;
}

/* Tridiagonal coefficients, src/poisson.f90:741-757. */
static void tridiag_coefficients(orc_ctx *c, int n, int inflow_outflow)
{
    const double d2 = c->delta * c->delta;
    c->ta = (double *)malloc(sizeof(double) * (size_t)n);
    c->tb = (double *)malloc(sizeof(double) * (size_t)n);
    c->tc = (double *)malloc(sizeof(double) * (size_t)n);
    for (int k = 0; k < n; k++) { c->ta[k] = 1.0 / d2; c->tb[k] = -(2.0 / d2); c->tc[k] = 1.0 / d2; }
    c->tb[0] = c->tb[0] + c->ta[0];
    if (inflow_outflow) c->tb[n - 1] = c->tb[n - 1] - c->tc[n - 1];
    else c->tb[n - 1] = c->tb[n - 1] + c->tc[n - 1];
    c->ta[0] = 0.0;
    c->tc[n - 1] = 0.0;
}

/* Thomas sweep, src/poisson.f90:1094-1135 (3-D form, multiply by the
 * reciprocal pivot except on the last row, which divides). */
static void thomas_column(int n, const double *a, const double *b, const double *cc, double mx, double my,
                          double *re, double *im, size_t s, double *c1, double *dr, double *di)
{
    double factor = 1.0 / ((b[0] + mx) + my);
    c1[0] = cc[0] * factor;
    dr[0] = re[0] * factor; di[0] = im[0] * factor;
    for (int k = 1; k < n - 1; k++) {
        factor = 1.0 / (((b[k] + mx) + my) - a[k] * c1[k - 1]);
        c1[k] = cc[k] * factor;
        dr[k] = (re[(size_t)k * s] - a[k] * dr[k - 1]) * factor;
        di[k] = (im[(size_t)k * s] - a[k] * di[k - 1]) * factor;
    }
    factor = ((b[n - 1] + mx) + my) - a[n - 1] * c1[n - 2];
    if (factor != 0.0) {
        dr[n - 1] = (re[(size_t)(n - 1) * s] - a[n - 1] * dr[n - 2]) / factor;
        di[n - 1] = (im[(size_t)(n - 1) * s] - a[n - 1] * di[n - 2]) / factor;
    } else {
        dr[n - 1] = 0.0; di[n - 1] = 0.0;
    }
    re[(size_t)(n - 1) * s] = dr[n - 1]; im[(size_t)(n - 1) * s] = di[n - 1];
    for (int k = n - 2; k >= 0; k--) {
        re[(size_t)k * s] = dr[k] - c1[k] * re[(size_t)(k + 1) * s];
        im[(size_t)k * s] = di[k] - c1[k] * im[(size_t)(k + 1) * s];
    }
}

void orc_thomas(int n, const double *a, const double *b, const double *cc, double lam_x, double lam_y,
                double *re, double *im)
{
    double *w = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    thomas_column(n, a, b, cc, lam_x, lam_y, re, im, 1, w, w + n, w + 2 * n);
    free(w);
}

/* Common first stage: x r2c of every interior line of f into the spectrum. */
static void spectrum_x_forward(orc_ctx *c, const double *f, int gl)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, hx = nx / 2 + 1;
    const size_t sx = gl ? GX(c) : (size_t)nx, sy = gl ? GY(c) : (size_t)ny;
#pragma omp parallel
    {
        double *t = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
#pragma omp for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                const double *line = f + (size_t)gl + sx * ((size_t)(j - 1 + gl) + sy * (size_t)(k - 1 + gl));
                const size_t o = (size_t)hx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
                fft_r2c_line(&c->fx, line, 1, c->sre + o, c->sim + o, 1, t, t + nx);
            }
        free(t);
    }
}

static void spectrum_x_backward(orc_ctx *c, double *f, int gl)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, hx = nx / 2 + 1;
    const size_t sx = gl ? GX(c) : (size_t)nx, sy = gl ? GY(c) : (size_t)ny;
#pragma omp parallel
    {
        double *t = (double *)malloc(sizeof(double) * 2 * (size_t)nx);
#pragma omp for collapse(2) schedule(static)
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                double *line = f + (size_t)gl + sx * ((size_t)(j - 1 + gl) + sy * (size_t)(k - 1 + gl));
                const size_t o = (size_t)hx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
                fft_c2r_line(&c->fx, c->sre + o, c->sim + o, 1, line, 1, t, t + nx);
            }
        free(t);
    }
}

/* c2c along y (dir 1) or z (dir 2) of every spectrum line, sign -1/+1. */
static void spectrum_c2c(orc_ctx *c, int dir, int sign)
{
    const int ny = c->ny, nz = c->nz, hx = c->nx / 2 + 1;
    const offt *p = dir == 1 ? &c->fy : &c->fz;
    const int n = p->n, nother = dir == 1 ? nz : ny;
    const size_t stride = dir == 1 ? (size_t)hx : (size_t)hx * (size_t)ny;
#pragma omp parallel
    {
        double *t = (double *)malloc(sizeof(double) * 2 * (size_t)n);
#pragma omp for collapse(2) schedule(static)
        for (int o = 0; o < nother; o++)
            for (int kx = 0; kx < hx; kx++) {
                const size_t base = dir == 1 ? (size_t)kx + (size_t)hx * (size_t)ny * (size_t)o
                                             : (size_t)kx + (size_t)hx * (size_t)o;
                for (int l = 0; l < n; l++) { t[l] = c->sre[base + stride * l]; t[n + l] = c->sim[base + stride * l]; }
                fft_line(p, t, t + n, sign);
                for (int l = 0; l < n; l++) { c->sre[base + stride * l] = t[l]; c->sim[base + stride * l] = t[n + l]; }
            }
        free(t);
    }
}

static void spectrum_scale(orc_ctx *c, double s)
{
    const size_t n = (size_t)(c->nx / 2 + 1) * c->ny * c->nz;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n; q++) { c->sre[q] = c->sre[q] / s; c->sim[q] = c->sim[q] / s; }
}

/* Sum of the interior in Fortran loop order (k, j, i), sequential. */
static double interior_sum(const orc_ctx *c, const double *f, int gl)
{
    const size_t sx = gl ? GX(c) : (size_t)c->nx, sy = gl ? GY(c) : (size_t)c->ny;
    double s = 0.0;
    for (int k = 1; k <= c->nz; k++)
        for (int j = 1; j <= c->ny; j++) {
            const double *line = f + (size_t)gl + sx * ((size_t)(j - 1 + gl) + sy * (size_t)(k - 1 + gl));
            for (int i = 0; i < c->nx; i++) s = s + line[i];
        }
    return s;
}

static size_t array_len(const orc_ctx *c, int gl) { return gl ? ghosted_len(c) : interior_len(c); }

/* poisson_solver_ppp, src/poisson.f90:941-1034. */
static void poisson_ppp(orc_ctx *c, double *f, int gl)
{
    const int hx = c->nx / 2 + 1, ny = c->ny, nz = c->nz;
    spectrum_x_forward(c, f, gl);
    spectrum_c2c(c, 1, -1);
    spectrum_c2c(c, 2, -1);
    spectrum_scale(c, (double)(float)(c->nx * c->ny * c->nz));                 /* :992 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < hx; i++) {                                      /* :995-1005 */
                const size_t q = (size_t)i + (size_t)hx * ((size_t)j + (size_t)ny * (size_t)k);
                const double s = (c->mwx[i] + c->mwy[j]) + c->mwz[k];
                if (s == 0.0) { c->sre[q] = 0.0; c->sim[q] = 0.0; }
                else { c->sre[q] = c->sre[q] / s; c->sim[q] = c->sim[q] / s; }
            }
    spectrum_c2c(c, 2, +1);
    spectrum_c2c(c, 1, +1);
    spectrum_x_backward(c, f, gl);
}

/* poisson_solver_pp, src/poisson.f90:416-505. */
static void poisson_pp(orc_ctx *c, double *f, int gl)
{
    const int hx = c->nx / 2 + 1, ny = c->ny;
    spectrum_x_forward(c, f, gl);
    spectrum_c2c(c, 1, -1);
    spectrum_scale(c, (double)(float)(c->nx * c->ny));                         /* :458 */
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < hx; i++) {                                          /* :461-471 */
            const size_t q = (size_t)i + (size_t)hx * (size_t)j;
            const double s = c->mwx[i] + c->mwy[j];
            if (s == 0.0) { c->sre[q] = 0.0; c->sim[q] = 0.0; }
            else { c->sre[q] = c->sre[q] / s; c->sim[q] = c->sim[q] / s; }
        }
    spectrum_c2c(c, 1, +1);
    spectrum_x_backward(c, f, gl);
    const double mean = interior_sum(c, f, gl);                                 /* :491-498 */
    const double sub = mean / (double)(float)(c->nx * c->ny * c->nz);
    const size_t n = array_len(c, gl);
    for (size_t q = 0; q < n; q++) f[q] = f[q] - sub;                           /* :503 whole array */
}

/* poisson_solver_ppn, src/poisson.f90:1038-1173. */
static void poisson_ppn(orc_ctx *c, double *f, int gl)
{
    const int hx = c->nx / 2 + 1, ny = c->ny, nz = c->nz;
    spectrum_x_forward(c, f, gl);
    spectrum_scale(c, (double)(float)c->nx);                                    /* :1074 */
    spectrum_c2c(c, 1, -1);
    spectrum_scale(c, (double)(float)c->ny);                                    /* :1087 */
#pragma omp parallel
    {
        double *w = (double *)malloc(sizeof(double) * 3 * (size_t)nz);
#pragma omp for collapse(2) schedule(static)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < hx; i++) {
                const size_t q = (size_t)i + (size_t)hx * (size_t)j;
                thomas_column(nz, c->ta, c->tb, c->tc, c->mwx[i], c->mwy[j], c->sre + q, c->sim + q,
                              (size_t)hx * (size_t)ny, w, w + nz, w + 2 * nz);
            }
        free(w);
    }
    spectrum_c2c(c, 1, +1);
    const size_t n = array_len(c, gl);
    memset(f, 0, n * sizeof(double));                                           /* :1151 */
    spectrum_x_backward(c, f, gl);
    const double mean = interior_sum(c, f, gl);                                 /* :1159-1166 */
    const double sub = mean / (double)(float)(c->nx * c->ny * c->nz);
    for (size_t q = 0; q < n; q++) f[q] = f[q] - sub;                           /* :1171 */
}

/* cos(pi m / (2n)) for m < 4n, in long double. */
static double *dct_table(int n)
{
    double *t = (double *)malloc(sizeof(double) * 4 * (size_t)n);
    const long double pi = acosl(-1.0L);
    for (int m = 0; m < 4 * n; m++) t[m] = (double)cosl(pi * (long double)m / (2.0L * (long double)n));
    return t;
}

/* FFTW_REDFT10 (DCT-II), unnormalised, by its definition:
 * y_k = 2 sum_j x_j cos(pi k (2j+1) / (2n)). */
static void dct2_line(int n, const double *T, const double *x, size_t xs, double *y, size_t ys)
{
    for (int k = 0; k < n; k++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += x[(size_t)j * xs] * T[((size_t)k * (2 * j + 1)) % (4 * (size_t)n)];
        y[(size_t)k * ys] = 2.0 * s;
    }
}

/* FFTW_REDFT01 (DCT-III), unnormalised: y_j = x_0 + 2 sum_{k>=1} x_k cos(pi k (2j+1) / (2n)). */
static void dct3_line(int n, const double *T, const double *x, size_t xs, double *y, size_t ys)
{
    for (int j = 0; j < n; j++) {
        double s = 0.0;
        for (int k = 1; k < n; k++) s += x[(size_t)k * xs] * T[((size_t)k * (2 * j + 1)) % (4 * (size_t)n)];
        y[(size_t)j * ys] = x[0] + 2.0 * s;
    }
}

/* Thomas sweep of the 2-D solvers, src/poisson.f90:346-384: divisions,
 * and the c1 and d1 denominators associated differently. */
static void thomas_2d(int n, const double *a, const double *b, const double *cc, double mx, double *re, double *im,
                      size_t s, double *c1, double *dr, double *di)
{
    c1[0] = cc[0] / (b[0] + mx);
    dr[0] = re[0] / (b[0] + mx);
    if (im) di[0] = im[0] / (b[0] + mx);
    for (int j = 1; j < n - 1; j++) {
        c1[j] = cc[j] / ((b[j] - a[j] * c1[j - 1]) + mx);
        const double den = (b[j] + mx) - a[j] * c1[j - 1];
        dr[j] = (re[(size_t)j * s] - a[j] * dr[j - 1]) / den;
        if (im) di[j] = (im[(size_t)j * s] - a[j] * di[j - 1]) / den;
    }
    const double frac = (b[n - 1] + mx) - a[n - 1] * c1[n - 2];
    if (frac != 0.0) {
        dr[n - 1] = (re[(size_t)(n - 1) * s] - a[n - 1] * dr[n - 2]) / frac;
        if (im) di[n - 1] = (im[(size_t)(n - 1) * s] - a[n - 1] * di[n - 2]) / frac;
    } else {
        dr[n - 1] = 0.0;
        if (im) di[n - 1] = 0.0;
    }
    re[(size_t)(n - 1) * s] = dr[n - 1];
    if (im) im[(size_t)(n - 1) * s] = di[n - 1];
    for (int j = n - 2; j >= 0; j--) {
        re[(size_t)j * s] = dr[j] - c1[j] * re[(size_t)(j + 1) * s];
        if (im) im[(size_t)j * s] = di[j] - c1[j] * im[(size_t)(j + 1) * s];
    }
}

/* Thomas sweep of the 3-D solvers on a real column (nnn), as thomas_column. */
static void thomas_real(int n, const double *a, const double *b, const double *cc, double mx, double my, double *x,
                        size_t s, double *c1, double *d)
{
    double factor = 1.0 / ((b[0] + mx) + my);
    c1[0] = cc[0] * factor;
    d[0] = x[0] * factor;
    for (int k = 1; k < n - 1; k++) {
        factor = 1.0 / (((b[k] + mx) + my) - a[k] * c1[k - 1]);
        c1[k] = cc[k] * factor;
        d[k] = (x[(size_t)k * s] - a[k] * d[k - 1]) * factor;
    }
    factor = ((b[n - 1] + mx) + my) - a[n - 1] * c1[n - 2];
    d[n - 1] = factor != 0.0 ? (x[(size_t)(n - 1) * s] - a[n - 1] * d[n - 2]) / factor : 0.0;
    x[(size_t)(n - 1) * s] = d[n - 1];
    for (int k = n - 2; k >= 0; k--) x[(size_t)k * s] = d[k] - c1[k] * x[(size_t)(k + 1) * s];
}

/* Pointer to interior line (j,k) of f with ghost level gl. */
static double *row(const orc_ctx *c, double *f, int gl, int j, int k)
{
    const size_t sx = gl ? GX(c) : (size_t)c->nx, sy = gl ? GY(c) : (size_t)c->ny;
    return f + (size_t)gl + sx * ((size_t)(j - 1 + gl) + sy * (size_t)(k - 1 + gl));
}

/* poisson_solver_pn, src/poisson.f90:306-412 (2-D: x periodic, y Neumann). */
static void poisson_pn(orc_ctx *c, double *f, int gl)
{
    const int hx = c->nx / 2 + 1, ny = c->ny;
    spectrum_x_forward(c, f, gl);
    spectrum_scale(c, (double)(float)c->nx);                                    /* :338 */
    double *w = (double *)malloc(sizeof(double) * 3 * (size_t)ny);
    for (int i = 0; i < hx; i++)                                                /* :344-384 */
        thomas_2d(ny, c->ta, c->tb, c->tc, c->mwx[i], c->sre + i, c->sim + i, (size_t)hx, w, w + ny, w + 2 * ny);
    free(w);
    spectrum_x_backward(c, f, gl);
    const double mean = interior_sum(c, f, gl);                                 /* :397-404 */
    const double sub = mean / (double)(float)(c->nx * c->ny * c->nz);
    for (int k = 1; k <= c->nz; k++)                                            /* :408: interior only */
        for (int j = 1; j <= ny; j++) {
            double *r = row(c, f, gl, j, k);
            for (int i = 0; i < c->nx; i++) r[i] = r[i] - sub;
        }
}

/* poisson_solver_nn, src/poisson.f90:509-596 (2-D: x and y Neumann). */
static void poisson_nn(orc_ctx *c, double *f, int gl)
{
    const int nx = c->nx, ny = c->ny;
    double *R = c->sre;                                                         /* outr_x, [i][j] */
    for (int j = 1; j <= ny; j++) dct2_line(nx, c->cosx, row(c, f, gl, j, 1), 1, R + (size_t)nx * (j - 1), 1);
    double *w = (double *)malloc(sizeof(double) * 3 * (size_t)ny);
    for (int i = 0; i < nx; i++)                                                /* :540-580 */
        thomas_2d(ny, c->ta, c->tb, c->tc, c->mwx[i], R + i, NULL, (size_t)nx, w, w + ny, w + 2 * ny);
    free(w);
    for (int j = 1; j <= ny; j++) dct3_line(nx, c->cosx, R + (size_t)nx * (j - 1), 1, row(c, f, gl, j, 1), 1);
    const size_t n = array_len(c, gl);
    const double s = (double)(float)(nx * 2);
    for (size_t q = 0; q < n; q++) f[q] = f[q] / s;                             /* :594: whole array */
}

/* poisson_solver_npn, src/poisson.f90:1177-1312 (x Neumann, y periodic, z Neumann). */
static void poisson_npn(orc_ctx *c, double *f, int gl)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, hy = ny / 2 + 1;
    double *R = (double *)malloc(sizeof(double) * (size_t)nx * ny * nz);      /* outr_x, [i][j][k] */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            dct2_line(nx, c->cosx, row(c, f, gl, j, k), 1, R + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)), 1);
    const size_t nr = (size_t)nx * ny * nz;
    for (size_t q = 0; q < nr; q++) R[q] = R[q] / (double)(nx * 2);            /* :1215 real(nx*2, dp) */
    double *t = (double *)malloc(sizeof(double) * 2 * (size_t)ny);
    for (int k = 0; k < nz; k++)                                                /* :1221-1228: y r2c */
        for (int i = 0; i < nx; i++) {
            const size_t o = (size_t)i + (size_t)nx * (size_t)hy * k;
            fft_r2c_line(&c->fy, R + (size_t)i + (size_t)nx * (size_t)ny * k, (size_t)nx, c->sre + o, c->sim + o,
                         (size_t)nx, t, t + ny);
        }
    const size_t ns = (size_t)nx * hy * nz;
    for (size_t q = 0; q < ns; q++) { c->sre[q] = c->sre[q] / (double)ny; c->sim[q] = c->sim[q] / (double)ny; }
    double *w = (double *)malloc(sizeof(double) * 3 * (size_t)nz);
    for (int jy = 0; jy < hy; jy++)                                             /* :1236-1280 */
        for (int i = 0; i < nx; i++) {
            const size_t q = (size_t)i + (size_t)nx * jy;
            thomas_column(nz, c->ta, c->tb, c->tc, c->mwx[i], c->mwy[jy], c->sre + q, c->sim + q,
                          (size_t)nx * hy, w, w + nz, w + 2 * nz);
        }
    free(w);
    for (int k = 0; k < nz; k++)                                                /* :1286-1291: y c2r */
        for (int i = 0; i < nx; i++) {
            const size_t o = (size_t)i + (size_t)nx * (size_t)hy * k;
            fft_c2r_line(&c->fy, c->sre + o, c->sim + o, (size_t)nx, R + (size_t)i + (size_t)nx * (size_t)ny * k,
                         (size_t)nx, t, t + ny);
        }
    free(t);
    memset(f, 0, array_len(c, gl) * sizeof(double));                           /* :1297 */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            dct3_line(nx, c->cosx, R + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)), 1, row(c, f, gl, j, k), 1);
    free(R);                                                                    /* no mean removal (:1310) */
}

/* poisson_solver_nnn, src/poisson.f90:1316-1451 (x, y, z Neumann). */
static void poisson_nnn(orc_ctx *c, double *f, int gl)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    double *R = (double *)malloc(sizeof(double) * (size_t)nx * ny * nz);
    double *Y = (double *)malloc(sizeof(double) * (size_t)nx * ny * nz);
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            dct2_line(nx, c->cosx, row(c, f, gl, j, k), 1, R + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)), 1);
    const size_t nr = (size_t)nx * ny * nz;
    for (size_t q = 0; q < nr; q++) R[q] = R[q] / (double)(nx * 2);            /* :1354 */
    for (int k = 0; k < nz; k++)                                                /* :1360-1365: y DCT-II */
        for (int i = 0; i < nx; i++) {
            const size_t o = (size_t)i + (size_t)nx * (size_t)ny * k;
            dct2_line(ny, c->cosy, R + o, (size_t)nx, Y + o, (size_t)nx);
        }
    for (size_t q = 0; q < nr; q++) Y[q] = Y[q] / (double)(ny * 2);            /* :1368 */
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)nz);
    for (int j = 0; j < ny; j++)                                                /* :1375-1419 */
        for (int i = 0; i < nx; i++)
            thomas_real(nz, c->ta, c->tb, c->tc, c->mwx[i], c->mwy[j], Y + (size_t)i + (size_t)nx * j,
                        (size_t)nx * ny, w, w + nz);
    free(w);
    for (int k = 0; k < nz; k++)                                                /* :1425-1430: y DCT-III */
        for (int i = 0; i < nx; i++) {
            const size_t o = (size_t)i + (size_t)nx * (size_t)ny * k;
            dct3_line(ny, c->cosy, Y + o, (size_t)nx, R + o, (size_t)nx);
        }
    memset(f, 0, array_len(c, gl) * sizeof(double));                           /* :1436 */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            dct3_line(nx, c->cosx, R + (size_t)nx * ((j - 1) + (size_t)ny * (k - 1)), 1, row(c, f, gl, j, k), 1);
    free(R);
    free(Y);                                                                    /* no mean removal (:1449) */
}

void orc_poisson_solve(orc_ctx *c, double *f, int gl)
{
    switch (c->solver) {
    case ORC_PP: poisson_pp(c, f, gl); break;
    case ORC_PPP: poisson_ppp(c, f, gl); break;
    case ORC_PPN: poisson_ppn(c, f, gl); break;
    case ORC_PN: poisson_pn(c, f, gl); break;
    case ORC_NN: poisson_nn(c, f, gl); break;
    case ORC_NPN: poisson_npn(c, f, gl); break;
    case ORC_NNN: poisson_nnn(c, f, gl); break;
    default: break;
    }
}

/* -------------------------------------------------------------------------- */
/* Setup                                                                      */
/* -------------------------------------------------------------------------- */

/* BC mapping of allocate_navier_stokes_fields, src/navier_stokes.f90:779-1017.
 * Returns the field bc type for field id on face s given physical code pc. */
static int field_bc_type(int field, int face, int pc, int *ok)
{
    const int vel = (field == ORC_U || field == ORC_V || field == ORC_W);
    *ok = 1;
    switch (pc) {
    case ORC_BC_PERIODIC: return 0;
    case ORC_BC_WALL: return vel ? 1 : 2;
    case ORC_BC_INFLOW:
        if (face == 5) { *ok = 0; return 0; }          /* back face has no Inflow branch (:989-1016) */
        if (!vel) return 2;
        if (face < 2) return field == ORC_U ? 1 : 2;     /* left/right (:800-809, :845-854) */
        return 1;
    case ORC_BC_OUTFLOW:
        if (face == 4) { *ok = 0; return 0; }          /* front face has no Outflow branch (:960-987) */
        return vel ? 2 : 1;
    }
    *ok = 0;
    return 0;
}

orc_ctx *orc_create(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                    const double *x0, const int *bc)
{
    if (dim == 2) nz = 1;
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->dim = dim; c->nx = nx; c->ny = ny; c->nz = nz;
    c->lx = lx; c->ly = ly; c->lz = lz;
    for (int d = 0; d < 3; d++) c->x0[d] = x0 ? x0[d] : 0.0;
    for (int s = 0; s < 6; s++) c->bc[s] = (dim == 3 || s < 4) ? bc[s] : ORC_BC_PERIODIC;
    c->delta = lx / (double)(float)nx;                                          /* src/grid.f90:140 */
    c->periodic[0] = c->bc[0] == ORC_BC_PERIODIC && c->bc[1] == ORC_BC_PERIODIC;   /* :105-121 */
    c->periodic[1] = c->bc[2] == ORC_BC_PERIODIC && c->bc[3] == ORC_BC_PERIODIC;
    c->periodic[2] = dim == 3 ? (c->bc[4] == ORC_BC_PERIODIC && c->bc[5] == ORC_BC_PERIODIC) : 1;

    /* Solver dispatch, src/poisson.f90:70-111. */
    const int *pb = c->periodic;
    if (dim == 3) {
        if (pb[0] && pb[1] && pb[2]) c->solver = ORC_PPP;
        else if (pb[0] && pb[1] && !pb[2]) c->solver = ORC_PPN;
        else if (!pb[0] && pb[1] && !pb[2]) c->solver = ORC_NPN;
        else if (!pb[0] && !pb[1] && !pb[2]) c->solver = ORC_NNN;
        else { free(c); return NULL; }   /* the reference stops (:91-94) */
    } else {
        if (pb[0] && pb[1]) c->solver = ORC_PP;
        else if (pb[0] && !pb[1]) c->solver = ORC_PN;
        else if (!pb[0] && !pb[1]) c->solver = ORC_NN;
        else { free(c); return NULL; }
    }

    /* allocate_navier_stokes_fields defaults (src/navier_stokes.f90:18-24). */
    c->density = 1.0; c->viscosity = 1.0; c->cfl = 1.0;
    field_alloc(c, &c->p, 'c');
    field_alloc(c, &c->phi, 'c');
    field_alloc(c, &c->u, 'x');
    field_alloc(c, &c->v, 'y');
    field_alloc(c, &c->w, 'z');
    for (int d = 0; d < 3; d++) {
        c->dv[d] = (double *)calloc(interior_len(c), sizeof(double));
        c->dvo[d] = (double *)calloc(interior_len(c), sizeof(double));
        c->S[d] = (double *)calloc(interior_len(c), sizeof(double));
        c->gp[d] = (double *)calloc(interior_len(c), sizeof(double));
    }
    ofield *fl[5] = {&c->p, &c->phi, &c->u, &c->v, &c->w};
    const int ids[5] = {ORC_P, ORC_PHI, ORC_U, ORC_V, ORC_W};
    for (int q = 0; q < 5; q++)
        for (int s = 0; s < (dim == 3 ? 6 : 4); s++) {
            int ok;
            fl[q]->type[s] = field_bc_type(ids[q], s, c->bc[s], &ok);
        }

    /* init_poisson_* (src/poisson.f90:117-172, 601-763). */
    fft_init(&c->fx, nx);
    fft_init(&c->fy, ny);
    fft_init(&c->fz, nz);
    /* Neumann directions use DCT wavenumbers (factor pi), :265,794,881,899 */
    c->mwx = modified_wavenumbers(nx, c->delta, pb[0] ? 2.0 : 1.0);
    c->mwy = modified_wavenumbers(ny, c->delta, pb[1] ? 2.0 : 1.0);
    if (c->solver == ORC_PPP) c->mwz = modified_wavenumbers(nz, c->delta, 2.0);
    if (c->solver == ORC_PPN || c->solver == ORC_NPN || c->solver == ORC_NNN)
        tridiag_coefficients(c, nz, c->bc[4] == ORC_BC_INFLOW && c->bc[5] == ORC_BC_OUTFLOW);
    if (c->solver == ORC_PN || c->solver == ORC_NN)                             /* along y, :214-226 */
        tridiag_coefficients(c, ny, c->bc[2] == ORC_BC_INFLOW && c->bc[3] == ORC_BC_OUTFLOW);
    if (!pb[0]) c->cosx = dct_table(nx);
    if (!pb[1]) c->cosy = dct_table(ny);
    size_t ns = (size_t)(nx / 2 + 1) * ny * nz;
    if ((size_t)nx * (ny / 2 + 1) * nz > ns) ns = (size_t)nx * (ny / 2 + 1) * nz;
    if ((size_t)nx * ny * nz > ns) ns = (size_t)nx * ny * nz;
    c->sre = (double *)calloc(ns, sizeof(double));
    c->sim = (double *)calloc(ns, sizeof(double));
    return c;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    if (c->mf) {
        ofield *mf[12] = {&c->vof, &c->h, &c->dd, &c->curv, &c->nrx, &c->nry, &c->lx2, &c->ly2,
                          &c->rho, &c->mu, &c->p_hat, &c->p_o};
        for (int q = 0; q < 12; q++) field_free(mf[q]);
    }
    field_free(&c->p); field_free(&c->phi); field_free(&c->u); field_free(&c->v); field_free(&c->w);
    for (int d = 0; d < 3; d++) { free(c->dv[d]); free(c->dvo[d]); free(c->S[d]); free(c->gp[d]); }
    fft_free(&c->fx); fft_free(&c->fy); fft_free(&c->fz);
    free(c->mwx); free(c->mwy); free(c->mwz);
    free(c->ta); free(c->tb); free(c->tc);
    free(c->sre); free(c->sim); free(c->cosx); free(c->cosy);
    free(c);
}

int orc_solver_kind(const orc_ctx *c) { return c->solver; }
double orc_delta(const orc_ctx *c) { return c->delta; }

void orc_set_params(orc_ctx *c, double density, double viscosity, const double *g, double cfl, int constant_cfl)
{
    c->density = density; c->viscosity = viscosity;
    for (int d = 0; d < 3; d++) c->g[d] = g ? g[d] : 0.0;
    c->cfl = cfl; c->constant_cfl = constant_cfl;
}

/* set_timestep, src/navier_stokes.f90:623-666 (single-phase branch). */
double orc_set_timestep(orc_ctx *c, double U)
{
    const double d = c->delta;
    c->dt_conv = c->cfl * d / U;
    c->dt_visc = 0.125 * d * d * c->density / c->viscosity;
    if (c->dim == 3) c->dt_visc = (1.0 / 6.0) * d * d * c->density / c->viscosity;
    const double dt = c->dt_conv < c->dt_visc ? c->dt_conv : c->dt_visc;
    c->dt_o = dt;
    return dt;
}

void orc_set_dt_o(orc_ctx *c, double dt_o) { c->dt_o = dt_o; }
double orc_get_dt_o(const orc_ctx *c) { return c->dt_o; }

void orc_get_dt_limits(const orc_ctx *c, double *dt_visc, double *dt_conv)
{
    if (dt_visc) *dt_visc = c->dt_visc;
    if (dt_conv) *dt_conv = c->dt_conv;
}

static ofield *ghosted(orc_ctx *c, int field)
{
    switch (field) {
    case ORC_P: return &c->p;
    case ORC_PHI: return &c->phi;
    case ORC_U: return &c->u;
    case ORC_V: return &c->v;
    case ORC_W: return &c->w;
    case ORC_VOF: return c->mf ? &c->vof : NULL;
    case ORC_RHO: return c->mf ? &c->rho : NULL;
    case ORC_MU: return c->mf ? &c->mu : NULL;
    }
    return NULL;
}

double *orc_field(orc_ctx *c, int field)
{
    ofield *f = ghosted(c, field);
    if (f) return f->f;
    if (field >= ORC_DVO_X && field <= ORC_DVO_Z) return c->dvo[field - ORC_DVO_X];
    if (field >= ORC_S_X && field <= ORC_S_Z) return c->S[field - ORC_S_X];
    return NULL;
}

size_t orc_field_len(const orc_ctx *c, int field)
{
    return (field <= ORC_W || field >= ORC_VOF) ? ghosted_len(c) : interior_len(c);
}

double *orc_bc_plane(orc_ctx *c, int field, int face)
{
    ofield *f = ghosted(c, field);
    return f ? f->plane[face] : NULL;
}

size_t orc_bc_plane_len(const orc_ctx *c, int face) { return plane_len(c, face); }

int orc_bc_type(const orc_ctx *c, int field, int face)
{
    ofield *f = ghosted((orc_ctx *)c, field);
    return f ? f->type[face] : -99;
}

void orc_set_bc_type(orc_ctx *c, int field, int face, int type)
{
    ofield *f = ghosted(c, field);
    if (f) f->type[face] = type;
}

void orc_update_ghost_nodes(orc_ctx *c, int field)
{
    ofield *f = ghosted(c, field);
    if (f) ghost_fill(c, f);
}

/* -------------------------------------------------------------------------- */
/* Operators (src/fields.f90) and the momentum right-hand side                */
/* -------------------------------------------------------------------------- */

/* add_advection, src/navier_stokes.f90:261-353; R is interior-sized and is
 * accumulated into (R - ... ). */
static void add_advection(const orc_ctx *c, const double *u, const double *v, const double *w,
                          double *rx, double *ry, double *rz)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, d3 = c->dim == 3;
    const double id = 1.0 / c->delta;
    const ptrdiff_t sj = (ptrdiff_t)GX(c), sk = (ptrdiff_t)(GX(c) * GY(c));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t g = GI(c, i, j, k);
                const size_t q = II(c, i, j, k);
                const double *U = u + g, *V = v + g, *W = w + g;
                /* x-momentum (:297-310) */
                double uuip = 0.25 * ((U[1] + U[0]) * (U[1] + U[0]));
                double uuim = 0.25 * ((U[-1] + U[0]) * (U[-1] + U[0]));
                double uvjp = ((U[sj] + U[0]) * (V[1] + V[0])) * 0.25;
                double uvjm = ((U[0] + U[-sj]) * (V[1 - sj] + V[-sj])) * 0.25;
                double r = (rx[q] - (uuip - uuim) * id) - (uvjp - uvjm) * id;
                if (d3) {
                    double uwkp = ((U[sk] + U[0]) * (W[1] + W[0])) * 0.25;
                    double uwkm = ((U[0] + U[-sk]) * (W[1 - sk] + W[-sk])) * 0.25;
                    r = r - (uwkp - uwkm) * id;
                }
                rx[q] = r;
                /* y-momentum (:315-327) */
                double vuip = ((V[1] + V[0]) * (U[sj] + U[0])) * 0.25;
                double vuim = ((V[0] + V[-1]) * (U[-1 + sj] + U[-1])) * 0.25;
                double vvjp = 0.25 * ((V[sj] + V[0]) * (V[sj] + V[0]));
                double vvjm = 0.25 * ((V[-sj] + V[0]) * (V[-sj] + V[0]));
                r = (ry[q] - (vuip - vuim) * id) - (vvjp - vvjm) * id;
                if (d3) {
                    double vwkp = ((V[sk] + V[0]) * (W[sj] + W[0])) * 0.25;
                    double vwkm = ((V[0] + V[-sk]) * (W[sj - sk] + W[-sk])) * 0.25;
                    r = r - (vwkp - vwkm) * id;
                }
                ry[q] = r;
                if (d3) {
                    /* z-momentum (:334-347) */
                    double wuip = ((W[0] + W[1]) * (U[0] + U[sk])) * 0.25;
                    double wuim = ((W[0] + W[-1]) * (U[-1] + U[-1 + sk])) * 0.25;
                    double wvjp = ((W[0] + W[sj]) * (V[0] + V[sk])) * 0.25;
                    double wvjm = ((W[0] + W[-sj]) * (V[-sj] + V[-sj + sk])) * 0.25;
                    double wwkp = ((W[0] + W[sk]) * (W[0] + W[sk])) * 0.25;
                    double wwkm = ((W[0] + W[-sk]) * (W[0] + W[-sk])) * 0.25;
                    rz[q] = ((rz[q] - (wuip - wuim) * id) - (wvjp - wvjm) * id) - (wwkp - wwkm) * id;
                }
            }
}

void orc_add_advection(orc_ctx *c, double *rx, double *ry, double *rz)
{
    add_advection(c, c->u.f, c->v.f, c->w.f, rx, ry, rz);
}

/* add_diffusion, constant-viscosity branch (src/navier_stokes.f90:384-404) with
 * laplacian_of_vector (src/fields.f90:298-343), then the body-force term
 * (src/navier_stokes.f90:245-255).  rhof == 0.5*(rho+rho) on a constant
 * density field (src/fields.f90:175-206). */
static void add_diffusion_and_source(const orc_ctx *c, double rhof, double *rx, double *ry, double *rz)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, d3 = c->dim == 3;
    const double id2 = 1.0 / (c->delta * c->delta);
    const double mu = c->viscosity;
    const ptrdiff_t sj = (ptrdiff_t)GX(c), sk = (ptrdiff_t)(GX(c) * GY(c));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t g = GI(c, i, j, k);
                const size_t q = II(c, i, j, k);
                const double *U = c->u.f + g, *V = c->v.f + g, *W = c->w.f + g;
                double lu = (((U[1] - 2.0 * U[0]) + U[-1]) + ((U[sj] - 2.0 * U[0]) + U[-sj])) * id2;
                double lv = (((V[1] - 2.0 * V[0]) + V[-1]) + ((V[sj] - 2.0 * V[0]) + V[-sj])) * id2;
                double lw = 0.0;
                if (d3) {
                    lu = lu + ((U[sk] - 2.0 * U[0]) + U[-sk]) * id2;
                    lv = lv + ((V[sk] - 2.0 * V[0]) + V[-sk]) * id2;
                    lw = ((((W[1] - 2.0 * W[0]) + W[-1]) + ((W[sj] - 2.0 * W[0]) + W[-sj])) +
                          ((W[sk] - 2.0 * W[0]) + W[-sk])) * id2;
                }
                rx[q] = rx[q] + mu * lu / rhof;
                ry[q] = ry[q] + mu * lv / rhof;
                if (d3) rz[q] = rz[q] + mu * lw / rhof;
                rx[q] = rx[q] + c->S[0][q] / rhof;
                ry[q] = ry[q] + c->S[1][q] / rhof;
                if (d3) rz[q] = rz[q] + c->S[2][q] / rhof;
            }
}

/* gradient_of_scalar, src/fields.f90:31-64. */
static void gradient(const orc_ctx *c, const double *s, double *gx, double *gy, double *gz)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, d3 = c->dim == 3;
    const double id = 1.0 / c->delta;
    const ptrdiff_t sj = (ptrdiff_t)GX(c), sk = (ptrdiff_t)(GX(c) * GY(c));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const double *S = s + GI(c, i, j, k);
                const size_t q = II(c, i, j, k);
                gx[q] = (S[1] - S[0]) * id;
                gy[q] = (S[sj] - S[0]) * id;
                if (d3) gz[q] = (S[sk] - S[0]) * id;
            }
}

void orc_gradient(orc_ctx *c, const double *s, double *gx, double *gy, double *gz)
{
    gradient(c, s, gx, gy, gz);
}

/* divergence_of_vector, src/fields.f90:120-153; out is interior-sized. */
static void divergence(const orc_ctx *c, const double *u, const double *v, const double *w, double *out)
{
    const int nx = c->nx, ny = c->ny, nz = c->nz, d3 = c->dim == 3;
    const double id = 1.0 / c->delta;
    const ptrdiff_t sj = (ptrdiff_t)GX(c), sk = (ptrdiff_t)(GX(c) * GY(c));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t g = GI(c, i, j, k);
                double d = (u[g] - u[g - 1]) * id + (v[g] - v[g - sj]) * id;
                if (d3) d = d + (w[g] - w[g - sk]) * id;
                out[II(c, i, j, k)] = d;
            }
}

void orc_divergence(orc_ctx *c, const double *u, const double *v, const double *w, double *div)
{
    divergence(c, u, v, w, div);
}

/* -------------------------------------------------------------------------- */
/* navier_stokes_solver, src/navier_stokes.f90:50-136                         */
/* -------------------------------------------------------------------------- */

static double max_velocity_sum(const orc_ctx *c)
{
    double mv = 0.0;
    for (int k = 1; k <= c->nz; k++)
        for (int j = 1; j <= c->ny; j++)
            for (int i = 1; i <= c->nx; i++) {
                const size_t g = GI(c, i, j, k);
                double vel = fabs(c->u.f[g]) + fabs(c->v.f[g]);
                if (c->dim == 3) vel = vel + fabs(c->w.f[g]);
                if (vel > mv) mv = vel;
            }
    return mv;
}

/* update_timestep, src/navier_stokes.f90:670-730: dt = min(dt_conv, dt_visc)
 * (:725), or min(dt_conv, dt_visc, dt_surf) in the MF build (:723); then
 * the 1.1 dt_o cap (:728). */
static void update_timestep(orc_ctx *c, double *dt)
{
    c->dt_o = *dt;
    const double mv = max_velocity_sum(c);
    c->dt_conv = mv > 0.0 ? c->cfl * c->delta / mv : 1.0;
    *dt = c->dt_conv < c->dt_visc ? c->dt_conv : c->dt_visc;
    if (c->mf && c->dt_surf < *dt) *dt = c->dt_surf;
    if (*dt > 1.1 * c->dt_o) *dt = 1.1 * c->dt_o;
}

static void mf_step(orc_ctx *c, double dt, double *maxdiv, double *maxcfl);

void orc_ns_step(orc_ctx *c, int step, double *dt_io, double *maxdiv, double *maxcfl)
{
    (void)step;
    if (c->mf) {   /* the MF build's step */
        if (c->constant_cfl) update_timestep(c, dt_io);                        /* :78 */
        mf_step(c, *dt_io, maxdiv, maxcfl);
        return;
    }
    const int nx = c->nx, ny = c->ny, nz = c->nz, d3 = c->dim == 3;
    if (c->constant_cfl) update_timestep(c, dt_io);                            /* :78 */
    const double dt = *dt_io;
    const double rhof = 0.5 * (c->density + c->density);                       /* center_to_face */
    const size_t ni = interior_len(c);

    /* predicted_velocity_field (:140-213) */
    const double A = 1.0 + 0.5 * dt / c->dt_o;
    const double B = -(0.5 * dt / c->dt_o);
    for (int d = 0; d < 3; d++) memset(c->dv[d], 0, ni * sizeof(double));     /* :227-231 */
    add_advection(c, c->u.f, c->v.f, c->w.f, c->dv[0], c->dv[1], c->dv[2]);   /* :234 */
    add_diffusion_and_source(c, rhof, c->dv[0], c->dv[1], c->dv[2]);          /* :237, :245-255 */
    gradient(c, c->p.f, c->gp[0], c->gp[1], c->gp[2]);                         /* :165 */
    double *vel[3] = {c->u.f, c->v.f, c->w.f};
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t g = GI(c, i, j, k), q = II(c, i, j, k);
                for (int d = 0; d < (d3 ? 3 : 2); d++) {
                    const double rhs = ((-(c->gp[d][q] / rhof) + A * c->dv[d][q]) + B * c->dvo[d][q]) + c->g[d];
                    vel[d][g] = vel[d][g] + dt * rhs;                          /* :187-198 */
                }
            }
    for (int d = 0; d < (d3 ? 3 : 2); d++) memcpy(c->dvo[d], c->dv[d], ni * sizeof(double));  /* :201-205 */
    ghost_fill(c, &c->u); ghost_fill(c, &c->v); if (d3) ghost_fill(c, &c->w); /* :208 */

    /* Poisson right-hand side (:111-121) */
    double *div = c->gp[0];   /* scratch */
    divergence(c, c->u.f, c->v.f, c->w.f, div);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++)
                c->phi.f[GI(c, i, j, k)] = div[II(c, i, j, k)] * c->density / dt;
    orc_poisson_solve(c, c->phi.f, 1);                                          /* :123 */
    ghost_fill(c, &c->phi);                                                     /* :124 */

    /* correct_velocity_field (:505-546) */
    gradient(c, c->phi.f, c->gp[0], c->gp[1], c->gp[2]);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t g = GI(c, i, j, k), q = II(c, i, j, k);
                for (int d = 0; d < (d3 ? 3 : 2); d++)
                    vel[d][g] = vel[d][g] - c->gp[d][q] * dt / rhof;
            }
    ghost_fill(c, &c->u); ghost_fill(c, &c->v); if (d3) ghost_fill(c, &c->w);

    /* update_pressure (:550-566): whole array, then ghosts */
    const size_t ng = ghosted_len(c);
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < ng; q++) c->p.f[q] = c->p.f[q] + c->phi.f[q];
    ghost_fill(c, &c->p);

    /* checks (:570-619) */
    divergence(c, c->u.f, c->v.f, c->w.f, div);
    double md = -HUGE_VAL;
    for (size_t q = 0; q < ni; q++) if (div[q] > md) md = div[q];
    c->maxdiv = md;
    c->maxcfl = dt * max_velocity_sum(c) / c->delta;
    if (maxdiv) *maxdiv = c->maxdiv;
    if (maxcfl) *maxcfl = c->maxcfl;
}

/* -------------------------------------------------------------------------- */
/* Two-phase solver (MF build), 2-D                                            */
/* -------------------------------------------------------------------------- */

static const double kSmall = 1.0e-14;   /* global_mod small (src/global.f90:16) */
static const double kCut = 1.0e-8;      /* cut (src/volume_of_fluid.f90:37) */

/* Two-point Gauss nodes on [0,1] (src/volume_of_fluid.f90:33-34). */
static double gauss_rp(void) { return 0.5 * (1.0 + 1.0 / sqrt(3.0)); }
static double gauss_rm(void) { return 0.5 * (1.0 - 1.0 / sqrt(3.0)); }

/* The quadratic surface of one cell (module variables cx, cy, a10, a01, a20, a02). */
typedef struct { double cx, cy, a10, a01, a20, a02; } osurf;

/* P(x, y), Eq. 9 of Ii et al. 2012 (src/volume_of_fluid.f90:387-398). */
static double surf_P(const osurf *q, double x, double y)
{
    return ((q->cx * q->a20 * (x * x) + q->cy * q->a02 * (y * y)) + q->a10 * x) + q->a01 * y;
}

/* The surface coefficients of cell g (src/volume_of_fluid.f90:260-273, 586-609). */
static osurf cell_surface(const orc_ctx *c, size_t g)
{
    osurf q;
    const double nx = c->nrx.f[g], ny = c->nry.f[g];
    if (fabs(nx) == fmax(fabs(nx), fabs(ny))) { q.cx = 0.0; q.cy = 1.0; }
    else { q.cx = 1.0; q.cy = 0.0; }
    q.a10 = nx - 0.5 * q.cx * c->lx2.f[g];
    q.a01 = ny - 0.5 * q.cy * c->ly2.f[g];
    q.a20 = 0.5 * q.cx * c->lx2.f[g];
    q.a02 = 0.5 * q.cy * c->ly2.f[g];
    return q;
}

static double solve_quadratic(double a, double b, double cc)   /* :402-416 */
{
    const double x1 = (-b + sqrt(b * b - 4.0 * a * cc)) / (2.0 * a);
    const double x2 = (-b - sqrt(b * b - 4.0 * a * cc)) / (2.0 * a);
    return fmax(x1, x2);
}

/* compute_norm, src/volume_of_fluid.f90:299-383 (Young's method). */
static void compute_norm(orc_ctx *c)
{
    const double delta = c->delta, id = 1.0 / delta, id2 = 1.0 / (delta * delta);
    const ptrdiff_t sj = (ptrdiff_t)GX(c);
    const double *f = c->vof.f;
    for (int j = 1; j <= c->ny; j++)
        for (int i = 1; i <= c->nx; i++) {
            const size_t g = GI(c, i, j, 1);
            const double *F = f + g;
#define VF(di, dj) F[(di) + (dj) * sj]
            double mx[4], my[4], nxs[4], nys[4];
            mx[0] = 0.5 * (VF(0, -1) + VF(0, 0) - VF(-1, -1) - VF(-1, 0)) * id;
            mx[1] = 0.5 * (VF(0, 0) + VF(0, 1) - VF(-1, 0) - VF(-1, 1)) * id;
            mx[2] = 0.5 * (VF(1, 0) + VF(1, 1) - VF(0, 0) - VF(0, 1)) * id;
            mx[3] = 0.5 * (VF(1, -1) + VF(1, 0) - VF(0, -1) - VF(0, 0)) * id;
            const double mxc = 0.25 * (mx[0] + mx[1] + mx[2] + mx[3]);
            my[0] = 0.5 * (VF(-1, 0) + VF(0, 0) - VF(-1, -1) - VF(0, -1)) * id;
            my[1] = 0.5 * (VF(-1, 1) + VF(0, 1) - VF(-1, 0) - VF(0, 0)) * id;
            my[2] = 0.5 * (VF(0, 1) + VF(1, 1) - VF(0, 0) - VF(1, 0)) * id;
            my[3] = 0.5 * (VF(0, 0) + VF(1, 0) - VF(0, -1) - VF(1, -1)) * id;
            const double myc = 0.25 * (my[0] + my[1] + my[2] + my[3]);
#undef VF
            for (int q = 0; q < 4; q++) {
                nxs[q] = mx[q] / sqrt(mx[q] * mx[q] + my[q] * my[q] + kSmall);
                nys[q] = my[q] / sqrt(mx[q] * mx[q] + my[q] * my[q] + kSmall);
            }
            c->nrx.f[g] = mxc / sqrt(mxc * mxc + myc * myc + kSmall);
            c->nry.f[g] = myc / sqrt(mxc * mxc + myc * myc + kSmall);
            c->lx2.f[g] = 0.5 * delta * (nxs[3] + nxs[2] - nxs[1] - nxs[0]);   /* quadratic = .true. */
            c->ly2.f[g] = 0.5 * delta * (nys[1] + nys[2] - nys[0] - nys[3]);
            c->curv.f[g] = -(c->lx2.f[g] + c->ly2.f[g]) * id2;
        }
    ghost_fill(c, &c->curv);
    ghost_fill(c, &c->nrx); ghost_fill(c, &c->nry);
    ghost_fill(c, &c->lx2); ghost_fill(c, &c->ly2);
}

/* get_h_from_vof, src/volume_of_fluid.f90:223-295. */
static void get_h_from_vof(orc_ctx *c)
{
    compute_norm(c);
    const double beta = c->beta, rp = gauss_rp(), rm = gauss_rm();
    for (int j = 1; j <= c->ny; j++)
        for (int i = 1; i <= c->nx; i++) {
            const size_t g = GI(c, i, j, 1);
            const double f = c->vof.f[g];
            if (f <= kCut || f >= (1.0 - kCut)) {
                c->h.f[g] = f;
                c->dd.f[g] = 0.0;
                continue;
            }
            const osurf q = cell_surface(c, g);
            const double A = (1.0 - q.cx) * exp(2.0 * beta * q.a10) + (1.0 - q.cy) * exp(2.0 * beta * q.a01);
            const double Bp = (1.0 - q.cx) * exp(2.0 * beta * surf_P(&q, 0.0, rp)) +
                              (1.0 - q.cy) * exp(2.0 * beta * surf_P(&q, rp, 0.0));
            const double Bm = (1.0 - q.cx) * exp(2.0 * beta * surf_P(&q, 0.0, rm)) +
                              (1.0 - q.cy) * exp(2.0 * beta * surf_P(&q, rm, 0.0));
            const double Q = (1.0 - q.cx) * exp(2.0 * beta * q.a10 * (2.0 * f - 1.0)) +
                             (1.0 - q.cy) * exp(2.0 * beta * q.a01 * (2.0 * f - 1.0));
            const double aa = A * Bm * Bp * (A - Q);
            const double bb = A * (Bp + Bm) * (1.0 - Q);
            const double cc = 1.0 - A * Q;
            c->dd.f[g] = log(solve_quadratic(aa, bb, cc)) / (2.0 * beta);
            c->h.f[g] = 0.5 * (1.0 + tanh(beta * (surf_P(&q, 0.5, 0.5) + c->dd.f[g])));
        }
    ghost_fill(c, &c->h);
    ghost_fill(c, &c->dd);
}

/* An_Int and Num_Int, src/volume_of_fluid.f90:636-661. */
static double an_int(const osurf *q, double beta, double a, double b, double xa, double xb, double ya, double yb,
                     double cf, double d0)
{
    return 0.5 * (b - a + 1.0 / (cf * beta) * log(cosh(beta * (surf_P(q, xb, yb) + d0)) /
                                                   cosh(beta * (surf_P(q, xa, ya) + d0))));
}

static double num_int(double a, double b, double qrm, double qrp) { return 0.5 * (qrm + qrp) * (b - a); }

/* compute_flux, src/volume_of_fluid.f90:548-632: flux through the high face of
 * cell (i, j) along dir (1 x, 2 y) with face velocity u. */
static double compute_flux(const orc_ctx *c, int dir, int i, int j, double u, double dt)
{
    const double delta = c->delta, beta = c->beta, rp = gauss_rp(), rm = gauss_rm();
    double xa, xb, ya, yb;
    int sgn, ii = i, jj = j;
    if (dir == 1) {
        if (u >= 0.0) { xa = 1.0 - dt * u / delta; xb = 1.0; sgn = 1; }
        else { xa = 0.0; xb = -dt * u / delta; sgn = -1; ii = i + 1; }
        ya = 0.0; yb = 1.0;
    } else {
        xa = 0.0; xb = 1.0;
        if (u >= 0.0) { ya = 1.0 - dt * u / delta; yb = 1.0; sgn = 1; }
        else { ya = 0.0; yb = -dt * u / delta; sgn = -1; jj = j + 1; }
    }
    const size_t g = GI(c, ii, jj, 1);
    const double f = c->vof.f[g];
    if (f <= kCut || f >= (1.0 - kCut)) return (double)sgn * delta * f * (xb - xa) * (yb - ya);
    const osurf q = cell_surface(c, g);
    const double d0 = c->dd.f[g];
    if (q.cx == 0.0)   /* |norm x| dominant: analytical in x, Gauss in y */
        return (double)sgn * delta *
               num_int(ya, yb, an_int(&q, beta, xa, xb, xa, xb, rm * (ya + yb), rm * (yb + ya), q.a10, d0),
                       an_int(&q, beta, xa, xb, xa, xb, rp * (ya + yb), rp * (yb + ya), q.a10, d0));
    return (double)sgn * delta *
           num_int(xa, xb, an_int(&q, beta, ya, yb, rm * (xa + xb), rm * (xa + xb), ya, yb, q.a01, d0),
                   an_int(&q, beta, ya, yb, rp * (xa + xb), rp * (xa + xb), ya, yb, q.a01, d0));
}

/* One directional sweep: out = (in - (fp - fm)/delta) / (1 - dt (u_h - u_l)/delta),
 * fluxes from the current reconstruction of c->vof. */
static void vof_sweep(orc_ctx *c, int dir, const double *in, double *out, double dt)
{
    const double delta = c->delta;
    const double *vel = dir == 1 ? c->u.f : c->v.f;
    const ptrdiff_t sh = dir == 1 ? 1 : (ptrdiff_t)GX(c);
    for (int j = 1; j <= c->ny; j++)
        for (int i = 1; i <= c->nx; i++) {
            const size_t g = GI(c, i, j, 1);
            const double fp = compute_flux(c, dir, i, j, vel[g], dt);
            const double fm = dir == 1 ? compute_flux(c, 1, i - 1, j, vel[g - sh], dt)
                                       : compute_flux(c, 2, i, j - 1, vel[g - sh], dt);
            out[g] = (in[g] - (fp - fm) / delta) / (1.0 - dt * (vel[g] - vel[g - sh]) / delta);
        }
}

/* advect_vof, src/volume_of_fluid.f90:420-544 (direction order alternates). */
void orc_advect_vof(orc_ctx *c, double dt)
{
    if (!c->mf) return;
    const size_t ng = ghosted_len(c);
    const double delta = c->delta;
    double *vof1 = (double *)calloc(ng, sizeof(double)), *vof2 = (double *)calloc(ng, sizeof(double));
    get_h_from_vof(c);
    const int first = c->x_first ? 1 : 2, second = c->x_first ? 2 : 1;
    vof_sweep(c, first, c->vof.f, vof1, dt);
    memcpy(c->vof.f, vof1, ng * sizeof(double));   /* vof = vof1 (ghosts of vof1 are zero) */
    ghost_fill(c, &c->vof);
    get_h_from_vof(c);
    vof_sweep(c, second, vof1, vof2, dt);
    const ptrdiff_t sj = (ptrdiff_t)GX(c);
    const double *u = c->u.f, *v = c->v.f;
    for (int j = 1; j <= c->ny; j++)
        for (int i = 1; i <= c->nx; i++) {
            const size_t g = GI(c, i, j, 1);
            const double du = u[g] - u[g - 1], dv = v[g] - v[g - sj];
            if (c->x_first)
                c->vof.f[g] = vof2[g] - dt * (vof1[g] * du / delta + vof2[g] * dv / delta);
            else
                c->vof.f[g] = vof2[g] - dt * (vof2[g] * du / delta + vof1[g] * dv / delta);
        }
    c->x_first = !c->x_first;
    ghost_fill(c, &c->vof);
    free(vof1);
    free(vof2);
}

void orc_vof_reconstruct(orc_ctx *c) { if (c->mf) get_h_from_vof(c); }

double *orc_vof_aux(orc_ctx *c, int which)
{
    if (!c->mf) return NULL;
    ofield *a[7] = {&c->h, &c->dd, &c->curv, &c->nrx, &c->nry, &c->lx2, &c->ly2};
    return which >= 0 && which < 7 ? a[which]->f : NULL;
}

/* update_material_properties, src/multiphase.f90:126-140 (whole arrays, then ghosts). */
void orc_update_material_properties(orc_ctx *c)
{
    if (!c->mf) return;
    const size_t ng = ghosted_len(c);
    for (size_t q = 0; q < ng; q++) {
        c->rho.f[q] = c->rho1 * c->vof.f[q] + c->rho0 * (1.0 - c->vof.f[q]);
        c->mu.f[q] = c->mu1 * c->vof.f[q] + c->mu0 * (1.0 - c->vof.f[q]);
    }
    ghost_fill(c, &c->rho);
    ghost_fill(c, &c->mu);
}

int orc_set_multiphase(orc_ctx *c, double rho0, double rho1, double mu0, double mu1, double sigma, double beta)
{
    if (c->dim != 2 || c->mf) return -1;
    for (int s = 0; s < 4; s++)
        if (c->bc[s] != ORC_BC_PERIODIC && c->bc[s] != ORC_BC_WALL) return -1;   /* :54-219 */
    c->mf = 1;
    c->x_first = 1;
    c->rho0 = rho0; c->rho1 = rho1; c->mu0 = mu0; c->mu1 = mu1; c->sigma = sigma; c->beta = beta;
    c->rhomin = fmin(rho0, rho1);                                 /* src/solver.f90:91-92 */
    c->irhomin = 1.0 / c->rhomin;
    ofield *f[12] = {&c->vof, &c->h, &c->dd, &c->curv, &c->nrx, &c->nry, &c->lx2, &c->ly2,
                     &c->rho, &c->mu, &c->p_hat, &c->p_o};
    for (int q = 0; q < 12; q++) {
        field_alloc(c, f[q], 'c');
        /* Periodic -> 0, Wall -> Neumann (vof family :54-219, rho/mu/p_hat/p_o as p) */
        for (int s = 0; s < 4; s++) f[q]->type[s] = c->bc[s] == ORC_BC_PERIODIC ? 0 : 2;
    }
    return 0;
}

/* set_timestep, MF branch (src/navier_stokes.f90:653-660). */
double orc_mf_set_timestep(orc_ctx *c, double U)
{
    const double pi = acos(-1.0);
    c->dt_conv = c->cfl * c->delta / U;
    c->dt_visc = 0.125 * c->delta * c->delta * fmin(c->rho0 / c->mu0, c->rho1 / c->mu1);
    double dt = fmin(c->dt_conv, c->dt_visc);
    if (c->sigma > 0.0) {
        const double dt_surf = sqrt(0.5 * (c->rho0 + c->rho1) * (c->delta * c->delta * c->delta) /
                                    (pi * c->sigma + 1.0e-16));
        dt = fmin(dt, dt_surf);
        c->dt_surf = dt_surf;
        c->have_dt_surf = 1;
    }
    c->dt_o = dt;
    return dt;
}

/* The MF step, src/navier_stokes.f90:50-136 with MF defined (2-D, fixed dt). */
static void mf_step(orc_ctx *c, double dt, double *maxdiv, double *maxcfl)
{
    const int nx = c->nx, ny = c->ny;
    const size_t ni = interior_len(c), ng = ghosted_len(c);
    const ptrdiff_t sj = (ptrdiff_t)GX(c);
    const double id = 1.0 / c->delta;

    orc_advect_vof(c, dt);                                                      /* :82 */
    orc_update_material_properties(c);                                          /* :85 */
    if (c->constant_cfl) {                                                      /* :88 */
        for (size_t q = 0; q < ng; q++)
            c->p_hat.f[q] = c->p_o.f[q] + (dt + c->dt_o) * (c->p.f[q] - c->p_o.f[q]) / c->dt_o;
    } else {
        for (size_t q = 0; q < ng; q++) c->p_hat.f[q] = 2.0 * c->p.f[q] - c->p_o.f[q];   /* :91 */
    }
    ghost_fill(c, &c->p_hat);

    /* predicted_velocity_field (:140-213) */
    const double A = 1.0 + 0.5 * dt / c->dt_o;
    const double B = -(0.5 * dt / c->dt_o);
    double *rfx = (double *)malloc(ni * sizeof(double)), *rfy = (double *)malloc(ni * sizeof(double));
    for (int j = 1; j <= ny; j++)                                               /* center_to_face */
        for (int i = 1; i <= nx; i++) {
            const size_t g = GI(c, i, j, 1), q = II(c, i, j, 1);
            rfx[q] = 0.5 * (c->rho.f[g + 1] + c->rho.f[g]);
            rfy[q] = 0.5 * (c->rho.f[g + sj] + c->rho.f[g]);
        }
    for (int d = 0; d < 3; d++) memset(c->dv[d], 0, ni * sizeof(double));
    add_advection(c, c->u.f, c->v.f, c->w.f, c->dv[0], c->dv[1], c->dv[2]);
    /* add_diffusion, variable viscosity (:405-452), then add_surface_tension (:458-501)
     * and the body force S/rhof (:245-255) */
    const double *U = c->u.f, *V = c->v.f, *M = c->mu.f, *vof = c->vof.f, *K = c->curv.f;
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t g = GI(c, i, j, 1), q = II(c, i, j, 1);
            const ptrdiff_t ip = 1, im = -1, jp = sj, jm = -sj;
            const double txxp = 2.0 * M[g + ip] * (U[g + ip] - U[g]) * id;
            const double txxm = 2.0 * M[g] * (U[g] - U[g + im]) * id;
            const double dtxx = (txxp - txxm) * id;
            const double txyp = 0.25 * (M[g] + M[g + ip] + M[g + jp] + M[g + ip + jp]) *
                                ((U[g + jp] - U[g]) * id + (V[g + ip] - V[g]) * id);
            const double txym = 0.25 * (M[g + jm] + M[g + ip + jm] + M[g] + M[g + ip]) *
                                ((U[g] - U[g + jm]) * id + (V[g + ip + jm] - V[g + jm]) * id);
            const double dtxy = (txyp - txym) * id;
            c->dv[0][q] = c->dv[0][q] + (dtxx + dtxy) / rfx[q];
            const double tyxp = txyp;
            const double tyxm = 0.25 * (M[g + im] + M[g] + M[g + im + jp] + M[g + jp]) *
                                ((U[g + im + jp] - U[g + im]) * id + (V[g] - V[g + im]) * id);
            const double dtyx = (tyxp - tyxm) * id;
            const double tyyp = 2.0 * M[g + jp] * (V[g + jp] - V[g]) * id;
            const double tyym = 2.0 * M[g] * (V[g] - V[g + jm]) * id;
            const double dtyy = (tyyp - tyym) * id;
            c->dv[1][q] = c->dv[1][q] + (dtyx + dtyy) / rfy[q];
        }
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t g = GI(c, i, j, 1), q = II(c, i, j, 1);
            c->dv[0][q] = c->dv[0][q] + c->sigma * 0.5 * (K[g + 1] + K[g]) * (vof[g + 1] - vof[g]) * id / rfx[q];
            c->dv[1][q] = c->dv[1][q] + c->sigma * 0.5 * (K[g + sj] + K[g]) * (vof[g + sj] - vof[g]) * id / rfy[q];
        }
    for (size_t q = 0; q < ni; q++) {
        c->dv[0][q] = c->dv[0][q] + c->S[0][q] / rfx[q];
        c->dv[1][q] = c->dv[1][q] + c->S[1][q] / rfy[q];
    }
    gradient(c, c->p.f, c->gp[0], c->gp[1], c->gp[2]);
    double *gphx = (double *)malloc(ni * sizeof(double)), *gphy = (double *)malloc(ni * sizeof(double));
    gradient(c, c->p_hat.f, gphx, gphy, NULL);
    double *vel[2] = {c->u.f, c->v.f};
    double *rf[2] = {rfx, rfy}, *gh[2] = {gphx, gphy};
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t g = GI(c, i, j, 1), q = II(c, i, j, 1);
            for (int d = 0; d < 2; d++) {
                const double gp = c->gp[d][q], r = rf[d][q];
                double rhs = ((-(gp / r) + A * c->dv[d][q]) + B * c->dvo[d][q]) + c->g[d];   /* :174-176 */
                rhs = ((rhs + gp / r) - c->irhomin * gp) - (1.0 / r - c->irhomin) * gh[d][q]; /* :178-184 */
                vel[d][g] = vel[d][g] + dt * rhs;
            }
        }
    for (int d = 0; d < 2; d++) memcpy(c->dvo[d], c->dv[d], ni * sizeof(double));
    ghost_fill(c, &c->u); ghost_fill(c, &c->v);

    /* Poisson right-hand side with rho_min (:111-112) */
    double *div = c->gp[0];
    divergence(c, c->u.f, c->v.f, c->w.f, div);
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) c->phi.f[GI(c, i, j, 1)] = div[II(c, i, j, 1)];
    for (size_t q = 0; q < ng; q++) c->phi.f[q] = c->phi.f[q] * c->rhomin / dt;   /* whole array */
    orc_poisson_solve(c, c->phi.f, 1);
    ghost_fill(c, &c->phi);

    /* correct_velocity_field, MF (:526-531) */
    gradient(c, c->phi.f, c->gp[0], c->gp[1], c->gp[2]);
    for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
            const size_t g = GI(c, i, j, 1), q = II(c, i, j, 1);
            for (int d = 0; d < 2; d++) vel[d][g] = vel[d][g] - c->gp[d][q] * dt * c->irhomin;
        }
    ghost_fill(c, &c->u); ghost_fill(c, &c->v);

    /* update_pressure (:550-566): p_o = p, p = p + phi */
    memcpy(c->p_o.f, c->p.f, ng * sizeof(double));
    for (size_t q = 0; q < ng; q++) c->p.f[q] = c->p.f[q] + c->phi.f[q];
    ghost_fill(c, &c->p);

    divergence(c, c->u.f, c->v.f, c->w.f, div);
    double md = -HUGE_VAL;
    for (size_t q = 0; q < ni; q++) if (div[q] > md) md = div[q];
    c->maxdiv = md;
    c->maxcfl = dt * max_velocity_sum(c) / c->delta;
    if (maxdiv) *maxdiv = c->maxdiv;
    if (maxcfl) *maxcfl = c->maxcfl;
    free(rfx); free(rfy); free(gphx); free(gphy);
}

/* -------------------------------------------------------------------------- */
/* Discrete adjoint of one step (no counterpart in the reference)             */
/* -------------------------------------------------------------------------- */

static int wrap(int i, int n) { return i < 1 ? i + n : (i > n ? i - n : i); }

/* interior index of (i,j,k) with periodic wrap */
static size_t IW(const orc_ctx *c, int i, int j, int k)
{
    return II(c, wrap(i, c->nx), wrap(j, c->ny), wrap(k, c->nz));
}

/* state value of a ghosted field at (i,j,k) with periodic wrap */
static double SV(const orc_ctx *c, const double *f, int i, int j, int k)
{
    return f[GI(c, wrap(i, c->nx), wrap(j, c->ny), wrap(k, c->nz))];
}

int orc_adjoint_step(orc_ctx *c, double dt, double *ub, double *vb, double *wb, double *pb, double *dvobx,
                     double *dvoby, double *dvobz)
{
    if (c->dim != 3 || c->solver != ORC_PPP) return -1;
    const int nx = c->nx, ny = c->ny, nz = c->nz;
    const size_t ni = interior_len(c);
    const double id = 1.0 / c->delta, id2 = 1.0 / (c->delta * c->delta);
    const double rho = c->density, rhof = 0.5 * (c->density + c->density), mu = c->viscosity;
    const double A = 1.0 + 0.5 * dt / c->dt_o, B = -(0.5 * dt / c->dt_o);
    double *bar[3] = {ub, vb, wb}, *dvob[3] = {dvobx, dvoby, dvobz};
    double *psi = (double *)calloc(ni, sizeof(double));
    double *star[3], *dvb[3], *out[3];
    for (int d = 0; d < 3; d++) {
        star[d] = (double *)malloc(ni * sizeof(double));
        dvb[d] = (double *)malloc(ni * sizeof(double));
        out[d] = (double *)calloc(ni, sizeof(double));
    }

    /* projection + pressure update, transposed:
     *   phibar = pbar' + (dt/rhof) D ubar',  rhsbar = S phibar,
     *   ubar* = ubar' - (rho/dt) G rhsbar  (G^T = -D on the periodic grid) */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t q = II(c, i, j, k);
                const double div = ((ub[q] - ub[IW(c, i - 1, j, k)]) * id + (vb[q] - vb[IW(c, i, j - 1, k)]) * id) +
                                   (wb[q] - wb[IW(c, i, j, k - 1)]) * id;
                psi[q] = pb[q] + (dt / rhof) * div;
            }
    orc_poisson_solve(c, psi, 0);   /* S is symmetric */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t q = II(c, i, j, k);
                const double s = rho / dt;
                star[0][q] = ub[q] - s * (psi[IW(c, i + 1, j, k)] - psi[q]) * id;
                star[1][q] = vb[q] - s * (psi[IW(c, i, j + 1, k)] - psi[q]) * id;
                star[2][q] = wb[q] - s * (psi[IW(c, i, j, k + 1)] - psi[q]) * id;
            }

    /* predictor u* = u + dt(-G p/rhof + A dv + B dv_o + g), dv_o' = dv, transposed */
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t q = II(c, i, j, k);
                const double div = ((star[0][q] - star[0][IW(c, i - 1, j, k)]) * id +
                                    (star[1][q] - star[1][IW(c, i, j - 1, k)]) * id) +
                                   (star[2][q] - star[2][IW(c, i, j, k - 1)]) * id;
                pb[q] = pb[q] + (dt / rhof) * div;
            }
    for (int d = 0; d < 3; d++)
        for (size_t q = 0; q < ni; q++) {
            dvb[d][q] = dt * A * star[d][q] + dvob[d][q];
            dvob[d][q] = dt * B * star[d][q];
            out[d][q] = star[d][q];
        }

    /* dv = adv(u) + mu lap(u)/rhof, transposed at the state u^n: each forward
     * term's partial derivatives times the adjoint of its output face */
    const double *U = c->u.f, *V = c->v.f, *W = c->w.f;
#define ACC(d, ii, jj, kk, val) (out[d][IW(c, (ii), (jj), (kk))] += (val))
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const size_t q = II(c, i, j, k);
                /* x-momentum at the u face (i,j,k) */
                {
                    const double a = dvb[0][q];
                    const double u0 = SV(c, U, i, j, k);
                    double t;
                    t = -a * id * 0.5 * (SV(c, U, i + 1, j, k) + u0);          /* uuip */
                    ACC(0, i + 1, j, k, t); ACC(0, i, j, k, t);
                    t = a * id * 0.5 * (SV(c, U, i - 1, j, k) + u0);           /* uuim */
                    ACC(0, i - 1, j, k, t); ACC(0, i, j, k, t);
                    t = -a * id * 0.25 * (SV(c, V, i + 1, j, k) + SV(c, V, i, j, k));   /* uvjp: d/du */
                    ACC(0, i, j + 1, k, t); ACC(0, i, j, k, t);
                    t = -a * id * 0.25 * (SV(c, U, i, j + 1, k) + u0);                  /* uvjp: d/dv */
                    ACC(1, i + 1, j, k, t); ACC(1, i, j, k, t);
                    t = a * id * 0.25 * (SV(c, V, i + 1, j - 1, k) + SV(c, V, i, j - 1, k));   /* uvjm */
                    ACC(0, i, j, k, t); ACC(0, i, j - 1, k, t);
                    t = a * id * 0.25 * (u0 + SV(c, U, i, j - 1, k));
                    ACC(1, i + 1, j - 1, k, t); ACC(1, i, j - 1, k, t);
                    t = -a * id * 0.25 * (SV(c, W, i + 1, j, k) + SV(c, W, i, j, k));   /* uwkp */
                    ACC(0, i, j, k + 1, t); ACC(0, i, j, k, t);
                    t = -a * id * 0.25 * (SV(c, U, i, j, k + 1) + u0);
                    ACC(2, i + 1, j, k, t); ACC(2, i, j, k, t);
                    t = a * id * 0.25 * (SV(c, W, i + 1, j, k - 1) + SV(c, W, i, j, k - 1));   /* uwkm */
                    ACC(0, i, j, k, t); ACC(0, i, j, k - 1, t);
                    t = a * id * 0.25 * (u0 + SV(c, U, i, j, k - 1));
                    ACC(2, i + 1, j, k - 1, t); ACC(2, i, j, k - 1, t);
                }
                /* y-momentum at the v face (i,j,k) */
                {
                    const double b = dvb[1][q];
                    const double v0 = SV(c, V, i, j, k);
                    double t;
                    t = -b * id * 0.25 * (SV(c, U, i, j + 1, k) + SV(c, U, i, j, k));   /* vuip: d/dv */
                    ACC(1, i + 1, j, k, t); ACC(1, i, j, k, t);
                    t = -b * id * 0.25 * (SV(c, V, i + 1, j, k) + v0);                  /* vuip: d/du */
                    ACC(0, i, j + 1, k, t); ACC(0, i, j, k, t);
                    t = b * id * 0.25 * (SV(c, U, i - 1, j + 1, k) + SV(c, U, i - 1, j, k));   /* vuim */
                    ACC(1, i, j, k, t); ACC(1, i - 1, j, k, t);
                    t = b * id * 0.25 * (v0 + SV(c, V, i - 1, j, k));
                    ACC(0, i - 1, j + 1, k, t); ACC(0, i - 1, j, k, t);
                    t = -b * id * 0.5 * (SV(c, V, i, j + 1, k) + v0);                   /* vvjp */
                    ACC(1, i, j + 1, k, t); ACC(1, i, j, k, t);
                    t = b * id * 0.5 * (SV(c, V, i, j - 1, k) + v0);                    /* vvjm */
                    ACC(1, i, j - 1, k, t); ACC(1, i, j, k, t);
                    t = -b * id * 0.25 * (SV(c, W, i, j + 1, k) + SV(c, W, i, j, k));   /* vwkp */
                    ACC(1, i, j, k + 1, t); ACC(1, i, j, k, t);
                    t = -b * id * 0.25 * (SV(c, V, i, j, k + 1) + v0);
                    ACC(2, i, j + 1, k, t); ACC(2, i, j, k, t);
                    t = b * id * 0.25 * (SV(c, W, i, j + 1, k - 1) + SV(c, W, i, j, k - 1));   /* vwkm */
                    ACC(1, i, j, k, t); ACC(1, i, j, k - 1, t);
                    t = b * id * 0.25 * (v0 + SV(c, V, i, j, k - 1));
                    ACC(2, i, j + 1, k - 1, t); ACC(2, i, j, k - 1, t);
                }
                /* z-momentum at the w face (i,j,k) */
                {
                    const double e = dvb[2][q];
                    const double w0 = SV(c, W, i, j, k);
                    double t;
                    t = -e * id * 0.25 * (SV(c, U, i, j, k) + SV(c, U, i, j, k + 1));   /* wuip: d/dw */
                    ACC(2, i, j, k, t); ACC(2, i + 1, j, k, t);
                    t = -e * id * 0.25 * (w0 + SV(c, W, i + 1, j, k));                  /* wuip: d/du */
                    ACC(0, i, j, k, t); ACC(0, i, j, k + 1, t);
                    t = e * id * 0.25 * (SV(c, U, i - 1, j, k) + SV(c, U, i - 1, j, k + 1));   /* wuim */
                    ACC(2, i, j, k, t); ACC(2, i - 1, j, k, t);
                    t = e * id * 0.25 * (w0 + SV(c, W, i - 1, j, k));
                    ACC(0, i - 1, j, k, t); ACC(0, i - 1, j, k + 1, t);
                    t = -e * id * 0.25 * (SV(c, V, i, j, k) + SV(c, V, i, j, k + 1));   /* wvjp */
                    ACC(2, i, j, k, t); ACC(2, i, j + 1, k, t);
                    t = -e * id * 0.25 * (w0 + SV(c, W, i, j + 1, k));
                    ACC(1, i, j, k, t); ACC(1, i, j, k + 1, t);
                    t = e * id * 0.25 * (SV(c, V, i, j - 1, k) + SV(c, V, i, j - 1, k + 1));   /* wvjm */
                    ACC(2, i, j, k, t); ACC(2, i, j - 1, k, t);
                    t = e * id * 0.25 * (w0 + SV(c, W, i, j - 1, k));
                    ACC(1, i, j - 1, k, t); ACC(1, i, j - 1, k + 1, t);
                    t = -e * id * 0.5 * (w0 + SV(c, W, i, j, k + 1));                   /* wwkp */
                    ACC(2, i, j, k, t); ACC(2, i, j, k + 1, t);
                    t = e * id * 0.5 * (w0 + SV(c, W, i, j, k - 1));                    /* wwkm */
                    ACC(2, i, j, k, t); ACC(2, i, j, k - 1, t);
                }
                /* diffusion: mu/rhof times the (symmetric) periodic Laplacian */
                for (int d = 0; d < 3; d++) {
                    const double t = mu * dvb[d][q] * id2 / rhof;
                    ACC(d, i + 1, j, k, t); ACC(d, i - 1, j, k, t);
                    ACC(d, i, j + 1, k, t); ACC(d, i, j - 1, k, t);
                    ACC(d, i, j, k + 1, t); ACC(d, i, j, k - 1, t);
                    ACC(d, i, j, k, -6.0 * t);
                }
            }
#undef ACC
    for (int d = 0; d < 3; d++) {
        memcpy(bar[d], out[d], ni * sizeof(double));
        free(star[d]); free(dvb[d]); free(out[d]);
    }
    free(psi);
    return 0;
}
