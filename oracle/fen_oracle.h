/*
 * fen_oracle.h -- CPU restatement of FEN's single-phase fractional-step
 * Navier-Stokes time advance, used ONLY as the correctness checker and as the
 * CPU baseline.  Nothing in the product library links or calls this code.
 *
 * All arrays are Fortran-ordered (i fastest) real(dp) with one ghost layer,
 * exactly as FEN allocates them (src/scalar.f90:79-81): a ghosted field has
 * (nx+2)*(ny+2)*(nz+2) doubles, index (i,j,k) -> i + (nx+2)*(j + (ny+2)*k)
 * with i in 0..nx+1.  In 2-D nz == 1 and k spans 0..2.
 */
#ifndef FEN_ORACLE_H
#define FEN_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* Field ids (same numbering as the product C-ABI). */
enum {
    ORC_P = 0, ORC_PHI = 1, ORC_U = 2, ORC_V = 3, ORC_W = 4,
    ORC_DVO_X = 5, ORC_DVO_Y = 6, ORC_DVO_Z = 7,
    ORC_S_X = 8, ORC_S_Y = 9, ORC_S_Z = 10,
    ORC_VOF = 11, ORC_RHO = 12, ORC_MU = 13,   /* two-phase (MF) fields, ghosted */
    ORC_NFIELDS = 14
};

/* Physical boundary codes (grid%boundary_conditions strings). */
enum { ORC_BC_PERIODIC = 0, ORC_BC_WALL = 1, ORC_BC_INFLOW = 2, ORC_BC_OUTFLOW = 3 };

/* Poisson solver variants (src/poisson.f90:70-111). */
enum { ORC_PP = 0, ORC_PN = 1, ORC_NN = 2, ORC_PPP = 3, ORC_PPN = 4, ORC_NPN = 5, ORC_NNN = 6 };

/* grid%setup + init_solver (allocate_navier_stokes_fields + init_poisson_solver).
 * Returns NULL for an unsupported Poisson boundary combination. */
orc_ctx *orc_create(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                    const double *x0, const int *bc);
void orc_destroy(orc_ctx *c);
int orc_solver_kind(const orc_ctx *c);
double orc_delta(const orc_ctx *c);

void orc_set_params(orc_ctx *c, double density, double viscosity, const double *g,
                    double cfl, int constant_cfl);
double orc_set_timestep(orc_ctx *c, double U);       /* set_timestep: returns dt, sets dt_o */
void orc_set_dt_o(orc_ctx *c, double dt_o);
double orc_get_dt_o(const orc_ctx *c);
/* The module variables dt_visc and dt_conv (src/navier_stokes.f90:27). */
void orc_get_dt_limits(const orc_ctx *c, double *dt_visc, double *dt_conv);

/* Direct access to the storage of a field (ghosted for p, phi, u, v, w). */
double *orc_field(orc_ctx *c, int field);
size_t orc_field_len(const orc_ctx *c, int field);
/* Boundary value plane of a ghosted field; face 0..5 = left,right,bottom,top,front,back. */
double *orc_bc_plane(orc_ctx *c, int field, int face);
size_t orc_bc_plane_len(const orc_ctx *c, int face);
int orc_bc_type(const orc_ctx *c, int field, int face);
void orc_set_bc_type(orc_ctx *c, int field, int face, int type);

void orc_update_ghost_nodes(orc_ctx *c, int field);

/* One navier_stokes_solver call. */
void orc_ns_step(orc_ctx *c, int step, double *dt, double *maxdiv, double *maxcfl);

/* Two-phase (MF build) solver, 2-D: MTHINC volume of fluid (src/volume_of_fluid.f90),
 * material properties rho, mu from the vof (src/multiphase.f90:126-140), CSF surface
 * tension and the Dodd & Ferrante pressure splitting (src/navier_stokes.f90:80-96,
 * 174-184, 526-531).  orc_set_multiphase switches the context to it (returns 0, or -1
 * for 3-D or for faces other than Periodic/Wall); after the vof is written,
 * orc_update_material_properties forms rho and mu as init_solver does
 * (src/solver.f90:85-93).  orc_ns_step then runs the MF step; with constant_CFL
 * it takes update_timestep's MF branch (src/navier_stokes.f90:723) and the
 * p_hat extrapolation with dt_o (:88-89).  The reference leaves dt_surf unset
 * when sigma == 0 (:29, 657), so dt would become min(..., 0) there: that
 * combination is left to the caller (the GPU path refuses it). */
int orc_set_multiphase(orc_ctx *c, double rho0, double rho1, double mu0, double mu1, double sigma,
                       double beta);
void orc_update_material_properties(orc_ctx *c);
double orc_mf_set_timestep(orc_ctx *c, double U);    /* set_timestep, MF branch */
void orc_advect_vof(orc_ctx *c, double dt);          /* advect_vof with the resident u, v */
void orc_vof_reconstruct(orc_ctx *c);                /* get_h_from_vof (with compute_norm) */
/* Reconstruction fields: 0 h, 1 d, 2 curv, 3 norm x, 4 norm y, 5 l x, 6 l y (ghosted). */
double *orc_vof_aux(orc_ctx *c, int which);

/* Operators on host arrays shaped like the context's fields. */
void orc_add_advection(orc_ctx *c, double *rx, double *ry, double *rz);   /* interior-sized (gl=0) */
void orc_divergence(orc_ctx *c, const double *u, const double *v, const double *w, double *div); /* div gl=0 */
void orc_gradient(orc_ctx *c, const double *s, double *gx, double *gy, double *gz);               /* g gl=0 */

/* solve_poisson on an array with ghost level gl (0 or 1), in place. */
void orc_poisson_solve(orc_ctx *c, double *f, int gl);

/* 1-D transforms (for pinning against brute force). sign -1 forward, +1 backward. */
void orc_fft_c2c(int n, double *re, double *im, int sign);
void orc_fft_r2c(int n, const double *x, double *re, double *im);      /* n/2+1 outputs */
void orc_fft_c2r(int n, const double *re, const double *im, double *x);

/* Thomas sweep of poisson_solver_ppn for one complex column (length n). */
void orc_thomas(int n, const double *a, const double *b, const double *cc, double lam_x, double lam_y,
                double *re, double *im);

int orc_num_threads(void);
void orc_set_num_threads(int n);   /* OpenMP threads of the following calls */

/* Discrete adjoint of one orc_ns_step (3-D, triply periodic, ppp, no source,
 * fixed dt; no counterpart in the reference).  The context's u, v, w hold the
 * state u^n the step started from.  In: ub/vb/wb = adjoint of the new
 * velocity, pb = adjoint of the new pressure, dvob[3] = adjoint of the new
 * dv_o; out (in place): the adjoints of the step's inputs u^n, p^n, dv_o^n.
 * All arrays interior-sized (nx*ny*nz).  Returns 0, or -1 when the context
 * is not of that kind.  Written as the transposed partial derivatives of
 * each forward expression, scattered term by term, serially. */
int orc_adjoint_step(orc_ctx *c, double dt, double *ub, double *vb, double *wb, double *pb, double *dvobx,
                     double *dvoby, double *dvobz);

#ifdef __cplusplus
}
#endif
#endif
