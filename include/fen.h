/*
 * fen.h -- C-ABI of libfen, the MI355X (gfx950) implementation of FEN's
 * single-phase fractional-step Navier-Stokes time advance.
 *
 * One context owns one grid, the navier_stokes_mod fields and the Poisson
 * solver tables, all resident in HBM.  Every entry point returns an int
 * status (FEN_OK == 0); on failure fen_last_error() describes it.  No C++
 * exception crosses this boundary.  Arrays passed in or out are host
 * arrays in FEN's own layout: Fortran order (i fastest), real(8), and for
 * ghosted fields the full (nx+2)*(ny+2)*(nz+2) extent of
 * f(lo-1:hi+1, ...) (reference src/scalar.f90:79-81).  In 2-D nz == 1 and
 * the ghosted k extent is 0:2.
 *
 * Threading: one host thread per context; all device work is queued on the
 * context's own HIP stream.  Entry points that return scalars computed on
 * the device synchronise that stream; fen_ns_step does so only when asked
 * for maxdiv/maxcfl or when the time step is CFL-controlled.
 */
#ifndef FEN_H
#define FEN_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FEN_OK 0
#define FEN_ERR_ARG (-1)          /* bad argument (size, id, pointer) */
#define FEN_ERR_UNSUPPORTED (-2)  /* boundary combination / size not supported */
#define FEN_ERR_HIP (-3)          /* HIP runtime failure */
#define FEN_ERR_NONFINITE (-4)    /* NaN/Inf met by the step's checks */
#define FEN_ERR_STATE (-5)        /* call made in the wrong state (e.g. dt_o unset) */

typedef struct fen_ctx fen_ctx;

/* Field ids: module variables of navier_stokes_mod
 * (reference src/navier_stokes.f90:36-37). */
enum {
    FEN_P = 0,      /* pressure p, cell centre, ghosted */
    FEN_PHI = 1,    /* pressure increment phi, cell centre, ghosted */
    FEN_U = 2,      /* v%x, x faces, ghosted */
    FEN_V = 3,      /* v%y, y faces, ghosted */
    FEN_W = 4,      /* v%z, z faces, ghosted (3-D only) */
    FEN_DVO_X = 5,  /* dv_o%x, interior only (gl = 0) */
    FEN_DVO_Y = 6,
    FEN_DVO_Z = 7,
    FEN_S_X = 8,    /* volumetric source S%x, interior only */
    FEN_S_Y = 9,
    FEN_S_Z = 10,
    FEN_VOF = 11,   /* two-phase mode: volume of fluid (volume_of_fluid_mod vof), ghosted */
    FEN_RHO = 12,   /* two-phase mode: density rho (navier_stokes_mod rho), ghosted, read-only */
    FEN_MU = 13     /* two-phase mode: viscosity mu, ghosted, read-only */
};

/* Physical boundary codes: the strings of grid%boundary_conditions
 * (reference src/grid.f90:41-46, mapped in src/navier_stokes.f90:779-1017). */
enum { FEN_BC_PERIODIC = 0, FEN_BC_WALL = 1, FEN_BC_INFLOW = 2, FEN_BC_OUTFLOW = 3 };

/* Faces, in FEN's fill order (reference src/scalar.f90:254-388). */
enum { FEN_LEFT = 0, FEN_RIGHT = 1, FEN_BOTTOM = 2, FEN_TOP = 3, FEN_FRONT = 4, FEN_BACK = 5 };

/* Arithmetic of the step. */
enum {
    FEN_FP64 = 0,        /* everything in FP64, the reference's real(dp) */
    FEN_FP32_REFINE = 1  /* fields and spectrum stored in FP32; momentum stencils and
                            FFTs computed in FP32; the divergence residual (Poisson
                            right-hand side and checks), the projection update, the
                            eigenvalue divides and the tridiagonal elimination in FP64,
                            rounded once to FP32.  Each step's solve takes the FP64
                            residual of the previous step as part of its right-hand
                            side (div u* includes div u^n): the defect correction is
                            carried by the next solve, not run as a separate one. */
};

/* Poisson solver kinds reported by fen_solver_kind (reference src/poisson.f90:70-111). */
enum { FEN_SOLVER_PP = 0, FEN_SOLVER_PN = 1, FEN_SOLVER_NN = 2, FEN_SOLVER_PPP = 3,
       FEN_SOLVER_PPN = 4, FEN_SOLVER_NPN = 5, FEN_SOLVER_NNN = 6 };

/* Replaces grid%setup (src/grid.f90:67-200) + init_solver (src/solver.f90:34-99):
 * allocate_navier_stokes_fields (src/navier_stokes.f90:752-1067) and
 * init_poisson_solver (src/poisson.f90:57-113).  dim is 2 or 3 (CPP DIM);
 * bc[6] holds FEN_BC_* codes for left,right,bottom,top,front,back; the
 * periodic tuple selects the solver: pp, pn, nn (2-D), ppp, ppn, npn, nnn
 * (3-D).  Grid sizes must be powers of two (8..1024) in the transform
 * directions (and in y for pn); the tridiagonal direction needs >= 2 cells.
 * pn, nn, npn and nnn are FP64 only.  Returns FEN_ERR_UNSUPPORTED for those
 * limits and where the reference stops (src/poisson.f90:91-94, 106-109). */
int fen_create(fen_ctx **ctx, int dim, int nx, int ny, int nz, double lx, double ly, double lz,
               const double *x0, const int *bc, int precision);

/* Replaces destroy_solver (src/solver.f90:333-358). */
int fen_destroy(fen_ctx *ctx);

/* Message of the last failure on this context ("" if none). */
const char *fen_last_error(const fen_ctx *ctx);

/* Sets the module variables density, viscosity, g, CFL, constant_CFL
 * (src/navier_stokes.f90:18-42). */
int fen_set_params(fen_ctx *ctx, double density, double viscosity, const double *g, double cfl,
                   int constant_cfl);

/* Replaces set_timestep(comp_grid, dt, U) (src/navier_stokes.f90:623-666):
 * returns dt and sets dt_o = dt. */
int fen_set_timestep(fen_ctx *ctx, double U, double *dt);

/* Assigns the module variable dt_o (src/navier_stokes.f90:33). */
int fen_set_dt_o(fen_ctx *ctx, double dt_o);

/* Reads back maxdiv, maxCFL (src/navier_stokes.f90:24), dt_o, dt_visc and
 * dt_conv.  Synchronises.  Returns FEN_ERR_NONFINITE if the last step's
 * checks met a NaN or Inf.  maxdiv and maxcfl read 0 until a step has run
 * after the last velocity upload.  Any pointer may be NULL. */
int fen_get_scalars(fen_ctx *ctx, double *maxdiv, double *maxcfl, double *dt_o, double *dt_visc,
                    double *dt_conv);

/* Number of doubles of a field in host layout (ghosted or interior). */
int fen_field_count(const fen_ctx *ctx, int field, size_t *count);

/* Host <-> HBM copies of a whole field (synchronous).  Uploading a source
 * component switches the source term S/rhof on (src/navier_stokes.f90:245-255). */
int fen_upload(fen_ctx *ctx, int field, const double *host, size_t count);
int fen_download(fen_ctx *ctx, int field, double *host, size_t count);

/* Boundary value plane of a ghosted field: f%bc%left ... f%bc%back
 * (src/scalar.f90:28-37).  Left/right planes are (0:ny+1,0:nz+1), bottom/top
 * (0:nx+1,0:nz+1), front/back (0:nx+1,0:ny+1), first index fastest. */
int fen_set_bc_plane(fen_ctx *ctx, int field, int face, const double *plane, size_t count);

/* Overrides a field's bc type (f%bc%type_*: 0 periodic, 1 Dirichlet,
 * 2 Neumann; src/scalar.f90:17-26), as user programs do
 * (test/small_test/poisson/projection/projection.f90:68-77). */
int fen_set_bc_type(fen_ctx *ctx, int field, int face, int type);
int fen_get_bc_type(const fen_ctx *ctx, int field, int face, int *type);

/* Replaces scalar%update_ghost_nodes (src/scalar.f90:223-396) and, on one
 * GPU, update_halos (src/halo.f90:12-43) for a resident field. */
int fen_update_ghost_nodes(fen_ctx *ctx, int field);

/* Replaces navier_stokes_solver(comp_grid, step, dt) (src/navier_stokes.f90:50-136).
 * dt is in/out (it changes only when constant_CFL is set).  maxdiv and
 * maxcfl may be NULL; then the step is queued without a host sync and the
 * values are fetched later with fen_get_scalars. */
int fen_ns_step(fen_ctx *ctx, int step, double *dt, double *maxdiv, double *maxcfl);

/* Runs nsteps consecutive steps (step0+1 .. step0+nsteps) without host
 * synchronisation; for fixed-dt runs and benchmarks.  On a 3-D grid periodic
 * in every direction, periodic in x and y with no-slip walls at z, or (FP64)
 * a grid of the npn or nnn solver whose faces are periodic pairs or walls
 * (Dirichlet/Neumann), at fixed dt, each step's projection and pressure
 * update are fused with the next step's momentum (the state between the
 * steps of the run is not stored); the result equals nsteps fen_ns_step
 * calls bit for bit.  fen_ns_step on those grids leaves its projection
 * pending in the same way until the state is read or changed.
 * FEN_FUSED_RUN=0 in the environment at fen_create turns the fusion off. */
int fen_ns_run(fen_ctx *ctx, int step0, int nsteps, double *dt);

/* Two-phase mode (FEN's MF build), 2-D, FP64, fixed dt: replaces
 * allocate_vof_fields + allocate_multiphase_fields and the rho_0, rho_1,
 * mu_0, mu_1, sigma (multiphase_mod, src/multiphase.f90:13-20) and beta
 * (volume_of_fluid_mod, src/volume_of_fluid.f90:24) settings of init_solver
 * (src/solver.f90:81-94).  After it, fen_ns_step runs the MF step
 * (src/navier_stokes.f90:80-96 with MF): MTHINC advection of the vof,
 * material properties, p_hat, momentum with variable rho and mu and CSF
 * surface tension, the Poisson solve with rho_min, projection with 1/rho_min,
 * p_o = p.  fen_set_timestep takes the MF branch (:653-660).  Faces must be
 * Periodic or Wall (the reference stops otherwise).  With constant CFL the
 * step takes update_timestep's MF branch, min(dt_conv, dt_visc, dt_surf)
 * (:723), and extrapolates p_hat with dt_o (:88-89); the reference sets
 * dt_surf only when sigma > 0 (:657), so constant CFL needs sigma > 0 and a
 * fen_set_timestep call after this one (FEN_ERR_UNSUPPORTED otherwise).
 * Returns FEN_ERR_UNSUPPORTED for 3-D or FP32. */
int fen_set_multiphase(fen_ctx *ctx, double rho0, double rho1, double mu0, double mu1, double sigma,
                       double beta);

/* update_material_properties (src/multiphase.f90:126-140): rho and mu from the
 * resident vof; call after uploading FEN_VOF, as init_solver does after
 * get_vof_from_distance (src/solver.f90:88-89). */
int fen_update_material_properties(fen_ctx *ctx);

/* advect_vof(v, dt) alone (src/volume_of_fluid.f90:420-544) with the resident
 * velocity, as the reference's volume-of-fluid tests call it. */
int fen_advect_vof(fen_ctx *ctx, double dt);

/* The order of the vof's directional sweeps: volume_of_fluid_mod's x_first
 * (src/volume_of_fluid.f90:30), x sweep first when 1.  Every advection
 * (fen_ns_step, fen_ns_run, fen_advect_vof) flips it (:488, :532); it starts
 * at 1 with fen_set_multiphase and is not part of the restart file, so a
 * restarted run sets it as the reference's restart test does
 * (test/small_test/io/test_MF.f90:75-80).  FEN_ERR_STATE outside the
 * two-phase mode. */
int fen_set_vof_x_first(fen_ctx *ctx, int x_first);
int fen_get_vof_x_first(fen_ctx *ctx, int *x_first);

/* Replaces the solve_poisson procedure pointer (src/poisson.f90:21-29):
 * solves lap(phi) = f in place on a host array with ghost level gl (0 or 1),
 * with the same whole-array semantics as the selected reference solver. */
int fen_poisson_solve(fen_ctx *ctx, double *f, int gl);

/* fen_poisson_solve without any allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight loops, harnesses that forbid allocation).  Same arguments,
 * results and whole-array semantics as fen_poisson_solve; the block sums and
 * the mean of the pp/pn/ppn solvers go to a fixed device array of 256 KiB
 * (32768 block sums: ny*nz/(4096/nx) must not exceed that, else
 * FEN_ERR_UNSUPPORTED), of which each GPU holds one copy until the process
 * ends; nothing is allocated on the device or the host.  The context's
 * fields and spectrum (allocated by fen_create) are used as by
 * fen_poisson_solve, except that phi's pending mean is left alone.  FP64
 * contexts only.  Calls that overlap on one GPU (two host threads, or
 * contexts on different streams) share that array and race. */
int fen_poisson_solve_noalloc(fen_ctx *ctx, double *f, int gl);

/* Replaces add_advection(comp_grid, RHS) (src/navier_stokes.f90:261-353) on
 * the resident velocity: rx/ry/rz are interior-sized host arrays updated as
 * RHS = RHS - (advective terms). rz may be NULL in 2-D. */
int fen_add_advection(fen_ctx *ctx, double *rx, double *ry, double *rz);

/* Replace save_state / load_state (src/solver.f90:160-329): the restart
 * file at `path` holds the interiors of p, u, v, dv_o%x, dv_o%y, then in 3-D
 * w, dv_o%z, then in the two-phase mode vof, rho, mu, p_o, each as raw
 * float64 in Fortran order (the reference's order and format).  Loading
 * replaces those interiors and refills the ghosts of p and v (and of the
 * two-phase fields), as the reference does; a file whose size does not
 * match this context's grid and mode is refused with FEN_ERR_ARG. */
int fen_save_state(fen_ctx *ctx, const char *path);
int fen_load_state(fen_ctx *ctx, const char *path);

/* Replaces print_navier_stokes_solver_status (src/navier_stokes.f90:734-748):
 * formats the status line (no newline) into buf. */
int fen_format_status(fen_ctx *ctx, int step, double time, double dt, char *buf, size_t len);

/* Writes grid.json exactly as grid%print_json (src/grid.f90:233-265). */
int fen_write_grid_json(const fen_ctx *ctx, const char *path);

/* The same two texts without a context: the status line of the given values
 * (no newline), and grid.json's whole text (with its newlines).  Return
 * FEN_ERR_ARG if buf is too small.  fen_format_status and
 * fen_write_grid_json write these texts. */
int fen_status_text(int step, double time, double dt, double maxdiv, double maxcfl, char *buf, size_t len);
int fen_grid_json_text(int nx, int ny, int nz, const double *x0, double lx, double ly, double lz, char *buf,
                       size_t len);

/* Case file: a JSON description of a run's setup, read by fen_case_read.
 * FEN has no input file: its programs call grid%setup (src/grid.f90:67),
 * init_solver (src/solver.f90:34), assign navier_stokes_mod's module
 * variables (src/navier_stokes.f90:17-42), call set_timestep (:623) and
 * write the initial fields, then loop over advance_solution
 * (src/solver.f90:28) with print_solver_status, save_state and save_fields.
 * The case file holds that setup as data.  Its "Grid" object has the keys of
 * the grid.json that grid%print_json writes (src/grid.f90:246-258):
 *
 *   { "Grid": { "Nx": 256, "Ny": 256, "Nz": 256, "origin": [0, 0, 0],
 *               "Lx": 6.283185307179586, "Ly": ..., "Lz": ...,
 *               "boundary_conditions": ["Periodic", ... 4 (2-D) or 6 faces],
 *               "name": "grid" },
 *     "Physics":  { "density": 1, "viscosity": 1, "g": [0, 0, 0], "CFL": 1,
 *                   "constant_CFL": false },
 *     "Timestep": { "dt": 1e-3 }  or  { "U": 2.0 }   (set_timestep(comp_grid, dt, U)),
 *     "Precision": "FP64" | "FP32_REFINE",
 *     "Initial":  { "u": "u.raw", "v": ..., "w": ..., "p": ..., "vof": ... },
 *     "Restart":  { "step": 5, "dir": "data" },
 *     "Run":      { "steps": 100, "status_every": 1, "save_state": [50, 100],
 *                   "save_fields": [100], "data_dir": "data", "write_final": false },
 *     "Multiphase": { "rho_0": .., "rho_1": .., "mu_0": .., "mu_1": .., "sigma": 0,
 *                     "interface": { "level": .., "amplitude": .., "wavelength": .. } } }
 *
 * "Grid", "Timestep" and "Run" are required; omitted physics take FEN's
 * module defaults (density 1, viscosity 1, g 0, CFL 1, constant_CFL off);
 * omitted boundaries are Periodic, as in grid%setup.  Nz == 1 makes the case
 * 2-D.  Initial files are raw float64 interiors in Fortran order (the format
 * of scalar%read/write); fields without one start at zero.  Unknown keys,
 * wrong types and out-of-range values are errors.  A Multiphase interface is
 * the zero level of y - level - amplitude cos(2 pi (x - x0) / wavelength),
 * the heavy fluid (rho_1) above it.  No counterpart in the reference (it has
 * no input file; src/json.f90 is an unused output writer). */
#define FEN_CASE_MAXLIST 32
#define FEN_CASE_PATHLEN 256
typedef struct fen_case {
    int dim;                                /* 2 (Nz == 1) or 3 */
    int n[3];                               /* Nx, Ny, Nz */
    int bc[6];                              /* FEN_BC_* per face */
    int constant_cfl;
    int precision;                          /* FEN_FP64 or FEN_FP32_REFINE */
    int restart;                            /* step of the state to load, or -1 */
    int steps;
    int status_every;                       /* 0: no status lines */
    int n_save_state, save_state[FEN_CASE_MAXLIST];
    int n_save_fields, save_fields[FEN_CASE_MAXLIST];
    int write_final;                        /* write u.raw v.raw [w.raw] p.raw [vof.raw] at the end */
    int multiphase;                         /* the "Multiphase" object is present */
    double origin[3], L[3];
    double density, viscosity, g[3], cfl;
    double dt;                              /* fixed dt, or 0 */
    double U;                               /* velocity scale of set_timestep, or 0 */
    double rho0, rho1, mu0, mu1, sigma;
    double iface[3];                        /* interface level, amplitude, wavelength */
    char name[64];                          /* grid name: <name>.json */
    char data_dir[FEN_CASE_PATHLEN];        /* where save_state and save_fields write */
    char restart_dir[FEN_CASE_PATHLEN];     /* where the restart state is read */
    char init[5][FEN_CASE_PATHLEN];         /* u, v, w, p, vof ("" = none) */
} fen_case;

/* Parses a case file into *out.  On failure returns FEN_ERR_ARG and writes
 * the reason (file, line and column, or the key) into err. */
int fen_case_read(const char *path, fen_case *out, char *err, size_t errlen);

/* Reports the Poisson solver kind (FEN_SOLVER_*) and the grid spacing. */
int fen_solver_kind(const fen_ctx *ctx, int *kind);
int fen_grid_delta(const fen_ctx *ctx, double *delta);

/* Waits for all work queued on the context's stream. */
int fen_synchronize(fen_ctx *ctx);

/* Per-kernel timing with HIP events recorded on the context's stream.
 * mask selects kernels by bit (1 << k, k < fen_profile_count()); -1 times
 * all of them, 0 none.  An event is recorded after every launch of a
 * selected kernel; a launch's duration is the time from the event of the
 * launch just before it on the stream (which must be selected too) to its own.
 * fen_profile_read returns the summed milliseconds and count of the timed
 * launches and the kernel's algorithmic HBM bytes per launch;
 * fen_profile_predecessor names the kernel launched just before kernel k.
 * fen_profile_sample(ctx, n), n > 1, switches to sampling: only every n-th
 * time step is timed, with one event just before and one just after each
 * launch of a selected kernel (the events drain the stream; sampling keeps
 * that cost off the other steps).  n = 1 restores the default. */
int fen_profile_enable(fen_ctx *ctx, int mask);
int fen_profile_sample(fen_ctx *ctx, int every);
int fen_profile_reset(fen_ctx *ctx);
int fen_profile_count(void);
const char *fen_profile_name(int kernel);
int fen_profile_read(fen_ctx *ctx, int kernel, double *total_ms, long long *launches, double *bytes);
int fen_profile_predecessor(fen_ctx *ctx, int kernel, int *prev);

/* Discrete adjoint of the time step -- no counterpart in the reference; for
 * callers that differentiate a run (sensitivities, optimisation).  FP64, 3-D
 * triply periodic (ppp), no source term, fixed dt (constant_CFL off).
 *
 * fen_adjoint_record runs nsteps steps from the resident state, exactly as
 * fen_ns_run(ctx, step0, nsteps, dt) would, keeping a copy of the state every
 * `every` steps in device memory (ceil(nsteps/every) + 1 copies of u, v, w, p
 * and dv_o).  Returns FEN_ERR_UNSUPPORTED for other contexts. */
int fen_adjoint_record(fen_ctx *ctx, int step0, int nsteps, int every, double *dt);

/* Adjoint of the recorded run: given seed = dL/du(final) (interior arrays of
 * nx*ny*nz doubles for u, v, w), returns dL/du(initial) in grad_u/v/w and,
 * if grad_p is not NULL, dL/dp(initial) (interior arrays).  Each segment is
 * recomputed from its copy; the resident state is left at the end of the
 * recorded run.  FEN_ERR_STATE if nothing was recorded. */
int fen_adjoint_backward(fen_ctx *ctx, const double *seed_u, const double *seed_v, const double *seed_w,
                         double *grad_u, double *grad_v, double *grad_w, double *grad_p);

#ifdef __cplusplus
}
#endif
#endif
