// fen_internal.hpp -- shared definitions of libfen's host code and kernels.
//
// Data layout in HBM.  Element type R is double (FEN_FP64) or float
// (FEN_FP32_REFINE) for every field and the spectrum; eigenvalue tables,
// tridiagonal coefficients, the Thomas scratch and reductions are FP64.
//  * ghosted fields (p, phi, u, v, w and the predicted velocity u*, v*, w*):
//    FEN's own Fortran layout, (nx+2)*(ny+2)*(nz+2), i fastest, so host
//    uploads/downloads are plain copies.  In 2-D the k extent is 0:2 and only
//    the k = 1 plane is live.
//  * interior-only fields (dv_o, S): nx*ny*nz, i fastest.
//  * half spectrum of the Poisson solve: complex (double2 / float2), laid out
//    [kx][k][j] with j fastest, kx in 0..nx/2.  Every pass of the solver
//    then reads and writes runs of >= 128 contiguous bytes: x-lines are
//    transformed in LDS and scattered as j-runs, y-lines are contiguous,
//    z-lines are strided by ny and are read as tiles of adjacent j.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/fen.h"

namespace fen {

// Complex element of the spectrum for a real element type.
template <typename T> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };

// Grid extents as seen by kernels.
struct Box {
    int nx, ny, nz, dim;
    long gx, gxy;   // ghosted strides in j and k
    __host__ __device__ long g(int i, int j, int k) const { return i + gx * j + gxy * k; }
    __host__ __device__ long q(int i, int j, int k) const {
        return (i - 1) + (long)nx * ((j - 1) + (long)ny * (k - 1));
    }
};

// Ghost-fill description of one field (scalar%bc, src/scalar.f90:23-47).
template <class R> struct GhostDesc {
    R *f;
    const R *plane[6];
    int type[6];
    int loc;      // 'c', 'x', 'y', 'z'
    int groups;   // ghost groups to fill: bit 0 x faces, bit 1 y faces, bit 2 z faces
};

// Axes whose ghost layers a producer kernel writes itself: the longest
// prefix of (x, y, z) along which the field is periodic on both faces.
// The remaining axes are left to the ghost kernel, which runs after the
// producer and so sees the periodic images already in place.
__host__ __device__ inline int image_axes(const int *type, int dim)
{
    int f = 0;
    if (type[0] == 0 && type[1] == 0) f |= 1; else return f;
    if (type[2] == 0 && type[3] == 0) f |= 2; else return f;
    if (dim == 3 && type[4] == 0 && type[5] == 0) f |= 4;
    return f;
}

// Writes v to every periodic image of interior cell (i,j,k) along the axes
// in `ax` (bit 0 x, 1 y, 2 z): the ghost cells whose final value is a copy
// of this cell (src/scalar.f90:254-388, periodic branches).
template <class R>
__device__ __forceinline__ void write_images(R *f, const Box &b, int ax, int i, int j, int k, R v)
{
    const int ii = (ax & 1) ? (i == 1 ? b.nx + 1 : (i == b.nx ? 0 : -1)) : -1;
    const int jj = (ax & 2) ? (j == 1 ? b.ny + 1 : (j == b.ny ? 0 : -1)) : -1;
    const int kk = (ax & 4) ? (k == 1 ? b.nz + 1 : (k == b.nz ? 0 : -1)) : -1;
    if (ii < 0 && jj < 0 && kk < 0) return;
    if (ii >= 0) f[b.g(ii, j, k)] = v;
    if (jj >= 0) f[b.g(i, jj, k)] = v;
    if (kk >= 0) f[b.g(i, j, kk)] = v;
    if (ii >= 0 && jj >= 0) f[b.g(ii, jj, k)] = v;
    if (ii >= 0 && kk >= 0) f[b.g(ii, j, kk)] = v;
    if (jj >= 0 && kk >= 0) f[b.g(i, jj, kk)] = v;
    if (ii >= 0 && jj >= 0 && kk >= 0) f[b.g(ii, jj, kk)] = v;
}

// write_images() for callers whose j and k are the same across a wave (a
// k-marching tile row, an x line of the c2r): only the x test varies across
// lanes; `xedge` says whether the wave can hold i == 1 or i == nx.  Writes
// exactly the cells write_images() writes.
template <class R>
__device__ __forceinline__ void march_images(R *f, const Box &b, int ax, int i, int j, int k, R v, bool xedge)
{
    const bool xi = (ax & 1) && xedge && (i == 1 || i == b.nx);
    const int ii = i == 1 ? b.nx + 1 : 0;
    if (xi) f[b.g(ii, j, k)] = v;
    const bool yj = (ax & 2) && (j == 1 || j == b.ny);
    const bool zk = (ax & 4) && (k == 1 || k == b.nz);
    if (yj || zk) {
        const int jj = j == 1 ? b.ny + 1 : 0, kk = k == 1 ? b.nz + 1 : 0;
        if (yj) {
            f[b.g(i, jj, k)] = v;
            if (xi) f[b.g(ii, jj, k)] = v;
        }
        if (zk) {
            f[b.g(i, j, kk)] = v;
            if (xi) f[b.g(ii, j, kk)] = v;
            if (yj) {
                f[b.g(i, jj, kk)] = v;
                if (xi) f[b.g(ii, jj, kk)] = v;
            }
        }
    }
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X; this
// renumbering gives every XCD one contiguous range of the grid (a k-slab),
// so the stencil neighbours a workgroup needs are in its own XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb)
{
    if (nb & 7) return b;
    const int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

// Kernels whose launches are timed by the profiler.
enum KernelId {
    K_MOMENTUM = 0,   // advection + diffusion + grad p + AB2 + update + dv_o
    K_GHOST,          // ghost-node fill (any field set)
    K_DIV_R2C,        // divergence * rho/dt fused with the x r2c transform
    K_Y_FWD,          // y forward transform of the half spectrum
    K_Z_SOLVE,        // z forward + eigenvalue divide + z inverse (ppp)
    K_THOMAS,         // batched tridiagonal solve along z (ppn)
    K_Y_INV,          // y inverse transform
    K_Y_SOLVE,        // 2-D: y forward + divide + y inverse (pp)
    K_C2R,            // x c2r -> phi (+ partial sums for the mean)
    K_PROJECT,        // projection u = u* - dt grad(phi)/rho, p = p + phi
    K_CHECKS,         // max div(u), max(|u|+|v|+|w|)
    K_REDUCE,         // final small reductions
    K_PROJ_MOM,       // projection + checks of one step fused with the next step's momentum
    K_COUNT
};

// One profiled launch: an event recorded on the stream right after it.  A
// launch's duration is the time between its event and the event of the launch
// just before it in the stream (seq - 1), when that one was profiled too.
struct ProfMark {
    int kernel;
    long long seq;
    hipEvent_t ev;
    bool begin;   // recorded just before the launch (sampled mode)
};

struct ProfSlot {
    double bytes = 0.0;   // algorithmic HBM bytes per launch (last launch)
};

struct AdjointTape;

}  // namespace fen

// Two-phase (MF) state: volume_of_fluid_mod and multiphase_mod fields (FP64,
// ghosted, the host's (nx+2)(ny+2)(nz+2) layout).
struct MFState {
    double rho0 = 1, rho1 = 1, mu0 = 1, mu1 = 1, sigma = 0, beta = 1, rhomin = 1, irhomin = 1;
    double dt_surf = 0;          // set by set_timestep when sigma > 0 (src/navier_stokes.f90:657)
    bool have_dt_surf = false;
    int x_first = 1;
    double *vof = nullptr, *vof1 = nullptr, *vof2 = nullptr, *h = nullptr, *dd = nullptr, *curv = nullptr;
    double *nrx = nullptr, *nry = nullptr, *lx = nullptr, *ly = nullptr;
    double *rho = nullptr, *mu = nullptr, *p_hat = nullptr, *p_o = nullptr;
};

struct fen_ctx {
    int dim = 3, nx = 0, ny = 0, nz = 0;
    double lx = 0, ly = 0, lz = 0, x0[3] = {0, 0, 0}, delta = 0;
    int bc_phys[6] = {0, 0, 0, 0, 0, 0};
    int periodic[3] = {1, 1, 1};
    int solver = FEN_SOLVER_PPP;
    int precision = FEN_FP64;

    // navier_stokes_mod scalars (src/navier_stokes.f90:17-45)
    double density = 1.0, viscosity = 1.0, g[3] = {0, 0, 0}, cfl = 1.0;
    int constant_cfl = 0;
    double dt_o = 0.0, dt_visc = 0.0, dt_conv = 0.0;
    bool have_dt_o = false, have_dt_visc = false;

    fen::Box box{};
    size_t n_ghost = 0, n_int = 0;

    // bc types and device bc planes (element type R) of p, phi, u, v, w
    int bc_type[5][6] = {};
    void *d_plane[5][6] = {};
    void **d_plane_tab = nullptr;   // d_plane of u, v, w, p (6 each) in device memory
    void *d_wallbuf = nullptr;      // the general-wall fused kernel's face buffer (allocated on first use)

    // resident fields, element type R (fen::as<R>(...) at the use site)
    void *d_u[3] = {nullptr, nullptr, nullptr};    // velocity (state)
    void *d_us[3] = {nullptr, nullptr, nullptr};   // predicted velocity u*
    void *d_p = nullptr, *d_phi = nullptr;
    void *d_p2 = nullptr;   // second pressure array of a fused run (allocated on first use)
    MFState *mf = nullptr;  // two-phase mode (fen_set_multiphase)
    void *d_dvo[3] = {nullptr, nullptr, nullptr};
    void *d_S[3] = {nullptr, nullptr, nullptr};
    bool has_source = false;

    // Poisson solver state
    int hx = 0;                  // nx/2 + 1
    int num_cu = 256;            // compute units of the device (queried at create)
    void *d_spec = nullptr;      // complex R, [kx][k][j] (npn, nnn: real, [k][kx][j])
    double *d_c1 = nullptr;      // Thomas c1, computed at create (layout of the sweep's columns)
    double2 *d_dd = nullptr;     // FP64 Thomas right-hand side (ppn, FP32 form), [kx][k][j]
    double2 *d_spec2 = nullptr;  // npn: y half spectrum of the real x-DCT lines, [k][kx][jy]
    double2 *d_mkx = nullptr, *d_mky = nullptr;   // (cos, sin)(pi k/(2n)), k < n: DCTs along x, y
    double *d_mwx = nullptr, *d_mwy = nullptr, *d_mwz = nullptr;
    double *d_ta = nullptr, *d_tb = nullptr, *d_tc = nullptr;
    void *d_twx = nullptr, *d_twy = nullptr, *d_twz = nullptr;   // complex R, W_n^m, m < n
    double phi_shift = 0.0;      // mean/N removed lazily (pp, ppn)
    double *d_shift = nullptr;   // device copy of the subtracted mean
    bool shift_active = false;

    int img_vel = 0, img_p = 0, img_phi = 0;   // image_axes of (u,v,w), p, phi

    // reductions
    double *d_part = nullptr;    // per-block partials
    double *d_noalloc = nullptr; // device address of the allocation-free entry point's scratch (this device)
    size_t n_part = 0;
    double *d_scal = nullptr;    // [0] maxdiv [1] max|u|+|v|+|w| [2] nonfinite flag [3] mean shift
    double *h_scal = nullptr;    // pinned mirror
    double last_dt = 0.0;        // dt of the last step (for maxCFL)
    bool checks_valid = false;

    // Deferred projection (fen_ns_step / fen_ns_run on the grids the fused
    // projection + momentum kernel supports): a step ends after its Poisson
    // solve with u*, v*, w* in d_us, p before the step in d_p and phi in
    // d_phi; the next step's fused kernel, or a flush before anything reads
    // or changes the state, projects it.  The state arrays d_u and the
    // second pressure array d_p2 swap roles with d_us and d_p after every
    // fused step, so a pending state always sits in the same places.
    bool pending = false;
    bool pending_checked = false;   // d_scal holds the checks of the pending step
    double pend_dt = 0.0;           // dt of the pending step

    hipStream_t stream = nullptr;
    std::string err;

    fen::AdjointTape *tape = nullptr;       // recorded run of the discrete adjoint

    // One fixed-dt step captured as a HIP graph and replayed (fen_ns_step /
    // fen_ns_run without constant_CFL or profiling).  The graph bakes in the
    // kernel arguments: dt, dt_o and everything `config_version` counts
    // (parameters, bc types, the source term); a change of any rebuilds it.
    bool use_graph = true;                  // FEN_STEP_GRAPH=0 at create turns it off
    bool fused_run = true;                  // FEN_FUSED_RUN=0 at create turns fen_ns_run's fusion off
    hipGraphExec_t step_graph = nullptr;
    double graph_dt = 0.0, graph_dt_o = 0.0;
    int graph_version = -1;
    int config_version = 0;

    unsigned prof_mask = 0;
    int prof_every = 1;                     // > 1: time only every prof_every-th step, begin/end pairs
    long long step_count = 0;               // steps issued (advance calls)
    long long prof_base = 0;                // step_count when sampling was (re)started
    long long prof_seq = 0;                 // launches issued on the stream
    std::vector<fen::ProfMark> prof_marks;  // chronological
    std::vector<hipEvent_t> prof_pool;      // events owned by the context
    size_t prof_used = 0;
    fen::ProfSlot prof[fen::K_COUNT];
};

namespace fen {

template <class R> inline R *as(void *p) { return static_cast<R *>(p); }

// Launchers (defined in the kernel translation units, instantiated for
// R = double and R = float).  All enqueue on ctx->stream.
template <class R> void launch_ghost_desc(fen_ctx *c, const GhostDesc<R> *descs, int n);
template <class R> void launch_momentum(fen_ctx *c, double dt, double A, double B);
// vel == nullptr: the right-hand side is read from direct_src (strides sx, sy, ghost level gl)
template <class R>
void launch_div_r2c_impl(fen_ctx *c, const R *const *vel, const R *direct_src, long sx, long sy, int gl, double rho,
                         double dt, double post);
template <class R> void launch_y_fft(fen_ctx *c, int sign, double post_scale);
template <class R> void launch_y_solve_2d(fen_ctx *c, double inv_n);
template <class R> void launch_z_solve(fen_ctx *c, double inv_n);
template <class R> void launch_thomas(fen_ctx *c);
long c2r_blocks(const fen_ctx *c);
// part/n_part/shift: where the block sums and the mean go (default: the
// context's d_part, n_part, d_shift).
template <class R> void launch_c2r(fen_ctx *c, R *phi, int gl_out, long sx, long sy, bool want_sum, int img,
                                   double *part = nullptr, size_t n_part = 0, double *shift = nullptr);
// store == false: the checks of the projection only (max div, max |u|+|v|+|w|
// of the projected velocity), nothing written but the reduction (3-D).
template <class R> void launch_project(fen_ctx *c, double dt, double *checks_out, bool store = true);
bool fused_z_walls(const fen_ctx *c);
void mf_advect_vof(fen_ctx *c, double dt);
// p_hat = 2 p - p_o, or with `dt` > 0 (constant CFL) p_o + (dt + dt_o)(p - p_o)/dt_o.
void mf_update_fields(fen_ctx *c, bool with_phat, double dt = 0.0);
void mf_momentum(fen_ctx *c, double dt);
void mf_ghost_field(fen_ctx *c, double *f);
// npn and nnn (FP64) whose state fields keep the standard wall rules: the fused
// kernel's general-wall form applies.
bool fused_general_walls(const fen_ctx *c);
template <class R>
void launch_project_momentum(fen_ctx *c, double dt_p, double dt_m, double A, double B, void *const *us_in,
                             const void *p_in, void *const *us_out, void *p_out, void *const *save = nullptr);
template <class R> void launch_checks(fen_ctx *c, const R *const *vel, double *out2);
long stencil_blocks(const Box &b);   // workgroups of the momentum / projection grids
// Neumann-direction solver passes (neumann_kernels.hip, FP64).
void launch_dct2_x(fen_ctx *c, const double *const *vel, const double *direct, long sx, long sy, int gl, double rho,
                   double dt, double post);
// Whether every ghost of phi is a copy of an interior value (x Neumann;
// y, z periodic or Neumann), so that the x DCT-III can write them.
bool dct3_writes_ghosts(const fen_ctx *c);
void launch_dct3_x(fen_ctx *c, double *phi, int gl, long sx, long sy, double post, bool img = false);
void launch_dct_y(fen_ctx *c, int inverse, double post);
void launch_rfft_y(fen_ctx *c, int inverse, double post);
void launch_thomas_neumann(fen_ctx *c);
void launch_thomas_c1(fen_ctx *c);
template <class R> void launch_apply_shift(fen_ctx *c, R *f, size_t n);
void launch_advection_only(fen_ctx *c, double *rx, double *ry, double *rz);   // FP64 contexts

// One FP64 time step (the step fen_ns_step runs, without the CFL update).
void advance_fp64(fen_ctx *c, double dt);
// The same step reading the velocity from `in` (ghosted u, v, w) and
// writing the new velocity to `out` instead of the context's arrays, which
// keep their contents (p and dv_o are the context's, updated in place).
void advance_fp64_into(fen_ctx *c, double dt, double *const *in, double *const *out);
// The velocities of n consecutive FP64 states of a fixed-dt run that starts
// from the context's (completed) state: seg[s] receives u, v, w (ghosted,
// one after the other) of the state before step s, after n - 1 steps, fused
// where fen_ns_run would fuse them; p and dv_o advance in the context.
void states_fp64(fen_ctx *c, double dt, int n, double *const *seg);
// nsteps FP64 steps as fen_ns_run takes them (fused when the case allows;
// bitwise equal to nsteps calls of advance_fp64).
void run_fp64(fen_ctx *c, int nsteps, double dt);
// Completes a pending step (the state is then where a single step leaves it).
void flush_state(fen_ctx *c);

// Discrete adjoint (adjoint.hip).
const char *adjoint_unsupported(const fen_ctx *c);   // nullptr when the context is supported
void adjoint_record(fen_ctx *c, int step0, int nsteps, int every, double dt);
void adjoint_backward(fen_ctx *c, const double *seed[3], double *grad[3], double *grad_p);
void adjoint_free(fen_ctx *c);

// Profiling helpers.
void prof_begin(fen_ctx *c, int k);
void prof_end(fen_ctx *c, int k, double bytes);

}  // namespace fen

#define FEN_HIP_CHECK(expr)                                                        \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + \
                                                       hipGetErrorString(e_));     \
    } while (0)
