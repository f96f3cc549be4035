// fen_api.hip -- the extern "C" boundary of libfen (declared in include/fen.h).
//
// Orchestrates one navier_stokes_solver step (src/navier_stokes.f90:50-136)
// as a fixed sequence of kernels on the context's stream:
//   K1 momentum (u -> u*, dv_o)        ghost(u*)
//   K2 div(u*)*rho/dt + x r2c or DCT-II
//   K3..K5 y / z passes: FFT, eigenvalue divide or Thomas per solver
//   K6 x c2r or DCT-III -> phi (+ mean, pp/ppn/pn)   ghost(phi)
//   K7 projection + p update           ghost(u, p)
//   K8 checks
// On the 3-D grids the fused kernel supports (periodic, or periodic in x, y
// with walls at z) a step ends after K6 with K7-K8 pending; the next step
// runs them fused with its K1 (step_pending), and anything that reads or
// changes the state first completes them (flush_pending).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "fen_internal.hpp"

using fen::Box;
using fen::GhostDesc;
using fen::as;

// ---------------------------------------------------------------------------
// Scratch of fen_poisson_solve_noalloc: the c2r block sums and the subtracted
// mean of one solve.  kNoAllocParts block sums cover ny*nz/(4096/nx) <= 32768
// (every grid up to 512 x 512 x 256, 256^3, 1024 x 1024 x 128).  The array
// exists once per GPU for the life of the process; calls on one GPU share it.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocParts = 32768;
constexpr size_t kNoAllocDoubles = kNoAllocParts + 32;   // + the mean, on its own 256-byte line
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace fen {

static hipEvent_t prof_event(fen_ctx *c)
{
    if (c->prof_used == c->prof_pool.size()) {
        hipEvent_t e;
        FEN_HIP_CHECK(hipEventCreate(&e));
        c->prof_pool.push_back(e);
    }
    return c->prof_pool[c->prof_used++];
}

// Timed steps are the 1st, (n+1)-th, ... after fen_profile_sample / fen_profile_reset.
static bool prof_sampled(const fen_ctx *c) { return (c->step_count - 1 - c->prof_base) % c->prof_every == 0; }

// Sampled mode only: an event just before the launch of a selected kernel.
void prof_begin(fen_ctx *c, int k)
{
    if (c->prof_every <= 1 || !(c->prof_mask & (1u << k)) || !prof_sampled(c)) return;
    hipEvent_t e = prof_event(c);
    FEN_HIP_CHECK(hipEventRecord(e, c->stream));
    c->prof_marks.push_back(ProfMark{k, c->prof_seq, e, true});
}

void prof_end(fen_ctx *c, int k, double bytes)
{
    const long long seq = c->prof_seq++;
    if (!(c->prof_mask & (1u << k))) return;
    if (c->prof_every > 1 && !prof_sampled(c)) return;
    hipEvent_t e = prof_event(c);
    FEN_HIP_CHECK(hipEventRecord(e, c->stream));
    c->prof_marks.push_back(ProfMark{k, seq, e, false});
    c->prof[k].bytes = bytes;
}
}  // namespace fen

namespace {

const char *kKernelNames[fen::K_COUNT] = {
    "momentum", "ghost", "div_r2c", "y_fwd", "z_solve", "thomas", "y_inv", "y_solve", "c2r", "project",
    "checks", "reduce", "project_momentum"};

int fail(fen_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

bool pow2(int n) { return n >= 8 && n <= 1024 && (n & (n - 1)) == 0; }

int field_slot(int field)   // index into bc_type / d_plane
{
    return field >= FEN_P && field <= FEN_W ? field : -1;
}

char field_loc(int field)
{
    switch (field) {
    case FEN_U: return 'x';
    case FEN_V: return 'y';
    case FEN_W: return 'z';
    default: return 'c';
    }
}

// BC mapping of allocate_navier_stokes_fields (src/navier_stokes.f90:779-1017).
int map_bc(int field, int face, int pc, bool *ok)
{
    const bool vel = field == FEN_U || field == FEN_V || field == FEN_W;
    *ok = true;
    switch (pc) {
    case FEN_BC_PERIODIC: return 0;
    case FEN_BC_WALL: return vel ? 1 : 2;
    case FEN_BC_INFLOW:
        if (face == FEN_BACK) { *ok = false; return 0; }
        if (!vel) return 2;
        if (face == FEN_LEFT || face == FEN_RIGHT) return field == FEN_U ? 1 : 2;
        return 1;
    case FEN_BC_OUTFLOW:
        if (face == FEN_FRONT) { *ok = false; return 0; }
        return vel ? 2 : 1;
    }
    *ok = false;
    return 0;
}

// Modified wavenumbers (src/poisson.f90:141,628,646,664), glibc cos as gfortran's dcos.
std::vector<double> modified_wavenumbers(int n, double delta, double factor)
{
    std::vector<double> m(n);
    const double pi = std::acos(-1.0);
    for (int i = 1; i <= n; i++)
        m[i - 1] = 2.0 * (std::cos(factor * pi * ((double)i - 1.0) / (double)(float)n) - 1.0) / (delta * delta);
    return m;
}

template <class R> std::vector<typename fen::Cplx<R>::type> twiddles(int n)
{
    std::vector<typename fen::Cplx<R>::type> w(n);
    const long double tpi = 2.0L * std::acos(-1.0L);
    for (int m = 0; m < n; m++) {
        const long double a = tpi * (long double)m / (long double)n;
        w[m] = {(R)std::cos(a), (R)(-std::sin(a))};
    }
    return w;
}

// (cos, sin)(pi k / (2n)), k < n, for Makhoul's DCT post/pre-twiddle, in long double.
std::vector<double2> dct_twiddles(int n)
{
    std::vector<double2> w(n);
    const long double pi = std::acos(-1.0L);
    for (int k = 0; k < n; k++) {
        const long double a = pi * (long double)k / (2.0L * (long double)n);
        w[k] = double2{(double)std::cos(a), (double)std::sin(a)};
    }
    return w;
}

// Element size of the context's fields.
size_t elem(const fen_ctx *c) { return c->precision == FEN_FP64 ? sizeof(double) : sizeof(float); }

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress with that device current) and kept: in host
// code the array's name is not a device address.
double *noalloc_scratch()
{
    static std::mutex mu;
    static std::vector<double *> addr;
    int dev = 0;
    FEN_HIP_CHECK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((size_t)dev >= addr.size()) addr.resize((size_t)dev + 1, nullptr);
    if (!addr[dev]) {
        void *p = nullptr;
        FEN_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        addr[dev] = static_cast<double *>(p);
    }
    return addr[dev];
}

void *dev_alloc_bytes(size_t bytes)
{
    void *p = nullptr;
    FEN_HIP_CHECK(hipMalloc(&p, bytes));
    FEN_HIP_CHECK(hipMemset(p, 0, bytes));
    return p;
}

// Calls f(R{}) with R the context's field element type.
template <class F> void with_type(const fen_ctx *c, F f)
{
    if (c->precision == FEN_FP64) f(double{});
    else f(float{});
}

// Host FP64 <-> device R copies (FP32: converted through a host buffer, so
// every value is rounded to nearest once).
void copy_in(fen_ctx *c, void *dst, const double *src, size_t n)
{
    if (c->precision == FEN_FP64) {
        FEN_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
        return;
    }
    std::vector<float> h(n);
    for (size_t q = 0; q < n; q++) h[q] = (float)src[q];
    FEN_HIP_CHECK(hipMemcpyAsync(dst, h.data(), n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
}

void copy_out(fen_ctx *c, double *dst, const void *src, size_t n)
{
    if (c->precision == FEN_FP64) {
        FEN_HIP_CHECK(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
        return;
    }
    std::vector<float> h(n);
    FEN_HIP_CHECK(hipMemcpyAsync(h.data(), src, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < n; q++) dst[q] = (double)h[q];
}

template <typename T> T *dev_alloc(size_t n)
{
    void *p = nullptr;
    FEN_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    FEN_HIP_CHECK(hipMemset(p, 0, n * sizeof(T)));
    return static_cast<T *>(p);
}

template <typename T> T *dev_copy(const std::vector<T> &h)
{
    T *d = dev_alloc<T>(h.size());
    FEN_HIP_CHECK(hipMemcpy(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return d;
}

size_t plane_count(const fen_ctx *c, int face)
{
    const size_t gx = c->nx + 2, gy = c->ny + 2, gz = c->nz + 2;
    if (face < 2) return gy * gz;
    if (face < 4) return gx * gz;
    return gx * gy;
}

template <class R> GhostDesc<R> make_desc(fen_ctx *c, int field, void *arr)
{
    GhostDesc<R> d{};
    d.f = as<R>(arr);
    const int s = field_slot(field);
    for (int f = 0; f < 6; f++) {
        d.type[f] = c->bc_type[s][f];
        d.plane[f] = as<R>(c->d_plane[s][f]);
    }
    d.loc = field_loc(field);
    d.groups = 7;
    return d;
}

// Ghost groups left after a producer that wrote the images along `img` axes.
int remaining_groups(const fen_ctx *c, int img) { return ~img & (c->dim == 3 ? 7 : 3); }

void refresh_images(fen_ctx *c)
{
    c->img_vel = fen::image_axes(c->bc_type[FEN_U], c->dim) & fen::image_axes(c->bc_type[FEN_V], c->dim);
    if (c->dim == 3) c->img_vel &= fen::image_axes(c->bc_type[FEN_W], c->dim);
    c->img_p = fen::image_axes(c->bc_type[FEN_P], c->dim);
    c->img_phi = fen::image_axes(c->bc_type[FEN_PHI], c->dim);
}

void *field_ptr(fen_ctx *c, int field)
{
    switch (field) {
    case FEN_P: return c->d_p;
    case FEN_PHI: return c->d_phi;
    case FEN_U: return c->d_u[0];
    case FEN_V: return c->d_u[1];
    case FEN_W: return c->d_u[2];
    case FEN_DVO_X: case FEN_DVO_Y: case FEN_DVO_Z: return c->d_dvo[field - FEN_DVO_X];
    case FEN_S_X: case FEN_S_Y: case FEN_S_Z: return c->d_S[field - FEN_S_X];
    case FEN_VOF: return c->mf ? c->mf->vof : nullptr;
    case FEN_RHO: return c->mf ? c->mf->rho : nullptr;
    case FEN_MU: return c->mf ? c->mf->mu : nullptr;
    }
    return nullptr;
}

void free_all(fen_ctx *c)
{
    fen::adjoint_free(c);
    if (c->step_graph) (void)hipGraphExecDestroy(c->step_graph);
    auto f = [](void *p) { if (p) (void)hipFree(p); };
    for (int d = 0; d < 3; d++) { f(c->d_u[d]); f(c->d_us[d]); f(c->d_dvo[d]); f(c->d_S[d]); }
    f(c->d_p); f(c->d_p2); f(c->d_phi); f(c->d_spec); f(c->d_c1); f(c->d_dd);
    f(c->d_mwx); f(c->d_mwy); f(c->d_mwz); f(c->d_ta); f(c->d_tb); f(c->d_tc);
    f(c->d_twx); f(c->d_twy); f(c->d_twz); f(c->d_spec2); f(c->d_mkx); f(c->d_mky); f(c->d_part); f(c->d_scal);
    for (int s = 0; s < 5; s++) for (int p = 0; p < 6; p++) f(c->d_plane[s][p]);
    f(c->d_plane_tab);
    f(c->d_wallbuf);
    if (c->mf) {
        MFState &m = *c->mf;
        for (double *q : {m.vof, m.vof1, m.vof2, m.h, m.dd, m.curv, m.nrx, m.nry, m.lx, m.ly, m.rho, m.mu, m.p_hat,
                          m.p_o})
            f(q);
        delete c->mf;
        c->mf = nullptr;
    }
    if (c->h_scal) (void)hipHostFree(c->h_scal);
    for (auto e : c->prof_pool) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

void sync(fen_ctx *c) { FEN_HIP_CHECK(hipStreamSynchronize(c->stream)); }

// Whether the solver subtracts the mean of phi: pp, ppn and pn
// (src/poisson.f90:397-408, 495-503, 1159-1171); npn, nnn and nn do not.
bool removes_mean(const fen_ctx *c)
{
    return c->solver == FEN_SOLVER_PP || c->solver == FEN_SOLVER_PPN || c->solver == FEN_SOLVER_PN;
}

// Whether x is a cosine (Neumann) direction: nn, npn, nnn.
bool x_cosine(const fen_ctx *c)
{
    return c->solver == FEN_SOLVER_NN || c->solver == FEN_SOLVER_NPN || c->solver == FEN_SOLVER_NNN;
}

// The Poisson solve up to the x backward transform: the x transform of
// div(u*) rho/dt (vel) or of `direct` (strides sx, sy, ghost level gl), the
// scalings of the reference solver, and the solve in the other directions.
// Leaves phi's x spectrum in the spectrum buffer.
template <class R>
void solve_to_x_spectrum(fen_ctx *c, const R *const *vel, const R *direct, long sx, long sy, int gl, double rho,
                         double dt)
{
    using namespace fen;
    const double nx = (double)c->nx, ny = (double)c->ny;
    switch (c->solver) {
    case FEN_SOLVER_PPP:
        launch_div_r2c_impl<R>(c, vel, direct, sx, sy, gl, rho, dt, 1.0);
        launch_y_fft<R>(c, -1, 1.0);
        launch_z_solve<R>(c, 1.0 / (double)(float)((long)c->nx * c->ny * c->nz));
        launch_y_fft<R>(c, +1, 1.0);
        return;
    case FEN_SOLVER_PPN:
        launch_div_r2c_impl<R>(c, vel, direct, sx, sy, gl, rho, dt, 1.0 / nx);
        launch_y_fft<R>(c, -1, 1.0 / ny);
        launch_thomas<R>(c);
        launch_y_fft<R>(c, +1, 1.0);
        return;
    case FEN_SOLVER_PP:
        launch_div_r2c_impl<R>(c, vel, direct, sx, sy, gl, rho, dt, 1.0);
        launch_y_solve_2d<R>(c, 1.0 / (double)(float)((long)c->nx * c->ny));
        return;
    default: break;
    }
    if constexpr (std::is_same_v<R, double>) {
        // scalings are powers of two: multiplying by the reciprocal equals the
        // reference's division
        switch (c->solver) {
        case FEN_SOLVER_PN:   // src/poisson.f90:330-384
            launch_div_r2c_impl<double>(c, vel, direct, sx, sy, gl, rho, dt, 1.0 / nx);
            launch_thomas_neumann(c);
            return;
        case FEN_SOLVER_NN:   // :530-580; the 1/(2nx) follows the x DCT-III
            launch_dct2_x(c, vel, direct, sx, sy, gl, rho, dt, 1.0);
            launch_thomas_neumann(c);
            return;
        case FEN_SOLVER_NPN:  // :1205-1291
            launch_dct2_x(c, vel, direct, sx, sy, gl, rho, dt, 1.0 / (2.0 * nx));
            launch_rfft_y(c, 0, 1.0 / ny);
            launch_thomas_neumann(c);
            launch_rfft_y(c, 1, 1.0);
            return;
        case FEN_SOLVER_NNN:  // :1344-1430
            launch_dct2_x(c, vel, direct, sx, sy, gl, rho, dt, 1.0 / (2.0 * nx));
            launch_dct_y(c, 0, 1.0 / (2.0 * ny));
            launch_thomas_neumann(c);
            launch_dct_y(c, 1, 1.0);
            return;
        default: break;
        }
    }
    throw std::runtime_error("solver not available for this precision");
}

// x backward transform of the spectrum into the interior of phi; sums the
// interior for the mean when `mean`; writes the periodic images in `img`.
template <class R> void x_backward(fen_ctx *c, R *phi, int gl, long sx, long sy, bool mean, int img,
                                   double *part = nullptr, size_t n_part = 0, double *shift = nullptr)
{
    if constexpr (std::is_same_v<R, double>) {
        if (x_cosine(c)) {
            // nn divides by 2nx after the DCT-III (src/poisson.f90:594)
            fen::launch_dct3_x(c, phi, gl, sx, sy, c->solver == FEN_SOLVER_NN ? 1.0 / (2.0 * (double)c->nx) : 1.0,
                               img != 0);
            return;
        }
    }
    fen::launch_c2r<R>(c, phi, gl, sx, sy, mean, img, part, n_part, shift);
}

// K2..K6 of a step: the Poisson solve on u*, v*, w* (`us`) up to phi with
// its ghosts; see the kernel sequence at the top of this file.
template <class R> void solve_phi(fen_ctx *c, double dt, void *const *us)
{
    using namespace fen;
    const R *vs[3] = {as<R>(us[0]), as<R>(us[1]), as<R>(us[2])};
    // phi = div(u*) rho / dt; the MF build scales by rho_min (src/navier_stokes.f90:111-121)
    solve_to_x_spectrum<R>(c, vs, nullptr, 0, 0, 0, c->mf ? c->mf->rhomin : c->density, dt);
    const bool mean = removes_mean(c);
    bool dirichlet = false;
    for (int f = 0; f < 6; f++) dirichlet = dirichlet || c->bc_type[FEN_PHI][f] == 1;
    // images written by the c2r only when no shift has to be applied first;
    // the DCT-III writes every ghost when they are all copies, else none
    const int img_phi = x_cosine(c) ? (fen::dct3_writes_ghosts(c) ? 7 : 0)
                                    : ((mean && dirichlet) ? 0 : c->img_phi);
    x_backward<R>(c, as<R>(c->d_phi), 1, c->box.gx, c->box.gxy, mean, img_phi);
    c->shift_active = mean;
    if (mean && dirichlet) {   // the shift does not commute with 2*bc - f: apply it first
        launch_apply_shift<R>(c, as<R>(c->d_phi), c->n_ghost);
        c->shift_active = false;
    }
    if (const int rest = remaining_groups(c, img_phi)) {
        GhostDesc<R> g = make_desc<R>(c, FEN_PHI, c->d_phi);
        g.groups = rest;
        launch_ghost_desc<R>(c, &g, 1);
    }
}

// K7..K8 of a step: projection, pressure update, checks and ghosts of u, p.
template <class R> void finish_step(fen_ctx *c, double dt)
{
    using namespace fen;
    const bool d3 = c->dim == 3;
    launch_project<R>(c, dt, c->d_scal);
    if (const int rest = remaining_groups(c, c->img_vel & c->img_p)) {
        GhostDesc<R> g[4] = {make_desc<R>(c, FEN_U, c->d_u[0]), make_desc<R>(c, FEN_V, c->d_u[1]),
                             d3 ? make_desc<R>(c, FEN_W, c->d_u[2]) : make_desc<R>(c, FEN_P, c->d_p),
                             make_desc<R>(c, FEN_P, c->d_p)};
        for (auto &d : g) d.groups = rest;
        launch_ghost_desc<R>(c, g, d3 ? 4 : 3);
    }
    c->checks_valid = true;
}

// K1 of a step: momentum into u*, v*, w* and their ghosts.
template <class R> void predict(fen_ctx *c, double dt)
{
    using namespace fen;
    const double A = 1.0 + 0.5 * dt / c->dt_o;          // src/navier_stokes.f90:157-158
    const double B = -(0.5 * dt / c->dt_o);
    c->last_dt = dt;
    c->step_count++;
    launch_momentum<R>(c, dt, A, B);
    if (const int rest = remaining_groups(c, c->img_vel)) {
        GhostDesc<R> g[3] = {make_desc<R>(c, FEN_U, c->d_us[0]), make_desc<R>(c, FEN_V, c->d_us[1]),
                             make_desc<R>(c, FEN_W, c->d_us[2])};
        for (auto &d : g) d.groups = rest;
        launch_ghost_desc<R>(c, g, c->dim == 3 ? 3 : 2);
    }
}

// One time step; see the kernel sequence at the top of this file.
template <class R> void advance(fen_ctx *c, double dt)
{
    predict<R>(c, dt);
    solve_phi<R>(c, dt, c->d_us);
    finish_step<R>(c, dt);
}

// The MF build's step (src/navier_stokes.f90:50-136 with MF, fixed dt), 2-D FP64.
void advance_mf(fen_ctx *c, double dt)
{
    using namespace fen;
    c->last_dt = dt;
    c->step_count++;
    mf_advect_vof(c, dt);          // :82
    // :85 and p_hat (:88-92), with ghosts; the constant-CFL form extrapolates with dt_o
    mf_update_fields(c, true, c->constant_cfl ? dt : 0.0);
    mf_momentum(c, dt);            // predicted_velocity_field with MF (:140-213)
    if (const int rest = remaining_groups(c, c->img_vel)) {
        GhostDesc<double> g[2] = {make_desc<double>(c, FEN_U, c->d_us[0]), make_desc<double>(c, FEN_V, c->d_us[1])};
        for (auto &d : g) d.groups = rest;
        launch_ghost_desc<double>(c, g, 2);
    }
    solve_phi<double>(c, dt, c->d_us);
    // update_pressure with MF: p_o = p before p = p + phi (:553-559)
    FEN_HIP_CHECK(hipMemcpyAsync(c->mf->p_o, c->d_p, c->n_ghost * sizeof(double), hipMemcpyDeviceToDevice,
                                 c->stream));
    finish_step<double>(c, dt);
}

// Whether a step may end with its projection pending, to be fused into the
// next step's momentum (launch_project_momentum): 3-D, every field periodic
// in every direction (ppp), x, y periodic with no-slip walls at z (ppn), or
// npn/nnn in FP64 with the standard wall rules.
bool deferrable(const fen_ctx *c)
{
    if (!c->fused_run || c->dim != 3 || c->mf) return false;
    if (c->solver == FEN_SOLVER_PPP) return c->img_vel == 7 && c->img_p == 7 && c->img_phi == 7;
    if (c->solver == FEN_SOLVER_PPN) return fen::fused_z_walls(c);
    return fen::fused_general_walls(c);   // npn, nnn (FP64)
}

// One step that ends pending (fen_ctx::pending).  If the previous step is
// pending, its projection and pressure update run fused with this step's
// momentum (its checks, superseded by this step's, are not formed); the
// new u*, v*, w* land in the state arrays and p in the
// second pressure array, and the arrays then swap roles, so a pending step
// always leaves u* in d_us and p (before its update) in d_p.  The fused
// kernel computes the same values as the separate projection and momentum,
// so a sequence of pending steps equals the same number of advance() calls
// bit for bit.
template <class R> void step_pending(fen_ctx *c, double dt, void *const *save = nullptr)
{
    using namespace fen;
    if (!c->pending) {
        predict<R>(c, dt);
    } else {
        if (!c->d_p2) c->d_p2 = dev_alloc_bytes(c->n_ghost * elem(c));
        const double A = 1.0 + 0.5 * dt / c->dt_o, B = -(0.5 * dt / c->dt_o);   // src/navier_stokes.f90:157-158
        c->last_dt = dt;
        c->step_count++;
        void *out[3] = {c->d_u[0], c->d_u[1], c->d_u[2]};
        launch_project_momentum<R>(c, c->pend_dt, dt, A, B, c->d_us, c->d_p, out, c->d_p2, save);
        for (int d = 0; d < 3; d++) std::swap(c->d_u[d], c->d_us[d]);
        std::swap(c->d_p, c->d_p2);
        c->config_version++;   // a captured step graph holds the old array roles
        // ghosts of u*, v*, w* the momentum did not write as images (walls)
        if (const int fill = remaining_groups(c, c->img_vel)) {
            GhostDesc<R> g[3] = {make_desc<R>(c, FEN_U, c->d_us[0]), make_desc<R>(c, FEN_V, c->d_us[1]),
                                 make_desc<R>(c, FEN_W, c->d_us[2])};
            for (auto &d : g) d.groups = fill;
            launch_ghost_desc<R>(c, g, 3);
        }
    }
    solve_phi<R>(c, dt, c->d_us);
    c->pending = true;
    c->pending_checked = false;
    c->pend_dt = dt;
    c->checks_valid = true;
}

// Completes a pending step (projection, pressure update, checks, ghosts),
// leaving the state where advance() leaves it.  Every entry point that reads
// or changes the state, the fields' boundary data or the parameters the
// projection uses calls this first.
void flush_pending(fen_ctx *c)
{
    if (!c->pending) return;
    c->pending = false;
    with_type(c, [&](auto tag) { finish_step<decltype(tag)>(c, c->pend_dt); });
}

// The checks of a pending step (max div and max |u|+|v|+|w| of its projected
// velocity) into d_scal, without forming the state.
void pending_checks(fen_ctx *c)
{
    if (!c->pending || c->pending_checked) return;
    with_type(c, [&](auto tag) { fen::launch_project<decltype(tag)>(c, c->pend_dt, c->d_scal, false); });
    c->pending_checked = true;
}
}  // namespace

void fen::advance_fp64(fen_ctx *c, double dt) { advance<double>(c, dt); }

void fen::advance_fp64_into(fen_ctx *c, double dt, double *const *in, double *const *out)
{
    void *own[3] = {c->d_u[0], c->d_u[1], c->d_u[2]};
    for (int d = 0; d < 3; d++) c->d_u[d] = in[d];
    predict<double>(c, dt);
    solve_phi<double>(c, dt, c->d_us);
    for (int d = 0; d < 3; d++) c->d_u[d] = out[d];   // the projection writes the new state here
    finish_step<double>(c, dt);
    for (int d = 0; d < 3; d++) c->d_u[d] = own[d];
}

void fen::states_fp64(fen_ctx *c, double dt, int n, double *const *seg)
{
    if (c->pending) throw std::runtime_error("states_fp64 needs a completed state");
    const size_t ng = c->n_ghost;
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(seg[0] + d * ng, c->d_u[d], ng * sizeof(double), hipMemcpyDeviceToDevice,
                                     c->stream));
    if (n < 2) return;
    if (!deferrable(c) || c->has_source) {
        for (int s = 0; s + 1 < n; s++) {
            double *in[3], *out[3];
            for (int d = 0; d < 3; d++) {
                in[d] = seg[s] + d * ng;
                out[d] = seg[s + 1] + d * ng;
            }
            advance_fp64_into(c, dt, in, out);
        }
        return;
    }
    // fused: each step's projection, run with the next step's momentum,
    // saves the state it forms into its slot; the last state is the
    // completed projection of the last step, written into the last slot
    step_pending<double>(c, dt);
    for (int s = 1; s + 1 < n; s++) {
        void *save[3] = {seg[s], seg[s] + ng, seg[s] + 2 * ng};
        step_pending<double>(c, dt, save);
    }
    void *own[3] = {c->d_u[0], c->d_u[1], c->d_u[2]};
    for (int d = 0; d < 3; d++) c->d_u[d] = seg[n - 1] + d * ng;
    flush_pending(c);
    for (int d = 0; d < 3; d++) c->d_u[d] = own[d];
}

void fen::run_fp64(fen_ctx *c, int nsteps, double dt)
{
    if (deferrable(c)) {
        for (int s = 0; s < nsteps; s++) step_pending<double>(c, dt);
        flush_pending(c);
    } else {
        for (int s = 0; s < nsteps; s++) advance<double>(c, dt);
    }
}

void fen::flush_state(fen_ctx *c) { flush_pending(c); }

namespace {

// One step through the cached graph, capturing it first if dt, dt_o or the
// configuration changed.  The host-side effects of a step (last_dt, the
// pending-shift flag, checks_valid) are those advance() leaves, which do not
// depend on the state before the step.
void advance_graph(fen_ctx *c, double dt)
{
    if (!c->step_graph || c->graph_dt != dt || c->graph_dt_o != c->dt_o || c->graph_version != c->config_version) {
        if (c->step_graph) {
            FEN_HIP_CHECK(hipGraphExecDestroy(c->step_graph));
            c->step_graph = nullptr;
        }
        hipGraph_t g = nullptr;
        FEN_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        try {
            with_type(c, [&](auto tag) { advance<decltype(tag)>(c, dt); });
        } catch (...) {
            (void)hipStreamEndCapture(c->stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        FEN_HIP_CHECK(hipStreamEndCapture(c->stream, &g));
        const hipError_t e = hipGraphInstantiate(&c->step_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        FEN_HIP_CHECK(e);
        c->graph_dt = dt;
        c->graph_dt_o = c->dt_o;
        c->graph_version = c->config_version;
    } else {
        c->last_dt = dt;
        c->step_count++;
        const bool mean = removes_mean(c);
        bool dirichlet = false;
        for (int f = 0; f < 6; f++) dirichlet = dirichlet || c->bc_type[FEN_PHI][f] == 1;
        c->shift_active = mean && !dirichlet;
        c->checks_valid = true;
    }
    FEN_HIP_CHECK(hipGraphLaunch(c->step_graph, c->stream));
}

}  // namespace

#define FEN_TRY(ctx, ...)                                              \
    try {                                                              \
        __VA_ARGS__;                                                   \
    } catch (const std::exception &e_) {                               \
        return fail(ctx, FEN_ERR_HIP, e_.what());                      \
    }

extern "C" {

int fen_create(fen_ctx **out, int dim, int nx, int ny, int nz, double lx, double ly, double lz, const double *x0,
               const int *bc, int precision)
{
    if (!out) return FEN_ERR_ARG;
    *out = nullptr;
    if (dim != 2 && dim != 3) return FEN_ERR_ARG;
    if (!bc) return FEN_ERR_ARG;
    if (dim == 2) nz = 1;
    if (precision != FEN_FP64 && precision != FEN_FP32_REFINE) return FEN_ERR_UNSUPPORTED;
    fen_ctx *c = new fen_ctx();
    c->dim = dim; c->nx = nx; c->ny = ny; c->nz = nz;
    c->lx = lx; c->ly = ly; c->lz = lz;
    for (int d = 0; d < 3; d++) c->x0[d] = x0 ? x0[d] : 0.0;
    for (int s = 0; s < 6; s++) c->bc_phys[s] = (dim == 3 || s < 4) ? bc[s] : FEN_BC_PERIODIC;
    c->precision = precision;
    if (const char *g = std::getenv("FEN_STEP_GRAPH")) c->use_graph = std::atoi(g) != 0;
    if (const char *g = std::getenv("FEN_FUSED_RUN")) c->fused_run = std::atoi(g) != 0;
    c->delta = lx / (double)(float)nx;                                   // src/grid.f90:140
    c->periodic[0] = c->bc_phys[0] == FEN_BC_PERIODIC && c->bc_phys[1] == FEN_BC_PERIODIC;
    c->periodic[1] = c->bc_phys[2] == FEN_BC_PERIODIC && c->bc_phys[3] == FEN_BC_PERIODIC;
    c->periodic[2] = dim == 3 ? (c->bc_phys[4] == FEN_BC_PERIODIC && c->bc_phys[5] == FEN_BC_PERIODIC) : 1;
    const int *pb = c->periodic;
    // src/poisson.f90:70-111; other combinations stop in the reference (:91-94)
    if (dim == 3) {
        if (pb[0] && pb[1] && pb[2]) c->solver = FEN_SOLVER_PPP;
        else if (pb[0] && pb[1] && !pb[2]) c->solver = FEN_SOLVER_PPN;
        else if (!pb[0] && pb[1] && !pb[2]) c->solver = FEN_SOLVER_NPN;
        else if (!pb[0] && !pb[1] && !pb[2]) c->solver = FEN_SOLVER_NNN;
        else { delete c; return FEN_ERR_UNSUPPORTED; }
    } else {
        if (pb[0] && pb[1]) c->solver = FEN_SOLVER_PP;
        else if (pb[0] && !pb[1]) c->solver = FEN_SOLVER_PN;
        else if (!pb[0] && !pb[1]) c->solver = FEN_SOLVER_NN;
        else { delete c; return FEN_ERR_UNSUPPORTED; }
    }
    const int sv = c->solver;
    const bool neumann_solver = sv == FEN_SOLVER_PN || sv == FEN_SOLVER_NN || sv == FEN_SOLVER_NPN ||
                                sv == FEN_SOLVER_NNN;
    // transformed directions need power-of-two lengths, and so do the rows of
    // the x r2c / c2r (pn: its Thomas direction y); Thomas directions >= 2
    const bool y_transformed = !(sv == FEN_SOLVER_PN || sv == FEN_SOLVER_NN);
    if (!pow2(nx) || ((y_transformed || sv == FEN_SOLVER_PN) && !pow2(ny)) || (sv == FEN_SOLVER_PPP && !pow2(nz)) || (dim == 3 && nz < 2) ||
        ny < 2 || (neumann_solver && precision != FEN_FP64)) {
        delete c;
        return FEN_ERR_UNSUPPORTED;
    }
    for (int s = 0; s < 5; s++)
        for (int f = 0; f < 6; f++) {
            bool ok = true;
            c->bc_type[s][f] = (dim == 3 || f < 4) ? map_bc(s, f, c->bc_phys[f], &ok) : 0;
            if (!ok) { delete c; return FEN_ERR_UNSUPPORTED; }
        }
    refresh_images(c);
    Box &b = c->box;
    b.nx = nx; b.ny = ny; b.nz = nz; b.dim = dim;
    b.gx = nx + 2;
    b.gxy = (long)(nx + 2) * (ny + 2);
    c->n_ghost = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    c->n_int = (size_t)nx * ny * nz;
    c->hx = nx / 2 + 1;
    try {
        FEN_HIP_CHECK(hipStreamCreate(&c->stream));
        c->d_noalloc = noalloc_scratch();
        {
            int dev = 0;
            FEN_HIP_CHECK(hipGetDevice(&dev));
            FEN_HIP_CHECK(hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
        }
        const size_t es = elem(c);
        for (int d = 0; d < 3; d++) {
            c->d_u[d] = dev_alloc_bytes(c->n_ghost * es);
            c->d_us[d] = dev_alloc_bytes(c->n_ghost * es);
            c->d_dvo[d] = dev_alloc_bytes(c->n_int * es);
        }
        c->d_p = dev_alloc_bytes(c->n_ghost * es);
        c->d_phi = dev_alloc_bytes(c->n_ghost * es);
        for (int s = 0; s < 5; s++)
            for (int f = 0; f < 6; f++) c->d_plane[s][f] = dev_alloc_bytes(plane_count(c, f) * es);
        {   // the planes of u, v, w, p in a device table (the general-wall fused kernel reads them there)
            void *tab[24];
            const int fid[4] = {FEN_U, FEN_V, FEN_W, FEN_P};
            for (int q = 0; q < 4; q++)
                for (int f = 0; f < 6; f++) tab[6 * q + f] = c->d_plane[fid[q]][f];
            c->d_plane_tab = static_cast<void **>(dev_alloc_bytes(sizeof tab));
            FEN_HIP_CHECK(hipMemcpy(c->d_plane_tab, tab, sizeof tab, hipMemcpyHostToDevice));
        }
        const size_t nspec = (size_t)c->hx * ny * nz;
        c->d_spec = dev_alloc_bytes(nspec * 2 * es);
        const size_t hy = (size_t)ny / 2 + 1;
        if (c->solver == FEN_SOLVER_PPN) {
            c->d_c1 = dev_alloc<double>(nspec);
            if (precision != FEN_FP64) c->d_dd = dev_alloc<double2>(nspec);
        }
        if (c->solver == FEN_SOLVER_PN) c->d_c1 = dev_alloc<double>((size_t)c->hx * ny);
        if (c->solver == FEN_SOLVER_NN || c->solver == FEN_SOLVER_NNN) c->d_c1 = dev_alloc<double>(c->n_int);
        if (c->solver == FEN_SOLVER_NPN) {
            c->d_spec2 = dev_alloc<double2>((size_t)nx * nz * hy);
            c->d_c1 = dev_alloc<double>((size_t)nx * nz * hy);
        }
        // DCT directions use the pi-based wavenumbers (src/poisson.f90:265,794,881,899)
        c->d_mwx = dev_copy(modified_wavenumbers(nx, c->delta, pb[0] ? 2.0 : 1.0));
        c->d_mwy = dev_copy(modified_wavenumbers(ny, c->delta, pb[1] ? 2.0 : 1.0));
        if (c->solver == FEN_SOLVER_PPP) c->d_mwz = dev_copy(modified_wavenumbers(nz, c->delta, 2.0));
        // tridiagonal coefficients along z (ppn, npn, nnn) or y (pn, nn): src/poisson.f90:214-226,741-757
        const bool thomas_y = c->solver == FEN_SOLVER_PN || c->solver == FEN_SOLVER_NN;
        if (thomas_y || c->solver == FEN_SOLVER_PPN || c->solver == FEN_SOLVER_NPN || c->solver == FEN_SOLVER_NNN) {
            const int n = thomas_y ? ny : nz;
            const int lo_face = thomas_y ? FEN_BOTTOM : FEN_FRONT, hi_face = thomas_y ? FEN_TOP : FEN_BACK;
            const double d2 = c->delta * c->delta;
            std::vector<double> a(n, 1.0 / d2), bb(n, -(2.0 / d2)), cc(n, 1.0 / d2);
            bb[0] = bb[0] + a[0];
            if (c->bc_phys[lo_face] == FEN_BC_INFLOW && c->bc_phys[hi_face] == FEN_BC_OUTFLOW)
                bb[n - 1] = bb[n - 1] - cc[n - 1];
            else bb[n - 1] = bb[n - 1] + cc[n - 1];
            a[0] = 0.0;
            cc[n - 1] = 0.0;
            c->d_ta = dev_copy(a); c->d_tb = dev_copy(bb); c->d_tc = dev_copy(cc);
        }
        if (!pb[0]) c->d_mkx = dev_copy(dct_twiddles(nx));
        if (c->solver == FEN_SOLVER_NNN) c->d_mky = dev_copy(dct_twiddles(ny));
        with_type(c, [&](auto tag) {
            using R = decltype(tag);
            c->d_twx = dev_copy(twiddles<R>(nx));
            if (y_transformed) c->d_twy = dev_copy(twiddles<R>(ny));
            if (c->solver == FEN_SOLVER_PPP) c->d_twz = dev_copy(twiddles<R>(nz));
        });
        if (c->d_c1) {
            fen::launch_thomas_c1(c);
            FEN_HIP_CHECK(hipGetLastError());
        }
        c->n_part = std::max<size_t>(std::max<size_t>(3 * 1024, (size_t)ny * nz / 2 + 1),
                                     3 * (size_t)fen::stencil_blocks(c->box));
        c->d_part = dev_alloc<double>(c->n_part);
        c->d_scal = dev_alloc<double>(8);
        c->d_shift = c->d_scal + 3;
        FEN_HIP_CHECK(hipHostMalloc((void **)&c->h_scal, 8 * sizeof(double)));
        FEN_HIP_CHECK(hipDeviceSynchronize());
    } catch (const std::exception &) {
        free_all(c);
        delete c;
        return FEN_ERR_HIP;
    }
    *out = c;
    return FEN_OK;
}

int fen_destroy(fen_ctx *c)
{
    if (!c) return FEN_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    delete c;
    return FEN_OK;
}

const char *fen_last_error(const fen_ctx *c) { return c ? c->err.c_str() : "null context"; }

int fen_set_params(fen_ctx *c, double density, double viscosity, const double *g, double cfl, int constant_cfl)
{
    if (!c) return FEN_ERR_ARG;
    if (!(density > 0.0) || !(viscosity >= 0.0)) return fail(c, FEN_ERR_ARG, "density must be > 0, viscosity >= 0");
    const double gn[3] = {g ? g[0] : 0.0, g ? g[1] : 0.0, g ? g[2] : 0.0};
    // a caller that sets the same values every step (the Fortran layer does)
    // must not invalidate the captured step graph
    bool same = c->density == density && c->viscosity == viscosity && c->cfl == cfl &&
                (c->constant_cfl != 0) == (constant_cfl != 0);
    for (int d = 0; d < 3; d++) same = same && c->g[d] == gn[d];
    if (same) return FEN_OK;
    FEN_TRY(c, flush_pending(c));
    c->density = density;
    c->viscosity = viscosity;
    for (int d = 0; d < 3; d++) c->g[d] = gn[d];
    c->cfl = cfl;
    c->constant_cfl = constant_cfl;
    c->config_version++;
    return FEN_OK;
}

int fen_set_timestep(fen_ctx *c, double U, double *dt)
{
    if (!c || !dt) return FEN_ERR_ARG;
    // src/navier_stokes.f90:641-664 (single-phase branch)
    const double d = c->delta;
    c->dt_conv = c->cfl * d / U;
    c->dt_visc = 0.125 * d * d * c->density / c->viscosity;
    if (c->dim == 3) c->dt_visc = (1.0 / 6.0) * d * d * c->density / c->viscosity;
    *dt = c->dt_conv < c->dt_visc ? c->dt_conv : c->dt_visc;
    if (c->mf) {   // MF branch (:653-660)
        const MFState &m = *c->mf;
        c->dt_visc = 0.125 * d * d * std::fmin(m.rho0 / m.mu0, m.rho1 / m.mu1);
        *dt = std::fmin(c->dt_conv, c->dt_visc);
        if (m.sigma > 0.0) {
            const double pi = std::acos(-1.0);
            const double dt_surf = std::sqrt(0.5 * (m.rho0 + m.rho1) * (d * d * d) / (pi * m.sigma + 1.0e-16));
            *dt = std::fmin(*dt, dt_surf);
            c->mf->dt_surf = dt_surf;
            c->mf->have_dt_surf = true;
        }
    }
    c->dt_o = *dt;
    c->have_dt_o = c->have_dt_visc = true;
    return FEN_OK;
}

int fen_set_multiphase(fen_ctx *c, double rho0, double rho1, double mu0, double mu1, double sigma, double beta)
{
    if (!c) return FEN_ERR_ARG;
    if (c->dim != 2 || c->precision != FEN_FP64 || c->mf)
        return fail(c, FEN_ERR_UNSUPPORTED, "the two-phase mode is 2-D, FP64, set once");
    FEN_TRY(c, flush_pending(c));
    for (int f = 0; f < 4; f++)
        if (c->bc_phys[f] != FEN_BC_PERIODIC && c->bc_phys[f] != FEN_BC_WALL)
            return fail(c, FEN_ERR_UNSUPPORTED, "two-phase faces must be Periodic or Wall");
    if (!(rho0 > 0.0 && rho1 > 0.0 && mu0 > 0.0 && mu1 > 0.0 && beta > 0.0 && sigma >= 0.0))
        return fail(c, FEN_ERR_ARG, "two-phase properties must be positive");
    FEN_TRY(c, {
        MFState *m = new MFState();
        m->rho0 = rho0; m->rho1 = rho1; m->mu0 = mu0; m->mu1 = mu1; m->sigma = sigma; m->beta = beta;
        m->rhomin = std::fmin(rho0, rho1);   // src/solver.f90:91-92
        m->irhomin = 1.0 / m->rhomin;
        c->mf = m;
        for (double **q : {&m->vof, &m->vof1, &m->vof2, &m->h, &m->dd, &m->curv, &m->nrx, &m->nry, &m->lx, &m->ly,
                           &m->rho, &m->mu, &m->p_hat, &m->p_o})
            *q = dev_alloc<double>(c->n_ghost);
        c->config_version++;
    });
    return FEN_OK;
}

int fen_update_material_properties(fen_ctx *c)
{
    if (!c) return FEN_ERR_ARG;
    if (!c->mf) return fail(c, FEN_ERR_STATE, "not in two-phase mode");
    FEN_TRY(c, {
        flush_pending(c);
        fen::mf_update_fields(c, false);
        sync(c);
    });
    return FEN_OK;
}

int fen_advect_vof(fen_ctx *c, double dt)
{
    if (!c) return FEN_ERR_ARG;
    if (!c->mf) return fail(c, FEN_ERR_STATE, "not in two-phase mode");
    FEN_TRY(c, {
        flush_pending(c);
        fen::mf_advect_vof(c, dt);
        sync(c);
    });
    return FEN_OK;
}

int fen_set_vof_x_first(fen_ctx *c, int x_first)
{
    if (!c) return FEN_ERR_ARG;
    if (!c->mf) return fail(c, FEN_ERR_STATE, "not in two-phase mode");
    c->mf->x_first = x_first != 0;
    return FEN_OK;
}

int fen_get_vof_x_first(fen_ctx *c, int *x_first)
{
    if (!c || !x_first) return FEN_ERR_ARG;
    if (!c->mf) return fail(c, FEN_ERR_STATE, "not in two-phase mode");
    *x_first = c->mf->x_first ? 1 : 0;
    return FEN_OK;
}

int fen_set_dt_o(fen_ctx *c, double dt_o)
{
    if (!c) return FEN_ERR_ARG;
    c->dt_o = dt_o;
    c->have_dt_o = true;
    return FEN_OK;
}

int fen_get_scalars(fen_ctx *c, double *maxdiv, double *maxcfl, double *dt_o, double *dt_visc, double *dt_conv)
{
    if (!c) return FEN_ERR_ARG;
    FEN_TRY(c, {
        if (dt_o) *dt_o = c->dt_o;
        if (dt_visc) *dt_visc = c->dt_visc;
        if (dt_conv) *dt_conv = c->dt_conv;
        if ((maxdiv || maxcfl) && !c->checks_valid) {   // no step yet: nothing measured
            if (maxdiv) *maxdiv = 0.0;
            if (maxcfl) *maxcfl = 0.0;
        } else if (maxdiv || maxcfl) {
            pending_checks(c);
            FEN_HIP_CHECK(hipMemcpyAsync(c->h_scal, c->d_scal, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            sync(c);
            if (maxdiv) *maxdiv = c->h_scal[0];
            if (maxcfl) *maxcfl = c->last_dt * c->h_scal[1] / c->delta;    // src/navier_stokes.f90:617
            if (c->h_scal[2] != 0.0) return fail(c, FEN_ERR_NONFINITE, "non-finite velocity or divergence");
        }
    });
    return FEN_OK;
}

int fen_field_count(const fen_ctx *c, int field, size_t *count)
{
    if (!c || !count) return FEN_ERR_ARG;
    if (field < FEN_P || field > FEN_MU) return FEN_ERR_ARG;
    if (field >= FEN_VOF && !c->mf) return FEN_ERR_ARG;
    *count = (field <= FEN_W || field >= FEN_VOF) ? c->n_ghost : c->n_int;
    return FEN_OK;
}

int fen_upload(fen_ctx *c, int field, const double *host, size_t count)
{
    if (!c || !host) return FEN_ERR_ARG;
    size_t n = 0;
    if (fen_field_count(c, field, &n) != FEN_OK || n != count)
        return fail(c, FEN_ERR_ARG, "field id or element count does not match the grid");
    if (field == FEN_W && c->dim == 2) return fail(c, FEN_ERR_ARG, "no w component in 2-D");
    if (field == FEN_RHO || field == FEN_MU)
        return fail(c, FEN_ERR_ARG, "rho and mu follow from the vof (fen_update_material_properties)");
    FEN_TRY(c, {
        // a source term is read by the next momentum only; anything else
        // completes a pending step first
        if (field < FEN_S_X || field > FEN_S_Z) flush_pending(c);
        if (field >= FEN_S_X && field <= FEN_S_Z && !c->d_S[0]) {
            for (int d = 0; d < 3; d++) c->d_S[d] = dev_alloc_bytes(c->n_int * elem(c));
        }
        if (field >= FEN_S_X && field <= FEN_S_Z && !c->has_source) {
            c->has_source = true;
            c->config_version++;
        }
        copy_in(c, field_ptr(c, field), host, n);
        if (field == FEN_PHI) c->shift_active = false;
        if (field >= FEN_U && field <= FEN_W) c->checks_valid = false;
    });
    return FEN_OK;
}

int fen_download(fen_ctx *c, int field, double *host, size_t count)
{
    if (!c || !host) return FEN_ERR_ARG;
    size_t n = 0;
    if (fen_field_count(c, field, &n) != FEN_OK || n != count)
        return fail(c, FEN_ERR_ARG, "field id or element count does not match the grid");
    FEN_TRY(c, {
        flush_pending(c);
        const void *src = field_ptr(c, field);
        if (!src) {   // source never set: zero, as allocated by the reference
            std::memset(host, 0, n * sizeof(double));
            return FEN_OK;
        }
        copy_out(c, host, src, n);
        if (field == FEN_PHI && c->shift_active) {
            FEN_HIP_CHECK(hipMemcpyAsync(c->h_scal + 3, c->d_shift, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            sync(c);
            const double s = c->h_scal[3];
            if (c->precision == FEN_FP64) {
                for (size_t q = 0; q < n; q++) host[q] = host[q] - s;
            } else {   // as the device applies it: FP64 difference rounded to R
                for (size_t q = 0; q < n; q++) host[q] = (double)(float)(host[q] - s);
            }
        }
    });
    return FEN_OK;
}

int fen_set_bc_plane(fen_ctx *c, int field, int face, const double *plane, size_t count)
{
    if (!c || !plane) return FEN_ERR_ARG;
    const int s = field_slot(field);
    if (s < 0 || face < 0 || face > 5) return fail(c, FEN_ERR_ARG, "bad field or face");
    if (count != plane_count(c, face)) return fail(c, FEN_ERR_ARG, "plane size does not match the face");
    FEN_TRY(c, {
        flush_pending(c);
        copy_in(c, c->d_plane[s][face], plane, count);
    });
    return FEN_OK;
}

int fen_set_bc_type(fen_ctx *c, int field, int face, int type)
{
    if (!c) return FEN_ERR_ARG;
    const int s = field_slot(field);
    if (s < 0 || face < 0 || face > 5 || type < 0 || type > 2) return fail(c, FEN_ERR_ARG, "bad field, face or type");
    FEN_TRY(c, flush_pending(c));
    c->bc_type[s][face] = type;
    refresh_images(c);
    c->config_version++;
    return FEN_OK;
}

int fen_get_bc_type(const fen_ctx *c, int field, int face, int *type)
{
    if (!c || !type) return FEN_ERR_ARG;
    const int s = field_slot(field);
    if (s < 0 || face < 0 || face > 5) return FEN_ERR_ARG;
    *type = c->bc_type[s][face];
    return FEN_OK;
}

int fen_update_ghost_nodes(fen_ctx *c, int field)
{
    if (!c) return FEN_ERR_ARG;
    if (field == FEN_VOF && c->mf) {   // vof%update_ghost_nodes: Periodic or Neumann
        FEN_TRY(c, fen::mf_ghost_field(c, c->mf->vof));
        return FEN_OK;
    }
    if (field_slot(field) < 0) return fail(c, FEN_ERR_ARG, "field has no ghost nodes");
    if (field == FEN_W && c->dim == 2) return fail(c, FEN_ERR_ARG, "no w component in 2-D");
    FEN_TRY(c, {
        flush_pending(c);
        with_type(c, [&](auto tag) {
            using R = decltype(tag);
            if (field == FEN_PHI && c->shift_active) {
                fen::launch_apply_shift<R>(c, as<R>(c->d_phi), c->n_ghost);
                c->shift_active = false;
            }
            GhostDesc<R> g = make_desc<R>(c, field, field_ptr(c, field));
            fen::launch_ghost_desc<R>(c, &g, 1);
        });
        FEN_HIP_CHECK(hipGetLastError());
    });
    return FEN_OK;
}

int fen_ns_step(fen_ctx *c, int step, double *dt, double *maxdiv, double *maxcfl)
{
    (void)step;
    if (!c || !dt) return FEN_ERR_ARG;
    if (!c->have_dt_o) return fail(c, FEN_ERR_STATE, "dt_o is undefined: call fen_set_timestep or fen_set_dt_o");
    // the reference leaves dt_surf unset unless sigma > 0 (src/navier_stokes.f90:29, 657) and
    // takes min(dt_conv, dt_visc, dt_surf) with it (:723)
    if (c->mf && c->constant_cfl && !c->mf->have_dt_surf)
        return fail(c, FEN_ERR_UNSUPPORTED,
                    "constant CFL in the two-phase mode needs sigma > 0 and fen_set_timestep (dt_surf)");
    FEN_TRY(c, {
        if (c->constant_cfl) {
            // update_timestep (src/navier_stokes.f90:670-730): the max of
            // |u|+|v|+|w| of the state, from the last step's checks
            if (!c->have_dt_visc) return fail(c, FEN_ERR_STATE, "constant CFL needs fen_set_timestep first");
            if (c->pending) {
                pending_checks(c);
            } else if (!c->checks_valid) {
                with_type(c, [&](auto tag) {
                    using R = decltype(tag);
                    const R *vn[3] = {as<R>(c->d_u[0]), as<R>(c->d_u[1]), as<R>(c->d_u[2])};
                    fen::launch_checks<R>(c, vn, c->d_scal);
                });
            }
            FEN_HIP_CHECK(hipMemcpyAsync(c->h_scal, c->d_scal, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
            sync(c);
            c->dt_o = *dt;
            const double mv = c->h_scal[1];
            c->dt_conv = mv > 0.0 ? c->cfl * c->delta / mv : 1.0;
            *dt = c->dt_conv < c->dt_visc ? c->dt_conv : c->dt_visc;
            if (c->mf && c->mf->dt_surf < *dt) *dt = c->mf->dt_surf;   // MF build (:723)
            if (*dt > 1.1 * c->dt_o) *dt = 1.1 * c->dt_o;
        }
        const double dtv = *dt;
        if (deferrable(c)) {
            with_type(c, [&](auto tag) { step_pending<decltype(tag)>(c, dtv); });
        } else {
            flush_pending(c);
            if (c->mf) advance_mf(c, dtv);
            else if (c->use_graph && !c->constant_cfl && !c->prof_mask) advance_graph(c, dtv);
            else with_type(c, [&](auto tag) { advance<decltype(tag)>(c, dtv); });
        }
        FEN_HIP_CHECK(hipGetLastError());
    });
    if (maxdiv || maxcfl) return fen_get_scalars(c, maxdiv, maxcfl, nullptr, nullptr, nullptr);
    return FEN_OK;
}

int fen_ns_run(fen_ctx *c, int step0, int nsteps, double *dt)
{
    if (!c || !dt || nsteps < 0) return FEN_ERR_ARG;
    for (int s = 0; s < nsteps; s++) {
        const int rc = fen_ns_step(c, step0 + 1 + s, dt, nullptr, nullptr);
        if (rc != FEN_OK) return rc;
    }
    return FEN_OK;
}

namespace {

// The whole-array semantics of the reference's solvers that the device
// solve leaves to the host: the x backward transform writes the interior
// only, so with ghosts (gl = 1) ppn/npn/nnn leave them 0 (phi%f = 0 first,
// src/poisson.f90:1151, 1297, 1436) and nn divides them by 2nx with the rest
// (:594); the mean s is subtracted over the whole array (pp :503, ppn
// :1171) or the interior (pn :408).
void poisson_host_fixups(const fen_ctx *c, double *f, size_t n, int gl, bool mean, double s)
{
    const size_t gx = c->nx + 2, gy = c->ny + 2;
    auto interior = [&](size_t q) {
        const size_t i = q % gx, j = (q / gx) % gy, k = q / (gx * gy);
        return i >= 1 && i <= (size_t)c->nx && j >= 1 && j <= (size_t)c->ny && k >= 1 && k <= (size_t)c->nz;
    };
    const int sv = c->solver;
    if (gl && (sv == FEN_SOLVER_PPN || sv == FEN_SOLVER_NPN || sv == FEN_SOLVER_NNN)) {
        for (size_t q = 0; q < n; q++)
            if (!interior(q)) f[q] = 0.0;
    }
    if (gl && sv == FEN_SOLVER_NN) {
        const double s2 = (double)(float)(2 * c->nx);
        for (size_t q = 0; q < n; q++)
            if (!interior(q)) f[q] = f[q] / s2;
    }
    if (mean) {
        if (sv == FEN_SOLVER_PN) {
            for (size_t q = 0; q < n; q++)
                if (!gl || interior(q)) f[q] = f[q] - s;
        } else {
            for (size_t q = 0; q < n; q++) f[q] = f[q] - s;
        }
    }
}

}  // namespace

int fen_poisson_solve(fen_ctx *c, double *f, int gl)
{
    if (!c || !f || (gl != 0 && gl != 1)) return FEN_ERR_ARG;
    const size_t n = gl ? c->n_ghost : c->n_int;
    const long sx = gl ? c->nx + 2 : c->nx;
    const long sy = gl ? (long)(c->nx + 2) * (c->ny + 2) : (long)c->nx * c->ny;
    FEN_TRY(c, {
        flush_pending(c);
        // the solve writes its own mean into the shift slot: apply phi's
        // pending shift first
        if (c->shift_active) {
            with_type(c, [&](auto tag) {
                using R = decltype(tag);
                fen::launch_apply_shift<R>(c, as<R>(c->d_phi), c->n_ghost);
            });
            c->shift_active = false;
        }
        void *buf = c->d_us[0];   // scratch between steps
        copy_in(c, buf, f, n);
        const bool mean = removes_mean(c);
        with_type(c, [&](auto tag) {
            using R = decltype(tag);
            solve_to_x_spectrum<R>(c, nullptr, as<R>(buf), sx, sy, gl, 1.0, 1.0);
            x_backward<R>(c, as<R>(buf), gl, sx, sy, mean, 0);
        });
        copy_out(c, f, buf, n);
        if (mean) {
            FEN_HIP_CHECK(hipMemcpyAsync(c->h_scal + 3, c->d_shift, sizeof(double), hipMemcpyDeviceToHost, c->stream));
            sync(c);
        }
        poisson_host_fixups(c, f, n, gl, mean, mean ? c->h_scal[3] : 0.0);
    });
    return FEN_OK;
}

int fen_poisson_solve_noalloc(fen_ctx *c, double *f, int gl)
{
    if (!c || !f || (gl != 0 && gl != 1)) return FEN_ERR_ARG;
    if (c->precision != FEN_FP64)
        return fail(c, FEN_ERR_UNSUPPORTED, "fen_poisson_solve_noalloc is for FP64 contexts");
    if (!c->d_noalloc) return fail(c, FEN_ERR_STATE, "no scratch address for this device");
    const bool mean = removes_mean(c);
    if (mean && (size_t)fen::c2r_blocks(c) > kNoAllocParts)
        return fail(c, FEN_ERR_UNSUPPORTED,
                    "grid too large for fen_poisson_solve_noalloc: the mean needs " +
                        std::to_string(fen::c2r_blocks(c)) + " block sums, the scratch holds " +
                        std::to_string(kNoAllocParts) + " (ny*nz/(4096/nx) at most)");
    const size_t n = gl ? c->n_ghost : c->n_int;
    const long sx = gl ? c->nx + 2 : c->nx;
    const long sy = gl ? (long)(c->nx + 2) * (c->ny + 2) : (long)c->nx * c->ny;
    double *part = c->d_noalloc, *shift = c->d_noalloc + kNoAllocParts;
    FEN_TRY(c, {
        // a pending step is completed first (its projection reads u* in the
        // buffer used below); phi's own pending mean stays where it is
        flush_pending(c);
        double *buf = as<double>(c->d_us[0]);   // context memory between steps
        FEN_HIP_CHECK(hipMemcpyAsync(buf, f, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        solve_to_x_spectrum<double>(c, nullptr, buf, sx, sy, gl, 1.0, 1.0);
        x_backward<double>(c, buf, gl, sx, sy, mean, 0, part, kNoAllocParts, shift);
        FEN_HIP_CHECK(hipMemcpyAsync(f, buf, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        double s = 0.0;
        if (mean) FEN_HIP_CHECK(hipMemcpyAsync(&s, shift, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        sync(c);
        poisson_host_fixups(c, f, n, gl, mean, s);
    });
    return FEN_OK;
}

int fen_add_advection(fen_ctx *c, double *rx, double *ry, double *rz)
{
    if (!c || !rx || !ry || (c->dim == 3 && !rz)) return FEN_ERR_ARG;
    if (c->precision != FEN_FP64) return fail(c, FEN_ERR_UNSUPPORTED, "add_advection is provided for FP64 contexts");
    FEN_TRY(c, {
        flush_pending(c);
        double *r[3] = {as<double>(c->d_us[0]), as<double>(c->d_us[1]), as<double>(c->d_us[2])};
        double *h[3] = {rx, ry, rz};
        const int nd = c->dim == 3 ? 3 : 2;
        for (int d = 0; d < nd; d++)
            FEN_HIP_CHECK(hipMemcpyAsync(r[d], h[d], c->n_int * sizeof(double), hipMemcpyHostToDevice, c->stream));
        fen::launch_advection_only(c, r[0], r[1], r[2]);
        fen::prof_end(c, fen::K_MOMENTUM, 0.0);
        for (int d = 0; d < nd; d++)
            FEN_HIP_CHECK(hipMemcpyAsync(h[d], r[d], c->n_int * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        sync(c);
    });
    return FEN_OK;
}

int fen_format_status(fen_ctx *c, int step, double time, double dt, char *buf, size_t len)
{
    if (!c || !buf || len == 0) return FEN_ERR_ARG;
    double md = 0.0, mc = 0.0;
    int rc = fen_get_scalars(c, &md, &mc, nullptr, nullptr, nullptr);
    if (rc != FEN_OK && rc != FEN_ERR_NONFINITE) return rc;
    const int rt = fen_status_text(step, time, dt, md, mc, buf, len);
    return rt != FEN_OK ? rt : rc;
}

int fen_write_grid_json(const fen_ctx *c, const char *path)
{
    if (!c || !path) return FEN_ERR_ARG;
    char text[1024];
    const int rc = fen_grid_json_text(c->nx, c->ny, c->nz, c->x0, c->lx, c->ly, c->lz, text, sizeof text);
    if (rc != FEN_OK) return rc;
    FILE *fp = std::fopen(path, "w");
    if (!fp) return FEN_ERR_ARG;
    std::fputs(text, fp);
    std::fclose(fp);
    return FEN_OK;
}

int fen_solver_kind(const fen_ctx *c, int *kind)
{
    if (!c || !kind) return FEN_ERR_ARG;
    *kind = c->solver;
    return FEN_OK;
}

int fen_grid_delta(const fen_ctx *c, double *delta)
{
    if (!c || !delta) return FEN_ERR_ARG;
    *delta = c->delta;
    return FEN_OK;
}

int fen_synchronize(fen_ctx *c)
{
    if (!c) return FEN_ERR_ARG;
    FEN_TRY(c, sync(c));
    return FEN_OK;
}

int fen_profile_enable(fen_ctx *c, int mask)
{
    if (!c) return FEN_ERR_ARG;
    c->prof_mask = (unsigned)mask;
    return FEN_OK;
}

int fen_profile_sample(fen_ctx *c, int every)
{
    if (!c || every < 1) return FEN_ERR_ARG;
    c->prof_every = every;
    c->prof_base = c->step_count;
    return FEN_OK;
}

int fen_profile_reset(fen_ctx *c)
{
    if (!c) return FEN_ERR_ARG;
    FEN_TRY(c, {
        sync(c);
        c->prof_marks.clear();
        c->prof_used = 0;
        c->prof_base = c->step_count;
    });
    return FEN_OK;
}

int fen_profile_count(void) { return fen::K_COUNT; }

const char *fen_profile_name(int k) { return k >= 0 && k < fen::K_COUNT ? kKernelNames[k] : ""; }

int fen_profile_read(fen_ctx *c, int k, double *total_ms, long long *launches, double *bytes)
{
    if (!c || k < 0 || k >= fen::K_COUNT) return FEN_ERR_ARG;
    FEN_TRY(c, {
        sync(c);
        double ms = 0.0;
        long long n = 0;
        for (size_t m = 1; m < c->prof_marks.size(); m++) {
            const fen::ProfMark &cur = c->prof_marks[m], &prev = c->prof_marks[m - 1];
            if (cur.kernel != k || cur.begin) continue;
            const bool pair = prev.begin && prev.seq == cur.seq;             // sampled mode
            const bool chain = !prev.begin && prev.seq == cur.seq - 1;       // every launch
            if (!pair && !chain) continue;
            float e = 0.0f;
            FEN_HIP_CHECK(hipEventElapsedTime(&e, prev.ev, cur.ev));
            ms += e;
            n++;
        }
        if (total_ms) *total_ms = ms;
        if (launches) *launches = n;
        if (bytes) *bytes = c->prof[k].bytes;
    });
    return FEN_OK;
}

int fen_profile_predecessor(fen_ctx *c, int k, int *prev)
{
    if (!c || !prev || k < 0 || k >= fen::K_COUNT) return FEN_ERR_ARG;
    *prev = -1;
    for (size_t m = 1; m < c->prof_marks.size(); m++) {
        const fen::ProfMark &cur = c->prof_marks[m], &pm = c->prof_marks[m - 1];
        if (cur.kernel == k && !cur.begin && !pm.begin && pm.seq == cur.seq - 1) { *prev = pm.kernel; break; }
    }
    return FEN_OK;
}

int fen_adjoint_record(fen_ctx *c, int step0, int nsteps, int every, double *dt)
{
    if (!c || !dt || nsteps < 1 || every < 1) return FEN_ERR_ARG;
    if (const char *why = fen::adjoint_unsupported(c)) return fail(c, FEN_ERR_UNSUPPORTED, why);
    if (!c->have_dt_o) return fail(c, FEN_ERR_STATE, "dt_o is undefined: call fen_set_timestep or fen_set_dt_o");
    FEN_TRY(c, {
        flush_pending(c);
        fen::adjoint_record(c, step0, nsteps, every, *dt);
    });
    return FEN_OK;
}

int fen_adjoint_backward(fen_ctx *c, const double *seed_u, const double *seed_v, const double *seed_w,
                         double *grad_u, double *grad_v, double *grad_w, double *grad_p)
{
    if (!c || !seed_u || !seed_v || !seed_w || !grad_u || !grad_v || !grad_w) return FEN_ERR_ARG;
    if (!c->tape) return fail(c, FEN_ERR_STATE, "no recorded run: call fen_adjoint_record first");
    const double *seed[3] = {seed_u, seed_v, seed_w};
    double *grad[3] = {grad_u, grad_v, grad_w};
    FEN_TRY(c, {
        flush_pending(c);
        fen::adjoint_backward(c, seed, grad, grad_p);
    });
    return FEN_OK;
}

}  // extern "C"
