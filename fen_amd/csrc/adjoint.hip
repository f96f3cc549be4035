// adjoint.hip -- discrete adjoint of the FP64 time step (3-D, triply
// periodic, ppp, constant density and viscosity, fixed dt).  No counterpart
// in the reference: it differentiates navier_stokes_solver
// (src/navier_stokes.f90:50-136) as this library computes it.
//
// One forward step maps (u, p, dv_o) -> (u', p', dv_o'):
//   dv  = adv(u) + mu lap(u) / rho                 (advection in flux form)
//   u*  = u + dt (-G p / rho + A dv + B dv_o)       dv_o' = dv
//   phi = S (rho/dt) D u*                           (S: the FFT Poisson solve)
//   u'  = u* - (dt/rho) G phi                       p' = p + phi
// with D the MAC divergence, G the gradient (G^T = -D on the periodic grid)
// and S symmetric.  Its transpose, given (ub', pb', dvob'):
//   rb   = S (pb' + (dt/rho) D ub')       (the adjoint of the solve's right-hand side)
//   ub*  = ub' - (rho/dt) G rb
//   pb   = pb' + (dt/rho) D ub*
//   dvb  = dt A ub* + dvob'              dvob = dt B ub*
//   ub   = ub* + J_adv(u)^T dvb + (mu/rho) lap(dvb)
// J_adv^T is evaluated in gather form from the six advective fluxes
//   Qxx(i) = (u_i + u_{i+1})^2/4,  Rxy(i,j) = (u_{j} + u_{j+1})(v_i + v_{i+1})/4, ...
// each used by two momentum equations with opposite signs.  The adjoint of
// the forward state is needed at every step: a run is recorded with
// checkpoints every K steps and each segment is recomputed before its
// backward sweep.  Adjoint arrays are interior-sized (nx*ny*nz), i fastest;
// the recorded states are ghosted with their periodic images in place.
#include <algorithm>
#include <stdexcept>
#include <vector>

#include "fen_internal.hpp"

namespace fen {

namespace {

constexpr int kAX = 64, kAY = 4;

struct AdjDims {
    int nx, ny, nz;
    long gx, gxy;
    __device__ long q(int i, int j, int k) const { return i + (long)nx * (j + (long)ny * k); }   // 0-based
    __device__ long g(int i, int j, int k) const { return (i + 1) + gx * (j + 1) + gxy * (k + 1); }
};

__device__ __forceinline__ int up(int i, int n) { return i + 1 == n ? 0 : i + 1; }
__device__ __forceinline__ int dn(int i, int n) { return i == 0 ? n - 1 : i - 1; }

// r = pb + (dt/rho) D ub
__global__ __launch_bounds__(kAX * kAY) void adj_rhs_kernel(AdjDims d, const double *ub, const double *vb,
                                                           const double *wb, const double *pb, double *r,
                                                           double id, double dt_rho)
{
    const int i = blockIdx.x * kAX + threadIdx.x % kAX, j = blockIdx.y * kAY + threadIdx.x / kAX, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const long q = d.q(i, j, k);
    const double div = ((ub[q] - ub[d.q(dn(i, d.nx), j, k)]) * id + (vb[q] - vb[d.q(i, dn(j, d.ny), k)]) * id) +
                       (wb[q] - wb[d.q(i, j, dn(k, d.nz))]) * id;
    r[q] = pb[q] + dt_rho * div;
}

struct AdjProjArgs {
    const double *ub, *vb, *wb;   // ub'
    const double *psi;            // rb = S (pb' + (dt/rho) D ub')
    double *pb;                   // in pb', out pb
    double *dvo[3];               // in dvob', out dvob
    double *us[3];                // out ub*
    double *dvb[3];               // out dvb
    double id, s, dt_rho, dtA, dtB;   // s = rho/dt
};

// ub* = ub' - (rho/dt) G rb, pb += (dt/rho) D ub*, dvb = dtA ub* + dvob', dvob = dtB ub*
__global__ __launch_bounds__(kAX * kAY) void adj_project_kernel(AdjDims d, AdjProjArgs a)
{
    const int i = blockIdx.x * kAX + threadIdx.x % kAX, j = blockIdx.y * kAY + threadIdx.x / kAX, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const int ip = up(i, d.nx), im = dn(i, d.nx), jp = up(j, d.ny), jm = dn(j, d.ny);
    const int kp = up(k, d.nz), km = dn(k, d.nz);
    const long q = d.q(i, j, k);
    const double p0 = a.psi[q];
    const double sx = a.ub[q] - a.s * ((a.psi[d.q(ip, j, k)] - p0) * a.id);
    const double sy = a.vb[q] - a.s * ((a.psi[d.q(i, jp, k)] - p0) * a.id);
    const double sz = a.wb[q] - a.s * ((a.psi[d.q(i, j, kp)] - p0) * a.id);
    const long qxm = d.q(im, j, k), qym = d.q(i, jm, k), qzm = d.q(i, j, km);
    const double sxm = a.ub[qxm] - a.s * ((p0 - a.psi[qxm]) * a.id);
    const double sym = a.vb[qym] - a.s * ((p0 - a.psi[qym]) * a.id);
    const double szm = a.wb[qzm] - a.s * ((p0 - a.psi[qzm]) * a.id);
    const double div = ((sx - sxm) * a.id + (sy - sym) * a.id) + (sz - szm) * a.id;
    a.pb[q] = a.pb[q] + a.dt_rho * div;
    const double st[3] = {sx, sy, sz};
#pragma unroll
    for (int c = 0; c < 3; c++) {
        a.us[c][q] = st[c];
        a.dvb[c][q] = a.dtA * st[c] + a.dvo[c][q];
        a.dvo[c][q] = a.dtB * st[c];
    }
}

struct AdjMomArgs {
    const double *u, *v, *w;      // recorded state (ghosted)
    const double *us[3];          // ub*
    const double *dvb[3];         // dvb (a, b, e below)
    double *ub[3];                // out ub
    double id, id2, mu_rho;
};

// ub = ub* + J_adv^T dvb + (mu/rho) lap(dvb), gathered per face.
__global__ __launch_bounds__(kAX * kAY) void adj_momentum_kernel(AdjDims d, AdjMomArgs a)
{
    const int i = blockIdx.x * kAX + threadIdx.x % kAX, j = blockIdx.y * kAY + threadIdx.x / kAX, k = blockIdx.z;
    if (i >= d.nx || j >= d.ny) return;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int ip = up(i, nx), im = dn(i, nx), jp = up(j, ny), jm = dn(j, ny), kp = up(k, nz), km = dn(k, nz);
    const double *A = a.dvb[0], *Bv = a.dvb[1], *E = a.dvb[2];
    const double id = a.id;
    // state at offsets from (i,j,k); ghost layers hold the periodic images
    auto U = [&](int di, int dj, int dk) { return a.u[d.g(i + di, j + dj, k + dk)]; };
    auto V = [&](int di, int dj, int dk) { return a.v[d.g(i + di, j + dj, k + dk)]; };
    auto W = [&](int di, int dj, int dk) { return a.w[d.g(i + di, j + dj, k + dk)]; };
    // adjoint values with periodic wrap (0-based interior indices)
    auto X = [&](const double *f, int ii, int jj, int kk) { return f[d.q(ii, jj, kk)]; };
    // flux adjoints
    auto Qxx = [&](int ii, int jj, int kk) { return (X(A, up(ii, nx), jj, kk) - X(A, ii, jj, kk)) * id; };
    auto Qyy = [&](int ii, int jj, int kk) { return (X(Bv, ii, up(jj, ny), kk) - X(Bv, ii, jj, kk)) * id; };
    auto Qzz = [&](int ii, int jj, int kk) { return (X(E, ii, jj, up(kk, nz)) - X(E, ii, jj, kk)) * id; };
    auto Rxy = [&](int ii, int jj, int kk) {
        return (X(A, ii, up(jj, ny), kk) - X(A, ii, jj, kk)) * id + (X(Bv, up(ii, nx), jj, kk) - X(Bv, ii, jj, kk)) * id;
    };
    auto Rxz = [&](int ii, int jj, int kk) {
        return (X(A, ii, jj, up(kk, nz)) - X(A, ii, jj, kk)) * id + (X(E, up(ii, nx), jj, kk) - X(E, ii, jj, kk)) * id;
    };
    auto Ryz = [&](int ii, int jj, int kk) {
        return (X(Bv, ii, jj, up(kk, nz)) - X(Bv, ii, jj, kk)) * id + (X(E, ii, up(jj, ny), kk) - X(E, ii, jj, kk)) * id;
    };
    auto lap = [&](const double *f) {
        const double c = X(f, i, j, k);
        return ((X(f, ip, j, k) - 2.0 * c + X(f, im, j, k)) + (X(f, i, jp, k) - 2.0 * c + X(f, i, jm, k)) +
                (X(f, i, j, kp) - 2.0 * c + X(f, i, j, km))) * a.id2;
    };
    const long q = d.q(i, j, k);
    const double u0 = U(0, 0, 0), v0 = V(0, 0, 0), w0 = W(0, 0, 0);
    // u face
    double gu = Qxx(i, j, k) * 0.5 * (u0 + U(1, 0, 0)) + Qxx(im, j, k) * 0.5 * (U(-1, 0, 0) + u0);
    gu += Rxy(i, j, k) * 0.25 * (V(1, 0, 0) + v0) + Rxy(i, jm, k) * 0.25 * (V(1, -1, 0) + V(0, -1, 0));
    gu += Rxz(i, j, k) * 0.25 * (W(1, 0, 0) + w0) + Rxz(i, j, km) * 0.25 * (W(1, 0, -1) + W(0, 0, -1));
    // v face
    double gv = Qyy(i, j, k) * 0.5 * (v0 + V(0, 1, 0)) + Qyy(i, jm, k) * 0.5 * (V(0, -1, 0) + v0);
    gv += Rxy(i, j, k) * 0.25 * (U(0, 1, 0) + u0) + Rxy(im, j, k) * 0.25 * (U(-1, 1, 0) + U(-1, 0, 0));
    gv += Ryz(i, j, k) * 0.25 * (W(0, 1, 0) + w0) + Ryz(i, j, km) * 0.25 * (W(0, 1, -1) + W(0, 0, -1));
    // w face
    double gw = Qzz(i, j, k) * 0.5 * (w0 + W(0, 0, 1)) + Qzz(i, j, km) * 0.5 * (W(0, 0, -1) + w0);
    gw += Rxz(i, j, k) * 0.25 * (U(0, 0, 1) + u0) + Rxz(im, j, k) * 0.25 * (U(-1, 0, 1) + U(-1, 0, 0));
    gw += Ryz(i, j, k) * 0.25 * (V(0, 0, 1) + v0) + Ryz(i, jm, k) * 0.25 * (V(0, -1, 1) + V(0, -1, 0));
    a.ub[0][q] = a.us[0][q] + gu + a.mu_rho * lap(A);
    a.ub[1][q] = a.us[1][q] + gv + a.mu_rho * lap(Bv);
    a.ub[2][q] = a.us[2][q] + gw + a.mu_rho * lap(E);
}

}  // namespace

// Recorded run: checkpoints of the full state every `every` steps, plus the
// final state; segment buffer of the velocity at the start of each step.
struct AdjointTape {
    int nsteps = 0, every = 0, step0 = 0;
    double dt = 0.0, dt_o = 0.0;
    std::vector<double *> ckpt;    // [u v w p](ghosted) [dvo x3](interior), one block per checkpoint
    double *final_state = nullptr;
    std::vector<double *> seg;     // every x [u v w](ghosted)
    double *work = nullptr;        // 14 interior arrays of adjoint state and temporaries
};

namespace {

size_t ckpt_len(const fen_ctx *c) { return 4 * c->n_ghost + 3 * c->n_int; }

void save_state(fen_ctx *c, double *dst)
{
    const size_t ng = c->n_ghost * sizeof(double), ni = c->n_int * sizeof(double);
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(dst + d * c->n_ghost, c->d_u[d], ng, hipMemcpyDeviceToDevice, c->stream));
    FEN_HIP_CHECK(hipMemcpyAsync(dst + 3 * c->n_ghost, c->d_p, ng, hipMemcpyDeviceToDevice, c->stream));
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(dst + 4 * c->n_ghost + d * c->n_int, c->d_dvo[d], ni, hipMemcpyDeviceToDevice,
                                     c->stream));
}

void load_state(fen_ctx *c, const double *src)
{
    const size_t ng = c->n_ghost * sizeof(double), ni = c->n_int * sizeof(double);
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(c->d_u[d], src + d * c->n_ghost, ng, hipMemcpyDeviceToDevice, c->stream));
    FEN_HIP_CHECK(hipMemcpyAsync(c->d_p, src + 3 * c->n_ghost, ng, hipMemcpyDeviceToDevice, c->stream));
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(c->d_dvo[d], src + 4 * c->n_ghost + d * c->n_int, ni, hipMemcpyDeviceToDevice,
                                     c->stream));
}

double *dalloc(size_t n)
{
    void *p = nullptr;
    FEN_HIP_CHECK(hipMalloc(&p, n * sizeof(double)));
    return static_cast<double *>(p);
}

AdjDims dims(const fen_ctx *c) { return AdjDims{c->nx, c->ny, c->nz, c->box.gx, c->box.gxy}; }

dim3 cell_grid(const fen_ctx *c)
{
    return dim3((unsigned)((c->nx + kAX - 1) / kAX), (unsigned)((c->ny + kAY - 1) / kAY), (unsigned)c->nz);
}

// One adjoint step at the recorded state `st` ([u v w] ghosted).
void adjoint_step(fen_ctx *c, AdjointTape *t, const double *st, double **ub, double *pb, double **dvo)
{
    const size_t n = c->n_int;
    double *w = t->work + 7 * n;   // temporaries after the 7 adjoint-state arrays
    double *us[3] = {w, w + n, w + 2 * n}, *dvb[3] = {w + 3 * n, w + 4 * n, w + 5 * n}, *r = w + 6 * n;
    const AdjDims d = dims(c);
    const dim3 grid = cell_grid(c), blk(kAX * kAY);
    const double rho = c->density, dt = t->dt;
    const double id = 1.0 / c->delta;
    const double A = 1.0 + 0.5 * dt / t->dt_o, B = -(0.5 * dt / t->dt_o);
    // r = pb' + (dt/rho) D ub', then r = S r: the adjoint of the solve's right-hand side
    hipLaunchKernelGGL(adj_rhs_kernel, grid, blk, 0, c->stream, d, ub[0], ub[1], ub[2], pb, r, id, dt / rho);
    const long sx = c->nx, sy = (long)c->nx * c->ny;
    launch_div_r2c_impl<double>(c, nullptr, r, sx, sy, 0, 1.0, 1.0, 1.0);
    launch_y_fft<double>(c, -1, 1.0);
    launch_z_solve<double>(c, 1.0 / (double)(float)((long)c->nx * c->ny * c->nz));
    launch_y_fft<double>(c, +1, 1.0);
    launch_c2r<double>(c, r, 0, sx, sy, false, 0);
    AdjProjArgs pa{};
    pa.ub = ub[0]; pa.vb = ub[1]; pa.wb = ub[2];
    pa.psi = r;
    pa.pb = pb;
    for (int k = 0; k < 3; k++) { pa.dvo[k] = dvo[k]; pa.us[k] = us[k]; pa.dvb[k] = dvb[k]; }
    pa.id = id;
    pa.s = rho / dt;
    pa.dt_rho = dt / rho;
    pa.dtA = dt * A;
    pa.dtB = dt * B;
    hipLaunchKernelGGL(adj_project_kernel, grid, blk, 0, c->stream, d, pa);
    AdjMomArgs ma{};
    ma.u = st; ma.v = st + c->n_ghost; ma.w = st + 2 * c->n_ghost;
    for (int k = 0; k < 3; k++) { ma.us[k] = us[k]; ma.dvb[k] = dvb[k]; ma.ub[k] = ub[k]; }
    ma.id = id;
    ma.id2 = 1.0 / (c->delta * c->delta);
    ma.mu_rho = c->viscosity / rho;
    hipLaunchKernelGGL(adj_momentum_kernel, grid, blk, 0, c->stream, d, ma);
}

}  // namespace

void adjoint_free(fen_ctx *c)
{
    AdjointTape *t = c->tape;
    if (!t) return;
    for (double *p : t->ckpt) (void)hipFree(p);
    for (double *p : t->seg) (void)hipFree(p);
    if (t->final_state) (void)hipFree(t->final_state);
    if (t->work) (void)hipFree(t->work);
    delete t;
    c->tape = nullptr;
}

const char *adjoint_unsupported(const fen_ctx *c)
{
    if (c->precision != FEN_FP64) return "the adjoint is provided for FP64 contexts";
    if (c->dim != 3 || c->solver != FEN_SOLVER_PPP) return "the adjoint needs a 3-D triply periodic grid";
    if (c->has_source) return "the adjoint does not include a source term";
    if (c->constant_cfl) return "the adjoint needs a fixed dt (constant_CFL off)";
    return nullptr;
}

void adjoint_record(fen_ctx *c, int step0, int nsteps, int every, double dt)
{
    adjoint_free(c);
    AdjointTape *t = new AdjointTape();
    c->tape = t;
    t->nsteps = nsteps;
    t->every = every;
    t->step0 = step0;
    t->dt = dt;
    t->dt_o = c->dt_o;
    const int nseg = (nsteps + every - 1) / every;
    for (int s = 0; s < nseg; s++) t->ckpt.push_back(dalloc(ckpt_len(c)));
    t->final_state = dalloc(ckpt_len(c));
    for (int s = 0; s < nsteps; s += every) {
        save_state(c, t->ckpt[s / every]);
        run_fp64(c, std::min(every, nsteps - s), dt);
    }
    save_state(c, t->final_state);
    FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
}

void adjoint_backward(fen_ctx *c, const double *seed[3], double *grad[3], double *grad_p)
{
    AdjointTape *t = c->tape;
    const size_t n = c->n_int;
    if (!t->work) t->work = dalloc(14 * n);
    if (t->seg.empty())
        for (int s = 0; s < t->every; s++) t->seg.push_back(dalloc(3 * c->n_ghost));
    double *ub[3] = {t->work, t->work + n, t->work + 2 * n}, *pb = t->work + 3 * n;
    double *dvo[3] = {t->work + 4 * n, t->work + 5 * n, t->work + 6 * n};
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(ub[d], seed[d], n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    FEN_HIP_CHECK(hipMemsetAsync(pb, 0, 4 * n * sizeof(double), c->stream));   // pb and dvo
    const double dt_o_saved = c->dt_o;
    c->dt_o = t->dt_o;
    const int nseg = (t->nsteps + t->every - 1) / t->every;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int first = sg * t->every, len = std::min(t->every, t->nsteps - first);
        load_state(c, t->ckpt[sg]);
        // the segment's states, each formed straight into its slot (the
        // state after the segment's last step is not needed)
        states_fp64(c, t->dt, len, t->seg.data());
        for (int s = len - 1; s >= 0; s--) adjoint_step(c, t, t->seg[s], ub, pb, dvo);
    }
    for (int d = 0; d < 3; d++)
        FEN_HIP_CHECK(hipMemcpyAsync(grad[d], ub[d], n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad_p) FEN_HIP_CHECK(hipMemcpyAsync(grad_p, pb, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    load_state(c, t->final_state);   // leave the context at the end of the recorded run
    c->dt_o = dt_o_saved;
    c->checks_valid = false;
    FEN_HIP_CHECK(hipStreamSynchronize(c->stream));
}

}  // namespace fen
