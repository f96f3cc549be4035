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

// ub = ub* + J_adv^T dvb + (mu/rho) lap(dvb), gathered per face, as a
// k-march: a 64x4 tile of cells walks kMK planes of z with planes k-1, k,
// k+1 of u, v, w (ghosted, images in place) and of the three dvb arrays
// (periodic wrap) in an LDS ring, so each value is read from memory once per
// tile instead of once per use (up to 13 uses per cell; a one-thread-per-cell
// form reading every operand from memory took 87 us against 60 us at 128^3).
// Plane k+2 is loaded into registers while plane k is computed.
constexpr int kMX = 64, kMY = 4, kMK = 16;
constexpr int kMLX = kMX + 2, kMN = kMLX * (kMY + 2);      // haloed tile
constexpr int kMPER = (kMN + kMX * kMY - 1) / (kMX * kMY);   // tile elements per thread

__global__ __launch_bounds__(kMX * kMY) void adj_momentum_march_kernel(AdjDims d, AdjMomArgs a)
{
    __shared__ double lds[6 * 3 * kMN];   // [field u v w A B E][slot][tile]
    const int i0 = blockIdx.x * kMX, j0 = blockIdx.y * kMY, k0 = blockIdx.z * kMK;
    const int kend = min(k0 + kMK, d.nz);
    const int tx = threadIdx.x % kMX, ty = threadIdx.x / kMX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = i < d.nx && j < d.ny;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    // tile element e at (i0 - 1 + e % kMLX, j0 - 1 + e / kMLX); past the grid
    // clamped to the last ghost (values no active cell reads)
    long og[kMPER], oq[kMPER];   // ghosted and interior (wrapped) plane offsets
#pragma unroll
    for (int m = 0; m < kMPER; m++) {
        const int e = min((int)threadIdx.x + m * kMX * kMY, kMN - 1);
        const int ii = min(i0 - 1 + e % kMLX, nx), jj = min(j0 - 1 + e / kMLX, ny);
        og[m] = (ii + 1) + d.gx * (jj + 1);
        oq[m] = (ii < 0 ? nx - 1 : (ii == nx ? 0 : ii)) + (long)nx * (jj < 0 ? ny - 1 : (jj == ny ? 0 : jj));
    }
    const double *Af = a.dvb[0], *Bf = a.dvb[1], *Ef = a.dvb[2];
    double raw[6][kMPER];
    auto fetch = [&](int kk) {   // plane kk in [-1, nz]
        const long kg = d.gxy * (kk + 1), kq = (long)nx * ny * (kk < 0 ? nz - 1 : (kk == nz ? 0 : kk));
#pragma unroll
        for (int m = 0; m < kMPER; m++) {
            raw[0][m] = a.u[og[m] + kg];
            raw[1][m] = a.v[og[m] + kg];
            raw[2][m] = a.w[og[m] + kg];
            raw[3][m] = Af[oq[m] + kq];
            raw[4][m] = Bf[oq[m] + kq];
            raw[5][m] = Ef[oq[m] + kq];
        }
    };
    auto commit = [&](int slot) {
#pragma unroll
        for (int m = 0; m < kMPER; m++) {
            const int e = (int)threadIdx.x + m * kMX * kMY;
            if (e < kMN) {
#pragma unroll
                for (int f = 0; f < 6; f++) lds[(f * 3 + slot) * kMN + e] = raw[f][m];
            }
        }
    };
    fetch(k0 - 1);
    commit(0);
    fetch(k0);
    commit(1);
    fetch(k0 + 1);
    int sm = 0, sc = 1, sn = 2;
    const int c = (ty + 1) * kMLX + tx + 1;
    const double id = a.id;
    for (int k = k0; k < kend; k++) {
        commit(sn);   // plane k + 1
        __syncthreads();
        if (k + 2 <= nz && k + 2 <= kend) fetch(k + 2);
        if (active) {
            auto at = [&](int f, int di, int dj, int dk) {
                const int sl = dk < 0 ? sm : (dk > 0 ? sn : sc);
                return lds[(f * 3 + sl) * kMN + c + dj * kMLX + di];
            };
            auto U = [&](int di, int dj, int dk) { return at(0, di, dj, dk); };
            auto V = [&](int di, int dj, int dk) { return at(1, di, dj, dk); };
            auto W = [&](int di, int dj, int dk) { return at(2, di, dj, dk); };
            auto A = [&](int di, int dj, int dk) { return at(3, di, dj, dk); };
            auto B = [&](int di, int dj, int dk) { return at(4, di, dj, dk); };
            auto E = [&](int di, int dj, int dk) { return at(5, di, dj, dk); };
            // flux adjoints at offsets from (i, j, k)
            auto Qxx = [&](int di, int dj, int dk) { return (A(di + 1, dj, dk) - A(di, dj, dk)) * id; };
            auto Qyy = [&](int di, int dj, int dk) { return (B(di, dj + 1, dk) - B(di, dj, dk)) * id; };
            auto Qzz = [&](int di, int dj, int dk) { return (E(di, dj, dk + 1) - E(di, dj, dk)) * id; };
            auto Rxy = [&](int di, int dj, int dk) {
                return (A(di, dj + 1, dk) - A(di, dj, dk)) * id + (B(di + 1, dj, dk) - B(di, dj, dk)) * id;
            };
            auto Rxz = [&](int di, int dj, int dk) {
                return (A(di, dj, dk + 1) - A(di, dj, dk)) * id + (E(di + 1, dj, dk) - E(di, dj, dk)) * id;
            };
            auto Ryz = [&](int di, int dj, int dk) {
                return (B(di, dj, dk + 1) - B(di, dj, dk)) * id + (E(di, dj + 1, dk) - E(di, dj, dk)) * id;
            };
            auto lap = [&](int f) {
                const double c0 = at(f, 0, 0, 0);
                return ((at(f, 1, 0, 0) - 2.0 * c0 + at(f, -1, 0, 0)) + (at(f, 0, 1, 0) - 2.0 * c0 + at(f, 0, -1, 0)) +
                        (at(f, 0, 0, 1) - 2.0 * c0 + at(f, 0, 0, -1))) * a.id2;
            };
            const long q = d.q(i, j, k);
            const double u0 = U(0, 0, 0), v0 = V(0, 0, 0), w0 = W(0, 0, 0);
            double gu = Qxx(0, 0, 0) * 0.5 * (u0 + U(1, 0, 0)) + Qxx(-1, 0, 0) * 0.5 * (U(-1, 0, 0) + u0);
            gu += Rxy(0, 0, 0) * 0.25 * (V(1, 0, 0) + v0) + Rxy(0, -1, 0) * 0.25 * (V(1, -1, 0) + V(0, -1, 0));
            gu += Rxz(0, 0, 0) * 0.25 * (W(1, 0, 0) + w0) + Rxz(0, 0, -1) * 0.25 * (W(1, 0, -1) + W(0, 0, -1));
            double gv = Qyy(0, 0, 0) * 0.5 * (v0 + V(0, 1, 0)) + Qyy(0, -1, 0) * 0.5 * (V(0, -1, 0) + v0);
            gv += Rxy(0, 0, 0) * 0.25 * (U(0, 1, 0) + u0) + Rxy(-1, 0, 0) * 0.25 * (U(-1, 1, 0) + U(-1, 0, 0));
            gv += Ryz(0, 0, 0) * 0.25 * (W(0, 1, 0) + w0) + Ryz(0, 0, -1) * 0.25 * (W(0, 1, -1) + W(0, 0, -1));
            double gw = Qzz(0, 0, 0) * 0.5 * (w0 + W(0, 0, 1)) + Qzz(0, 0, -1) * 0.5 * (W(0, 0, -1) + w0);
            gw += Rxz(0, 0, 0) * 0.25 * (U(0, 0, 1) + u0) + Rxz(-1, 0, 0) * 0.25 * (U(-1, 0, 1) + U(-1, 0, 0));
            gw += Ryz(0, 0, 0) * 0.25 * (V(0, 0, 1) + v0) + Ryz(0, -1, 0) * 0.25 * (V(0, -1, 1) + V(0, -1, 0));
            a.ub[0][q] = a.us[0][q] + gu + a.mu_rho * lap(3);
            a.ub[1][q] = a.us[1][q] + gv + a.mu_rho * lap(4);
            a.ub[2][q] = a.us[2][q] + gw + a.mu_rho * lap(5);
        }
        __syncthreads();   // every read of slot sm precedes the next commit into it
        const int t = sm;
        sm = sc; sc = sn; sn = t;
    }
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
    const dim3 mgrid((unsigned)((c->nx + kMX - 1) / kMX), (unsigned)((c->ny + kMY - 1) / kMY),
                     (unsigned)((c->nz + kMK - 1) / kMK));
    hipLaunchKernelGGL(adj_momentum_march_kernel, mgrid, dim3(kMX * kMY), 0, c->stream, d, ma);
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
