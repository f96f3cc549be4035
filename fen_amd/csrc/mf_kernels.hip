// mf_kernels.hip -- the two-phase step of FEN's MF build on the GPU, 2-D,
// FP64: MTHINC volume of fluid (Ii et al. 2012; reference
// src/volume_of_fluid.f90), material properties (src/multiphase.f90:126-140),
// and the momentum right-hand side with variable density and viscosity, CSF
// surface tension and the Dodd & Ferrante pressure splitting
// (src/navier_stokes.f90:140-213, 384-501).  One thread per cell; the
// expressions keep the reference's evaluation order (the library is built
// with -ffp-contract=off).
#include <stdexcept>

#include "fen_internal.hpp"

namespace fen {

namespace {

constexpr int kBX = 16, kBY = 16;
constexpr double kSmall = 1.0e-14;   // global_mod small (src/global.f90:16)
constexpr double kCut = 1.0e-8;      // cut (src/volume_of_fluid.f90:37)

__device__ __forceinline__ double gauss_rp() { return 0.5 * (1.0 + 1.0 / sqrt(3.0)); }
__device__ __forceinline__ double gauss_rm() { return 0.5 * (1.0 - 1.0 / sqrt(3.0)); }

struct Surf {
    double cx, cy, a10, a01, a20, a02;
};

// P(x, y), Eq. 9 of Ii et al. (src/volume_of_fluid.f90:387-398)
__device__ __forceinline__ double surf_p(const Surf &q, double x, double y)
{
    return ((q.cx * q.a20 * (x * x) + q.cy * q.a02 * (y * y)) + q.a10 * x) + q.a01 * y;
}

__device__ __forceinline__ Surf cell_surface(double nx, double ny, double lx, double ly)
{
    Surf q;
    if (fabs(nx) == fmax(fabs(nx), fabs(ny))) { q.cx = 0.0; q.cy = 1.0; }
    else { q.cx = 1.0; q.cy = 0.0; }
    q.a10 = nx - 0.5 * q.cx * lx;
    q.a01 = ny - 0.5 * q.cy * ly;
    q.a20 = 0.5 * q.cx * lx;
    q.a02 = 0.5 * q.cy * ly;
    return q;
}

struct VofArgs {
    Box b;
    double delta, id, id2, beta, dt;
    const double *vof;
    double *h, *dd, *curv, *nrx, *nry, *lx, *ly;
};

// compute_norm (src/volume_of_fluid.f90:299-383) and the cell loop of
// get_h_from_vof (:223-295) in one pass: the normalisation needs only the
// cell's own normal and curvature terms.
__global__ __launch_bounds__(kBX *kBY) void vof_reconstruct_kernel(VofArgs a)
{
    const int i = blockIdx.x * kBX + threadIdx.x % kBX + 1, j = blockIdx.y * kBY + threadIdx.x / kBX + 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, 1), sj = a.b.gx;
    const double *F = a.vof + g;
    auto V = [&](int di, int dj) { return F[di + dj * sj]; };
    const double id = a.id;
    double mx[4], my[4];
    mx[0] = 0.5 * (V(0, -1) + V(0, 0) - V(-1, -1) - V(-1, 0)) * id;
    mx[1] = 0.5 * (V(0, 0) + V(0, 1) - V(-1, 0) - V(-1, 1)) * id;
    mx[2] = 0.5 * (V(1, 0) + V(1, 1) - V(0, 0) - V(0, 1)) * id;
    mx[3] = 0.5 * (V(1, -1) + V(1, 0) - V(0, -1) - V(0, 0)) * id;
    const double mxc = 0.25 * (mx[0] + mx[1] + mx[2] + mx[3]);
    my[0] = 0.5 * (V(-1, 0) + V(0, 0) - V(-1, -1) - V(0, -1)) * id;
    my[1] = 0.5 * (V(-1, 1) + V(0, 1) - V(-1, 0) - V(0, 0)) * id;
    my[2] = 0.5 * (V(0, 1) + V(1, 1) - V(0, 0) - V(1, 0)) * id;
    my[3] = 0.5 * (V(0, 0) + V(1, 0) - V(0, -1) - V(1, -1)) * id;
    const double myc = 0.25 * (my[0] + my[1] + my[2] + my[3]);
    double nxs[4], nys[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        nxs[q] = mx[q] / sqrt(mx[q] * mx[q] + my[q] * my[q] + kSmall);
        nys[q] = my[q] / sqrt(mx[q] * mx[q] + my[q] * my[q] + kSmall);
    }
    const double nx = mxc / sqrt(mxc * mxc + myc * myc + kSmall);
    const double ny = myc / sqrt(mxc * mxc + myc * myc + kSmall);
    const double lx = 0.5 * a.delta * (nxs[3] + nxs[2] - nxs[1] - nxs[0]);   // quadratic reconstruction
    const double ly = 0.5 * a.delta * (nys[1] + nys[2] - nys[0] - nys[3]);
    a.nrx[g] = nx;
    a.nry[g] = ny;
    a.lx[g] = lx;
    a.ly[g] = ly;
    a.curv[g] = -(lx + ly) * a.id2;

    const double f = V(0, 0);
    if (f <= kCut || f >= (1.0 - kCut)) {
        a.h[g] = f;
        a.dd[g] = 0.0;
        return;
    }
    const Surf q = cell_surface(nx, ny, lx, ly);
    const double beta = a.beta, rp = gauss_rp(), rm = gauss_rm();
    const double A = (1.0 - q.cx) * exp(2.0 * beta * q.a10) + (1.0 - q.cy) * exp(2.0 * beta * q.a01);
    const double Bp = (1.0 - q.cx) * exp(2.0 * beta * surf_p(q, 0.0, rp)) +
                      (1.0 - q.cy) * exp(2.0 * beta * surf_p(q, rp, 0.0));
    const double Bm = (1.0 - q.cx) * exp(2.0 * beta * surf_p(q, 0.0, rm)) +
                      (1.0 - q.cy) * exp(2.0 * beta * surf_p(q, rm, 0.0));
    const double Q = (1.0 - q.cx) * exp(2.0 * beta * q.a10 * (2.0 * f - 1.0)) +
                     (1.0 - q.cy) * exp(2.0 * beta * q.a01 * (2.0 * f - 1.0));
    const double aa = A * Bm * Bp * (A - Q);
    const double bb = A * (Bp + Bm) * (1.0 - Q);
    const double cc = 1.0 - A * Q;
    const double x1 = (-bb + sqrt(bb * bb - 4.0 * aa * cc)) / (2.0 * aa);   // solve_quadratic (:402-416)
    const double x2 = (-bb - sqrt(bb * bb - 4.0 * aa * cc)) / (2.0 * aa);
    const double d = log(fmax(x1, x2)) / (2.0 * beta);
    a.dd[g] = d;
    a.h[g] = 0.5 * (1.0 + tanh(beta * (surf_p(q, 0.5, 0.5) + d)));
}

// An_Int and Num_Int (src/volume_of_fluid.f90:636-661)
__device__ __forceinline__ double an_int(const Surf &q, double beta, double a, double b, double xa, double xb,
                                         double ya, double yb, double cf, double d0)
{
    return 0.5 * (b - a + 1.0 / (cf * beta) * log(cosh(beta * (surf_p(q, xb, yb) + d0)) /
                                                   cosh(beta * (surf_p(q, xa, ya) + d0))));
}

__device__ __forceinline__ double num_int(double a, double b, double qrm, double qrp)
{
    return 0.5 * (qrm + qrp) * (b - a);
}

struct SweepArgs {
    Box b;
    double delta, beta, dt;
    const double *vof, *dd, *nrx, *nry, *lx, *ly;   // the current reconstruction
    const double *vel;                              // u (dir 1) or v (dir 2)
    const double *in;
    double *out;
};

// compute_flux (src/volume_of_fluid.f90:548-632): flux through the high face
// of cell (i, j) along DIR with face velocity u.
template <int DIR> __device__ double face_flux(const SweepArgs &a, int i, int j, double u)
{
    const double delta = a.delta, dt = a.dt, beta = a.beta;
    double xa, xb, ya, yb;
    int sgn, ii = i, jj = j;
    if (DIR == 1) {
        if (u >= 0.0) { xa = 1.0 - dt * u / delta; xb = 1.0; sgn = 1; }
        else { xa = 0.0; xb = -dt * u / delta; sgn = -1; ii = i + 1; }
        ya = 0.0; yb = 1.0;
    } else {
        xa = 0.0; xb = 1.0;
        if (u >= 0.0) { ya = 1.0 - dt * u / delta; yb = 1.0; sgn = 1; }
        else { ya = 0.0; yb = -dt * u / delta; sgn = -1; jj = j + 1; }
    }
    const long g = a.b.g(ii, jj, 1);
    const double f = a.vof[g];
    if (f <= kCut || f >= (1.0 - kCut)) return (double)sgn * delta * f * (xb - xa) * (yb - ya);
    const Surf q = cell_surface(a.nrx[g], a.nry[g], a.lx[g], a.ly[g]);
    const double d0 = a.dd[g], rp = gauss_rp(), rm = gauss_rm();
    if (q.cx == 0.0)
        return (double)sgn * delta *
               num_int(ya, yb, an_int(q, beta, xa, xb, xa, xb, rm * (ya + yb), rm * (yb + ya), q.a10, d0),
                       an_int(q, beta, xa, xb, xa, xb, rp * (ya + yb), rp * (yb + ya), q.a10, d0));
    return (double)sgn * delta *
           num_int(xa, xb, an_int(q, beta, ya, yb, rm * (xa + xb), rm * (xa + xb), ya, yb, q.a01, d0),
                   an_int(q, beta, ya, yb, rp * (xa + xb), rp * (xa + xb), ya, yb, q.a01, d0));
}

// One directional sweep of advect_vof (src/volume_of_fluid.f90:437-461).
template <int DIR> __global__ __launch_bounds__(kBX *kBY) void vof_sweep_kernel(SweepArgs a)
{
    const int i = blockIdx.x * kBX + threadIdx.x % kBX + 1, j = blockIdx.y * kBY + threadIdx.x / kBX + 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, 1), sh = DIR == 1 ? 1 : a.b.gx;
    const double fp = face_flux<DIR>(a, i, j, a.vel[g]);
    const double fm = DIR == 1 ? face_flux<1>(a, i - 1, j, a.vel[g - sh]) : face_flux<2>(a, i, j - 1, a.vel[g - sh]);
    a.out[g] = (a.in[g] - (fp - fm) / a.delta) / (1.0 - a.dt * (a.vel[g] - a.vel[g - sh]) / a.delta);
}

// vof^{n+1} = vof2 - dt (vof_x du + vof_y dv) with the sweep-order weights (:463-470, :509-516).
__global__ __launch_bounds__(kBX *kBY) void vof_finish_kernel(Box b, double delta, double dt, int x_first,
                                                                  const double *vof1, const double *vof2,
                                                                  const double *u, const double *v, double *vof)
{
    const int i = blockIdx.x * kBX + threadIdx.x % kBX + 1, j = blockIdx.y * kBY + threadIdx.x / kBX + 1;
    if (i > b.nx || j > b.ny) return;
    const long g = b.g(i, j, 1);
    const double du = u[g] - u[g - 1], dv = v[g] - v[g - b.gx];
    if (x_first) vof[g] = vof2[g] - dt * (vof1[g] * du / delta + vof2[g] * dv / delta);
    else vof[g] = vof2[g] - dt * (vof2[g] * du / delta + vof1[g] * dv / delta);
}

// rho, mu from the vof over the whole array (src/multiphase.f90:133-134), and
// p_hat in the same pass: 2 p - p_o at fixed dt (src/navier_stokes.f90:91),
// p_o + (dt + dt_o)(p - p_o)/dt_o with constant CFL (:89; `dt` > 0).
__global__ __launch_bounds__(256) void mf_fields_kernel(long n, const double *vof, double rho0, double rho1,
                                                        double mu0, double mu1, double *rho, double *mu,
                                                        const double *p, const double *p_o, double *p_hat,
                                                        double dt, double dt_o)
{
    const long q = (long)blockIdx.x * 256 + threadIdx.x;
    if (q >= n) return;
    const double f = vof[q];
    rho[q] = rho1 * f + rho0 * (1.0 - f);
    mu[q] = mu1 * f + mu0 * (1.0 - f);
    if (p_hat) {
        if (dt > 0.0) p_hat[q] = p_o[q] + (dt + dt_o) * (p[q] - p_o[q]) / dt_o;
        else p_hat[q] = 2.0 * p[q] - p_o[q];
    }
}

struct MFMomArgs {
    Box b;
    double id, dt, A, B, g0, g1, sigma, irhomin;
    const double *u, *v, *p, *p_hat, *rho, *mu, *vof, *curv, *s0, *s1;
    double *us, *vs, *dvo0, *dvo1;
    int img;
};

// predicted_velocity_field with MF (src/navier_stokes.f90:140-213): advection
// (:261-353), variable-viscosity diffusion (:405-452), surface tension
// (:458-501), body force (:245-255), -grad p/rhof + A dv + B dv_o + g, the
// Dodd & Ferrante correction (:178-184), v* = v + dt RHS, dv_o = dv.
template <int SRC> __global__ __launch_bounds__(kBX *kBY) void mf_momentum_kernel(MFMomArgs a)
{
    const int i = blockIdx.x * kBX + threadIdx.x % kBX + 1, j = blockIdx.y * kBY + threadIdx.x / kBX + 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, 1), q = a.b.q(i, j, 1), sj = a.b.gx;
    const double id = a.id;
    auto U = [&](int di, int dj) { return a.u[g + di + dj * sj]; };
    auto V = [&](int di, int dj) { return a.v[g + di + dj * sj]; };
    auto M = [&](int di, int dj) { return a.mu[g + di + dj * sj]; };
    const double rfx = 0.5 * (a.rho[g + 1] + a.rho[g]);   // center_to_face (src/fields.f90:210-235)
    const double rfy = 0.5 * (a.rho[g + sj] + a.rho[g]);
    const double u0 = U(0, 0), v0 = V(0, 0);

    // advection, R = 0 - ...
    double rx, ry;
    {
        const double up = U(1, 0), um = U(-1, 0);
        const double uuip = 0.25 * ((up + u0) * (up + u0));
        const double uuim = 0.25 * ((um + u0) * (um + u0));
        const double uvjp = ((U(0, 1) + u0) * (V(1, 0) + v0)) * 0.25;
        const double uvjm = ((u0 + U(0, -1)) * (V(1, -1) + V(0, -1))) * 0.25;
        rx = (0.0 - (uuip - uuim) * id) - (uvjp - uvjm) * id;
        const double vuip = ((V(1, 0) + v0) * (U(0, 1) + u0)) * 0.25;
        const double vuim = ((v0 + V(-1, 0)) * (U(-1, 1) + U(-1, 0))) * 0.25;
        const double vp = V(0, 1), vm = V(0, -1);
        const double vvjp = 0.25 * ((vp + v0) * (vp + v0));
        const double vvjm = 0.25 * ((vm + v0) * (vm + v0));
        ry = (0.0 - (vuip - vuim) * id) - (vvjp - vvjm) * id;
    }
    // diffusion, div(2 mu D) with mu at cell centres
    {
        const double txxp = 2.0 * M(1, 0) * (U(1, 0) - U(0, 0)) * id;
        const double txxm = 2.0 * M(0, 0) * (U(0, 0) - U(-1, 0)) * id;
        const double dtxx = (txxp - txxm) * id;
        const double txyp = 0.25 * (M(0, 0) + M(1, 0) + M(0, 1) + M(1, 1)) *
                            ((U(0, 1) - U(0, 0)) * id + (V(1, 0) - V(0, 0)) * id);
        const double txym = 0.25 * (M(0, -1) + M(1, -1) + M(0, 0) + M(1, 0)) *
                            ((U(0, 0) - U(0, -1)) * id + (V(1, -1) - V(0, -1)) * id);
        const double dtxy = (txyp - txym) * id;
        rx = rx + (dtxx + dtxy) / rfx;
        const double tyxp = txyp;
        const double tyxm = 0.25 * (M(-1, 0) + M(0, 0) + M(-1, 1) + M(0, 1)) *
                            ((U(-1, 1) - U(-1, 0)) * id + (V(0, 0) - V(-1, 0)) * id);
        const double dtyx = (tyxp - tyxm) * id;
        const double tyyp = 2.0 * M(0, 1) * (V(0, 1) - V(0, 0)) * id;
        const double tyym = 2.0 * M(0, 0) * (V(0, 0) - V(0, -1)) * id;
        const double dtyy = (tyyp - tyym) * id;
        ry = ry + (dtyx + dtyy) / rfy;
    }
    // surface tension (CSF)
    {
        const double *K = a.curv + g, *F = a.vof + g;
        rx = rx + a.sigma * 0.5 * (K[1] + K[0]) * (F[1] - F[0]) * id / rfx;
        ry = ry + a.sigma * 0.5 * (K[sj] + K[0]) * (F[sj] - F[0]) * id / rfy;
    }
    if constexpr (SRC) {
        rx = rx + a.s0[q] / rfx;
        ry = ry + a.s1[q] / rfy;
    }
    const double p0 = a.p[g], ph0 = a.p_hat[g];
    {
        const double gp = (a.p[g + 1] - p0) * id, gh = (a.p_hat[g + 1] - ph0) * id;
        double rhs = ((-(gp / rfx) + a.A * rx) + a.B * a.dvo0[q]) + a.g0;
        rhs = ((rhs + gp / rfx) - a.irhomin * gp) - (1.0 / rfx - a.irhomin) * gh;
        const double val = u0 + a.dt * rhs;
        a.us[g] = val;
        write_images(a.us, a.b, a.img, i, j, 1, val);
        a.dvo0[q] = rx;
    }
    {
        const double gp = (a.p[g + sj] - p0) * id, gh = (a.p_hat[g + sj] - ph0) * id;
        double rhs = ((-(gp / rfy) + a.A * ry) + a.B * a.dvo1[q]) + a.g1;
        rhs = ((rhs + gp / rfy) - a.irhomin * gp) - (1.0 / rfy - a.irhomin) * gh;
        const double val = v0 + a.dt * rhs;
        a.vs[g] = val;
        write_images(a.vs, a.b, a.img, i, j, 1, val);
        a.dvo1[q] = ry;
    }
}

dim3 cell_grid(const Box &b) { return dim3((unsigned)((b.nx + kBX - 1) / kBX), (unsigned)((b.ny + kBY - 1) / kBY)); }

// Ghost fill of MF fields: Periodic -> images, Wall -> Neumann (the vof
// family and rho, mu, p_hat, p_o; src/volume_of_fluid.f90:54-219,
// src/multiphase.f90:44-120).
void mf_ghost(fen_ctx *c, double *const *f, int n)
{
    for (int s = 0; s < n; s += 8) {
        GhostDesc<double> d[8];
        const int m = n - s < 8 ? n - s : 8;
        for (int q = 0; q < m; q++) {
            d[q] = GhostDesc<double>{};
            d[q].f = f[s + q];
            for (int face = 0; face < 6; face++) {
                d[q].type[face] = (c->dim == 3 || face < 4) ? (c->bc_phys[face] == FEN_BC_PERIODIC ? 0 : 2) : 0;
                d[q].plane[face] = nullptr;
            }
            d[q].loc = 'c';
            d[q].groups = 3;
        }
        launch_ghost_desc<double>(c, d, m);
    }
}

void vof_reconstruct(fen_ctx *c)
{
    MFState &m = *c->mf;
    VofArgs a{};
    a.b = c->box;
    a.delta = c->delta;
    a.id = 1.0 / c->delta;
    a.id2 = 1.0 / (c->delta * c->delta);
    a.beta = m.beta;
    a.vof = m.vof;
    a.h = m.h; a.dd = m.dd; a.curv = m.curv; a.nrx = m.nrx; a.nry = m.nry; a.lx = m.lx; a.ly = m.ly;
    hipLaunchKernelGGL(vof_reconstruct_kernel, cell_grid(c->box), dim3(kBX * kBY), 0, c->stream, a);
    double *f[7] = {m.curv, m.nrx, m.nry, m.lx, m.ly, m.h, m.dd};
    mf_ghost(c, f, 7);
}

template <int DIR> void vof_sweep(fen_ctx *c, const double *in, double *out, double dt)
{
    MFState &m = *c->mf;
    SweepArgs a{};
    a.b = c->box;
    a.delta = c->delta;
    a.beta = m.beta;
    a.dt = dt;
    a.vof = m.vof; a.dd = m.dd; a.nrx = m.nrx; a.nry = m.nry; a.lx = m.lx; a.ly = m.ly;
    a.vel = as<double>(DIR == 1 ? c->d_u[0] : c->d_u[1]);
    a.in = in;
    a.out = out;
    hipLaunchKernelGGL((vof_sweep_kernel<DIR>), cell_grid(c->box), dim3(kBX * kBY), 0, c->stream, a);
}

}  // namespace

void mf_ghost_field(fen_ctx *c, double *f) { mf_ghost(c, &f, 1); }

// advect_vof (src/volume_of_fluid.f90:420-544) with the resident velocity.
void mf_advect_vof(fen_ctx *c, double dt)
{
    MFState &m = *c->mf;
    const size_t bytes = c->n_ghost * sizeof(double);
    vof_reconstruct(c);
    if (m.x_first) vof_sweep<1>(c, m.vof, m.vof1, dt);
    else vof_sweep<2>(c, m.vof, m.vof1, dt);
    // vof = vof1 over the whole array (vof1's ghosts stay zero), then its ghosts
    FEN_HIP_CHECK(hipMemcpyAsync(m.vof, m.vof1, bytes, hipMemcpyDeviceToDevice, c->stream));
    double *vf[1] = {m.vof};
    mf_ghost(c, vf, 1);
    vof_reconstruct(c);
    if (m.x_first) vof_sweep<2>(c, m.vof1, m.vof2, dt);
    else vof_sweep<1>(c, m.vof1, m.vof2, dt);
    hipLaunchKernelGGL(vof_finish_kernel, cell_grid(c->box), dim3(kBX * kBY), 0, c->stream, c->box, c->delta, dt,
                       m.x_first, (const double *)m.vof1, (const double *)m.vof2, (const double *)as<double>(c->d_u[0]),
                       (const double *)as<double>(c->d_u[1]), m.vof);
    m.x_first = !m.x_first;
    mf_ghost(c, vf, 1);
}

// update_material_properties (and p_hat when `with_phat`).
void mf_update_fields(fen_ctx *c, bool with_phat, double dt)
{
    MFState &m = *c->mf;
    const long n = (long)c->n_ghost;
    hipLaunchKernelGGL(mf_fields_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       (const double *)m.vof, m.rho0, m.rho1, m.mu0, m.mu1, m.rho, m.mu,
                       (const double *)as<double>(c->d_p), (const double *)m.p_o, with_phat ? m.p_hat : nullptr,
                       dt, c->dt_o);
    double *f[3] = {m.rho, m.mu, m.p_hat};
    mf_ghost(c, f, with_phat ? 3 : 2);
}

void mf_momentum(fen_ctx *c, double dt)
{
    MFState &m = *c->mf;
    MFMomArgs a{};
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.dt = dt;
    a.A = 1.0 + 0.5 * dt / c->dt_o;   // src/navier_stokes.f90:157-158
    a.B = -(0.5 * dt / c->dt_o);
    a.g0 = c->g[0];
    a.g1 = c->g[1];
    a.sigma = m.sigma;
    a.irhomin = m.irhomin;
    a.u = as<double>(c->d_u[0]); a.v = as<double>(c->d_u[1]); a.p = as<double>(c->d_p);
    a.p_hat = m.p_hat; a.rho = m.rho; a.mu = m.mu; a.vof = m.vof; a.curv = m.curv;
    a.s0 = as<double>(c->d_S[0]); a.s1 = as<double>(c->d_S[1]);
    a.us = as<double>(c->d_us[0]); a.vs = as<double>(c->d_us[1]);
    a.dvo0 = as<double>(c->d_dvo[0]); a.dvo1 = as<double>(c->d_dvo[1]);
    a.img = c->img_vel;
    prof_begin(c, K_MOMENTUM);
    if (c->has_source) hipLaunchKernelGGL(mf_momentum_kernel<1>, cell_grid(c->box), dim3(kBX * kBY), 0, c->stream, a);
    else hipLaunchKernelGGL(mf_momentum_kernel<0>, cell_grid(c->box), dim3(kBX * kBY), 0, c->stream, a);
    prof_end(c, K_MOMENTUM, 8.0 * 16.0 * c->n_int);
}

}  // namespace fen
