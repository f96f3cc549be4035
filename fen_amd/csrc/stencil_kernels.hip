// stencil_kernels.hip -- FP64 staggered-grid kernels of the time step.
//
// Every expression keeps the reference's evaluation order (Fortran left to
// right, explicit parentheses) and the library is compiled with
// -ffp-contract=off, so these kernels reproduce the CPU arithmetic bit for bit.
#include <stdexcept>

#include "fen_internal.hpp"

namespace fen {

constexpr int kTX = 64, kTY = 4;

// ---------------------------------------------------------------------------
// Ghost fill: scalar%update_ghost_nodes (src/scalar.f90:223-396) evaluated
// as a pure function of the interior.  The reference fills left, right,
// bottom, top, front, back over full planes, so a ghost's final value is the
// last axis rule applied to the state left by the earlier axes:
//   S1 = x rules on the interior S0, S2 = y rules on S1, S3 = z rules on S2.
// A Dirichlet rule on a field's normal face at the high end also overwrites
// the last interior plane (f(N) = bc).  Every cell written here is computed
// from interior cells that no thread writes, so one launch needs no ordering.
// ---------------------------------------------------------------------------
struct GhostLaunch {
    GhostDesc d[4];
    Box b;
    long nx_cells, ny_cells, nz_cells;   // cells per group, per field
    int ngx, ngy, ngz;                    // planes per group (2 or 3)
};

__device__ __forceinline__ double plane_at(const double *p, long idx) { return p[idx]; }

__device__ double ghost_s1(const GhostDesc &d, const Box &b, int i, int j, int k)
{
    const long pj = j + (long)(b.ny + 2) * k;
    if (i == 0) {
        const int t = d.type[0];
        if (t == 0) return d.f[b.g(b.nx, j, k)];
        if (t == 1) {
            const double bc = plane_at(d.plane[0], pj);
            return d.loc == 'x' ? bc : 2.0 * bc - d.f[b.g(1, j, k)];
        }
        if (t == 2) return d.f[b.g(1, j, k)];
        return d.f[b.g(i, j, k)];
    }
    if (i == b.nx + 1) {
        const int t = d.type[1];
        if (t == 0) return d.f[b.g(1, j, k)];
        if (t == 1) {
            const double bc = plane_at(d.plane[1], pj);
            return d.loc == 'x' ? bc : 2.0 * bc - d.f[b.g(b.nx, j, k)];
        }
        if (t == 2) return d.f[b.g(b.nx, j, k)];
        return d.f[b.g(i, j, k)];
    }
    if (i == b.nx && d.type[1] == 1 && d.loc == 'x') return plane_at(d.plane[1], pj);
    return d.f[b.g(i, j, k)];
}

__device__ double ghost_s2(const GhostDesc &d, const Box &b, int i, int j, int k)
{
    const long pi = i + (long)(b.nx + 2) * k;
    if (j == 0) {
        const int t = d.type[2];
        if (t == 0) return ghost_s1(d, b, i, b.ny, k);
        if (t == 1) {
            const double bc = plane_at(d.plane[2], pi);
            return d.loc == 'y' ? bc : 2.0 * bc - ghost_s1(d, b, i, 1, k);
        }
        if (t == 2) return ghost_s1(d, b, i, 1, k);
        return ghost_s1(d, b, i, j, k);
    }
    if (j == b.ny + 1) {
        const int t = d.type[3];
        if (t == 0) return ghost_s1(d, b, i, 1, k);
        if (t == 1) {
            const double bc = plane_at(d.plane[3], pi);
            return d.loc == 'y' ? bc : 2.0 * bc - ghost_s1(d, b, i, b.ny, k);
        }
        if (t == 2) return ghost_s1(d, b, i, b.ny, k);
        return ghost_s1(d, b, i, j, k);
    }
    if (j == b.ny && d.type[3] == 1 && d.loc == 'y') return plane_at(d.plane[3], pi);
    return ghost_s1(d, b, i, j, k);
}

__device__ double ghost_s3(const GhostDesc &d, const Box &b, int i, int j, int k)
{
    if (b.dim == 3) {
        const long pij = i + (long)(b.nx + 2) * j;
        if (k == 0) {
            const int t = d.type[4];
            if (t == 0) return ghost_s2(d, b, i, j, b.nz);
            if (t == 1) {
                const double bc = plane_at(d.plane[4], pij);
                return d.loc == 'z' ? bc : 2.0 * bc - ghost_s2(d, b, i, j, 1);
            }
            if (t == 2) return ghost_s2(d, b, i, j, 1);
            return ghost_s2(d, b, i, j, k);
        }
        if (k == b.nz + 1) {
            const int t = d.type[5];
            if (t == 0) return ghost_s2(d, b, i, j, 1);
            if (t == 1) {
                const double bc = plane_at(d.plane[5], pij);
                return d.loc == 'z' ? bc : 2.0 * bc - ghost_s2(d, b, i, j, b.nz);
            }
            if (t == 2) return ghost_s2(d, b, i, j, b.nz);
            return ghost_s2(d, b, i, j, k);
        }
        if (k == b.nz && d.type[5] == 1 && d.loc == 'z') return plane_at(d.plane[5], pij);
    }
    return ghost_s2(d, b, i, j, k);
}

// Map group-local plane index p to the coordinate: 0 -> 0, 1 -> n+1, 2 -> n.
__device__ __forceinline__ int plane_coord(int p, int n) { return p == 0 ? 0 : (p == 1 ? n + 1 : n); }

__global__ __launch_bounds__(256) void ghost_kernel(GhostLaunch a)
{
    const GhostDesc &d = a.d[blockIdx.y];
    const Box &b = a.b;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tx = a.nx_cells * a.ngx, ty = a.ny_cells * a.ngy, tz = a.nz_cells * a.ngz;
    const int k_lo = b.dim == 3 ? 1 : 1, k_n = b.nz;
    int i, j, k;
    if (t < tx) {
        // x group: i ghost (or overwritten normal plane), j and k interior
        const int p = (int)(t / a.nx_cells);
        const long r = t % a.nx_cells;
        i = plane_coord(p, b.nx);
        j = (int)(r % b.ny) + 1;
        k = (int)(r / b.ny) + k_lo;
    } else if ((t -= tx) < ty) {
        // y group: j ghost, all i, k interior
        const int p = (int)(t / a.ny_cells);
        const long r = t % a.ny_cells;
        j = plane_coord(p, b.ny);
        i = (int)(r % (b.nx + 2));
        k = (int)(r / (b.nx + 2)) + k_lo;
    } else if ((t -= ty) < tz) {
        // z group: k ghost, all i and j
        const int p = (int)(t / a.nz_cells);
        const long r = t % a.nz_cells;
        k = plane_coord(p, k_n);
        i = (int)(r % (b.nx + 2));
        j = (int)(r / (b.nx + 2));
    } else {
        return;
    }
    const double val = ghost_s3(d, b, i, j, k);
    d.f[b.g(i, j, k)] = val;
}

static int planes_for(const GhostDesc *d, int n, int face_hi, int loc)
{
    // 3 planes if any field has a Dirichlet rule on its normal high face.
    for (int q = 0; q < n; q++)
        if (d[q].type[face_hi] == 1 && d[q].loc == loc) return 3;
    return 2;
}

void launch_ghost_desc(fen_ctx *c, const GhostDesc *descs, int n)
{
    if (n <= 0 || n > 4) throw std::runtime_error("ghost fill of 1..4 fields");
    GhostLaunch a{};
    for (int q = 0; q < n; q++) a.d[q] = descs[q];
    a.b = c->box;
    const Box &b = c->box;
    a.ngx = planes_for(descs, n, FEN_RIGHT, 'x');
    a.ngy = planes_for(descs, n, FEN_TOP, 'y');
    a.ngz = b.dim == 3 ? planes_for(descs, n, FEN_BACK, 'z') : 0;
    a.nx_cells = (long)b.ny * b.nz;
    a.ny_cells = (long)(b.nx + 2) * b.nz;
    a.nz_cells = (long)(b.nx + 2) * (b.ny + 2);
    const long total = a.nx_cells * a.ngx + a.ny_cells * a.ngy + a.nz_cells * a.ngz;
    const dim3 grid((unsigned)((total + 255) / 256), (unsigned)n);
    prof_begin(c, K_GHOST);
    hipLaunchKernelGGL(ghost_kernel, grid, dim3(256), 0, c->stream, a);
    prof_end(c, K_GHOST, 16.0 * total * n);
}

// ---------------------------------------------------------------------------
// K1: momentum.  dv = advection (src/navier_stokes.f90:261-353) + mu lap(v)/rhof
// (:384-404, src/fields.f90:298-343) + S/rhof (:245-255); then
// RHS = -grad(p)/rhof + A dv + B dv_o + g and v* = v + dt RHS (:157-198);
// dv_o = dv (:201-205).  rhof == density exactly for a constant density
// field (0.5*(rho+rho) == rho), mu == viscosity.
// ---------------------------------------------------------------------------
struct MomArgs {
    const double *u, *v, *w, *p;
    double *us, *vs, *ws;
    double *dvo0, *dvo1, *dvo2;
    const double *s0, *s1, *s2;
    Box b;
    double id, id2, mu, rhof, dt, A, B, g0, g1, g2;
};

template <int D3, int SRC>
__global__ __launch_bounds__(kTX * kTY) void momentum_kernel(MomArgs a)
{
    const int i = blockIdx.x * kTX + threadIdx.x + 1;
    const int j = blockIdx.y * kTY + threadIdx.y + 1;
    const int k = D3 ? (int)blockIdx.z + 1 : 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, k);
    const long q = a.b.q(i, j, k);
    const long sj = a.b.gx, sk = a.b.gxy;
    const double *U = a.u + g, *V = a.v + g, *W = a.w + g;
    const double id = a.id, id2 = a.id2;

    // ---- x momentum
    double r;
    {
        const double uuip = 0.25 * ((U[1] + U[0]) * (U[1] + U[0]));
        const double uuim = 0.25 * ((U[-1] + U[0]) * (U[-1] + U[0]));
        const double uvjp = ((U[sj] + U[0]) * (V[1] + V[0])) * 0.25;
        const double uvjm = ((U[0] + U[-sj]) * (V[1 - sj] + V[-sj])) * 0.25;
        r = (0.0 - (uuip - uuim) * id) - (uvjp - uvjm) * id;
        if constexpr (D3) {
            const double uwkp = ((U[sk] + U[0]) * (W[1] + W[0])) * 0.25;
            const double uwkm = ((U[0] + U[-sk]) * (W[1 - sk] + W[-sk])) * 0.25;
            r = r - (uwkp - uwkm) * id;
        }
    }
    double lu = (((U[1] - 2.0 * U[0]) + U[-1]) + ((U[sj] - 2.0 * U[0]) + U[-sj])) * id2;
    if constexpr (D3) lu = lu + ((U[sk] - 2.0 * U[0]) + U[-sk]) * id2;
    r = r + a.mu * lu / a.rhof;
    if constexpr (SRC) r = r + a.s0[q] / a.rhof;
    else r = r + 0.0;
    const double dvx = r;

    // ---- y momentum
    {
        const double vuip = ((V[1] + V[0]) * (U[sj] + U[0])) * 0.25;
        const double vuim = ((V[0] + V[-1]) * (U[-1 + sj] + U[-1])) * 0.25;
        const double vvjp = 0.25 * ((V[sj] + V[0]) * (V[sj] + V[0]));
        const double vvjm = 0.25 * ((V[-sj] + V[0]) * (V[-sj] + V[0]));
        r = (0.0 - (vuip - vuim) * id) - (vvjp - vvjm) * id;
        if constexpr (D3) {
            const double vwkp = ((V[sk] + V[0]) * (W[sj] + W[0])) * 0.25;
            const double vwkm = ((V[0] + V[-sk]) * (W[sj - sk] + W[-sk])) * 0.25;
            r = r - (vwkp - vwkm) * id;
        }
    }
    double lv = (((V[1] - 2.0 * V[0]) + V[-1]) + ((V[sj] - 2.0 * V[0]) + V[-sj])) * id2;
    if constexpr (D3) lv = lv + ((V[sk] - 2.0 * V[0]) + V[-sk]) * id2;
    r = r + a.mu * lv / a.rhof;
    if constexpr (SRC) r = r + a.s1[q] / a.rhof;
    else r = r + 0.0;
    const double dvy = r;

    const double *P = a.p + g;
    const double A = a.A, B = a.B, dt = a.dt;
    {
        const double gp = (P[1] - P[0]) * id;
        const double rhs = ((-(gp / a.rhof) + A * dvx) + B * a.dvo0[q]) + a.g0;
        a.us[g] = U[0] + dt * rhs;
        a.dvo0[q] = dvx;
    }
    {
        const double gp = (P[sj] - P[0]) * id;
        const double rhs = ((-(gp / a.rhof) + A * dvy) + B * a.dvo1[q]) + a.g1;
        a.vs[g] = V[0] + dt * rhs;
        a.dvo1[q] = dvy;
    }
    if constexpr (D3) {
        // ---- z momentum
        const double wuip = ((W[0] + W[1]) * (U[0] + U[sk])) * 0.25;
        const double wuim = ((W[0] + W[-1]) * (U[-1] + U[-1 + sk])) * 0.25;
        const double wvjp = ((W[0] + W[sj]) * (V[0] + V[sk])) * 0.25;
        const double wvjm = ((W[0] + W[-sj]) * (V[-sj] + V[-sj + sk])) * 0.25;
        const double wwkp = ((W[0] + W[sk]) * (W[0] + W[sk])) * 0.25;
        const double wwkm = ((W[0] + W[-sk]) * (W[0] + W[-sk])) * 0.25;
        r = ((0.0 - (wuip - wuim) * id) - (wvjp - wvjm) * id) - (wwkp - wwkm) * id;
        const double lw = ((((W[1] - 2.0 * W[0]) + W[-1]) + ((W[sj] - 2.0 * W[0]) + W[-sj])) +
                           ((W[sk] - 2.0 * W[0]) + W[-sk])) * id2;
        r = r + a.mu * lw / a.rhof;
        if constexpr (SRC) r = r + a.s2[q] / a.rhof;
        else r = r + 0.0;
        const double gp = (P[sk] - P[0]) * id;
        const double rhs = ((-(gp / a.rhof) + A * r) + B * a.dvo2[q]) + a.g2;
        a.ws[g] = W[0] + dt * rhs;
        a.dvo2[q] = r;
    }
}

void launch_momentum(fen_ctx *c, double dt, double A, double B)
{
    MomArgs a{};
    a.u = c->d_u[0]; a.v = c->d_u[1]; a.w = c->d_u[2]; a.p = c->d_p;
    a.us = c->d_us[0]; a.vs = c->d_us[1]; a.ws = c->d_us[2];
    a.dvo0 = c->d_dvo[0]; a.dvo1 = c->d_dvo[1]; a.dvo2 = c->d_dvo[2];
    a.s0 = c->d_S[0]; a.s1 = c->d_S[1]; a.s2 = c->d_S[2];
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.id2 = 1.0 / (c->delta * c->delta);
    a.mu = c->viscosity;
    a.rhof = 0.5 * (c->density + c->density);
    a.dt = dt; a.A = A; a.B = B;
    a.g0 = c->g[0]; a.g1 = c->g[1]; a.g2 = c->g[2];
    const Box &b = c->box;
    const dim3 grid((unsigned)((b.nx + kTX - 1) / kTX), (unsigned)((b.ny + kTY - 1) / kTY),
                    (unsigned)(b.dim == 3 ? b.nz : 1));
    const int d3 = b.dim == 3;
    const double words = d3 ? (c->has_source ? 16.0 : 13.0) : (c->has_source ? 11.0 : 9.0);
    prof_begin(c, K_MOMENTUM);
    if (d3) {
        if (c->has_source) hipLaunchKernelGGL((momentum_kernel<1, 1>), grid, dim3(kTX, kTY), 0, c->stream, a);
        else hipLaunchKernelGGL((momentum_kernel<1, 0>), grid, dim3(kTX, kTY), 0, c->stream, a);
    } else {
        if (c->has_source) hipLaunchKernelGGL((momentum_kernel<0, 1>), grid, dim3(kTX, kTY), 0, c->stream, a);
        else hipLaunchKernelGGL((momentum_kernel<0, 0>), grid, dim3(kTX, kTY), 0, c->stream, a);
    }
    prof_end(c, K_MOMENTUM, words * 8.0 * c->n_int);
}

// add_advection alone on the resident velocity (reference test
// test/small_test/navier_stokes/advection/advection.f90:50): R -= adv.
template <int D3>
__global__ __launch_bounds__(kTX * kTY) void advection_kernel(const double *u, const double *v, const double *w,
                                                              double *rx, double *ry, double *rz, Box b, double id)
{
    const int i = blockIdx.x * kTX + threadIdx.x + 1;
    const int j = blockIdx.y * kTY + threadIdx.y + 1;
    const int k = D3 ? (int)blockIdx.z + 1 : 1;
    if (i > b.nx || j > b.ny) return;
    const long g = b.g(i, j, k), q = b.q(i, j, k);
    const long sj = b.gx, sk = b.gxy;
    const double *U = u + g, *V = v + g, *W = w + g;
    double uuip = 0.25 * ((U[1] + U[0]) * (U[1] + U[0]));
    double uuim = 0.25 * ((U[-1] + U[0]) * (U[-1] + U[0]));
    double uvjp = ((U[sj] + U[0]) * (V[1] + V[0])) * 0.25;
    double uvjm = ((U[0] + U[-sj]) * (V[1 - sj] + V[-sj])) * 0.25;
    double r = (rx[q] - (uuip - uuim) * id) - (uvjp - uvjm) * id;
    if constexpr (D3) {
        double uwkp = ((U[sk] + U[0]) * (W[1] + W[0])) * 0.25;
        double uwkm = ((U[0] + U[-sk]) * (W[1 - sk] + W[-sk])) * 0.25;
        r = r - (uwkp - uwkm) * id;
    }
    rx[q] = r;
    double vuip = ((V[1] + V[0]) * (U[sj] + U[0])) * 0.25;
    double vuim = ((V[0] + V[-1]) * (U[-1 + sj] + U[-1])) * 0.25;
    double vvjp = 0.25 * ((V[sj] + V[0]) * (V[sj] + V[0]));
    double vvjm = 0.25 * ((V[-sj] + V[0]) * (V[-sj] + V[0]));
    r = (ry[q] - (vuip - vuim) * id) - (vvjp - vvjm) * id;
    if constexpr (D3) {
        double vwkp = ((V[sk] + V[0]) * (W[sj] + W[0])) * 0.25;
        double vwkm = ((V[0] + V[-sk]) * (W[sj - sk] + W[-sk])) * 0.25;
        r = r - (vwkp - vwkm) * id;
    }
    ry[q] = r;
    if constexpr (D3) {
        double wuip = ((W[0] + W[1]) * (U[0] + U[sk])) * 0.25;
        double wuim = ((W[0] + W[-1]) * (U[-1] + U[-1 + sk])) * 0.25;
        double wvjp = ((W[0] + W[sj]) * (V[0] + V[sk])) * 0.25;
        double wvjm = ((W[0] + W[-sj]) * (V[-sj] + V[-sj + sk])) * 0.25;
        double wwkp = ((W[0] + W[sk]) * (W[0] + W[sk])) * 0.25;
        double wwkm = ((W[0] + W[-sk]) * (W[0] + W[-sk])) * 0.25;
        rz[q] = ((rz[q] - (wuip - wuim) * id) - (wvjp - wvjm) * id) - (wwkp - wwkm) * id;
    }
}

void launch_advection_only(fen_ctx *c, double *rx, double *ry, double *rz)
{
    const Box &b = c->box;
    const dim3 grid((unsigned)((b.nx + kTX - 1) / kTX), (unsigned)((b.ny + kTY - 1) / kTY),
                    (unsigned)(b.dim == 3 ? b.nz : 1));
    if (b.dim == 3)
        hipLaunchKernelGGL((advection_kernel<1>), grid, dim3(kTX, kTY), 0, c->stream, c->d_u[0], c->d_u[1],
                           c->d_u[2], rx, ry, rz, b, 1.0 / c->delta);
    else
        hipLaunchKernelGGL((advection_kernel<0>), grid, dim3(kTX, kTY), 0, c->stream, c->d_u[0], c->d_u[1],
                           c->d_u[2], rx, ry, rz, b, 1.0 / c->delta);
}

// ---------------------------------------------------------------------------
// K7: correct_velocity_field (src/navier_stokes.f90:505-546) and the interior
// of update_pressure (:550-566): v = v* - ((grad(phi)*dt)/rhof), p = p + phi.
// phi may carry a pending mean shift (pp, ppn) that is subtracted here,
// value by value, exactly where the reference has already subtracted it.
// ---------------------------------------------------------------------------
struct ProjArgs {
    const double *us, *vs, *ws, *phi;
    double *u, *v, *w, *p;
    const double *shift;
    Box b;
    double id, dt, rhof;
};

template <int D3, int SHIFT>
__global__ __launch_bounds__(kTX * kTY) void project_kernel(ProjArgs a)
{
    const int i = blockIdx.x * kTX + threadIdx.x + 1;
    const int j = blockIdx.y * kTY + threadIdx.y + 1;
    const int k = D3 ? (int)blockIdx.z + 1 : 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, k);
    const long sj = a.b.gx, sk = a.b.gxy;
    const double *F = a.phi + g;
    double s = 0.0;
    if constexpr (SHIFT) s = a.shift[0];
    auto ph = [&](long o) -> double {
        if constexpr (SHIFT) return F[o] - s;
        else return F[o];
    };
    const double f0 = ph(0);
    const double gx = (ph(1) - f0) * a.id;
    const double gy = (ph(sj) - f0) * a.id;
    a.u[g] = a.us[g] - gx * a.dt / a.rhof;
    a.v[g] = a.vs[g] - gy * a.dt / a.rhof;
    if constexpr (D3) {
        const double gz = (ph(sk) - f0) * a.id;
        a.w[g] = a.ws[g] - gz * a.dt / a.rhof;
    }
    a.p[g] = a.p[g] + f0;
}

void launch_project(fen_ctx *c, double dt)
{
    ProjArgs a{};
    a.us = c->d_us[0]; a.vs = c->d_us[1]; a.ws = c->d_us[2]; a.phi = c->d_phi;
    a.u = c->d_u[0]; a.v = c->d_u[1]; a.w = c->d_u[2]; a.p = c->d_p;
    a.shift = c->d_shift;
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.dt = dt;
    a.rhof = 0.5 * (c->density + c->density);
    const Box &b = c->box;
    const dim3 grid((unsigned)((b.nx + kTX - 1) / kTX), (unsigned)((b.ny + kTY - 1) / kTY),
                    (unsigned)(b.dim == 3 ? b.nz : 1));
    const bool sh = c->shift_active;
    prof_begin(c, K_PROJECT);
#define L_P(D, S) hipLaunchKernelGGL((project_kernel<D, S>), grid, dim3(kTX, kTY), 0, c->stream, a)
    if (b.dim == 3) { if (sh) L_P(1, 1); else L_P(1, 0); }
    else { if (sh) L_P(0, 1); else L_P(0, 0); }
#undef L_P
    prof_end(c, K_PROJECT, (b.dim == 3 ? 9.0 : 7.0) * 8.0 * c->n_int);
}

// ---------------------------------------------------------------------------
// K8: checks (src/navier_stokes.f90:570-619): signed max of div(v) and
// max(|u|+|v|+|w|) over the interior; flags NaN/Inf.  Max is exact, so the
// result does not depend on the reduction order.
// ---------------------------------------------------------------------------
constexpr int kCheckBlocks = 1024;

template <int D3>
__global__ __launch_bounds__(256) void checks_kernel(const double *u, const double *v, const double *w, Box b,
                                                     double id, double *part)
{
    __shared__ double sd[256], sv[256], sf[256];
    const long n = (long)b.nx * b.ny * b.nz;
    double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int i = (int)(t % b.nx) + 1;
        const long r = t / b.nx;
        const int j = (int)(r % b.ny) + 1, k = (int)(r / b.ny) + 1;
        const long g = b.g(i, j, k);
        double d = (u[g] - u[g - 1]) * id + (v[g] - v[g - b.gx]) * id;
        if constexpr (D3) d = d + (w[g] - w[g - b.gxy]) * id;
        double vel = fabs(u[g]) + fabs(v[g]);
        if constexpr (D3) vel = vel + fabs(w[g]);
        if (!(d == d) || !(vel == vel) || isinf(vel)) bad = 1.0;
        if (d > md) md = d;
        if (vel > mv) mv = vel;
    }
    sd[threadIdx.x] = md; sv[threadIdx.x] = mv; sf[threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (sd[threadIdx.x + s] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + s];
            if (sv[threadIdx.x + s] > sv[threadIdx.x]) sv[threadIdx.x] = sv[threadIdx.x + s];
            if (sf[threadIdx.x + s] > sf[threadIdx.x]) sf[threadIdx.x] = sf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = sd[0];
        part[3 * blockIdx.x + 1] = sv[0];
        part[3 * blockIdx.x + 2] = sf[0];
    }
}

__global__ __launch_bounds__(256) void checks_final_kernel(const double *part, int n, double *out)
{
    __shared__ double sd[256], sv[256], sf[256];
    double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
    for (int q = threadIdx.x; q < n; q += 256) {
        if (part[3 * q] > md) md = part[3 * q];
        if (part[3 * q + 1] > mv) mv = part[3 * q + 1];
        if (part[3 * q + 2] > bad) bad = part[3 * q + 2];
    }
    sd[threadIdx.x] = md; sv[threadIdx.x] = mv; sf[threadIdx.x] = bad;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (sd[threadIdx.x + s] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + s];
            if (sv[threadIdx.x + s] > sv[threadIdx.x]) sv[threadIdx.x] = sv[threadIdx.x + s];
            if (sf[threadIdx.x + s] > sf[threadIdx.x]) sf[threadIdx.x] = sf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sd[0]; out[1] = sv[0]; out[2] = sf[0]; }
}

void launch_checks(fen_ctx *c, const double *const *vel, double *out)
{
    const Box &b = c->box;
    const double id = 1.0 / c->delta;
    long nblk = ((long)c->n_int + 255) / 256;
    if (nblk > kCheckBlocks) nblk = kCheckBlocks;
    if ((size_t)(3 * nblk) > c->n_part) throw std::runtime_error("partial buffer too small");
    prof_begin(c, K_CHECKS);
    if (b.dim == 3)
        hipLaunchKernelGGL((checks_kernel<1>), dim3((unsigned)nblk), dim3(256), 0, c->stream, vel[0], vel[1], vel[2],
                           b, id, c->d_part);
    else
        hipLaunchKernelGGL((checks_kernel<0>), dim3((unsigned)nblk), dim3(256), 0, c->stream, vel[0], vel[1], vel[2],
                           b, id, c->d_part);
    prof_end(c, K_CHECKS, (b.dim == 3 ? 3.0 : 2.0) * 8.0 * c->n_int);
    prof_begin(c, K_REDUCE);
    hipLaunchKernelGGL(checks_final_kernel, dim3(1), dim3(256), 0, c->stream, (const double *)c->d_part, (int)nblk,
                       out);
    prof_end(c, K_REDUCE, 24.0 * nblk);
}

}  // namespace fen
