// stencil_kernels.hip -- staggered-grid kernels of the time step, for field
// element type R = double (FP64) and R = float (the FP32 form).
//
// Every expression keeps the reference's evaluation order (Fortran left to
// right, explicit parentheses) and the library is compiled with
// -ffp-contract=off, so the FP64 kernels reproduce the CPU arithmetic bit
// for bit.
//
// Producers (momentum, projection) write the periodic images of the boundary
// cells they compute, so a periodic ghost fill costs no extra pass; the ghost
// kernel then runs only over the axes that are not periodic.
#include <algorithm>
#include <cmath>
#include <stdexcept>
#include <type_traits>

#include "fen_internal.hpp"
#include "stencil_march.hpp"

namespace fen {

constexpr int kTX = 64, kTY = 4, kThreads = kTX * kTY;

// True when r is a power of two (its reciprocal is then exact).
template <class T> static bool pow2_value(T r)
{
    int e = 0;
    return r > T(0) && std::frexp(r, &e) == T(0.5);
}

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
constexpr int kGhostFields = 8;   // fields per ghost-fill launch
template <class R> struct GhostLaunch {
    GhostDesc<R> d[kGhostFields];
    Box b;
    int nx_cells, ny_cells, nz_cells;              // cells per plane of each group
    int ngx[kGhostFields], ngy[kGhostFields], ngz[kGhostFields];   // planes per group and field (0: skipped)
};

template <class R> __device__ __forceinline__ R ghost_s1(const GhostDesc<R> &d, const Box &b, int i, int j, int k)
{
    const R two = R(2.0);
    const long pj = j + (long)(b.ny + 2) * k;
    if (i == 0) {
        const int t = d.type[0];
        if (t == 0) return d.f[b.g(b.nx, j, k)];
        if (t == 1) {
            const R bc = d.plane[0][pj];
            return d.loc == 'x' ? bc : two * bc - d.f[b.g(1, j, k)];
        }
        if (t == 2) return d.f[b.g(1, j, k)];
        return d.f[b.g(i, j, k)];
    }
    if (i == b.nx + 1) {
        const int t = d.type[1];
        if (t == 0) return d.f[b.g(1, j, k)];
        if (t == 1) {
            const R bc = d.plane[1][pj];
            return d.loc == 'x' ? bc : two * bc - d.f[b.g(b.nx, j, k)];
        }
        if (t == 2) return d.f[b.g(b.nx, j, k)];
        return d.f[b.g(i, j, k)];
    }
    if (i == b.nx && d.type[1] == 1 && d.loc == 'x') return d.plane[1][pj];
    return d.f[b.g(i, j, k)];
}

template <class R> __device__ __forceinline__ R ghost_s2(const GhostDesc<R> &d, const Box &b, int i, int j, int k)
{
    const R two = R(2.0);
    const long pi = i + (long)(b.nx + 2) * k;
    if (j == 0) {
        const int t = d.type[2];
        if (t == 0) return ghost_s1(d, b, i, b.ny, k);
        if (t == 1) {
            const R bc = d.plane[2][pi];
            return d.loc == 'y' ? bc : two * bc - ghost_s1(d, b, i, 1, k);
        }
        if (t == 2) return ghost_s1(d, b, i, 1, k);
        return ghost_s1(d, b, i, j, k);
    }
    if (j == b.ny + 1) {
        const int t = d.type[3];
        if (t == 0) return ghost_s1(d, b, i, 1, k);
        if (t == 1) {
            const R bc = d.plane[3][pi];
            return d.loc == 'y' ? bc : two * bc - ghost_s1(d, b, i, b.ny, k);
        }
        if (t == 2) return ghost_s1(d, b, i, b.ny, k);
        return ghost_s1(d, b, i, j, k);
    }
    if (j == b.ny && d.type[3] == 1 && d.loc == 'y') return d.plane[3][pi];
    return ghost_s1(d, b, i, j, k);
}

template <class R> __device__ __forceinline__ R ghost_s3(const GhostDesc<R> &d, const Box &b, int i, int j, int k)
{
    const R two = R(2.0);
    if (b.dim == 3) {
        const long pij = i + (long)(b.nx + 2) * j;
        if (k == 0) {
            const int t = d.type[4];
            if (t == 0) return ghost_s2(d, b, i, j, b.nz);
            if (t == 1) {
                const R bc = d.plane[4][pij];
                return d.loc == 'z' ? bc : two * bc - ghost_s2(d, b, i, j, 1);
            }
            if (t == 2) return ghost_s2(d, b, i, j, 1);
            return ghost_s2(d, b, i, j, k);
        }
        if (k == b.nz + 1) {
            const int t = d.type[5];
            if (t == 0) return ghost_s2(d, b, i, j, 1);
            if (t == 1) {
                const R bc = d.plane[5][pij];
                return d.loc == 'z' ? bc : two * bc - ghost_s2(d, b, i, j, b.nz);
            }
            if (t == 2) return ghost_s2(d, b, i, j, b.nz);
            return ghost_s2(d, b, i, j, k);
        }
        if (k == b.nz && d.type[5] == 1 && d.loc == 'z') return d.plane[5][pij];
    }
    return ghost_s2(d, b, i, j, k);
}

// Group plane index p -> coordinate: 0 -> 0, 1 -> n+1, 2 -> n (normal overwrite).
__device__ __forceinline__ int plane_coord(int p, int n) { return p == 0 ? 0 : (p == 1 ? n + 1 : n); }

// One field per blockIdx.y.  The descriptor is selected with a compile-time
// index (Q) and the rule functions above are inlined: a run-time index into
// the by-value argument block, or its address passed to a call that is not
// inlined, makes the compiler copy the whole block (464 bytes) into per-lane
// scratch memory, which cost 55 us per field at 256^3.
template <class R, int Q> __device__ __forceinline__ void ghost_field(const GhostLaunch<R> &a)
{
    const GhostDesc<R> &d = a.d[Q];
    const Box &b = a.b;
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int tx = a.nx_cells * a.ngx[Q], ty = a.ny_cells * a.ngy[Q], tz = a.nz_cells * a.ngz[Q];
    int i, j, k;
    if (t < tx) {
        // x faces: i ghost (or the overwritten normal plane), j and k interior
        const int p = t / a.nx_cells, r = t - p * a.nx_cells;
        i = plane_coord(p, b.nx);
        k = r / b.ny;
        j = r - k * b.ny + 1;
        k += 1;
    } else if ((t -= tx) < ty) {
        // y faces: j ghost, every i, k interior
        const int p = t / a.ny_cells, r = t - p * a.ny_cells;
        j = plane_coord(p, b.ny);
        k = r / (b.nx + 2);
        i = r - k * (b.nx + 2);
        k += 1;
    } else if ((t -= ty) < tz) {
        // z faces: k ghost, every i and j
        const int p = t / a.nz_cells, r = t - p * a.nz_cells;
        k = plane_coord(p, b.nz);
        j = r / (b.nx + 2);
        i = r - j * (b.nx + 2);
    } else {
        return;
    }
    d.f[b.g(i, j, k)] = ghost_s3(d, b, i, j, k);
}

template <class R> __global__ __launch_bounds__(256) void ghost_kernel(GhostLaunch<R> a)
{
    switch (blockIdx.y) {
    case 0: ghost_field<R, 0>(a); break;
    case 1: ghost_field<R, 1>(a); break;
    case 2: ghost_field<R, 2>(a); break;
    case 3: ghost_field<R, 3>(a); break;
    case 4: ghost_field<R, 4>(a); break;
    case 5: ghost_field<R, 5>(a); break;
    case 6: ghost_field<R, 6>(a); break;
    default: ghost_field<R, 7>(a); break;
    }
}

template <class R> void launch_ghost_desc(fen_ctx *c, const GhostDesc<R> *descs, int n)
{
    if (n <= 0 || n > kGhostFields) throw std::runtime_error("ghost fill of 1..8 fields");
    GhostLaunch<R> a{};
    const Box &b = c->box;
    a.b = b;
    a.nx_cells = b.ny * b.nz;
    a.ny_cells = (b.nx + 2) * b.nz;
    a.nz_cells = (b.nx + 2) * (b.ny + 2);
    long most = 0;
    double cells = 0.0;
    for (int q = 0; q < n; q++) {
        const GhostDesc<R> &d = descs[q];
        a.d[q] = d;
        a.ngx[q] = (d.groups & 1) ? (d.type[FEN_RIGHT] == 1 && d.loc == 'x' ? 3 : 2) : 0;
        a.ngy[q] = (d.groups & 2) ? (d.type[FEN_TOP] == 1 && d.loc == 'y' ? 3 : 2) : 0;
        a.ngz[q] = (b.dim == 3 && (d.groups & 4)) ? (d.type[FEN_BACK] == 1 && d.loc == 'z' ? 3 : 2) : 0;
        const long tot = (long)a.nx_cells * a.ngx[q] + (long)a.ny_cells * a.ngy[q] + (long)a.nz_cells * a.ngz[q];
        most = tot > most ? tot : most;
        cells += (double)tot;
    }
    if (most == 0) return;
    if (most > INT32_MAX - 256) throw std::runtime_error("ghost fill: too many cells for 32-bit indices");
    const dim3 grid((unsigned)((most + 255) / 256), (unsigned)n);
    prof_begin(c, K_GHOST);
    hipLaunchKernelGGL(ghost_kernel<R>, grid, dim3(256), 0, c->stream, a);
    prof_end(c, K_GHOST, 2.0 * sizeof(R) * cells);
}

// 2-D: one cell per thread, neighbours straight from memory (the whole
// 128^2 problem fits in cache).
template <class R, int SRC>
__global__ __launch_bounds__(kThreads) void momentum_kernel_2d(MomArgs<R> a)
{
    const int i = blockIdx.x * kTX + (threadIdx.x % kTX) + 1;
    const int j = blockIdx.y * kTY + (threadIdx.x / kTX) + 1;
    if (i > a.b.nx || j > a.b.ny) return;
    const long g = a.b.g(i, j, 1), sj = a.b.gx;
    auto at = [&](const R *f) {
        return [=](int di, int dj, int dk) { return f[g + di + sj * dj]; };
    };
    const long q = a.b.q(i, j, 1);
    auto img = [&](R *f, R v) { write_images(f, a.b, a.img, i, j, 1, v); };
    momentum_cell<0, SRC>(a, i, j, 1, at(a.u), at(a.v), at(a.w), at(a.p), a.dvo0[q], a.dvo1[q], R(0), img);
}

// 3-D launch shapes (see stencil_march.hpp), chosen per element type by a
// sweep of tile height, planes per workgroup, ring unrolling and prefetch
// depth on MI355X at 256^3.
// (tile height TY, planes per workgroup KB, ring unrolling UN, prefetch depth
// PF, minimum waves per SIMD for the register allocation OCC; the fused
// kernel's tiles are FTX = 128 wide: two waves per row, so every row of a
// tile plane is a 1 KB (FP64) contiguous run -- 0.811 -> 0.804 ms FP64 and
// 0.464 -> 0.456 ms FP32 per step against 64 x 8 tiles; 256 x 2 lost 10%.
// 128-wide momentum and projection tiles gained 0.6% on the per-step path
// of the Neumann solvers and are not used: MTXW = PTXW = 64)
template <class R> struct Shape;
template <> struct Shape<double> {
    static constexpr int MTY = 8, MKB = 32, MUN = 1, MPF = 1, MOCC = 1, MTXW = 64;   // momentum
    static constexpr int PTY = 4, PKB = 16, PUN = 1, PPF = 1, POCC = 1, PTXW = 64;   // projection
    static constexpr int FTY = 4, FKB = 16, FOCC = 1, FTX = 128;          // projection + momentum
    static constexpr bool FLONG = true;   // march as many planes as keep every CU busy
};
template <> struct Shape<float> {
    static constexpr int MTY = 4, MKB = 16, MUN = 2, MPF = 2, MOCC = 1, MTXW = 64;
    static constexpr int PTY = 4, PKB = 16, PUN = 1, PPF = 1, POCC = 4, PTXW = 64;
    static constexpr int FTY = 4, FKB = 32, FOCC = 1, FTX = 128;
    static constexpr bool FLONG = false;   // two workgroups per CU: longer marches measured slower
};
constexpr int MTX = 64;

static void grid_blocks(const Box &b, int ty, int kb, int *nbx, int *nby, int *nbz, int tx = MTX)
{
    if (b.dim == 3) {
        *nbx = (b.nx + tx - 1) / tx;
        *nby = (b.ny + ty - 1) / ty;
        *nbz = (b.nz + kb - 1) / kb;
    } else {
        *nbx = (b.nx + kTX - 1) / kTX;
        *nby = (b.ny + kTY - 1) / kTY;
        *nbz = 1;
    }
}

// Workgroups of the kernels that write one partial each (projection, fused
// projection + momentum), either element type: the partial buffer's size.
long stencil_blocks(const Box &b)
{
    const int shapes[4][3] = {{Shape<double>::PTY, Shape<double>::PKB, Shape<double>::PTXW},
                              {Shape<float>::PTY, Shape<float>::PKB, Shape<float>::PTXW},
                              {Shape<double>::FTY, Shape<double>::FKB, Shape<double>::FTX},
                              {Shape<float>::FTY, Shape<float>::FKB, Shape<float>::FTX}};
    long most = 0;
    for (const auto &sh : shapes) {
        int nbx, nby, nbz;
        grid_blocks(b, sh[0], sh[1], &nbx, &nby, &nbz, sh[2]);
        most = std::max(most, (long)nbx * nby * nbz);
    }
    return most;
}

template <class R> MomArgs<R> momentum_args(fen_ctx *c, double dt, double A, double B)
{
    MomArgs<R> a{};
    a.u = as<R>(c->d_u[0]); a.v = as<R>(c->d_u[1]); a.w = as<R>(c->d_u[2]); a.p = as<R>(c->d_p);
    a.us = as<R>(c->d_us[0]); a.vs = as<R>(c->d_us[1]); a.ws = as<R>(c->d_us[2]);
    a.dvo0 = as<R>(c->d_dvo[0]); a.dvo1 = as<R>(c->d_dvo[1]); a.dvo2 = as<R>(c->d_dvo[2]);
    a.s0 = as<R>(c->d_S[0]); a.s1 = as<R>(c->d_S[1]); a.s2 = as<R>(c->d_S[2]);
    a.b = c->box;
    // coefficients are formed in FP64 and rounded once to R
    a.id = (R)(1.0 / c->delta);
    a.id2 = (R)(1.0 / (c->delta * c->delta));
    a.mu = (R)c->viscosity;
    a.rhof = (R)(0.5 * (c->density + c->density));
    a.rpow2 = pow2_value(a.rhof) ? 1 : 0;
    a.irhof = a.rpow2 ? R(1) / a.rhof : R(0);
    a.dt = (R)dt; a.A = (R)A; a.B = (R)B;
    a.g0 = (R)c->g[0]; a.g1 = (R)c->g[1]; a.g2 = (R)c->g[2];
    a.img = c->img_vel;
    return a;
}

template <class R> void launch_momentum(fen_ctx *c, double dt, double A, double B)
{
    MomArgs<R> a = momentum_args<R>(c, dt, A, B);
    int nbz;
    using SH = Shape<R>;
    grid_blocks(c->box, SH::MTY, SH::MKB, &a.nbx, &a.nby, &nbz, SH::MTXW);
    const int d3 = c->box.dim == 3;
    const double words = d3 ? (c->has_source ? 16.0 : 13.0) : (c->has_source ? 11.0 : 9.0);
    prof_begin(c, K_MOMENTUM);
    if (d3) {
        const dim3 grid((unsigned)(a.nbx * a.nby * nbz));
        if (c->has_source)
            hipLaunchKernelGGL((momentum_march_kernel<R, 1, SH::MTY, SH::MKB, SH::MUN, SH::MPF, SH::MOCC, SH::MTXW>),
                               grid, dim3(SH::MTXW * SH::MTY), 0, c->stream, a);
        else
            hipLaunchKernelGGL((momentum_march_kernel<R, 0, SH::MTY, SH::MKB, SH::MUN, SH::MPF, SH::MOCC, SH::MTXW>),
                               grid, dim3(SH::MTXW * SH::MTY), 0, c->stream, a);
    } else {
        const dim3 grid((unsigned)a.nbx, (unsigned)a.nby);
        if (c->has_source) hipLaunchKernelGGL((momentum_kernel_2d<R, 1>), grid, dim3(kThreads), 0, c->stream, a);
        else hipLaunchKernelGGL((momentum_kernel_2d<R, 0>), grid, dim3(kThreads), 0, c->stream, a);
    }
    prof_end(c, K_MOMENTUM, words * sizeof(R) * c->n_int);
}

// add_advection alone on the resident velocity (reference test
// test/small_test/navier_stokes/advection/advection.f90:50): R -= adv.
template <int D3>
__global__ __launch_bounds__(kThreads) void advection_kernel(const double *u, const double *v, const double *w,
                                                             double *rx, double *ry, double *rz, Box b, double id)
{
    const int i = blockIdx.x * kTX + (threadIdx.x % kTX) + 1;
    const int j = blockIdx.y * kTY + (threadIdx.x / kTX) + 1;
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
    const double *u = as<double>(c->d_u[0]), *v = as<double>(c->d_u[1]), *w = as<double>(c->d_u[2]);
    if (b.dim == 3)
        hipLaunchKernelGGL((advection_kernel<1>), grid, dim3(kThreads), 0, c->stream, u, v, w, rx, ry, rz, b,
                           1.0 / c->delta);
    else
        hipLaunchKernelGGL((advection_kernel<0>), grid, dim3(kThreads), 0, c->stream, u, v, w, rx, ry, rz, b,
                           1.0 / c->delta);
}

template <class R, int SHIFT>
__global__ __launch_bounds__(kThreads) void project_kernel_2d(ProjArgs<R> a)
{
    const int i = blockIdx.x * kTX + (threadIdx.x % kTX) + 1;
    const int j = blockIdx.y * kTY + (threadIdx.x / kTX) + 1;
    const Box &b = a.b;
    double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
    if (i <= b.nx && j <= b.ny) {
        const long g = b.g(i, j, 1), sj = b.gx;
        const double s = SHIFT ? a.shift[0] : 0.0;
        auto F = [&](int di, int dj, int dk) -> double {
            if constexpr (SHIFT) return (double)a.phi[g + di + sj * dj] - s;
            else return (double)a.phi[g + di + sj * dj];
        };
        auto img = [&](R *f, R v) { write_images(f, a.b, a.img, i, j, 1, v); };
        project_cell<0>(a, i, j, 1, F, a.us[g], a.vs[g], 0.0, a.p[g], a.us[g - 1], a.vs[g - sj], 0.0, md, mv, bad,
                        img);
    }
    block_max3<kThreads>(md, mv, bad, a.part, (long)blockIdx.y * gridDim.x + blockIdx.x);
}

// One workgroup reduces the per-workgroup partials.  Each thread issues the
// loads of B entries before comparing any, so the pass costs about one
// memory latency per B entries instead of one per entry.
constexpr int kFinalThreads = 1024;
__global__ __launch_bounds__(kFinalThreads) void checks_final_kernel(const double *part, long n, double *out)
{
    constexpr int NT = kFinalThreads, B = 4;
    __shared__ double sd[NT], sv[NT], sf[NT];
    double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
    for (long q0 = threadIdx.x; q0 < n; q0 += (long)NT * B) {
        double x[B][3];
#pragma unroll
        for (int m = 0; m < B; m++) {
            const long q = q0 + (long)m * NT;
            const long qq = q < n ? q : q0;   // a valid entry; its value is dropped below
            x[m][0] = part[3 * qq];
            x[m][1] = part[3 * qq + 1];
            x[m][2] = part[3 * qq + 2];
        }
#pragma unroll
        for (int m = 0; m < B; m++) {
            if (q0 + (long)m * NT >= n) continue;
            if (x[m][0] > md) md = x[m][0];
            if (x[m][1] > mv) mv = x[m][1];
            if (x[m][2] > bad) bad = x[m][2];
        }
    }
    sd[threadIdx.x] = md; sv[threadIdx.x] = mv; sf[threadIdx.x] = bad;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (sd[threadIdx.x + s] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + s];
            if (sv[threadIdx.x + s] > sv[threadIdx.x]) sv[threadIdx.x] = sv[threadIdx.x + s];
            if (sf[threadIdx.x + s] > sf[threadIdx.x]) sf[threadIdx.x] = sf[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[0] = sd[0]; out[1] = sv[0]; out[2] = sf[0]; }
}

template <class R> void launch_project(fen_ctx *c, double dt, double *checks_out, bool store)
{
    if (!store && c->box.dim != 3) throw std::runtime_error("checks-only projection is 3-D");
    ProjArgs<R> a{};
    a.us = as<R>(c->d_us[0]); a.vs = as<R>(c->d_us[1]); a.ws = as<R>(c->d_us[2]); a.phi = as<R>(c->d_phi);
    a.u = as<R>(c->d_u[0]); a.v = as<R>(c->d_u[1]); a.w = as<R>(c->d_u[2]); a.p = as<R>(c->d_p);
    a.shift = c->d_shift;
    a.part = c->d_part;
    const int lo[3] = {FEN_LEFT, FEN_BOTTOM, FEN_FRONT}, hi[3] = {FEN_RIGHT, FEN_TOP, FEN_BACK};
    for (int d = 0; d < 3; d++) {
        const int slot = FEN_U + d;
        a.lo_type[d] = c->bc_type[slot][lo[d]];
        a.hi_type[d] = c->bc_type[slot][hi[d]];
        a.lo_plane[d] = as<R>(c->d_plane[slot][lo[d]]);
        a.hi_plane[d] = as<R>(c->d_plane[slot][hi[d]]);
    }
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.dt = dt;
    a.rhof = 0.5 * (c->density + c->density);
    a.rpow2 = pow2_value(a.rhof) ? 1 : 0;
    a.irhof = a.rpow2 ? 1.0 / a.rhof : 0.0;
    if (c->mf) {   // MF: v - grad(phi) dt / rho_min as (grad(phi) dt) * irhomin (src/navier_stokes.f90:526-531)
        a.rhof = c->mf->rhomin;
        a.rpow2 = 1;
        a.irhof = c->mf->irhomin;
    }
    a.img = c->img_vel & c->img_p;
    int nbz;
    using SH = Shape<R>;
    grid_blocks(c->box, SH::PTY, SH::PKB, &a.nbx, &a.nby, &nbz, SH::PTXW);
    const long nblk = (long)a.nbx * a.nby * nbz;
    if ((size_t)(3 * nblk) > c->n_part) throw std::runtime_error("partial buffer too small");
    const bool sh = c->shift_active;
    const int d3 = c->box.dim == 3;
    const int kid = store ? K_PROJECT : K_CHECKS;
    prof_begin(c, kid);
    if (d3) {
        const dim3 grid((unsigned)nblk);
#define FEN_PJ(SH_, ST_)                                                                                            \
    hipLaunchKernelGGL((project_march_kernel<R, SH_, SH::PTY, SH::PKB, SH::PUN, SH::PPF, SH::POCC, SH::PTXW, ST_>), \
                       grid, dim3(SH::PTXW * SH::PTY), 0, c->stream, a)
        if (store) {
            if (sh) FEN_PJ(1, 1);
            else FEN_PJ(0, 1);
        } else {
            if (sh) FEN_PJ(1, 0);
            else FEN_PJ(0, 0);
        }
#undef FEN_PJ
    } else {
        const dim3 grid((unsigned)a.nbx, (unsigned)a.nby);
        if (sh) hipLaunchKernelGGL((project_kernel_2d<R, 1>), grid, dim3(kThreads), 0, c->stream, a);
        else hipLaunchKernelGGL((project_kernel_2d<R, 0>), grid, dim3(kThreads), 0, c->stream, a);
    }
    // the checks alone read u*, v*, w* and phi
    prof_end(c, kid, (store ? (d3 ? 9.0 : 7.0) : 4.0) * sizeof(R) * c->n_int);
    prof_begin(c, K_REDUCE);
    hipLaunchKernelGGL(checks_final_kernel, dim3(1), dim3(kFinalThreads), 0, c->stream, (const double *)c->d_part,
                       nblk, checks_out);
    prof_end(c, K_REDUCE, 24.0 * nblk);
}

// x and y periodic for every field, z faces no-slip walls (the ghost rules
// the fused kernel evaluates itself): u, v, w Dirichlet, p and phi Neumann.
bool fused_z_walls(const fen_ctx *c)
{
    if (c->box.dim != 3 || c->img_vel != 3 || c->img_p != 3 || c->img_phi != 3) return false;
    for (int f : {FEN_FRONT, FEN_BACK}) {
        if (c->bc_type[FEN_U][f] != 1 || c->bc_type[FEN_V][f] != 1 || c->bc_type[FEN_W][f] != 1) return false;
        if (c->bc_type[FEN_P][f] != 2 || c->bc_type[FEN_PHI][f] != 2) return false;
    }
    return true;
}

// The general-wall form of the fused kernel (3-D FP64, no pending mean
// shift; the Neumann-direction grids npn and nnn): every axis of every
// state field is periodic on both faces, or Dirichlet/Neumann on both
// (the tile formation evaluates the ghost fill's rule chain).  Grids the
// two dedicated forms take are excluded.
bool fused_general_walls(const fen_ctx *c)
{
    if (c->box.dim != 3 || c->precision != FEN_FP64) return false;
    if (c->solver != FEN_SOLVER_NPN && c->solver != FEN_SOLVER_NNN) return false;
    for (int ax = 0; ax < 3; ax++) {
        const bool per = c->periodic[ax] != 0;
        for (int f : {FEN_U, FEN_V, FEN_W, FEN_P, FEN_PHI})
            for (int side = 0; side < 2; side++) {
                const int t = c->bc_type[f][2 * ax + side];
                if (per ? t != 0 : (t != 1 && t != 2)) return false;
            }
    }
    return true;
}

// Step n's projection fused with step n+1's momentum (a grid periodic in
// every direction, x, y periodic with z walls, or the general walls of
// fused_general_walls).  Step n's checks are not formed: step n+1 ends
// pending, so its own checks are what a reader asks for.
// us_in/p_in: u*, v*, w* and p of step n;
// us_out/p_out: u*, v*, w* of step n+1 and p after step n (distinct arrays);
// save (FP64, periodic grid, no source), if given: u, v, w of step n with
// their periodic images.
template <class R>
void launch_project_momentum(fen_ctx *c, double dt_p, double dt_m, double A, double B, void *const *us_in,
                             const void *p_in, void *const *us_out, void *p_out, void *const *save)
{
    const bool periodic = c->img_vel == 7 && c->img_p == 7 && c->img_phi == 7 && !c->shift_active;
    const bool zwalls = fused_z_walls(c);
    const bool general = !periodic && !zwalls && std::is_same<R, double>::value && fused_general_walls(c) &&
                         !c->shift_active;
    if (c->box.dim != 3 || !(periodic || zwalls || general))
        throw std::runtime_error("fused projection + momentum needs a periodic grid, x, y periodic with z walls, "
                                 "or the npn/nnn walls");
    PMArgs<R> a{};
    if (general) {
        a.wall = (c->periodic[0] ? 0 : 1) | (c->periodic[1] ? 0 : 2) | (c->periodic[2] ? 0 : 4);
        // packed per field: face f's bc type in bits 2f..2f+1, location in bits 12..13
        const int fid[4] = {FEN_U, FEN_V, FEN_W, FEN_P};
        const int loc[4] = {1, 2, 3, 0};
        for (int q = 0; q < 4; q++) {
            a.gtype[q] = loc[q] << 12;
            for (int f = 0; f < 6; f++) a.gtype[q] |= c->bc_type[fid[q]][f] << (2 * f);
        }
        a.gplane = reinterpret_cast<const R *const *>(c->d_plane_tab);
        if (!c->d_wallbuf) FEN_HIP_CHECK(hipMalloc(&c->d_wallbuf, 4 * pm_wall_layout(c->box).per_field * sizeof(R)));
        a.wbuf = as<R>(c->d_wallbuf);
    }
    if (zwalls) {
        a.shift = c->shift_active ? c->d_shift : nullptr;
        for (int d = 0; d < 3; d++) {
            a.zlo[d] = as<R>(c->d_plane[FEN_U + d][FEN_FRONT]);
            a.zhi[d] = as<R>(c->d_plane[FEN_U + d][FEN_BACK]);
        }
    }
    a.m = momentum_args<R>(c, dt_m, A, B);
    a.m.us = as<R>(us_out[0]); a.m.vs = as<R>(us_out[1]); a.m.ws = as<R>(us_out[2]);
    a.us = as<R>(us_in[0]); a.vs = as<R>(us_in[1]); a.ws = as<R>(us_in[2]);
    a.phi = as<R>(c->d_phi);
    a.p_in = static_cast<const R *>(p_in);
    a.p_out = as<R>(p_out);
    a.pid = 1.0 / c->delta;
    a.pdt = dt_p;
    a.prhof = 0.5 * (c->density + c->density);
    a.prpow2 = pow2_value(a.prhof) ? 1 : 0;
    a.pirhof = a.prpow2 ? 1.0 / a.prhof : 0.0;
    using SH = Shape<R>;
    int nbz;
    grid_blocks(c->box, SH::FTY, SH::FKB, &a.m.nbx, &a.m.nby, &nbz, SH::FTX);
    // planes per workgroup: FKB, or with FLONG the longest march (each
    // workgroup loads two planes before its first) that leaves no CU idle
    a.kb = SH::FKB;
    if (SH::FLONG) {
        const long cols = (long)a.m.nbx * a.m.nby;
        while (2 * a.kb <= c->box.nz && cols * ((c->box.nz + 2 * a.kb - 1) / (2 * a.kb)) >= c->num_cu) a.kb *= 2;
    }
    nbz = (c->box.nz + a.kb - 1) / a.kb;
    const long nblk = (long)a.m.nbx * a.m.nby * nbz;
    if (general) {
        // the state at the wall positions, read by the march below
        const long nw = pm_wall_layout(c->box).per_field;
        hipLaunchKernelGGL(wall_state_kernel<R>, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, c->stream, a);
    }
    prof_begin(c, K_PROJ_MOM);
    const dim3 grid((unsigned)nblk), block(SH::FTX * SH::FTY);
#define FEN_PM(SRC, ZW) \
    hipLaunchKernelGGL((project_momentum_kernel<R, SRC, ZW, SH::FTY, SH::FOCC, SH::FTX>), grid, block, 0, c->stream, a)
    if (save) {
        if constexpr (std::is_same<R, double>::value) {
            if (!periodic || c->has_source) throw std::runtime_error("state saving needs a periodic grid and no source");
            for (int d = 0; d < 3; d++) a.save[d] = as<R>(save[d]);
            hipLaunchKernelGGL((project_momentum_kernel<R, 0, 0, SH::FTY, SH::FOCC, SH::FTX, 1>), grid, block, 0,
                               c->stream, a);
        } else {
            throw std::runtime_error("state saving is FP64");
        }
    } else if (c->has_source) {
        if (zwalls) FEN_PM(1, 1);
        else if (general) FEN_PM(1, 2);
        else FEN_PM(1, 0);
    } else {
        if (zwalls) FEN_PM(0, 1);
        else if (general) FEN_PM(0, 2);
        else FEN_PM(0, 0);
    }
#undef FEN_PM
    // reads u*, v*, w*, p, phi, dv_o (+ S); writes p, u*, v*, w*, dv_o
    prof_end(c, K_PROJ_MOM, (c->has_source ? 18.0 : 15.0) * sizeof(R) * c->n_int);
}

// ---------------------------------------------------------------------------
// Stand-alone checks on the resident velocity (used before the first step
// of a CFL-controlled run): same quantities as above, ghosts already filled.
// ---------------------------------------------------------------------------
constexpr int kCheckBlocks = 1024;

template <class R, int D3>
__global__ __launch_bounds__(256) void checks_kernel(const R *u, const R *v, const R *w, Box b, double id,
                                                     double *part)
{
    __shared__ double sd[256], sv[256], sf[256];
    const long n = (long)b.nx * b.ny * b.nz;
    double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
    for (long t = (long)blockIdx.x * 256 + threadIdx.x; t < n; t += (long)gridDim.x * 256) {
        const int i = (int)(t % b.nx) + 1;
        const long r = t / b.nx;
        const int j = (int)(r % b.ny) + 1, k = (int)(r / b.ny) + 1;
        const long g = b.g(i, j, k);
        const double ug = u[g], vg = v[g];
        double d = (ug - (double)u[g - 1]) * id + (vg - (double)v[g - b.gx]) * id;
        if constexpr (D3) d = d + ((double)w[g] - (double)w[g - b.gxy]) * id;
        double vel = fabs(ug) + fabs(vg);
        if constexpr (D3) vel = vel + fabs((double)w[g]);
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

template <class R> void launch_checks(fen_ctx *c, const R *const *vel, double *out)
{
    const Box &b = c->box;
    const double id = 1.0 / c->delta;
    long nblk = ((long)c->n_int + 255) / 256;
    if (nblk > kCheckBlocks) nblk = kCheckBlocks;
    if ((size_t)(3 * nblk) > c->n_part) throw std::runtime_error("partial buffer too small");
    prof_begin(c, K_CHECKS);
    if (b.dim == 3)
        hipLaunchKernelGGL((checks_kernel<R, 1>), dim3((unsigned)nblk), dim3(256), 0, c->stream, vel[0], vel[1],
                           vel[2], b, id, c->d_part);
    else
        hipLaunchKernelGGL((checks_kernel<R, 0>), dim3((unsigned)nblk), dim3(256), 0, c->stream, vel[0], vel[1],
                           vel[2], b, id, c->d_part);
    prof_end(c, K_CHECKS, (b.dim == 3 ? 3.0 : 2.0) * sizeof(R) * c->n_int);
    prof_begin(c, K_REDUCE);
    hipLaunchKernelGGL(checks_final_kernel, dim3(1), dim3(kFinalThreads), 0, c->stream, (const double *)c->d_part,
                       nblk, out);
    prof_end(c, K_REDUCE, 24.0 * nblk);
}

template void launch_ghost_desc<double>(fen_ctx *, const GhostDesc<double> *, int);
template void launch_ghost_desc<float>(fen_ctx *, const GhostDesc<float> *, int);
template void launch_momentum<double>(fen_ctx *, double, double, double);
template void launch_momentum<float>(fen_ctx *, double, double, double);
template void launch_project<double>(fen_ctx *, double, double *, bool);
template void launch_project<float>(fen_ctx *, double, double *, bool);
template void launch_project_momentum<double>(fen_ctx *, double, double, double, double, void *const *, const void *,
                                              void *const *, void *, void *const *);
template void launch_project_momentum<float>(fen_ctx *, double, double, double, double, void *const *, const void *,
                                             void *const *, void *, void *const *);
template void launch_checks<double>(fen_ctx *, const double *const *, double *);
template void launch_checks<float>(fen_ctx *, const float *const *, double *);

}  // namespace fen
