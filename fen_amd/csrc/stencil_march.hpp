// stencil_march.hpp -- k-marching launch shape of the 3-D momentum and
// projection kernels.
//
// A workgroup owns a 64 x TY tile of (i, j) and marches KB planes in k.
// Planes k-1, k, k+1 of the stencil inputs, each with a one-cell halo in i
// and j, sit in a 3-slot LDS ring, so every value is read from HBM once per
// tile plus the halo.  The ring is software-pipelined: the global loads of
// plane k+2 are issued into registers before plane k is computed and are
// written to LDS one iteration later, so HBM latency overlaps arithmetic.
// Inputs needed only at the cell itself (dv_o; w*, p in the projection) are
// prefetched one plane ahead in registers.  With UNROLL == 3 the k loop is
// unrolled by three so every ring slot is a compile-time LDS offset.
#pragma once

#include <type_traits>

#include "stencil_cells.hpp"

namespace fen {

template <int TY> struct TileShape {
    static constexpr int TX = 64, LX = TX + 2, LY = TY + 2, N = LX * LY, NT = TX * TY;
    static constexpr int PER = (N + NT - 1) / NT;
};

template <class S> struct TileIdx {
    long off[S::PER];   // offset in the ghosted array at k = 0, or -1 outside the tile/grid
};

template <class S> __device__ __forceinline__ TileIdx<S> tile_index(const Box &b, int i0, int j0)
{
    TileIdx<S> t;
#pragma unroll
    for (int q = 0; q < S::PER; q++) {
        const int e = threadIdx.x + q * S::NT;
        const int ri = e % S::LX, rj = e / S::LX;
        const int gi = i0 - 1 + ri, gj = j0 - 1 + rj;
        const bool ok = e < S::N && gi <= b.nx + 1 && gj <= b.ny + 1;
        t.off[q] = ok ? b.g(gi, gj, 0) : -1;
    }
    return t;
}

template <class S, int NF, class R> struct TileRegs {
    R v[NF][S::PER];
};

template <class S, int NF, class R>
__device__ __forceinline__ void tile_fetch(TileRegs<S, NF, R> &r, const R *const *src, const TileIdx<S> &t,
                                           long kofs)
{
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int q = 0; q < S::PER; q++)
            if (t.off[q] >= 0) r.v[f][q] = src[f][t.off[q] + kofs];
}

// LDS layout: [field][slot 0..2][S::N]
template <class S, int NF, class R, class SL>
__device__ __forceinline__ void tile_commit(R *lds, SL slot, const TileRegs<S, NF, R> &r, const TileIdx<S> &t)
{
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int q = 0; q < S::PER; q++)
            if (t.off[q] >= 0) lds[(f * 3 + (int)slot) * S::N + threadIdx.x + q * S::NT] = r.v[f][q];
}

template <int V> using IC = std::integral_constant<int, V>;

// ----------------------------------------------------------------- momentum
template <class R, int SRC, int TY, int KB, int UNROLL>
struct MomentumMarch {
    using S = TileShape<TY>;
    static constexpr int NT = S::NT;

    template <class SM, class SC, class SN>
    __device__ __forceinline__ static void plane(const MomArgs<R> &a, R *lds, TileRegs<S, 4, R> &r,
                                                 const TileIdx<S> &t, const R *const *src, int i, int j, int k,
                                                 int kend, bool active, int c, R (&o)[3], SM sm, SC sc, SN sn)
    {
        tile_commit<S, 4>(lds, sn, r, t);   // plane k+1
        __syncthreads();
        if (k + 2 <= kend + 1) tile_fetch<S, 4>(r, src, t, (long)(k + 2) * a.b.gxy);
        R on[3] = {R(0), R(0), R(0)};
        const long q = active ? a.b.q(i, j, k) : 0;
        const long qs = (long)a.b.nx * a.b.ny;
        if (active && k + 1 <= kend) {
            on[0] = a.dvo0[q + qs];
            on[1] = a.dvo1[q + qs];
            on[2] = a.dvo2[q + qs];
        }
        if (active) {
            auto at = [&](int f) {
                return [=](int di, int dj, int dk) {
                    const int slot = dk < 0 ? (int)sm : (dk > 0 ? (int)sn : (int)sc);
                    return lds[(f * 3 + slot) * S::N + c + dj * S::LX + di];
                };
            };
            momentum_cell<1, SRC>(a, i, j, k, at(0), at(1), at(2), at(3), o[0], o[1], o[2]);
        }
        o[0] = on[0]; o[1] = on[1]; o[2] = on[2];
        __syncthreads();
    }

    __device__ __forceinline__ static void run(const MomArgs<R> &a, R *lds)
    {
        const Box &b = a.b;
        const int bid = xcd_remap(blockIdx.x, gridDim.x);
        const int bx = bid % a.nbx, rest = bid / a.nbx;
        const int by = rest % a.nby, bz = rest / a.nby;
        const int tx = threadIdx.x % S::TX, ty = threadIdx.x / S::TX;
        const int i0 = bx * S::TX + 1, j0 = by * TY + 1, k0 = bz * KB + 1;
        const int kend = min(k0 + KB - 1, b.nz);
        const int i = i0 + tx, j = j0 + ty;
        const bool active = i <= b.nx && j <= b.ny;
        const int c = (ty + 1) * S::LX + tx + 1;
        const R *src[4] = {a.u, a.v, a.w, a.p};
        const TileIdx<S> t = tile_index<S>(b, i0, j0);
        TileRegs<S, 4, R> r;
        tile_fetch<S, 4>(r, src, t, (long)(k0 - 1) * b.gxy);
        tile_commit<S, 4>(lds, IC<0>{}, r, t);
        tile_fetch<S, 4>(r, src, t, (long)k0 * b.gxy);
        tile_commit<S, 4>(lds, IC<1>{}, r, t);
        tile_fetch<S, 4>(r, src, t, (long)(k0 + 1) * b.gxy);
        R o[3] = {R(0), R(0), R(0)};
        if (active) {
            const long q = b.q(i, j, k0);
            o[0] = a.dvo0[q]; o[1] = a.dvo1[q]; o[2] = a.dvo2[q];
        }
        if constexpr (UNROLL == 3) {
            for (int k = k0; k <= kend; k += 3) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, o, IC<0>{}, IC<1>{}, IC<2>{});
                if (k + 1 <= kend) plane(a, lds, r, t, src, i, j, k + 1, kend, active, c, o, IC<1>{}, IC<2>{}, IC<0>{});
                if (k + 2 <= kend) plane(a, lds, r, t, src, i, j, k + 2, kend, active, c, o, IC<2>{}, IC<0>{}, IC<1>{});
            }
        } else {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k++) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, o, sm, sc, sn);
                const int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
            }
        }
    }
};

template <class R, int SRC, int TY, int KB, int UNROLL>
__global__ __launch_bounds__(64 * TY) void momentum_march_kernel(MomArgs<R> a)
{
    __shared__ R lds[4 * 3 * TileShape<TY>::N];
    MomentumMarch<R, SRC, TY, KB, UNROLL>::run(a, lds);
}

// --------------------------------------------------------------- projection
// Ring fields: phi, u*, v* (u* and v* need the cell below in i and j for the
// checks); w* and p are prefetched one plane ahead, w* one plane below is
// carried in a register.
template <class R, int SHIFT, int TY, int KB, int UNROLL>
struct ProjectMarch {
    using S = TileShape<TY>;
    static constexpr int NT = S::NT;

    template <class SM, class SC, class SN>
    __device__ __forceinline__ static void plane(const ProjArgs<R> &a, R *lds, TileRegs<S, 3, R> &r,
                                                 const TileIdx<S> &t, const R *const *src, int i, int j, int k,
                                                 int kend, bool active, int c, double s, double &wsm, double &wsc,
                                                 double &pc, double &md, double &mv, double &bad, SM sm, SC sc, SN sn)
    {
        tile_commit<S, 3>(lds, sn, r, t);
        __syncthreads();
        if (k + 2 <= kend + 1) tile_fetch<S, 3>(r, src, t, (long)(k + 2) * a.b.gxy);
        double wnext = 0.0, pnext = 0.0;
        const long g = a.b.g(i, j, k);
        if (active && k + 1 <= kend) {
            wnext = a.ws[g + a.b.gxy];
            pnext = a.p[g + a.b.gxy];
        }
        if (active) {
            auto L = [&](int f, int di, int dj, int dk) -> double {
                const int slot = dk < 0 ? (int)sm : (dk > 0 ? (int)sn : (int)sc);
                return lds[(f * 3 + slot) * S::N + c + dj * S::LX + di];
            };
            auto F = [&](int di, int dj, int dk) {
                if constexpr (SHIFT) return L(0, di, dj, dk) - s;
                else return L(0, di, dj, dk);
            };
            project_cell<1>(a, i, j, k, F, L(1, 0, 0, 0), L(2, 0, 0, 0), wsc, pc, L(1, -1, 0, 0), L(2, 0, -1, 0),
                            wsm, md, mv, bad);
        }
        wsm = wsc;
        wsc = wnext;
        pc = pnext;
        __syncthreads();
    }

    __device__ __forceinline__ static void run(const ProjArgs<R> &a, R *lds)
    {
        const Box &b = a.b;
        const int bid = xcd_remap(blockIdx.x, gridDim.x);
        const int bx = bid % a.nbx, rest = bid / a.nbx;
        const int by = rest % a.nby, bz = rest / a.nby;
        const int tx = threadIdx.x % S::TX, ty = threadIdx.x / S::TX;
        const int i0 = bx * S::TX + 1, j0 = by * TY + 1, k0 = bz * KB + 1;
        const int kend = min(k0 + KB - 1, b.nz);
        const int i = i0 + tx, j = j0 + ty;
        const bool active = i <= b.nx && j <= b.ny;
        const int c = (ty + 1) * S::LX + tx + 1;
        const double s = SHIFT ? a.shift[0] : 0.0;
        double md = -HUGE_VAL, mv = 0.0, bad = 0.0;
        const R *src[3] = {a.phi, a.us, a.vs};
        const TileIdx<S> t = tile_index<S>(b, i0, j0);
        TileRegs<S, 3, R> r;
        tile_fetch<S, 3>(r, src, t, (long)(k0 - 1) * b.gxy);
        tile_commit<S, 3>(lds, IC<0>{}, r, t);
        tile_fetch<S, 3>(r, src, t, (long)k0 * b.gxy);
        tile_commit<S, 3>(lds, IC<1>{}, r, t);
        tile_fetch<S, 3>(r, src, t, (long)(k0 + 1) * b.gxy);
        double wsm = 0.0, wsc = 0.0, pc = 0.0;
        if (active) {
            wsm = a.ws[b.g(i, j, k0 - 1)];
            wsc = a.ws[b.g(i, j, k0)];
            pc = a.p[b.g(i, j, k0)];
        }
        if constexpr (UNROLL == 3) {
            for (int k = k0; k <= kend; k += 3) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, s, wsm, wsc, pc, md, mv, bad, IC<0>{}, IC<1>{},
                      IC<2>{});
                if (k + 1 <= kend)
                    plane(a, lds, r, t, src, i, j, k + 1, kend, active, c, s, wsm, wsc, pc, md, mv, bad, IC<1>{},
                          IC<2>{}, IC<0>{});
                if (k + 2 <= kend)
                    plane(a, lds, r, t, src, i, j, k + 2, kend, active, c, s, wsm, wsc, pc, md, mv, bad, IC<2>{},
                          IC<0>{}, IC<1>{});
            }
        } else {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k++) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, s, wsm, wsc, pc, md, mv, bad, sm, sc, sn);
                const int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
            }
        }
        block_max3<NT>(md, mv, bad, a.part, bid);
    }
};

template <class R, int SHIFT, int TY, int KB, int UNROLL>
__global__ __launch_bounds__(64 * TY) void project_march_kernel(ProjArgs<R> a)
{
    __shared__ R lds[3 * 3 * TileShape<TY>::N];
    ProjectMarch<R, SHIFT, TY, KB, UNROLL>::run(a, lds);
}

}  // namespace fen
