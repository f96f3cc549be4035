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
// unrolled by three so every ring slot is a compile-time LDS offset.  With
// PF == 2 two register sets alternate, so planes k+2 and k+3 are in flight
// together (UNROLL == 2, ring slots at run time): twice the bytes in flight
// per wave, which FP32 fields need to cover the same HBM latency.
#pragma once

#include <type_traits>

#include "stencil_cells.hpp"

namespace fen {

template <int TY, int TX_ = 64> struct TileShape {
    static constexpr int TX = TX_, LX = TX + 2, LY = TY + 2, N = LX * LY, NT = TX * TY;
    static constexpr int PER = (N + NT - 1) / NT;
};

// Per-thread tile elements.  Loads and LDS writes are unconditional: an
// element past the tile reads a valid address (the tile corner) and writes a
// scratch row after the ring, and halo cells past the grid are clamped onto
// it; neither is ever read by an active cell.
template <class S> struct TileIdx {
    int off[S::PER];   // element offset in the ghosted array at k = 0
};

template <class S> __device__ __forceinline__ TileIdx<S> tile_index(const Box &b, int i0, int j0)
{
    TileIdx<S> t;
#pragma unroll
    for (int q = 0; q < S::PER; q++) {
        const int e = threadIdx.x + q * S::NT;
        const int ri = e % S::LX, rj = e / S::LX;
        int gi = min(i0 - 1 + ri, b.nx + 1), gj = min(j0 - 1 + rj, b.ny + 1);
        if (e >= S::N) { gi = i0 - 1; gj = j0 - 1; }
        t.off[q] = gi + (int)b.gx * gj;
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
    for (int f = 0; f < NF; f++) {
        const R *plane = src[f] + kofs;
#pragma unroll
        for (int q = 0; q < S::PER; q++) r.v[f][q] = plane[t.off[q]];
    }
}

// LDS layout: [field][slot 0..2][S::N], then a scratch row of S::NT.
template <class S, int NF> constexpr int ring_size() { return NF * 3 * S::N + S::NT; }

template <class S, int NF, class R, class SL>
__device__ __forceinline__ void tile_commit(R *lds, SL slot, const TileRegs<S, NF, R> &r, const TileIdx<S> &t)
{
    (void)t;
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int q = 0; q < S::PER; q++) {
            const int e = threadIdx.x + q * S::NT;
            int li = (f * 3 + (int)slot) * S::N + e;
            if (q == S::PER - 1 && e >= S::N) li = NF * 3 * S::N + (int)threadIdx.x;
            lds[li] = r.v[f][q];
        }
}

template <int V> using IC = std::integral_constant<int, V>;

// ----------------------------------------------------------------- momentum
template <class R, int SRC, int TY, int KB, int UNROLL, int PF = 1, int TX = 64>
struct MomentumMarch {
    static_assert((PF == 1 && (UNROLL == 1 || UNROLL == 3)) || (PF == 2 && UNROLL == 2), "launch shape");
    using S = TileShape<TY, TX>;
    static constexpr int NT = S::NT;

    template <class SM, class SC, class SN>
    __device__ __forceinline__ static void plane(const MomArgs<R> &a, R *lds, TileRegs<S, 4, R> &r,
                                                 const TileIdx<S> &t, const R *const *src, int i, int j, int k,
                                                 int kend, bool active, int c, R (&o)[3], bool xedge, SM sm, SC sc,
                                                 SN sn)
    {
        tile_commit<S, 4>(lds, sn, r, t);   // plane k+1
        __syncthreads();
        if (k + 1 + PF <= kend + 1) tile_fetch<S, 4>(r, src, t, (long)(k + 1 + PF) * a.b.gxy);
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
            auto img = [&](R *f, R v) { march_images(f, a.b, a.img, i, j, k, v, xedge); };
            momentum_cell<1, SRC>(a, i, j, k, at(0), at(1), at(2), at(3), o[0], o[1], o[2], img);
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
        // a wave is one j row: ty is wave-uniform, so j-dependent branches are scalar
        const int tx = threadIdx.x % S::TX, ty = __builtin_amdgcn_readfirstlane(threadIdx.x / S::TX);
        const bool xedge = bx == 0 || bx == a.nbx - 1;
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
        TileRegs<S, 4, R> r2;
        if constexpr (PF == 2) {
            if (k0 + 2 <= kend + 1) tile_fetch<S, 4>(r2, src, t, (long)(k0 + 2) * b.gxy);
        }
        R o[3] = {R(0), R(0), R(0)};
        if (active) {
            const long q = b.q(i, j, k0);
            o[0] = a.dvo0[q]; o[1] = a.dvo1[q]; o[2] = a.dvo2[q];
        }
        if constexpr (PF == 2) {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k += 2) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, o, xedge, sm, sc, sn);
                int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
                if (k + 1 <= kend) {
                    plane(a, lds, r2, t, src, i, j, k + 1, kend, active, c, o, xedge, sm, sc, sn);
                    tmp = sm;
                    sm = sc; sc = sn; sn = tmp;
                }
            }
        } else if constexpr (UNROLL == 3) {
            for (int k = k0; k <= kend; k += 3) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, o, xedge, IC<0>{}, IC<1>{}, IC<2>{});
                if (k + 1 <= kend) plane(a, lds, r, t, src, i, j, k + 1, kend, active, c, o, xedge, IC<1>{}, IC<2>{}, IC<0>{});
                if (k + 2 <= kend) plane(a, lds, r, t, src, i, j, k + 2, kend, active, c, o, xedge, IC<2>{}, IC<0>{}, IC<1>{});
            }
        } else {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k++) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, o, xedge, sm, sc, sn);
                const int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
            }
        }
    }
};

// OCC: minimum waves per SIMD the register allocation must allow (1: no limit).
template <class R, int SRC, int TY, int KB, int UNROLL, int PF = 1, int OCC = 1, int TX = 64>
__global__ __launch_bounds__(TX * TY, OCC) void momentum_march_kernel(MomArgs<R> a)
{
    __shared__ R lds[ring_size<TileShape<TY, TX>, 4>()];
    MomentumMarch<R, SRC, TY, KB, UNROLL, PF, TX>::run(a, lds);
}

// --------------------------------------------------------------- projection
// Ring fields: phi, u*, v* (u* and v* need the cell below in i and j for the
// checks); w* and p are prefetched one plane ahead, w* one plane below is
// carried in a register.
// ST == 0: the checks of the projection only; nothing is stored and p is not read.
template <class R, int SHIFT, int TY, int KB, int UNROLL, int PF = 1, int TX = 64, int ST = 1>
struct ProjectMarch {
    static_assert((PF == 1 && (UNROLL == 1 || UNROLL == 3)) || (PF == 2 && UNROLL == 2), "launch shape");
    using S = TileShape<TY, TX>;
    static constexpr int NT = S::NT;

    template <class SM, class SC, class SN>
    __device__ __forceinline__ static void plane(const ProjArgs<R> &a, R *lds, TileRegs<S, 3, R> &r,
                                                 const TileIdx<S> &t, const R *const *src, int i, int j, int k,
                                                 int kend, bool active, int c, double s, double &wsm, double &wsc,
                                                 double &pc, double &md, double &mv, double &bad, bool xedge, SM sm,
                                                 SC sc, SN sn)
    {
        tile_commit<S, 3>(lds, sn, r, t);
        __syncthreads();
        if (k + 1 + PF <= kend + 1) tile_fetch<S, 3>(r, src, t, (long)(k + 1 + PF) * a.b.gxy);
        double wnext = 0.0, pnext = 0.0;
        const long g = a.b.g(i, j, k);
        if (active && k + 1 <= kend) {
            wnext = a.ws[g + a.b.gxy];
            if constexpr (ST) pnext = a.p[g + a.b.gxy];
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
            auto img = [&](R *f, R v) { march_images(f, a.b, a.img, i, j, k, v, xedge); };
            project_cell<1, R, decltype(F), decltype(img), ST>(a, i, j, k, F, L(1, 0, 0, 0), L(2, 0, 0, 0), wsc, pc,
                                                              L(1, -1, 0, 0), L(2, 0, -1, 0), wsm, md, mv, bad, img);
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
        // a wave is one j row: ty is wave-uniform, so j-dependent branches are scalar
        const int tx = threadIdx.x % S::TX, ty = __builtin_amdgcn_readfirstlane(threadIdx.x / S::TX);
        const bool xedge = bx == 0 || bx == a.nbx - 1;
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
        TileRegs<S, 3, R> r2;
        if constexpr (PF == 2) {
            if (k0 + 2 <= kend + 1) tile_fetch<S, 3>(r2, src, t, (long)(k0 + 2) * b.gxy);
        }
        double wsm = 0.0, wsc = 0.0, pc = 0.0;
        if (active) {
            wsm = a.ws[b.g(i, j, k0 - 1)];
            wsc = a.ws[b.g(i, j, k0)];
            if constexpr (ST) pc = a.p[b.g(i, j, k0)];
        }
        if constexpr (PF == 2) {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k += 2) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, sm, sc, sn);
                int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
                if (k + 1 <= kend) {
                    plane(a, lds, r2, t, src, i, j, k + 1, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, sm, sc,
                          sn);
                    tmp = sm;
                    sm = sc; sc = sn; sn = tmp;
                }
            }
        } else if constexpr (UNROLL == 3) {
            for (int k = k0; k <= kend; k += 3) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, IC<0>{}, IC<1>{},
                      IC<2>{});
                if (k + 1 <= kend)
                    plane(a, lds, r, t, src, i, j, k + 1, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, IC<1>{},
                          IC<2>{}, IC<0>{});
                if (k + 2 <= kend)
                    plane(a, lds, r, t, src, i, j, k + 2, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, IC<2>{},
                          IC<0>{}, IC<1>{});
            }
        } else {
            int sm = 0, sc = 1, sn = 2;
            for (int k = k0; k <= kend; k++) {
                plane(a, lds, r, t, src, i, j, k, kend, active, c, s, wsm, wsc, pc, md, mv, bad, xedge, sm, sc, sn);
                const int tmp = sm;
                sm = sc; sc = sn; sn = tmp;
            }
        }
        block_max3<NT>(md, mv, bad, a.part, bid);
    }
};

template <class R, int SHIFT, int TY, int KB, int UNROLL, int PF = 1, int OCC = 1, int TX = 64, int ST = 1>
__global__ __launch_bounds__(TX * TY, OCC) void project_march_kernel(ProjArgs<R> a)
{
    __shared__ R lds[ring_size<TileShape<TY, TX>, 3>()];
    ProjectMarch<R, SHIFT, TY, KB, UNROLL, PF, TX, ST>::run(a, lds);
}

// ------------------------------------------------- projection + momentum
// Step n's projection and pressure update fused with step n+1's momentum,
// for a grid periodic in every direction.  The state (u, v, w, p) after
// step n is formed tile by tile from u*, v*, w*, p and phi of step n with
// the projection's arithmetic (each value rounded once to R), placed in the
// momentum's 4-field LDS ring and advanced to u*, v*, w* of step n+1; only
// p of step n and the momentum's outputs are stored.  Step n's checks are
// not formed: step n+1 ends pending, and its checks are the ones a reader
// asks for.  Every
// ghost position of the state is the periodic image of an interior cell, so
// tile cells outside the interior are formed at the wrapped interior index:
// the same values the unfused projection writes and its images copy.
// u*, v*, w* and p are read at the cell, phi at the cell and one cell above
// along each axis (phi's face ghosts are its periodic images).
// ZW: the z faces are no-slip walls instead (channel: x, y periodic).  The
// state's ghost planes k = 0 and nz+1 then follow the ghost fill's z rules
// applied to the projected first and last planes: u, v = 2 bc - f (the bc
// planes are indexed at the cell's own, unwrapped (i, j)), w = bc (and the
// last interior w is overwritten with bc), p = the neighbouring interior
// value; a pending mean shift of phi is subtracted as the projection does.
template <class R> struct PMArgs {
    MomArgs<R> m;                  // momentum of step n+1 (m.u, m.v, m.w, m.p unused)
    const R *__restrict__ us;      // u*, v*, w* of step n
    const R *__restrict__ vs;
    const R *__restrict__ ws;
    const R *__restrict__ phi;
    const R *__restrict__ p_in;    // p before step n
    R *__restrict__ p_out;         // p after step n (interior)
    const double *shift;           // pending mean of phi (ZW), or null
    const R *zlo[3], *zhi[3];      // ZW: bc planes of u, v, w at the front and back faces
    double pid, pdt, prhof, pirhof;
    int prpow2;
    int kb;                        // planes marched per workgroup
    // ZW == 2 (general walls): non-periodic axes (bit 0 x, 1 y, 2 z); per
    // state field u, v, w, p the bc type of face f in bits 2f..2f+1 and the
    // location (0 centre, 1 x, 2 y, 3 z face) in bits 12..13; and the bc
    // value planes, 6 per field in that order, in device memory
    int wall;
    int gtype[4];
    const R *const *gplane;
    R *wbuf;   // ZW == 2: the state at the wall positions (pm_wall_layout)
    R *save[3];   // SAVE: u, v, w of step n (ghosted, with periodic images)
};

// The general-wall form (ZW == 2).  Along one wall axis a ghosted position
// is the low ghost (0), the high ghost (n + 1) or the high face's
// overwritten normal (n, f(N) = bc): pm_wall_pos gives 0, 1, 2 for these
// and -1 elsewhere.
__device__ __forceinline__ int pm_wall_pos(int g, int n)
{
    return g == 0 ? 0 : (g == n + 1 ? 1 : (g == n ? 2 : -1));
}

// The state at every wall position is formed before the march by
// wall_state_kernel into a face buffer, one block per field of
// pm_wall_layout().per_field values: x section [nz+2][ny+2][3] (kk, gj,
// position px fastest: px = 1 and 2 read the same source cell, so a wave
// reads each of its lines once), y section [3][nz+2][nx+2] (py, kk, gi),
// z section [3][ny+2][nx+2] (pz, gj, gi).  Positions on several walls
// (edges, corners) hold the same value in each section that has them.
struct PMWallLayout {
    long xs, ys, zs;   // section offsets
    long xp, yp, zp;   // words per position plane
    long per_field;
};
__host__ __device__ __forceinline__ PMWallLayout pm_wall_layout(const Box &b)
{
    PMWallLayout w;
    w.xp = (long)(b.ny + 2) * (b.nz + 2);
    w.yp = (long)(b.nx + 2) * (b.nz + 2);
    w.zp = (long)(b.nx + 2) * (b.ny + 2);
    w.xs = 0;
    w.ys = 3 * w.xp;
    w.zs = w.ys + 3 * w.yp;
    w.per_field = w.zs + 3 * w.zp;
    return w;
}

// The ghost value of one state field at ghosted position (gi, gj, kk),
// given v = the field's projected value at the interior cell the ghost fill
// reads, (ix, jy, kz) per axis: the ghost kernel's x -> y -> z composition
// (S1, S2, S3) applied to that value, each level with the value plane
// indexed at that level's coordinates, and the high-face normal overwrite
// f(N) = bc.  Periodic axes only move the index.  gt: packed types and
// location (PMArgs::gtype), pl: the field's 6 planes.
template <class R>
__device__ __forceinline__ R pm_ghost_chain(int gt, const R *const *pl, const Box &b, int wall, int gi, int gj,
                                            int kk, int jy, int kz, R v)
{
    const R two = R(2.0);
    const int loc = (gt >> 12) & 3;   // 1 x, 2 y, 3 z, 0 centre
    auto ty = [&](int f) { return (gt >> (2 * f)) & 3; };
    if (wall & 1) {   // S1(gi, jy, kz)
        const long pj = jy + (long)(b.ny + 2) * kz;
        if (gi == 0) {
            if (ty(0) == 1) v = loc == 1 ? pl[0][pj] : two * pl[0][pj] - v;
        } else if (gi == b.nx + 1) {
            if (ty(1) == 1) v = loc == 1 ? pl[1][pj] : two * pl[1][pj] - v;
        } else if (gi == b.nx && ty(1) == 1 && loc == 1) {
            v = pl[1][pj];
        }
    }
    if (wall & 2) {   // S2(gi, gj, kz)
        const long pi = gi + (long)(b.nx + 2) * kz;
        if (gj == 0) {
            if (ty(2) == 1) v = loc == 2 ? pl[2][pi] : two * pl[2][pi] - v;
        } else if (gj == b.ny + 1) {
            if (ty(3) == 1) v = loc == 2 ? pl[3][pi] : two * pl[3][pi] - v;
        } else if (gj == b.ny && ty(3) == 1 && loc == 2) {
            v = pl[3][pi];
        }
    }
    if (wall & 4) {   // S3(gi, gj, kk)
        const long pij = gi + (long)(b.nx + 2) * gj;
        if (kk == 0) {
            if (ty(4) == 1) v = loc == 3 ? pl[4][pij] : two * pl[4][pij] - v;
        } else if (kk == b.nz + 1) {
            if (ty(5) == 1) v = loc == 3 ? pl[5][pij] : two * pl[5][pij] - v;
        } else if (kk == b.nz && ty(5) == 1 && loc == 3) {
            v = pl[5][pij];
        }
    }
    return v;
}

// Source index of a ghosted coordinate along one axis: periodic wraps,
// a wall axis reads the adjacent interior cell.
__device__ __forceinline__ int pm_src(int g, int n, bool wall)
{
    if (g == 0) return wall ? 1 : n;
    if (g == n + 1) return wall ? n : 1;
    return g;
}

// SAVE (periodic grids): also store u, v, w of step n and their periodic
// images into a.save, for callers that keep every state (the adjoint's
// recomputed segments).
template <class R, int SRC, int TY, int ZW, int TX = 64, int SAVE = 0>
struct ProjectMomentumMarch {
    using S = TileShape<TY, TX>;
    static constexpr int NT = S::NT, NRAW = 8;
    struct Raw {
        R v[NRAW][S::PER];
        int wl;   // ZW == 2: bit q set when element q took its state from the face buffer
    };
    // ZW == 2: per tile element, its face buffer offset at plane 0 and the
    // stride per plane when it sits on an x or y wall position (ofs < 0
    // otherwise), and its offset in a z position plane
    struct WallIdx {
        int ofs[S::PER], str[S::PER], zo[S::PER];
    };
    struct PlaneIdx {   // ZW: bc plane index i + (nx+2) j of each tile element
        int v[S::PER];
    };

    __device__ __forceinline__ static PlaneIdx plane_index(const Box &b, int i0, int j0)
    {
        PlaneIdx pi{};
        if constexpr (ZW == 1) {
#pragma unroll
            for (int q = 0; q < S::PER; q++) {
                const int e = threadIdx.x + q * S::NT;
                const int ri = e % S::LX, rj = e / S::LX;
                int gi = min(i0 - 1 + ri, b.nx + 1), gj = min(j0 - 1 + rj, b.ny + 1);
                if (e >= S::N) { gi = i0; gj = j0; }
                pi.v[q] = gi + (b.nx + 2) * gj;
            }
        }
        return pi;
    }

    // Tile element offsets at k = 0 with every ghost index wrapped to its
    // periodic interior image (elements past the grid are clamped first).
    __device__ __forceinline__ static TileIdx<S> index(const Box &b, int i0, int j0, int wall)
    {
        TileIdx<S> t;
#pragma unroll
        for (int q = 0; q < S::PER; q++) {
            const int e = threadIdx.x + q * S::NT;
            const int ri = e % S::LX, rj = e / S::LX;
            int gi = min(i0 - 1 + ri, b.nx + 1), gj = min(j0 - 1 + rj, b.ny + 1);
            if (e >= S::N) { gi = i0; gj = j0; }
            if constexpr (ZW == 2) {
                gi = pm_src(gi, b.nx, wall & 1);
                gj = pm_src(gj, b.ny, wall & 2);
            } else {
                gi = gi == 0 ? b.nx : (gi == b.nx + 1 ? 1 : gi);
                gj = gj == 0 ? b.ny : (gj == b.ny + 1 ? 1 : gj);
            }
            t.off[q] = gi + (int)b.gx * gj;
        }
        return t;
    }

    __device__ __forceinline__ static void fetch(Raw &r, const PMArgs<R> &a, const TileIdx<S> &t,
                                                 const WallIdx &wi, int kk)
    {
        const Box &b = a.m.b;
        // the interior plane the state at plane kk is formed from
        const int kw = ZW == 1   ? (kk == 0 ? 1 : (kk == b.nz + 1 ? b.nz : kk))
                       : ZW == 2 ? pm_src(kk, b.nz, a.wall & 4)
                                 : (kk == 0 ? b.nz : (kk == b.nz + 1 ? 1 : kk));
        const long kofs = (long)kw * b.gxy, sj = b.gx, sk = b.gxy;
        // ZW == 2: elements on a wall position load u, v, w, p of the face
        // buffer instead (a z wall plane: all of them)
        [[maybe_unused]] int pz = -1;
        [[maybe_unused]] PMWallLayout L{};
        if constexpr (ZW == 2) {
            pz = (a.wall & 4) ? pm_wall_pos(kk, b.nz) : -1;
            L = pm_wall_layout(b);
            r.wl = 0;
        }
#pragma unroll
        for (int q = 0; q < S::PER; q++) {
            const long o = t.off[q] + kofs;
            const R *su = a.us + o, *sv = a.vs + o, *sw = a.ws + o, *sp = a.p_in + o;
            if constexpr (ZW == 2) {
                const bool wl = pz >= 0 || wi.ofs[q] >= 0;
                const long ow = pz >= 0 ? L.zs + pz * L.zp + wi.zo[q] : wi.ofs[q] + (long)wi.str[q] * kk;
                if (wl) {
                    const R *w0 = a.wbuf + ow;
                    su = w0; sv = w0 + L.per_field; sw = w0 + 2 * L.per_field; sp = w0 + 3 * L.per_field;
                    r.wl |= 1 << q;
                }
            }
            r.v[0][q] = *su;
            r.v[1][q] = *sv;
            r.v[2][q] = *sw;
            r.v[3][q] = *sp;
            r.v[4][q] = a.phi[o];
            r.v[5][q] = a.phi[o + 1];
            r.v[6][q] = a.phi[o + sj];
            r.v[7][q] = a.phi[o + sk];
        }
    }

    // The state at tile plane kk (project_cell's arithmetic, then the z wall
    // rules when ZW) into ring slot `slot`.
    template <class SL>
    __device__ __forceinline__ static void commit(R *lds, SL slot, const Raw &r, const PMArgs<R> &a,
                                                  const PlaneIdx &pi, int kk, double sh)
    {
        auto dr = [&](double x) { return over_rho(x, a.prhof, a.pirhof, a.prpow2); };
        const int nz = a.m.b.nz;
#pragma unroll
        for (int q = 0; q < S::PER; q++) {
            const int e = threadIdx.x + q * S::NT;
            double f0 = (double)r.v[4][q], fx = (double)r.v[5][q], fy = (double)r.v[6][q], fz = (double)r.v[7][q];
            if constexpr (ZW == 1) {
                f0 = f0 - sh; fx = fx - sh; fy = fy - sh; fz = fz - sh;
            }
            R un = (R)((double)r.v[0][q] - dr((fx - f0) * a.pid * a.pdt));
            R vn = (R)((double)r.v[1][q] - dr((fy - f0) * a.pid * a.pdt));
            R wn = (R)((double)r.v[2][q] - dr((fz - f0) * a.pid * a.pdt));
            R pn = (R)((double)r.v[3][q] + f0);
            if constexpr (ZW == 2) {
                if (r.wl & (1 << q)) {   // the face buffer's final state
                    un = r.v[0][q]; vn = r.v[1][q]; wn = r.v[2][q]; pn = r.v[3][q];
                }
            }
            if constexpr (ZW == 1) {
                const R two = R(2.0);
                if (kk == 0) {
                    un = two * a.zlo[0][pi.v[q]] - un;
                    vn = two * a.zlo[1][pi.v[q]] - vn;
                    wn = a.zlo[2][pi.v[q]];
                } else if (kk >= nz) {
                    if (kk > nz) {
                        un = two * a.zhi[0][pi.v[q]] - un;
                        vn = two * a.zhi[1][pi.v[q]] - vn;
                    }
                    wn = a.zhi[2][pi.v[q]];
                }
            }
            const R vals[4] = {un, vn, wn, pn};
#pragma unroll
            for (int f = 0; f < 4; f++) {
                int li = (f * 3 + (int)slot) * S::N + e;
                if (q == S::PER - 1 && e >= S::N) li = 4 * 3 * S::N + (int)threadIdx.x;
                lds[li] = vals[f];
            }
        }
    }

    // ZW == 2: the face buffer offsets of this thread's tile elements.
    __device__ __forceinline__ static WallIdx wall_index(const PMArgs<R> &a, int i0, int j0)
    {
        WallIdx wi{};
        if constexpr (ZW == 2) {
            const Box &b = a.m.b;
            const PMWallLayout L = pm_wall_layout(b);
#pragma unroll
            for (int q = 0; q < S::PER; q++) {
                const int e = threadIdx.x + q * S::NT;
                const int ri = e % S::LX, rj = e / S::LX;
                int gi = min(i0 - 1 + ri, b.nx + 1), gj = min(j0 - 1 + rj, b.ny + 1);
                if (e >= S::N) { gi = i0; gj = j0; }
                const int px = (a.wall & 1) ? pm_wall_pos(gi, b.nx) : -1;
                const int py = (a.wall & 2) ? pm_wall_pos(gj, b.ny) : -1;
                wi.zo[q] = gi + (b.nx + 2) * gj;
                if (px >= 0) {
                    wi.ofs[q] = (int)(L.xs + px + 3 * gj);
                    wi.str[q] = 3 * (b.ny + 2);
                } else if (py >= 0) {
                    wi.ofs[q] = (int)(L.ys + py * L.yp + gi);
                    wi.str[q] = b.nx + 2;
                } else {
                    wi.ofs[q] = -1;
                    wi.str[q] = 0;
                }
            }
        }
        return wi;
    }

    template <class SM, class SC, class SN>
    __device__ __forceinline__ static void plane(const PMArgs<R> &a, R *lds, Raw &r, const TileIdx<S> &t,
                                                 const PlaneIdx &pi, double sh, int i, int j, int k, int kend,
                                                 bool active, int c, R (&o)[3], bool xedge, SM sm, SC sc, SN sn,
                                                 const WallIdx &wi)
    {
        const MomArgs<R> &m = a.m;
        commit(lds, sn, r, a, pi, k + 1, sh);   // state at plane k+1
        __syncthreads();
        if (k + 2 <= kend + 1) {
            fetch(r, a, t, wi, k + 2);
        }
        R on[3] = {R(0), R(0), R(0)};
        const long q = active ? m.b.q(i, j, k) : 0;
        const long qs = (long)m.b.nx * m.b.ny;
        if (active && k + 1 <= kend) {
            on[0] = load_stream(&m.dvo0[q + qs]);
            on[1] = load_stream(&m.dvo1[q + qs]);
            on[2] = load_stream(&m.dvo2[q + qs]);
        }
        if (active) {
            // p of step n; its checks are not formed here: the step ends
            // pending, and a read of the scalars runs the checks-only pass
            store_stream(&a.p_out[m.b.g(i, j, k)], lds[(3 * 3 + (int)sc) * S::N + c]);
            if constexpr (SAVE) {
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    const R st = lds[(f * 3 + (int)sc) * S::N + c];
                    a.save[f][m.b.g(i, j, k)] = st;
                    march_images(a.save[f], m.b, 7, i, j, k, st, xedge);
                }
            }
            // momentum of step n+1
            auto at = [&](int f) {
                return [=](int di, int dj, int dk) {
                    const int slotv = dk < 0 ? (int)sm : (dk > 0 ? (int)sn : (int)sc);
                    return lds[(f * 3 + slotv) * S::N + c + dj * S::LX + di];
                };
            };
            auto img = [&](R *f, R v) { march_images(f, m.b, m.img, i, j, k, v, xedge); };
            momentum_cell<1, SRC>(m, i, j, k, at(0), at(1), at(2), at(3), o[0], o[1], o[2], img);
        }
        o[0] = on[0]; o[1] = on[1]; o[2] = on[2];
        __syncthreads();
    }

    __device__ __forceinline__ static void run(const PMArgs<R> &a, R *lds)
    {
        const MomArgs<R> &m = a.m;
        const Box &b = m.b;
        const int bid = xcd_remap(blockIdx.x, gridDim.x);
        const int bx = bid % m.nbx, rest = bid / m.nbx;
        const int by = rest % m.nby, bz = rest / m.nby;
        const int tx = threadIdx.x % S::TX, ty = __builtin_amdgcn_readfirstlane(threadIdx.x / S::TX);
        const bool xedge = bx == 0 || bx == m.nbx - 1;
        const int i0 = bx * S::TX + 1, j0 = by * TY + 1, k0 = bz * a.kb + 1;
        const int kend = min(k0 + a.kb - 1, b.nz);
        const int i = i0 + tx, j = j0 + ty;
        const bool active = i <= b.nx && j <= b.ny;
        const int c = (ty + 1) * S::LX + tx + 1;
        const TileIdx<S> t = index(b, i0, j0, a.wall);
        const PlaneIdx pi = plane_index(b, i0, j0);
        const double sh = (ZW == 1 && a.shift) ? a.shift[0] : 0.0;
        Raw r;
        const WallIdx wi = wall_index(a, i0, j0);
        fetch(r, a, t, wi, k0 - 1);
        commit(lds, IC<0>{}, r, a, pi, k0 - 1, sh);
        fetch(r, a, t, wi, k0);
        commit(lds, IC<1>{}, r, a, pi, k0, sh);
        fetch(r, a, t, wi, k0 + 1);
        R o[3] = {R(0), R(0), R(0)};
        if (active) {
            const long q = b.q(i, j, k0);
            o[0] = load_stream(&m.dvo0[q]); o[1] = load_stream(&m.dvo1[q]); o[2] = load_stream(&m.dvo2[q]);
        }
        int sm = 0, sc = 1, sn = 2;
        for (int k = k0; k <= kend; k++) {
            plane(a, lds, r, t, pi, sh, i, j, k, kend, active, c, o, xedge, sm, sc, sn, wi);
            const int tmp = sm;
            sm = sc; sc = sn; sn = tmp;
        }
    }
};

template <class R, int SRC, int ZW, int TY, int OCC = 1, int TX = 64, int SAVE = 0>
__global__ __launch_bounds__(TX * TY, OCC) void project_momentum_kernel(PMArgs<R> a)
{
    __shared__ R lds[ring_size<TileShape<TY, TX>, 4>()];
    ProjectMomentumMarch<R, SRC, TY, ZW, TX, SAVE>::run(a, lds);
}

// The general-wall form's face buffer: u, v, w, p of step n at every wall
// position (pm_wall_layout), one thread per position.  Each is the
// projection of the interior cell the ghost fill reads (commit()'s
// arithmetic, so interior positions such as the overwritten face n match
// the march's own values bit for bit), then the ghost fill's rule chain.
template <class R>
__global__ __launch_bounds__(256) void wall_state_kernel(PMArgs<R> a)
{
    const Box &b = a.m.b;
    const PMWallLayout L = pm_wall_layout(b);
    const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= L.per_field) return;
    auto coord = [](int pos, int n) { return pos == 0 ? 0 : (pos == 1 ? n + 1 : n); };
    int gi, gj, kk;
    if (idx < L.ys) {
        if (!(a.wall & 1)) return;
        const long pos = idx % 3, rem = idx / 3;
        gi = coord((int)pos, b.nx); gj = (int)(rem % (b.ny + 2)); kk = (int)(rem / (b.ny + 2));
    } else if (idx < L.zs) {
        if (!(a.wall & 2)) return;
        const long pos = (idx - L.ys) / L.yp, rem = (idx - L.ys) % L.yp;
        gj = coord((int)pos, b.ny); gi = (int)(rem % (b.nx + 2)); kk = (int)(rem / (b.nx + 2));
    } else {
        if (!(a.wall & 4)) return;
        const long pos = (idx - L.zs) / L.zp, rem = (idx - L.zs) % L.zp;
        kk = coord((int)pos, b.nz); gi = (int)(rem % (b.nx + 2)); gj = (int)(rem / (b.nx + 2));
    }
    const int ix = pm_src(gi, b.nx, a.wall & 1), jy = pm_src(gj, b.ny, a.wall & 2), kz = pm_src(kk, b.nz, a.wall & 4);
    const long o = ix + (long)b.gx * jy + (long)b.gxy * kz;
    auto dr = [&](double x) { return over_rho(x, a.prhof, a.pirhof, a.prpow2); };
    const double f0 = (double)a.phi[o], fx = (double)a.phi[o + 1], fy = (double)a.phi[o + b.gx],
                 fz = (double)a.phi[o + b.gxy];
    R st[4];
    st[0] = (R)((double)a.us[o] - dr((fx - f0) * a.pid * a.pdt));
    st[1] = (R)((double)a.vs[o] - dr((fy - f0) * a.pid * a.pdt));
    st[2] = (R)((double)a.ws[o] - dr((fz - f0) * a.pid * a.pdt));
    st[3] = (R)((double)a.p_in[o] + f0);
#pragma unroll
    for (int f = 0; f < 4; f++)
        a.wbuf[f * L.per_field + idx] = pm_ghost_chain(a.gtype[f], a.gplane + 6 * f, b, a.wall, gi, gj, kk, jy, kz, st[f]);
}


}  // namespace fen
