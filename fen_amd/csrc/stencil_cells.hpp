// stencil_cells.hpp -- per-cell arithmetic of the momentum and projection
// kernels, shared by every launch shape.  Expressions keep the reference's
// evaluation order (Fortran left to right, explicit parentheses); the library
// is built with -ffp-contract=off so they round exactly like the CPU.
//
// R is the element type of the fields.  The momentum stencils compute in R
// (FP32 stencils in the FP32 form); the projection, its pressure update and
// the checks compute in FP64 from the stored values and round once to R.
#pragma once

#include "fen_internal.hpp"

namespace fen {

// x / rhof.  When rhof is a power of two its reciprocal is exact and
// x * (1/rhof) is the same correctly rounded value as x / rhof, bit for bit;
// the multiply replaces a ~10-instruction IEEE divide.
template <class T> __device__ __forceinline__ T over_rho(T x, T rhof, T irhof, int pow2)
{
    return pow2 ? x * irhof : x / rhof;
}

// ---------------------------------------------------------------------------
// K1: momentum.  dv = advection (src/navier_stokes.f90:261-353) + mu lap(v)/rhof
// (:384-404, src/fields.f90:298-343) + S/rhof (:245-255); then
// RHS = -grad(p)/rhof + A dv + B dv_o + g and v* = v + dt RHS (:157-198);
// dv_o = dv (:201-205).  rhof == density exactly for a constant density
// field (0.5*(rho+rho) == rho), mu == viscosity.
// ---------------------------------------------------------------------------
// dv_o and p are written by one step's momentum and read by the next one's,
// with the whole Poisson solve in between: streamed (non-temporal) accesses
// keep them from displacing u*, v*, w* and the spectrum in L2 and the
// Infinity Cache, which the solve reads next (256^3 FP64: 0.833 -> 0.811 ms
// per step; FP32: 0.476 -> 0.460 ms).
template <class R> __device__ __forceinline__ void store_stream(R *p, R v) { __builtin_nontemporal_store(v, p); }
template <class R> __device__ __forceinline__ R load_stream(const R *p) { return __builtin_nontemporal_load(p); }

template <class R> struct MomArgs {
    const R *__restrict__ u;
    const R *__restrict__ v;
    const R *__restrict__ w;
    const R *__restrict__ p;
    R *__restrict__ us;
    R *__restrict__ vs;
    R *__restrict__ ws;
    R *__restrict__ dvo0;
    R *__restrict__ dvo1;
    R *__restrict__ dvo2;
    const R *__restrict__ s0;
    const R *__restrict__ s1;
    const R *__restrict__ s2;
    Box b;
    R id, id2, mu, rhof, dt, A, B, g0, g1, g2;
    R irhof;     // 1/rhof, exact when rpow2
    int rpow2;   // rhof is a power of two
    int nbx, nby, img;
};

// One cell of the momentum update.  U, V, W, P are accessors returning the
// field value at offset (di, dj, dk) from cell (i, j, k); o0..o2 are the
// old dv_o values of the cell; IMG(f, v) writes the periodic images of the
// cell's value v in field f.
template <int D3, int SRC, class R, class FU, class FV, class FW, class FP, class IW>
__device__ __forceinline__ void momentum_cell(const MomArgs<R> &a, int i, int j, int k, FU U, FV V, FW W, FP P,
                                              R o0, R o1, R o2, IW IMG)
{
    const long g = a.b.g(i, j, k);
    const long q = a.b.q(i, j, k);
    const R id = a.id, id2 = a.id2;
    const R c0 = R(0.0), c025 = R(0.25), c2 = R(2.0);
    const R u0 = U(0, 0, 0), v0 = V(0, 0, 0);

    // ---- x momentum
    R r;
    {
        const R up = U(1, 0, 0), um = U(-1, 0, 0);
        const R uuip = c025 * ((up + u0) * (up + u0));
        const R uuim = c025 * ((um + u0) * (um + u0));
        const R uvjp = ((U(0, 1, 0) + u0) * (V(1, 0, 0) + v0)) * c025;
        const R uvjm = ((u0 + U(0, -1, 0)) * (V(1, -1, 0) + V(0, -1, 0))) * c025;
        r = (c0 - (uuip - uuim) * id) - (uvjp - uvjm) * id;
        if constexpr (D3) {
            const R uwkp = ((U(0, 0, 1) + u0) * (W(1, 0, 0) + W(0, 0, 0))) * c025;
            const R uwkm = ((u0 + U(0, 0, -1)) * (W(1, 0, -1) + W(0, 0, -1))) * c025;
            r = r - (uwkp - uwkm) * id;
        }
    }
    R lu = (((U(1, 0, 0) - c2 * u0) + U(-1, 0, 0)) + ((U(0, 1, 0) - c2 * u0) + U(0, -1, 0))) * id2;
    if constexpr (D3) lu = lu + ((U(0, 0, 1) - c2 * u0) + U(0, 0, -1)) * id2;
    r = r + over_rho(a.mu * lu, a.rhof, a.irhof, a.rpow2);
    if constexpr (SRC) r = r + over_rho(a.s0[q], a.rhof, a.irhof, a.rpow2);
    else r = r + c0;
    const R dvx = r;

    // ---- y momentum
    {
        const R vuip = ((V(1, 0, 0) + v0) * (U(0, 1, 0) + u0)) * c025;
        const R vuim = ((v0 + V(-1, 0, 0)) * (U(-1, 1, 0) + U(-1, 0, 0))) * c025;
        const R vp = V(0, 1, 0), vm = V(0, -1, 0);
        const R vvjp = c025 * ((vp + v0) * (vp + v0));
        const R vvjm = c025 * ((vm + v0) * (vm + v0));
        r = (c0 - (vuip - vuim) * id) - (vvjp - vvjm) * id;
        if constexpr (D3) {
            const R vwkp = ((V(0, 0, 1) + v0) * (W(0, 1, 0) + W(0, 0, 0))) * c025;
            const R vwkm = ((v0 + V(0, 0, -1)) * (W(0, 1, -1) + W(0, 0, -1))) * c025;
            r = r - (vwkp - vwkm) * id;
        }
    }
    R lv = (((V(1, 0, 0) - c2 * v0) + V(-1, 0, 0)) + ((V(0, 1, 0) - c2 * v0) + V(0, -1, 0))) * id2;
    if constexpr (D3) lv = lv + ((V(0, 0, 1) - c2 * v0) + V(0, 0, -1)) * id2;
    r = r + over_rho(a.mu * lv, a.rhof, a.irhof, a.rpow2);
    if constexpr (SRC) r = r + over_rho(a.s1[q], a.rhof, a.irhof, a.rpow2);
    else r = r + c0;
    const R dvy = r;

    const R A = a.A, B = a.B, dt = a.dt;
    const R p0 = P(0, 0, 0);
    {
        const R gp = (P(1, 0, 0) - p0) * id;
        const R rhs = ((-over_rho(gp, a.rhof, a.irhof, a.rpow2) + A * dvx) + B * o0) + a.g0;
        const R val = u0 + dt * rhs;
        a.us[g] = val;
        IMG(a.us, val);
        store_stream(&a.dvo0[q], dvx);
    }
    {
        const R gp = (P(0, 1, 0) - p0) * id;
        const R rhs = ((-over_rho(gp, a.rhof, a.irhof, a.rpow2) + A * dvy) + B * o1) + a.g1;
        const R val = v0 + dt * rhs;
        a.vs[g] = val;
        IMG(a.vs, val);
        store_stream(&a.dvo1[q], dvy);
    }
    if constexpr (D3) {
        // ---- z momentum
        const R w0 = W(0, 0, 0);
        const R wuip = ((w0 + W(1, 0, 0)) * (u0 + U(0, 0, 1))) * c025;
        const R wuim = ((w0 + W(-1, 0, 0)) * (U(-1, 0, 0) + U(-1, 0, 1))) * c025;
        const R wvjp = ((w0 + W(0, 1, 0)) * (v0 + V(0, 0, 1))) * c025;
        const R wvjm = ((w0 + W(0, -1, 0)) * (V(0, -1, 0) + V(0, -1, 1))) * c025;
        const R wp = W(0, 0, 1), wm = W(0, 0, -1);
        const R wwkp = ((w0 + wp) * (w0 + wp)) * c025;
        const R wwkm = ((w0 + wm) * (w0 + wm)) * c025;
        r = ((c0 - (wuip - wuim) * id) - (wvjp - wvjm) * id) - (wwkp - wwkm) * id;
        const R lw = ((((W(1, 0, 0) - c2 * w0) + W(-1, 0, 0)) + ((W(0, 1, 0) - c2 * w0) + W(0, -1, 0))) +
                           ((wp - c2 * w0) + wm)) * id2;
        r = r + over_rho(a.mu * lw, a.rhof, a.irhof, a.rpow2);
        if constexpr (SRC) r = r + over_rho(a.s2[q], a.rhof, a.irhof, a.rpow2);
        else r = r + c0;
        const R gp = (P(0, 0, 1) - p0) * id;
        const R rhs = ((-over_rho(gp, a.rhof, a.irhof, a.rpow2) + A * r) + B * o2) + a.g2;
        const R val = w0 + dt * rhs;
        a.ws[g] = val;
        IMG(a.ws, val);
        store_stream(&a.dvo2[q], r);
    }
}

// ---------------------------------------------------------------------------
// K7: correct_velocity_field (src/navier_stokes.f90:505-546), the interior of
// update_pressure (:550-566) and checks (:570-619) in one pass:
//   v = v* - ((grad(phi)*dt)/rhof),  p = p + phi,
//   maxdiv = signed max of div(v), max(|u|+|v|+|w|).
// phi may carry a pending mean shift (pp, ppn) that is subtracted here,
// value by value, exactly where the reference has already subtracted it.
// The divergence needs the new normal velocity one cell below; it is
// recomputed from u* and phi, and at non-periodic faces it takes the value
// the ghost fill gives it (bc for Dirichlet, the first cell for Neumann).
// A Dirichlet normal face at the high end overwrites the last cell with bc
// before the reference's checks run; the checks see that value too.
// ---------------------------------------------------------------------------
template <class R> struct ProjArgs {
    const R *__restrict__ us;
    const R *__restrict__ vs;
    const R *__restrict__ ws;
    const R *__restrict__ phi;
    R *__restrict__ u;
    R *__restrict__ v;
    R *__restrict__ w;
    R *__restrict__ p;
    const double *shift;
    double *part;
    const R *lo_plane[3];        // bc planes of the normal component: u left, v bottom, w front
    const R *hi_plane[3];        // u right, v top, w back
    int lo_type[3], hi_type[3];  // their bc types
    Box b;
    double id, dt, rhof;
    double irhof;   // 1/rhof, exact when rpow2
    int rpow2;      // rhof is a power of two
    int nbx, nby, img;
};

// One cell of the projection + checks.  F(di,dj,dk) returns phi (with the
// pending shift removed) around the cell; us0/vs0/ws0/p0 are u*, v*, w*, p at
// the cell and usm, vsm, wsm are u*, v*, w* one cell below along their axis.
// Arithmetic is FP64; each stored value is rounded once to R, and the
// checks see the stored (rounded) values, as a later read would.
// ST == 0: the checks only (nothing is stored; p0 is not used).
template <int D3, class R, class FF, class IW, int ST = 1>
__device__ __forceinline__ void project_cell(const ProjArgs<R> &a, int i, int j, int k, FF F, double us0, double vs0,
                                             double ws0, double p0, double usm, double vsm, double wsm, double &md,
                                             double &mv, double &bad, IW IMG)
{
    const Box &b = a.b;
    const long g = b.g(i, j, k);
    auto dr = [&](double x) { return over_rho(x, a.rhof, a.irhof, a.rpow2); };
    const double f0 = F(0, 0, 0);
    const R un = (R)(us0 - dr((F(1, 0, 0) - f0) * a.id * a.dt));
    const R vn = (R)(vs0 - dr((F(0, 1, 0) - f0) * a.id * a.dt));
    if constexpr (ST) {
        a.u[g] = un;
        a.v[g] = vn;
        IMG(a.u, un);
        IMG(a.v, vn);
    }
    R wn = R(0.0);
    if constexpr (D3) {
        wn = (R)(ws0 - dr((F(0, 0, 1) - f0) * a.id * a.dt));
        if constexpr (ST) {
            a.w[g] = wn;
            IMG(a.w, wn);
        }
    }
    if constexpr (ST) {
        const R pn = (R)(p0 + f0);
        a.p[g] = pn;
        IMG(a.p, pn);
    }

    // checks on the values the reference sees after its ghost fill
    const long pjk = j + (long)(b.ny + 2) * k, pik = i + (long)(b.nx + 2) * k, pij = i + (long)(b.nx + 2) * j;
    const double uc = (i == b.nx && a.hi_type[0] == 1) ? (double)a.hi_plane[0][pjk] : (double)un;
    const double vc = (j == b.ny && a.hi_type[1] == 1) ? (double)a.hi_plane[1][pik] : (double)vn;
    double um, vm;
    if (i > 1 || a.lo_type[0] == 0) um = (R)(usm - dr((f0 - F(-1, 0, 0)) * a.id * a.dt));
    else um = a.lo_type[0] == 1 ? (double)a.lo_plane[0][pjk] : (double)un;
    if (j > 1 || a.lo_type[1] == 0) vm = (R)(vsm - dr((f0 - F(0, -1, 0)) * a.id * a.dt));
    else vm = a.lo_type[1] == 1 ? (double)a.lo_plane[1][pik] : (double)vn;
    double d = (uc - um) * a.id + (vc - vm) * a.id;
    double vel = fabs(uc) + fabs(vc);
    if constexpr (D3) {
        const double wc = (k == b.nz && a.hi_type[2] == 1) ? (double)a.hi_plane[2][pij] : (double)wn;
        double wm;
        if (k > 1 || a.lo_type[2] == 0) wm = (R)(wsm - dr((f0 - F(0, 0, -1)) * a.id * a.dt));
        else wm = a.lo_type[2] == 1 ? (double)a.lo_plane[2][pij] : (double)wn;
        d = d + (wc - wm) * a.id;
        vel = vel + fabs(wc);
    }
    if (!(d == d) || !(vel == vel) || isinf(vel) || isinf(d)) bad = 1.0;
    if (d > md) md = d;
    if (vel > mv) mv = vel;
}

// Block-wide max of (maxdiv, maxvel, nonfinite flag) -> part[3*bid..].
template <int NT>
__device__ __forceinline__ void block_max3(double md, double mv, double bad, double *part, long bid)
{
    __shared__ double sd[NT], sv[NT], sf[NT];
    sd[threadIdx.x] = md; sv[threadIdx.x] = mv; sf[threadIdx.x] = bad;
    __syncthreads();
    for (int st = NT / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            if (sd[threadIdx.x + st] > sd[threadIdx.x]) sd[threadIdx.x] = sd[threadIdx.x + st];
            if (sv[threadIdx.x + st] > sv[threadIdx.x]) sv[threadIdx.x] = sv[threadIdx.x + st];
            if (sf[threadIdx.x + st] > sf[threadIdx.x]) sf[threadIdx.x] = sf[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * bid + 0] = sd[0];
        part[3 * bid + 1] = sv[0];
        part[3 * bid + 2] = sf[0];
    }
}

}  // namespace fen
