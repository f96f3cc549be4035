// poisson_kernels.hip -- fast direct Poisson solver passes, for element type
// R = double (FP64) and R = float (the FP32 form: FP32 transforms in LDS; the
// divergence, the eigenvalue divides and the Thomas sweep compute in FP64 and
// round once to R).
//
// Reference: poisson_solver_ppp (src/poisson.f90:941-1034),
// poisson_solver_ppn (:1038-1173), poisson_solver_pp (:416-505) and the
// right-hand side phi = div(u*)*rho/dt (src/navier_stokes.f90:111-121).
// Half-spectrum layout: spec[(kx*nz + (k-1))*ny + (j-1)], kx in 0..nx/2.
#include <stdexcept>

#include "fen_internal.hpp"
#include "fft_device.hpp"

namespace fen {

constexpr int kBlock = 256;

template <int N> constexpr int threads_per_line() { return N / 8; }
template <int N> constexpr int lines_per_block() { return kBlock / (N / 8); }

// ---------------------------------------------------------------------------
// K2: phi = (div(u*) * rho) / dt on TJ consecutive (j,k) rows, x r2c of row
// pairs packed as one complex line, spectrum scattered as j-runs.
// ---------------------------------------------------------------------------
template <class R> struct DivR2CArgs {
    using C = typename Cplx<R>::type;
    const R *u, *v, *w;
    C *spec;
    const C *tw;
    Box b;
    double id, rho, dt, post;
    long nrows, kxs;
    int direct;   // 1: `u` holds phi itself (interior values, strides sx/sy, ghost gl)
    long sx, sy;
    int gl;
    int ny_shift;   // log2(ny)
};

template <class R, int D3, int N, int DIRECT, int NT = kBlock>
__global__ __launch_bounds__(NT) void div_r2c_kernel(DivR2CArgs<R> a)
{
    using C = typename Cplx<R>::type;
    constexpr int T = threads_per_line<N>(), L = NT / T, TJ = 2 * L, LL = lds_line<N>();
    __shared__ C buf[L * LL];
    __shared__ C tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    const long row0 = (long)blockIdx.x * TJ;
    // thread (l, t) forms rows 2l and 2l+1 at x = t + q N/8 straight into
    // the first FFT stage's registers (the two real rows are the real and
    // imaginary parts of one complex line)
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    C v[8];
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const long r = row0 + 2 * l + h;
        const bool ok = r < a.nrows;
        const long rr = ok ? r : row0;   // a valid row: the loads are kept, their values dropped
        const int j = (int)(rr & (a.b.ny - 1)) + 1, k = (int)(rr >> a.ny_shift) + 1;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int i = t + q * (N / 8) + 1;
            double val;
            if constexpr (DIRECT) {
                val = a.u[(i - 1 + a.gl) + a.sx * (j - 1 + a.gl) + a.sy * (k - 1 + a.gl)];
            } else {
                const long g = a.b.g(i, j, k);
                double d = ((double)a.u[g] - (double)a.u[g - 1]) * a.id +
                           ((double)a.v[g] - (double)a.v[g - a.b.gx]) * a.id;
                if constexpr (D3) d = d + ((double)a.w[g] - (double)a.w[g - a.b.gxy]) * a.id;
                val = d * a.rho / a.dt;
            }
            if (h == 0) v[q].x = ok ? (R)val : R(0);
            else v[q].y = ok ? (R)val : R(0);
        }
    }
    C *line = buf + l * LL;
    fft_regs<N, -1>(v, line, t, tw);
    // the half spectrum of each row needs bins kx and N - kx: through LDS
#pragma unroll
    for (int q = 0; q < 8; q++) line[lds_pad(t + q * (N / 8))] = v[q];
    __syncthreads();
    for (int e = threadIdx.x; e < TJ * (N / 2 + 1); e += NT) {
        const int row = e % TJ, kx = e / TJ;
        const long r = row0 + row;
        if (r >= a.nrows) continue;
        const C *ln = buf + (row >> 1) * LL;
        const C zk = ln[lds_pad(kx)], zn = ln[lds_pad((N - kx) & (N - 1))];
        const R h = R(0.5), post = (R)a.post;   // post is 1 or 1/nx: exact in R
        C X;
        if ((row & 1) == 0) X = C{(zk.x + zn.x) * h, (zk.y - zn.y) * h};
        else X = C{(zk.y + zn.y) * h, (zn.x - zk.x) * h};
        X.x = X.x * post;
        X.y = X.y * post;
        a.spec[kx * a.kxs + r] = X;
    }
}

// ---------------------------------------------------------------------------
// K6: x c2r of row pairs -> phi rows (+ per-block sums for the mean).
// ---------------------------------------------------------------------------
template <class R> struct C2RArgs {
    using C = typename Cplx<R>::type;
    const C *spec;
    const C *tw;
    R *out;
    long nrows, kxs, sx, sy;
    int ny, gl, want_sum;
    int ny_shift;   // log2(ny)
    double *part;
    Box b;     // ghosted box of `out` when img != 0
    int img;   // periodic image axes to write (phi in the time step)
};

template <class R, int N, int NT = kBlock>
__global__ __launch_bounds__(NT) void c2r_kernel(C2RArgs<R> a)
{
    using C = typename Cplx<R>::type;
    constexpr int T = threads_per_line<N>(), L = NT / T, TJ = 2 * L, LL = lds_line<N>();
    static_assert(L * LL * sizeof(C) >= NT * sizeof(double), "the reduction reuses buf");
    __shared__ C buf[L * LL];
    __shared__ C tw[N];
    load_twiddles<N>(tw, a.tw);
    const long row0 = (long)blockIdx.x * TJ;
    // Staging: all of a thread's spectrum loads are issued before its LDS writes.
    constexpr int NE = L * (N / 2 + 1), PER = (NE + NT - 1) / NT;
    C X1[PER], X2[PER];
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int e = threadIdx.x + m * NT;
        const int ee = e < NE ? e : 0;
        const int l = ee % L, kx = ee / L;
        const long r = row0 + 2 * l;
        const bool ok1 = e < NE && r < a.nrows, ok2 = e < NE && r + 1 < a.nrows;
        const C v1 = a.spec[kx * a.kxs + (ok1 ? r : row0)];
        const C v2 = a.spec[kx * a.kxs + (ok2 ? r + 1 : row0)];
        X1[m] = ok1 ? v1 : C{R(0), R(0)};
        X2[m] = ok2 ? v2 : C{R(0), R(0)};
    }
#pragma unroll
    for (int m = 0; m < PER; m++) {
        const int e = threadIdx.x + m * NT;
        if (e >= NE) continue;
        const int l = e % L, kx = e / L;
        C *line = buf + l * LL;
        if (kx == 0 || kx == N / 2) {
            line[lds_pad(kx)] = C{X1[m].x, X2[m].x};   // Im of bins 0 and n/2 ignored
        } else {
            line[lds_pad(kx)] = C{X1[m].x - X2[m].y, X1[m].y + X2[m].x};
            line[lds_pad(N - kx)] = C{X1[m].x + X2[m].y, X2[m].x - X1[m].y};
        }
    }
    __syncthreads();
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    C *line = buf + l * LL;
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) v[q] = line[lds_pad(t + q * (N / 8))];
    __syncthreads();   // every thread holds its first-stage inputs before the stage writes
    fft_regs<N, 1>(v, line, t, tw);
    double acc = 0.0;
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const long r = row0 + 2 * l + h;
        if (r >= a.nrows) continue;
        const int j = (int)(r & (a.ny - 1)), k = (int)(r >> a.ny_shift);
        R *orow = a.out + a.gl + a.sx * (j + a.gl) + a.sy * (k + a.gl);
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int i = t + q * (N / 8);
            const R val = h ? v[q].y : v[q].x;
            orow[i] = val;
            if (a.img) write_images(a.out, a.b, a.img, i + 1, j + 1, k + 1, val);
            acc += (double)val;
        }
    }
    if (a.want_sum) {   // fft_regs has finished every read of buf: reuse it
        double *red = reinterpret_cast<double *>(buf);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
            __syncthreads();
        }
        if (threadIdx.x == 0) a.part[blockIdx.x] = red[0];
    }
}

// ---------------------------------------------------------------------------
// K3/K5: transforms along y; every y line spec[line*ny + 0..ny) is contiguous.
// mode 0: plain transform times `post`; mode 1 (pp, 2-D): forward, times
// inv_n, divide by mwx(kx)+mwy(j) (zero mode -> 0), inverse.
// ---------------------------------------------------------------------------
template <class R> struct YArgs {
    using C = typename Cplx<R>::type;
    C *spec;
    const C *tw;
    long nlines;
    double post, inv_n;
    const double *mwx, *mwy;
    int nz;
};

// Scales X by s1 and divides by s2 in FP64, rounding once to R (for R =
// double these are the reference's two roundings, in its order).
template <class C> __device__ __forceinline__ C scale_div(C X, double s1, double s2)
{
    using R = decltype(X.x);
    const double x = (double)X.x * s1, y = (double)X.y * s1;
    return C{(R)(x / s2), (R)(y / s2)};
}

template <class R, int N, int SIGN, int MODE>
__global__ __launch_bounds__(kBlock) void y_kernel(YArgs<R> a)
{
    using C = typename Cplx<R>::type;
    constexpr int T = threads_per_line<N>(), L = lines_per_block<N>(), LL = lds_line<N>();
    __shared__ C buf[L * LL];
    __shared__ C tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long ln = (long)blockIdx.x * L + l;
    const bool live = ln < a.nlines;
    C *row = a.spec + (live ? ln : (long)blockIdx.x * L) * N;   // a dead lane reads the block's first line
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const C x = row[t + q * (N / 8)];
        v[q] = live ? x : C{R(0), R(0)};
    }
    C *line = buf + l * LL;
    fft_regs<N, SIGN>(v, line, t, tw);
    if constexpr (MODE == 1) {
        const int kx = live ? (int)ln / (int)a.nz : 0;   // 32-bit division
        const double mx = a.mwx[kx];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const double s = mx + a.mwy[t + q * (N / 8)];
            v[q] = s == 0.0 ? C{R(0), R(0)} : scale_div(v[q], a.inv_n, s);
        }
        fft_regs<N, 1>(v, line, t, tw);
    }
    if (live) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            C X = v[q];
            X.x = (R)((double)X.x * a.post);
            X.y = (R)((double)X.y * a.post);
            row[t + q * (N / 8)] = X;
        }
    }
}

// ---------------------------------------------------------------------------
// K4 (ppp): z forward, /N, divide by (mwx+mwy)+mwz (zero mode -> 0), z inverse.
// A block owns L adjacent j lines of one kx plane; z lines have stride ny.
// ---------------------------------------------------------------------------
template <class R> struct ZArgs {
    using C = typename Cplx<R>::type;
    C *spec;
    const C *tw;
    int ny, nz;
    long kxs;
    double inv_n;
    const double *mwx, *mwy, *mwz;
};

template <class R, int N, int NT>
__global__ __launch_bounds__(NT) void z_solve_kernel(ZArgs<R> a)
{
    using C = typename Cplx<R>::type;
    constexpr int L = NT / (N / 8), LL = lds_line<N>();
    __shared__ C buf[L * LL];
    __shared__ C tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    const int kx = blockIdx.y;
    const int j0 = blockIdx.x * L;
    // adjacent threads own adjacent j lines: every k row of the block is one
    // contiguous run of L elements
    const int l = threadIdx.x % L, t = threadIdx.x / L;
    const bool live = j0 + l < a.ny;
    C *col = a.spec + kx * a.kxs + j0 + (live ? l : 0);   // a dead lane reads column j0, drops it
    C v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const C x = col[(long)(t + q * (N / 8)) * a.ny];
        v[q] = live ? x : C{R(0), R(0)};
    }
    double mz[8];
#pragma unroll
    for (int q = 0; q < 8; q++) mz[q] = a.mwz[t + q * (N / 8)];
    const double mxy = live ? a.mwx[kx] + a.mwy[j0 + l] : 0.0;
    C *line = buf + l * LL;
    fft_regs<N, -1>(v, line, t, tw);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double s = mxy + mz[q];
        v[q] = s == 0.0 ? C{R(0), R(0)} : scale_div(v[q], a.inv_n, s);
    }
    fft_regs<N, 1>(v, line, t, tw);
    if (live) {
#pragma unroll
        for (int q = 0; q < 8; q++) col[(long)(t + q * (N / 8)) * a.ny] = v[q];
    }
}

// The FP32 z-solve with two adjacent j lines per thread: each k row element
// of the pair is one 16-byte load, store and LDS access, so the transforms
// issue half the memory and LDS instructions of z_solve_kernel (whose
// FP32 form moves 8 bytes per LDS access).  Same arithmetic per line.
template <int N>
__global__ __launch_bounds__(kBlock) void z_solve_pair_kernel(ZArgs<float> a)
{
    using P = CPair<float2>;
    constexpr int L = kBlock / (N / 8), LL = lds_line<N>();
    __shared__ P buf[L * LL];
    __shared__ float2 tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    const int kx = blockIdx.y;
    const int l = threadIdx.x % L, t = threadIdx.x / L;
    const int j = blockIdx.x * 2 * L + 2 * l;   // lines j, j + 1 (ny is a multiple of 2L)
    float4 *col = reinterpret_cast<float4 *>(a.spec + kx * a.kxs + j);
    const long rs = a.ny / 2;   // k row stride in pairs
    P v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const float4 x = col[(long)(t + q * (N / 8)) * rs];
        v[q] = P{float2{x.x, x.y}, float2{x.z, x.w}};
    }
    double mz[8];
#pragma unroll
    for (int q = 0; q < 8; q++) mz[q] = a.mwz[t + q * (N / 8)];
    const double ma = a.mwx[kx] + a.mwy[j], mb = a.mwx[kx] + a.mwy[j + 1];
    P *line = buf + l * LL;
    fft_regs<N, -1>(v, line, t, tw);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const double sa = ma + mz[q], sb = mb + mz[q];
        v[q].a = sa == 0.0 ? float2{0.0f, 0.0f} : scale_div(v[q].a, a.inv_n, sa);
        v[q].b = sb == 0.0 ? float2{0.0f, 0.0f} : scale_div(v[q].b, a.inv_n, sb);
    }
    fft_regs<N, 1>(v, line, t, tw);
#pragma unroll
    for (int q = 0; q < 8; q++)
        col[(long)(t + q * (N / 8)) * rs] = float4{v[q].a.x, v[q].a.y, v[q].b.x, v[q].b.y};
}

// ---------------------------------------------------------------------------
// K4 (ppn): Thomas sweep along z for every (kx, j) column
// (src/poisson.f90:1094-1135); one thread per column, adjacent threads own
// adjacent j so every k-row access is coalesced.  The sweep is FP64 for
// either element type: c1 lives in an FP64 array of the spectrum's layout,
// computed once per context (launch_thomas_c1: it does not depend on the
// right-hand side), so the forward sweep recomputes its pivots in registers
// and only the back substitution reads it; d is kept in place in the spectrum when R is double, and in an
// FP64 scratch `dd` when R is float (the back substitution of a nearly
// singular column would amplify a rounded d by up to ~nz^2).
// ---------------------------------------------------------------------------
template <class R> struct ThomasArgs {
    typename Cplx<R>::type *spec;
    double *c1;
    double2 *dd;
    int ny, nz, hx;
    long kxs;
    const double *mwx, *mwy, *ta, *tb, *tc;
};

constexpr int kTU = 8;   // Thomas rows loaded ahead per batch

template <class R> __global__ __launch_bounds__(kBlock) void thomas_kernel(ThomasArgs<R> a)
{
    using C = typename Cplx<R>::type;
    constexpr bool IN_PLACE = sizeof(R) == sizeof(double);
    const long t = (long)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (long)a.hx * a.ny) return;
    const int kx = (int)(t / a.ny), j = (int)(t % a.ny);
    const long base = kx * a.kxs + j;
    const long s = a.ny;
    C *x = a.spec + base;
    const double *c1 = a.c1 + base;
    double2 *dd = IN_PLACE ? nullptr : a.dd + base;
    auto rhs = [&](long o) { const C v = x[o]; return double2{(double)v.x, (double)v.y}; };
    auto put_d = [&](long o, double2 v) {
        if constexpr (IN_PLACE) x[o] = C{(R)v.x, (R)v.y};
        else dd[o] = v;
    };
    auto get_d = [&](long o) {
        if constexpr (IN_PLACE) return rhs(o);
        else return dd[o];
    };
    const double mx = a.mwx[kx], my = a.mwy[j];
    const int n = a.nz;
    double factor = 1.0 / ((a.tb[0] + mx) + my);
    double c1p = a.tc[0] * factor;
    double2 d = rhs(0);
    d.x = d.x * factor;
    d.y = d.y * factor;
    put_d(0, d);
    auto fwd = [&](int k, double2 r) {
        factor = 1.0 / (((a.tb[k] + mx) + my) - a.ta[k] * c1p);
        c1p = a.tc[k] * factor;
        d.x = (r.x - a.ta[k] * d.x) * factor;
        d.y = (r.y - a.ta[k] * d.y) * factor;
        put_d(k * s, d);
    };
    // the loads of TU rows are issued together ahead of their sweep steps,
    // so TU memory latencies overlap instead of one per step
    int k = 1;
    for (; k + kTU <= n - 1; k += kTU) {
        double2 r[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) r[u] = rhs((long)(k + u) * s);
#pragma unroll
        for (int u = 0; u < kTU; u++) fwd(k + u, r[u]);
    }
    for (; k < n - 1; k++) fwd(k, rhs((long)k * s));
    factor = ((a.tb[n - 1] + mx) + my) - a.ta[n - 1] * c1p;
    const double2 r = rhs((n - 1) * s);
    if (factor != 0.0) {
        d.x = (r.x - a.ta[n - 1] * d.x) / factor;
        d.y = (r.y - a.ta[n - 1] * d.y) / factor;
    } else {
        d = double2{0.0, 0.0};
    }
    x[(n - 1) * s] = C{(R)d.x, (R)d.y};
    double2 xn = d;
    auto bwd = [&](int k, double2 dk, double ck) {
        xn.x = dk.x - ck * xn.x;
        xn.y = dk.y - ck * xn.y;
        x[k * s] = C{(R)xn.x, (R)xn.y};
    };
    k = n - 2;
    for (; k - kTU + 1 >= 0; k -= kTU) {
        double2 dk[kTU];
        double ck[kTU];
#pragma unroll
        for (int u = 0; u < kTU; u++) {
            dk[u] = get_d((long)(k - u) * s);
            ck[u] = c1[(long)(k - u) * s];
        }
#pragma unroll
        for (int u = 0; u < kTU; u++) bwd(k - u, dk[u], ck[u]);
    }
    for (; k >= 0; k--) bwd(k, get_d((long)k * s), c1[(long)k * s]);
}

// Sum of per-block partials in a fixed order, then shift = sum / N.  One
// workgroup of kShiftThreads: each thread adds its entries in batches of
// kShiftBatch loads issued together (one memory latency per batch), then a
// tree over the threads.
constexpr int kShiftThreads = 1024, kShiftBatch = 4;
__global__ __launch_bounds__(kShiftThreads) void mean_shift_kernel(const double *part, long n, double count,
                                                                   double *out)
{
    __shared__ double red[kShiftThreads];
    double acc = 0.0;
    for (long q0 = threadIdx.x; q0 < n; q0 += (long)kShiftThreads * kShiftBatch) {
        double x[kShiftBatch];
#pragma unroll
        for (int m = 0; m < kShiftBatch; m++) {
            const long q = q0 + (long)m * kShiftThreads;
            x[m] = part[q < n ? q : q0];
        }
#pragma unroll
        for (int m = 0; m < kShiftBatch; m++)
            if (q0 + (long)m * kShiftThreads < n) acc += x[m];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = kShiftThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = red[threadIdx.x] + red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0] / count;
}

template <class R> __global__ void apply_shift_kernel(R *f, long n, const double *shift)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) f[q] = (R)((double)f[q] - shift[0]);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define FEN_DISPATCH_N(n, MACRO)                         \
    switch (n) {                                         \
    case 8: MACRO(8); break;                             \
    case 16: MACRO(16); break;                           \
    case 32: MACRO(32); break;                           \
    case 64: MACRO(64); break;                           \
    case 128: MACRO(128); break;                         \
    case 256: MACRO(256); break;                         \
    case 512: MACRO(512); break;                         \
    case 1024: MACRO(1024); break;                       \
    default: throw std::runtime_error("unsupported transform length"); \
    }

static int ilog2(int n)   // n a power of two
{
    int l = 0;
    while ((1 << l) < n) l++;
    return l;
}

template <class R> static double spec_bytes(const fen_ctx *c)
{
    return (double)c->hx * c->ny * c->nz * 2.0 * sizeof(R);
}

template <class R> using CplxT = typename Cplx<R>::type;

// The x passes (div_r2c, c2r) at N = 512 run 512-thread workgroups: 16 rows
// per workgroup, so each access to a kx plane of the spectrum covers 256
// contiguous bytes instead of 128.
static int x_pass_threads(const fen_ctx *c) { return c->nx == 512 ? 512 : kBlock; }

template <class R, int D3, int N, int DIRECT>
static void launch_div_r2c_n(const fen_ctx *c, dim3 grid, const DivR2CArgs<R> &a)
{
    if constexpr (N == 512) hipLaunchKernelGGL((div_r2c_kernel<R, D3, N, DIRECT, 512>), grid, dim3(512), 0, c->stream, a);
    else hipLaunchKernelGGL((div_r2c_kernel<R, D3, N, DIRECT>), grid, dim3(kBlock), 0, c->stream, a);
}

// vel == nullptr: direct mode, the RHS is already in `direct_src` (phi array).
template <class R>
void launch_div_r2c_impl(fen_ctx *c, const R *const *vel, const R *direct_src, long sx, long sy, int gl, double rho,
                         double dt, double post)
{
    DivR2CArgs<R> a{};
    if (vel) { a.u = vel[0]; a.v = vel[1]; a.w = vel[2]; }
    else { a.u = direct_src; a.direct = 1; a.sx = sx; a.sy = sy; a.gl = gl; }
    a.spec = as<CplxT<R>>(c->d_spec);
    a.tw = as<CplxT<R>>(c->d_twx);
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.rho = rho;
    a.dt = dt;
    a.post = post;
    a.nrows = (long)c->ny * c->nz;
    a.kxs = (long)c->ny * c->nz;
    a.ny_shift = ilog2(c->ny);
    const int tj = 2 * (x_pass_threads(c) / (c->nx / 8));
    const dim3 grid((unsigned)((a.nrows + tj - 1) / tj));
    const double bytes = (vel ? (c->dim == 3 ? 3.0 : 2.0) : 1.0) * sizeof(R) * c->n_int + spec_bytes<R>(c);
    prof_begin(c, K_DIV_R2C);
#define L_DIV(NN)                                                                                        \
    if (a.direct) launch_div_r2c_n<R, 0, NN, 1>(c, grid, a);                                             \
    else if (c->dim == 3) launch_div_r2c_n<R, 1, NN, 0>(c, grid, a);                                     \
    else launch_div_r2c_n<R, 0, NN, 0>(c, grid, a);
    FEN_DISPATCH_N(c->nx, L_DIV)
#undef L_DIV
    prof_end(c, K_DIV_R2C, bytes);
}

template <class R> void launch_y_fft(fen_ctx *c, int sign, double post)
{
    YArgs<R> a{};
    a.spec = as<CplxT<R>>(c->d_spec);
    a.tw = as<CplxT<R>>(c->d_twy);
    a.nlines = (long)c->hx * c->nz;
    a.post = post;
    a.nz = c->nz;
    const int L = kBlock / (c->ny / 8);
    const dim3 grid((unsigned)((a.nlines + L - 1) / L));
    const int kid = sign < 0 ? K_Y_FWD : K_Y_INV;
    prof_begin(c, kid);
#define L_Y(NN)                                                                                  \
    if (sign < 0) hipLaunchKernelGGL((y_kernel<R, NN, -1, 0>), grid, dim3(kBlock), 0, c->stream, a); \
    else hipLaunchKernelGGL((y_kernel<R, NN, 1, 0>), grid, dim3(kBlock), 0, c->stream, a);
    FEN_DISPATCH_N(c->ny, L_Y)
#undef L_Y
    prof_end(c, kid, 2.0 * spec_bytes<R>(c));
}

template <class R> void launch_y_solve_2d(fen_ctx *c, double inv_n)
{
    YArgs<R> a{};
    a.spec = as<CplxT<R>>(c->d_spec);
    a.tw = as<CplxT<R>>(c->d_twy);
    a.nlines = (long)c->hx * c->nz;
    a.post = 1.0;
    a.inv_n = inv_n;
    a.mwx = c->d_mwx;
    a.mwy = c->d_mwy;
    a.nz = c->nz;
    const int L = kBlock / (c->ny / 8);
    const dim3 grid((unsigned)((a.nlines + L - 1) / L));
    prof_begin(c, K_Y_SOLVE);
#define L_YS(NN) hipLaunchKernelGGL((y_kernel<R, NN, -1, 1>), grid, dim3(kBlock), 0, c->stream, a);
    FEN_DISPATCH_N(c->ny, L_YS)
#undef L_YS
    prof_end(c, K_Y_SOLVE, 2.0 * spec_bytes<R>(c));
}

template <class R> void launch_z_solve(fen_ctx *c, double inv_n)
{
    ZArgs<R> a{};
    a.spec = as<CplxT<R>>(c->d_spec);
    a.tw = as<CplxT<R>>(c->d_twz);
    a.ny = c->ny;
    a.nz = c->nz;
    a.kxs = (long)c->ny * c->nz;
    a.inv_n = inv_n;
    a.mwx = c->d_mwx;
    a.mwy = c->d_mwy;
    a.mwz = c->d_mwz;
    // FP32: 512-thread workgroups hold 16 adjacent z lines when ny allows it,
    // so every k row of a workgroup is a 128-byte run instead of 64 bytes
    // (measured at 256^3: 61 -> 52 us).  FP64 keeps 256 threads: 512 would
    // need 74 KB of LDS per workgroup and leave 2 per CU (87 -> 100 us).
    const int nt = (sizeof(R) == sizeof(float) && c->ny % (512 / (c->nz / 8)) == 0) ? 512 : kBlock;
    const int L = nt / (c->nz / 8);
    const dim3 grid((unsigned)((c->ny + L - 1) / L), (unsigned)c->hx);
    prof_begin(c, K_Z_SOLVE);
    if constexpr (sizeof(R) == sizeof(float)) {
        // two lines per thread: 2 * (kBlock / (nz / 8)) lines per workgroup
        const int lp = 2 * (kBlock / (c->nz / 8));
        if (c->ny % lp == 0) {
            const dim3 gp((unsigned)(c->ny / lp), (unsigned)c->hx);
#define L_ZP(NN) hipLaunchKernelGGL((z_solve_pair_kernel<NN>), gp, dim3(kBlock), 0, c->stream, a);
            FEN_DISPATCH_N(c->nz, L_ZP)
#undef L_ZP
            prof_end(c, K_Z_SOLVE, 2.0 * spec_bytes<R>(c));
            return;
        }
    }
#define L_Z(NN)                                                                                       \
    if (nt == 512) hipLaunchKernelGGL((z_solve_kernel<R, NN, 512>), grid, dim3(512), 0, c->stream, a); \
    else hipLaunchKernelGGL((z_solve_kernel<R, NN, kBlock>), grid, dim3(kBlock), 0, c->stream, a);
    FEN_DISPATCH_N(c->nz, L_Z)
#undef L_Z
    prof_end(c, K_Z_SOLVE, 2.0 * spec_bytes<R>(c));
}

template <class R> void launch_thomas(fen_ctx *c)
{
    ThomasArgs<R> a{};
    a.spec = as<CplxT<R>>(c->d_spec);
    a.c1 = c->d_c1;
    a.dd = c->d_dd;
    if (sizeof(R) != sizeof(double) && !a.dd) throw std::runtime_error("FP64 Thomas scratch missing");
    a.ny = c->ny;
    a.nz = c->nz;
    a.hx = c->hx;
    a.kxs = (long)c->ny * c->nz;
    a.mwx = c->d_mwx;
    a.mwy = c->d_mwy;
    a.ta = c->d_ta;
    a.tb = c->d_tb;
    a.tc = c->d_tc;
    const long cols = (long)c->hx * c->ny;
    prof_begin(c, K_THOMAS);
    hipLaunchKernelGGL(thomas_kernel<R>, dim3((unsigned)((cols + kBlock - 1) / kBlock)), dim3(kBlock), 0, c->stream,
                       a);
    // spectrum read + written, c1 read (FP64), d written + read (in place for double, FP64 scratch for float)
    const double n_el = (double)c->hx * c->ny * c->nz;
    prof_end(c, K_THOMAS, 2.0 * spec_bytes<R>(c) + 8.0 * n_el + 32.0 * n_el);
}

// Workgroups of the x c2r (each writes one block sum when the mean is taken).
template <class R, int N> static void launch_c2r_n(const fen_ctx *c, long nblk, const C2RArgs<R> &a)
{
    if constexpr (N == 512) hipLaunchKernelGGL((c2r_kernel<R, N, 512>), dim3((unsigned)nblk), dim3(512), 0, c->stream, a);
    else hipLaunchKernelGGL((c2r_kernel<R, N>), dim3((unsigned)nblk), dim3(kBlock), 0, c->stream, a);
}

long c2r_blocks(const fen_ctx *c)
{
    const int tj = 2 * (x_pass_threads(c) / (c->nx / 8));
    return ((long)c->ny * c->nz + tj - 1) / tj;
}

template <class R> void launch_c2r(fen_ctx *c, R *phi, int gl_out, long sx, long sy, bool want_sum, int img,
                                   double *part, size_t n_part, double *shift)
{
    if (!part) {
        part = c->d_part;
        n_part = c->n_part;
        shift = c->d_shift;
    }
    C2RArgs<R> a{};
    a.spec = as<CplxT<R>>(c->d_spec);
    a.tw = as<CplxT<R>>(c->d_twx);
    a.out = phi;
    a.nrows = (long)c->ny * c->nz;
    a.kxs = a.nrows;
    a.sx = sx;
    a.sy = sy;
    a.ny = c->ny;
    a.gl = gl_out;
    a.ny_shift = ilog2(c->ny);
    a.want_sum = want_sum ? 1 : 0;
    a.part = part;
    a.b = c->box;
    a.img = img;
    const long nblk = c2r_blocks(c);
    if (want_sum && (size_t)nblk > n_part) throw std::runtime_error("partial-sum buffer too small");
    prof_begin(c, K_C2R);
#define L_C(NN) launch_c2r_n<R, NN>(c, nblk, a);
    FEN_DISPATCH_N(c->nx, L_C)
#undef L_C
    prof_end(c, K_C2R, spec_bytes<R>(c) + sizeof(R) * c->n_int);
    if (want_sum) {
        const double count = (double)(float)((long)c->nx * c->ny * c->nz);
        prof_begin(c, K_REDUCE);
        hipLaunchKernelGGL(mean_shift_kernel, dim3(1), dim3(kShiftThreads), 0, c->stream, (const double *)part, nblk,
                           count, shift);
        prof_end(c, K_REDUCE, 8.0 * nblk);
    }
}

template <class R> void launch_apply_shift(fen_ctx *c, R *f, size_t n)
{
    hipLaunchKernelGGL(apply_shift_kernel<R>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, f, (long)n,
                       (const double *)c->d_shift);
    prof_end(c, K_REDUCE, 2.0 * sizeof(R) * n);
}

#define FEN_INSTANTIATE(R)                                                                                     \
    template void launch_div_r2c_impl<R>(fen_ctx *, const R *const *, const R *, long, long, int, double, double, \
                                         double);                                                              \
    template void launch_y_fft<R>(fen_ctx *, int, double);                                                     \
    template void launch_y_solve_2d<R>(fen_ctx *, double);                                                     \
    template void launch_z_solve<R>(fen_ctx *, double);                                                        \
    template void launch_thomas<R>(fen_ctx *);                                                                 \
    template void launch_c2r<R>(fen_ctx *, R *, int, long, long, bool, int, double *, size_t, double *);       \
    template void launch_apply_shift<R>(fen_ctx *, R *, size_t);
FEN_INSTANTIATE(double)
FEN_INSTANTIATE(float)
#undef FEN_INSTANTIATE

}  // namespace fen
