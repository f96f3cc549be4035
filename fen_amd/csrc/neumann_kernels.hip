// neumann_kernels.hip -- FP64 passes of the Poisson solvers with Neumann
// (non-periodic) directions: pn and nn in 2-D, npn and nnn in 3-D
// (reference src/poisson.f90:306-596, 1177-1451).
//
// DCT-II (FFTW_REDFT10, y_k = 2 sum_j x_j cos(pi k (2j+1)/(2n))) and DCT-III
// (FFTW_REDFT01, y_j = x_0 + 2 sum_{k>=1} x_k cos(pi k (2j+1)/(2n))) use
// Makhoul's length-n complex FFT form, two real lines packed in one complex
// line:
//   forward: v = (x_0, x_2, ..., x_{n-2}, x_{n-1}, ..., x_3, x_1), V = FFT(v),
//            y_k = 2 Re(exp(-i pi k/(2n)) V_k);
//   inverse: U_k = (x_k - i x_{n-k}) exp(i pi k/(2n)) (x_n = 0), u = IFFT(U)
//            (unnormalised, real), y_{2m} = u_m, y_{2m+1} = u_{n-1-m}.
// The transforms run in LDS with the Stockham kernels of fft_device.hpp.
// Spectra are laid out [k][kx][j] (j fastest; npn's y half spectrum
// [k][kx][jy]): one k row of every z column is contiguous, so the Thomas
// sweeps along z read and write whole contiguous rows at each step.
#include <stdexcept>

#include "fen_internal.hpp"
#include "fft_device.hpp"

namespace fen {

namespace {

constexpr int kNB = 256;

template <int N> constexpr int nb_tpl() { return N / 8; }           // threads per complex line
template <int N> constexpr int nb_lines() { return kNB / (N / 8); } // complex lines per block

// Position of x_i in Makhoul's reordered sequence v.
__device__ __forceinline__ int mk_pos(int i, int n) { return (i & 1) ? n - 1 - (i >> 1) : (i >> 1); }

// ---------------------------------------------------------------------------
// DCT-II of 2L rows per block.  Source rows: SRC 0 = div(u*)*rho/dt on the
// ghosted velocity (x pass), 1 = interior rows of phi with strides (x pass,
// direct solve), 2 = contiguous real lines of length N (y pass).  Output
// element k of row r goes to out[k*ok + r*orr], times `post`.
// ---------------------------------------------------------------------------
struct DctArgs {
    const double *u, *v, *w;      // SRC 0: velocity; SRC 1/2: u is the source
    double *out;
    const double2 *tw;            // W_N^m, m < N
    const double2 *mk;            // (cos, sin)(pi k/(2N)), k < N
    Box b;
    double id, rho, dt, post;
    long nrows, ok, orr;          // rows, output strides
    long okz;                     // x pass: extra output stride per z (row r has k = r / ny)
    long sx, sy;                  // SRC 1: strides of the phi array
    int gl, d3;
};

// The DCT-II's last step: y_k = 2 Re(exp(-i pi k/(2N)) V_k) of the block's
// 2L rows from their packed spectra in LDS (row 2l in Re, 2l + 1 in Im of
// line l), times post.  Adjacent threads write adjacent addresses: rows
// (x pass, orr = 1) or k (y pass, ok = 1).
template <int N>
__device__ __forceinline__ void dct2_out(const DctArgs &a, const double2 *buf, const double2 *mks, long row0)
{
    constexpr int L = nb_lines<N>(), TJ = 2 * L, LL = lds_line<N>();
    const bool k_fast = a.ok == 1;
    if constexpr (kNB % TJ == 0) {
        if (!k_fast) {
            // x pass: a thread keeps one row and walks k
            const int row = threadIdx.x % TJ;
            const long r = row0 + row;
            if (r >= a.nrows) return;
            const long zo = a.okz ? (long)((int)r / a.b.ny) * a.okz : 0;   // rows < 2^31: 32-bit division
            double *o = a.out + r * a.orr + zo;
            const double2 *line = buf + (row >> 1) * LL;
            for (int k = threadIdx.x / TJ; k < N; k += kNB / TJ) {
                const double2 zk = line[lds_pad(k)], zn = line[lds_pad((N - k) & (N - 1))];
                const double c = mks[k].x, s = mks[k].y;
                double y;
                if ((row & 1) == 0) y = c * (zk.x + zn.x) + s * (zk.y - zn.y);
                else y = c * (zk.y + zn.y) + s * (zn.x - zk.x);
                o[k * a.ok] = y * a.post;
            }
            return;
        }
    }
    for (int e = threadIdx.x; e < TJ * N; e += kNB) {
        const int row = k_fast ? e / N : e % TJ, k = k_fast ? e % N : e / TJ;
        const long r = row0 + row;
        if (r >= a.nrows) continue;
        const double2 *line = buf + (row >> 1) * LL;
        const double2 zk = line[lds_pad(k)], zn = line[lds_pad((N - k) & (N - 1))];
        const double c = mks[k].x, s = mks[k].y;
        double y;
        if ((row & 1) == 0) y = c * (zk.x + zn.x) + s * (zk.y - zn.y);
        else y = c * (zk.y + zn.y) + s * (zn.x - zk.x);
        const long zo = a.okz ? (long)((int)r / a.b.ny) * a.okz : 0;   // rows < 2^31: 32-bit division
        a.out[k * a.ok + r * a.orr + zo] = y * a.post;
    }
}

template <int N, int SRC>
__global__ __launch_bounds__(kNB) void dct2_rows_kernel(DctArgs a)
{
    constexpr int T = nb_tpl<N>(), L = nb_lines<N>(), TJ = 2 * L, LL = lds_line<N>();
    __shared__ double2 buf[L * LL];
    __shared__ double2 tw[N];
    load_twiddles<N>(tw, a.tw);
    const double2 *mks = a.mk;
    const long row0 = (long)blockIdx.x * TJ;
    // staging: the loads of BATCH elements are issued before their LDS
    // writes; adjacent threads read adjacent x, so every load is coalesced
    // (Makhoul's reorder puts even and odd x in different halves of the
    // line, so the FFT's first stage cannot take them straight from memory)
    constexpr int PER = TJ * N / kNB, BATCH = PER < 8 ? PER : 8;
    static_assert(PER * kNB == TJ * N && PER % BATCH == 0, "staging batches");
#pragma unroll
    for (int b0 = 0; b0 < PER; b0 += BATCH) {
        double vals[BATCH];
#pragma unroll
        for (int m = 0; m < BATCH; m++) {
            const int e = threadIdx.x + (b0 + m) * kNB;
            const int row = e / N, i = e % N;
            const long r0 = row0 + row;
            const bool ok = r0 < a.nrows;
            const long r = ok ? r0 : row0;   // a valid row: the loads are kept, the value dropped
            double val;
            if constexpr (SRC == 2) {
                val = a.u[r * N + i];
            } else {
                // rows number ny*nz < 2^31: 32-bit division (a 64-bit one is a long routine)
                const int ri = (int)r, j = ri % a.b.ny + 1, k = ri / a.b.ny + 1;
                if constexpr (SRC == 1) {
                    val = a.u[(i + a.gl) + a.sx * (j - 1 + a.gl) + a.sy * (k - 1 + a.gl)];
                } else {
                    const long g = a.b.g(i + 1, j, k);
                    double d = (a.u[g] - a.u[g - 1]) * a.id + (a.v[g] - a.v[g - a.b.gx]) * a.id;
                    if (a.d3) d = d + (a.w[g] - a.w[g - a.b.gxy]) * a.id;
                    val = d * a.rho / a.dt;
                }
            }
            vals[m] = ok ? val : 0.0;
        }
#pragma unroll
        for (int m = 0; m < BATCH; m++) {
            const int e = threadIdx.x + (b0 + m) * kNB;
            const int row = e / N, i = e % N;
            double2 *line = buf + (row >> 1) * LL;
            if (row & 1) line[lds_pad(mk_pos(i, N))].y = vals[m];
            else line[lds_pad(mk_pos(i, N))].x = vals[m];
        }
    }
    __syncthreads();
    fft_lds<N, -1>(buf + (threadIdx.x / T) * LL, threadIdx.x % T, tw);
    dct2_out<N>(a, buf, mks, row0);
}

// The same DCT-II with the first Stockham stage's inputs loaded straight
// into registers (N <= 512, so a line's T = N/8 threads lie in one wave):
// SRC 0 (div(u*) rho/dt, x pass) or 2 (contiguous lines, y pass).  Thread j
// of a line needs Makhoul's v_p at p = j + rT, r = 0..7: the even x_{2j+2rT}
// (r < 4) and the odd x_{2rT+2T-1-2j} (v[7-r]).  Each thread loads aligned
// pairs of adjacent x with 16-byte loads, keeps the even element and passes
// the odd one through a cross-lane read (ds_bpermute) to the thread whose
// v[7-r] it is.  The values and the transform's arithmetic are
// dct2_rows_kernel's, so the results are bitwise equal.
//   SRC 2: pair (2rT+2j, +1); the odd x_{2rT+2j+1} goes to thread T-1-j.
//   SRC 0: ghosted pair (q, q+1), q = 2rT + 2j, holds elements q-1 (odd)
//   and q (even) (ghosted x index = element + 1); the odd one goes to thread
//   T-j, and thread 0 takes its own from chunk r+1 (lane 0 loads a fifth
//   pair, q = N, for x_{N-1}).  Needs 16-byte aligned fields and even gx.
// MKL: the (cos, sin) table of the last step in LDS (it costs the x pass an
// LDS-bound workgroup per CU, so only the y pass takes it).
template <int N, int SRC, int MKL>
__global__ __launch_bounds__(kNB) void dct2_pairs_kernel(DctArgs a)
{
    constexpr int T = nb_tpl<N>(), L = nb_lines<N>(), TJ = 2 * L, LL = lds_line<N>();
    static_assert(T <= 64 && (SRC == 0 || SRC == 2), "one wave per line; x pass or y pass");
    __shared__ double2 buf[L * LL];
    __shared__ double2 tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    __shared__ double2 mkl[MKL ? N : 1];
    const double2 *mks = a.mk;
    if constexpr (MKL) {
        load_twiddles<N>(mkl, a.mk);
        mks = mkl;
    }
    const long row0 = (long)blockIdx.x * TJ;
    const int l = threadIdx.x / T, j = threadIdx.x % T;
    auto ld2 = [](const double *p) { return *reinterpret_cast<const double2 *>(p); };
    double ev[2][4], od[2][5];   // [row][chunk]: the even element kept, the odd one passed on
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const long r0 = row0 + 2 * l + h;
        const bool ok = r0 < a.nrows;
        const long r = ok ? r0 : row0;   // a valid row: the loads are kept, the values dropped
        if constexpr (SRC == 2) {
            const double *in = a.u + r * N;
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const double2 p = ld2(in + 2 * c * T + 2 * j);
                ev[h][c] = ok ? p.x : 0.0;
                od[h][c] = ok ? p.y : 0.0;
            }
        } else {
            const int ri = (int)r, jy = ri % a.b.ny + 1, kz = ri / a.b.ny + 1;
            const long g0 = a.b.g(0, jy, kz);
            // elements q - 1 and q from the ghosted pair (q, q + 1)
            auto pair = [&](long gq, double &d_odd, double &d_even) {
                const double um = a.u[gq - 1];
                const double2 uu = ld2(a.u + gq), vv = ld2(a.v + gq), vs = ld2(a.v + gq - a.b.gx);
                double d0 = (uu.x - um) * a.id + (vv.x - vs.x) * a.id;
                double d1 = (uu.y - uu.x) * a.id + (vv.y - vs.y) * a.id;
                if (a.d3) {
                    const double2 ww = ld2(a.w + gq), ws = ld2(a.w + gq - a.b.gxy);
                    d0 = d0 + (ww.x - ws.x) * a.id;
                    d1 = d1 + (ww.y - ws.y) * a.id;
                }
                d_odd = ok ? d0 * a.rho / a.dt : 0.0;
                d_even = ok ? d1 * a.rho / a.dt : 0.0;
            };
#pragma unroll
            for (int c = 0; c < 4; c++) pair(g0 + 2 * c * T + 2 * j, od[h][c], ev[h][c]);
            od[h][4] = 0.0;
            if (j == 0) {
                double unused;
                pair(g0 + N, od[h][4], unused);
            }
        }
    }
    double2 v[8];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        v[c] = double2{ev[0][c], ev[1][c]};
        double o0, o1;
        if constexpr (SRC == 2) {
            o0 = __shfl(od[0][c], T - 1 - j, T);
            o1 = __shfl(od[1][c], T - 1 - j, T);
        } else {
            const double s0 = __shfl(od[0][c], T - j, T), s1 = __shfl(od[1][c], T - j, T);
            o0 = j == 0 ? od[0][c + 1] : s0;
            o1 = j == 0 ? od[1][c + 1] : s1;
        }
        v[7 - c] = double2{o0, o1};
    }
    double2 *line = buf + l * LL;
    fft_regs<N, -1>(v, line, j, tw);
#pragma unroll
    for (int q = 0; q < 8; q++) line[lds_pad(j + q * T)] = v[q];
    __syncthreads();
    dct2_out<N>(a, buf, mks, row0);
}

// ---------------------------------------------------------------------------
// DCT-III of 2L rows per block.  Input element k of row r at in[k*ik + r*ir].
// DST 0: interior rows of phi (strides sx, sy, ghost level gl), times post;
// DST 1: contiguous real lines of length N, times post.
// ---------------------------------------------------------------------------
struct Dct3Args {
    const double *in;
    double *out;
    const double2 *tw, *mk;
    long nrows, ik, ir;
    long ikz;   // x pass: extra input stride per z (row r has k = r / ny)
    long sx, sy;
    int gl, ny, nz;
    // x pass with gl == 1: also write the ghosts, when every ghost of phi
    // is a copy (Neumann x faces; y, z periodic or Neumann): each row's
    // x ghosts and the ghost rows that copy the row (yper/zper: periodic)
    int img, yper, zper;
    double post;
};

// The ghosted indices that hold row index J (interior 1..n) of an axis whose
// ghosts are copies: J itself, and the ghost rows 0 and n + 1 that copy it
// (Neumann: 0 copies 1 and n + 1 copies n; periodic: 0 copies n and n + 1
// copies 1).  Returns how many.
__device__ __forceinline__ int copy_rows(int J, int n, int per, int (&d)[3])
{
    int m = 0;
    d[m++] = J;
    if (J == (per ? n : 1)) d[m++] = 0;
    if (J == (per ? 1 : n)) d[m++] = n + 1;
    return m;
}

template <int N, int DST>
__global__ __launch_bounds__(kNB) void dct3_rows_kernel(Dct3Args a)
{
    constexpr int T = nb_tpl<N>(), L = nb_lines<N>(), LL = lds_line<N>();
    __shared__ double2 buf[L * LL];
    __shared__ double2 tw[N];
    load_twiddles<N>(tw, a.tw);
    const long row0 = (long)blockIdx.x * 2 * L;
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    double2 *line = buf + l * LL;
    // U_k of both rows packed, (x_k - i x_{N-k}) exp(i pi k/(2N)) (x_N = 0)
    auto packed = [&](double a1, double b1, double a2, double b2, int k) {
        const double c = a.mk[k].x, s = a.mk[k].y;
        const double2 U1 = double2{a1 * c + b1 * s, a1 * s - b1 * c};
        const double2 U2 = double2{a2 * c + b2 * s, a2 * s - b2 * c};
        return double2{U1.x - U2.y, U1.y + U2.x};
    };
    double2 v[8];
    if constexpr (DST == 1) {
        // y pass: k is the contiguous direction (ik = 1), so the first
        // stage's inputs k = t + q N/8 load straight into registers
        const long r1 = row0 + 2 * l, r2 = r1 + 1;
        const bool ok1 = r1 < a.nrows, ok2 = r2 < a.nrows;
        const double *in1 = a.in + (ok1 ? r1 : row0) * a.ir, *in2 = a.in + (ok2 ? r2 : row0) * a.ir;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int k = t + q * (N / 8);
            const int kn = k > 0 ? N - k : 0;
            const double a1 = in1[k * a.ik], b1 = in1[kn * a.ik], a2 = in2[k * a.ik], b2 = in2[kn * a.ik];
            v[q] = packed(ok1 ? a1 : 0.0, ok1 && k > 0 ? b1 : 0.0, ok2 ? a2 : 0.0, ok2 && k > 0 ? b2 : 0.0, k);
        }
    } else {
        // x pass: rows are the contiguous direction (ir = 1): staged through
        // LDS with adjacent threads on adjacent rows; all loads issue first
        constexpr int PER = L * N / kNB;
        static_assert(PER * kNB == L * N, "staging");
        double a1[PER], b1[PER], a2[PER], b2[PER];   // x_k and x_{N-k} of both rows
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = threadIdx.x + m * kNB;
            const int ls = e % L, k = e / L;
            const long r = row0 + 2 * ls;
            a1[m] = 0.0; b1[m] = 0.0; a2[m] = 0.0; b2[m] = 0.0;
            // rows r and r + 1 share k (3-D: r and ny even; 2-D: k = 0), so one offset serves both
            const long zo = a.ikz ? (long)((int)r / a.ny) * a.ikz : 0;
            if (r < a.nrows) {
                a1[m] = a.in[k * a.ik + r * a.ir + zo];
                if (k > 0) b1[m] = a.in[(N - k) * a.ik + r * a.ir + zo];
            }
            if (r + 1 < a.nrows) {
                a2[m] = a.in[k * a.ik + (r + 1) * a.ir + zo];
                if (k > 0) b2[m] = a.in[(N - k) * a.ik + (r + 1) * a.ir + zo];
            }
        }
#pragma unroll
        for (int m = 0; m < PER; m++) {
            const int e = threadIdx.x + m * kNB;
            const int ls = e % L, k = e / L;
            buf[ls * LL + lds_pad(k)] = packed(a1[m], b1[m], a2[m], b2[m], k);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = line[lds_pad(t + q * (N / 8))];
        __syncthreads();   // every thread holds its first-stage inputs before the stage writes
    }
    fft_regs<N, 1>(v, line, t, tw);
    // y_{2m} = u_m, y_{2m+1} = u_{N-1-m}: position p holds y_i,
    // i = 2p (p < N/2), 2(N-1-p)+1 (p >= N/2)
#pragma unroll
    for (int h = 0; h < 2; h++) {
        const long r = row0 + 2 * l + h;
        if (r >= a.nrows) continue;
        if constexpr (DST == 0) {
            const int ri = (int)r, j = ri % (int)a.ny, k = ri / (int)a.ny;   // 32-bit division
            int jd[3], kd[3], nj = 1, nk = 1;
            jd[0] = j + a.gl;
            kd[0] = k + a.gl;
            if (a.img) {
                nj = copy_rows(j + 1, a.ny, a.yper, jd);
                nk = copy_rows(k + 1, a.nz, a.zper, kd);
            }
            for (int dk = 0; dk < nk; dk++)
                for (int dj = 0; dj < nj; dj++) {
                    double *orow = a.out + a.gl + a.sx * jd[dj] + a.sy * kd[dk];
#pragma unroll
                    for (int q = 0; q < 8; q++) {
                        const int p = t + q * (N / 8);
                        const int i = p < N / 2 ? 2 * p : 2 * (N - 1 - p) + 1;
                        orow[i] = (h ? v[q].y : v[q].x) * a.post;
                    }
                    // Neumann x ghosts: phi(0) = phi(1) is y_0 (position 0),
                    // phi(N+1) = phi(N) is y_{N-1} (position N/2): thread 0's q = 0, 4
                    if (a.img && t == 0) {
                        orow[-1] = (h ? v[0].y : v[0].x) * a.post;
                        orow[N] = (h ? v[4].y : v[4].x) * a.post;
                    }
                }
        } else {
            double *orow = a.out + r * N;
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const int p = t + q * (N / 8);
                const int i = p < N / 2 ? 2 * p : 2 * (N - 1 - p) + 1;
                orow[i] = (h ? v[q].y : v[q].x) * a.post;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Real FFT along contiguous lines of length N: A[r*N + j] -> B[r*H + jy],
// H = N/2 + 1, times post (npn, y direction), and back (c2r, the imaginary
// parts of bins 0 and N/2 ignored).
// ---------------------------------------------------------------------------
struct RfftArgs {
    double *re;        // real lines
    double2 *cx;       // half-spectrum lines
    const double2 *tw;
    long nrows;
    double post;
};

template <int N>
__global__ __launch_bounds__(kNB) void rfft_rows_kernel(RfftArgs a)
{
    constexpr int T = nb_tpl<N>(), L = nb_lines<N>(), TJ = 2 * L, LL = lds_line<N>(), H = N / 2 + 1;
    __shared__ double2 buf[L * LL];
    __shared__ double2 tw[N];
    load_twiddles<N>(tw, a.tw);
    const long row0 = (long)blockIdx.x * TJ;
    // rows 2l, 2l+1 at the first stage's positions straight into registers;
    // past the end a row reads the block's first (value dropped)
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long r1 = row0 + 2 * l, r2 = r1 + 1;
    const bool ok1 = r1 < a.nrows, ok2 = r2 < a.nrows;
    const double *in1 = a.re + (ok1 ? r1 : row0) * N, *in2 = a.re + (ok2 ? r2 : row0) * N;
    double2 v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int j = t + q * (N / 8);
        const double x1 = in1[j], x2 = in2[j];
        v[q] = double2{ok1 ? x1 : 0.0, ok2 ? x2 : 0.0};
    }
    double2 *line = buf + l * LL;
    fft_regs<N, -1>(v, line, t, tw);
    // bins k and N - k of the packed line: through LDS
#pragma unroll
    for (int q = 0; q < 8; q++) line[lds_pad(t + q * (N / 8))] = v[q];
    __syncthreads();
    for (int e = threadIdx.x; e < TJ * H; e += kNB) {
        const int row = e / H, k = e % H;
        const long r = row0 + row;
        if (r >= a.nrows) continue;
        const double2 *ln = buf + (row >> 1) * LL;
        const double2 zk = ln[lds_pad(k)], zn = ln[lds_pad((N - k) & (N - 1))];
        double2 X;
        if ((row & 1) == 0) X = double2{(zk.x + zn.x) * 0.5, (zk.y - zn.y) * 0.5};
        else X = double2{(zk.y + zn.y) * 0.5, (zn.x - zk.x) * 0.5};
        a.cx[r * H + k] = double2{X.x * a.post, X.y * a.post};
    }
}

template <int N>
__global__ __launch_bounds__(kNB) void irfft_rows_kernel(RfftArgs a)
{
    constexpr int T = nb_tpl<N>(), L = nb_lines<N>(), LL = lds_line<N>(), H = N / 2 + 1;
    __shared__ double2 buf[L * LL];
    __shared__ double2 tw[N];
    load_twiddles<N>(tw, a.tw);   // first read after the first stage's barrier
    const long row0 = (long)blockIdx.x * 2 * L;
    // the packed line of rows 2l, 2l+1 at the first stage's positions p:
    // bin k = p (p <= N/2) or its mirror N - p, loaded straight from the
    // half spectra (contiguous in k); past the end a row reads the block's first
    const int l = threadIdx.x / T, t = threadIdx.x % T;
    const long r1 = row0 + 2 * l, r2 = r1 + 1;
    const bool ok1 = r1 < a.nrows, ok2 = r2 < a.nrows;
    const double2 *c1 = a.cx + (ok1 ? r1 : row0) * H, *c2 = a.cx + (ok2 ? r2 : row0) * H;
    double2 v[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int p = t + q * (N / 8);
        const int k = p <= N / 2 ? p : N - p;
        const double2 y1 = c1[k], y2 = c2[k];
        const double2 X1 = ok1 ? y1 : double2{0.0, 0.0};
        const double2 X2 = ok2 ? y2 : double2{0.0, 0.0};
        if (p == 0 || p == N / 2) v[q] = double2{X1.x, X2.x};   // Im of bins 0 and N/2 ignored
        else if (p < N / 2) v[q] = double2{X1.x - X2.y, X1.y + X2.x};
        else v[q] = double2{X1.x + X2.y, X2.x - X1.y};
    }
    double2 *line = buf + l * LL;
    fft_regs<N, 1>(v, line, t, tw);
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const int j = t + q * (N / 8);
        if (ok1) a.re[r1 * N + j] = v[q].x;
        if (ok2) a.re[r2 * N + j] = v[q].y;
    }
}

// ---------------------------------------------------------------------------
// Batched Thomas sweeps over strided columns.  Column t = o*ni + q starts at
// o*os + q and has stride s and length n; its diagonal shift is mwx[o]
// (+ mwy[q] in 3-D).  E is double (real spectra) or double2 (complex).
// ASSOC2D: the 2-D solvers' form (src/poisson.f90:346-384): divisions, c1
// denominator (b - a c1) + mx, d denominator (b + mx) - a c1.  Otherwise
// the 3-D form (:1094-1135, 1236-1280): reciprocal pivots, last row divides.
// c1 lives in an FP64 array laid out like the column elements.  It depends
// on the coefficients and the column's wavenumbers only, so it is computed
// once when the context is created (thomas_c1_kernel, the same expressions);
// the sweeps carry their own c1 recurrence in registers for the pivots and
// read the stored c1 in the back substitution.
// ---------------------------------------------------------------------------
struct ThomasGArgs {
    void *x;
    double *c1;
    long ncol, ni, os, s;
    int n;
    const double *mwx, *mwy, *ta, *tb, *tc;
};

__device__ __forceinline__ double sub_(double a, double b) { return a - b; }
__device__ __forceinline__ double2 sub_(double2 a, double2 b) { return double2{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ double mul_(double a, double f) { return a * f; }
__device__ __forceinline__ double2 mul_(double2 a, double f) { return double2{a.x * f, a.y * f}; }
__device__ __forceinline__ double div_(double a, double f) { return a / f; }
__device__ __forceinline__ double2 div_(double2 a, double f) { return double2{a.x / f, a.y / f}; }
template <class E> __device__ __forceinline__ E zero_();
template <> __device__ __forceinline__ double zero_<double>() { return 0.0; }
template <> __device__ __forceinline__ double2 zero_<double2>() { return double2{0.0, 0.0}; }

constexpr int kGU = 8;   // rows loaded ahead per batch in the strided sweeps (16 and 32 measured slower)

// SPLIT (E = double on complex columns): the real and imaginary parts of a
// column are swept by two adjacent threads.  The coefficients are real, so
// each part's recurrence is the complex one's for that part, bit for bit,
// and the sweep runs on twice as many threads (the npn columns number only
// nx (ny/2+1), half a wave per SIMD at 256^3).
template <class E, int ASSOC2D, int SPLIT = 0>
__global__ __launch_bounds__(kNB) void thomas_general_kernel(ThomasGArgs a)
{
    const long tt = (long)blockIdx.x * kNB + threadIdx.x;
    if (tt >= (SPLIT ? 2 : 1) * a.ncol) return;
    const long t = SPLIT ? tt >> 1 : tt;
    const long o = t / a.ni, q = t % a.ni;
    const long base = o * a.os + q, s = SPLIT ? 2 * a.s : a.s, cs = a.s;   // x and c1 strides
    E *x = SPLIT ? static_cast<E *>(a.x) + 2 * base + (tt & 1) : static_cast<E *>(a.x) + base;
    double *c1 = a.c1 + base;
    const double mx = a.mwx[o], my = ASSOC2D ? 0.0 : a.mwy[q];
    const int n = a.n;
    const double *A = a.ta, *B = a.tb, *C = a.tc;
    E d;
    double c1p;
    // one forward step on row k with right-hand side r
    auto fwd = [&](int k, E r) {
        if constexpr (ASSOC2D) {
            const double cn = C[k] / ((B[k] - A[k] * c1p) + mx);
            d = div_(sub_(r, mul_(d, A[k])), (B[k] + mx) - A[k] * c1p);
            c1p = cn;
        } else {
            const double factor = 1.0 / (((B[k] + mx) + my) - A[k] * c1p);
            c1p = C[k] * factor;
            d = mul_(sub_(r, mul_(d, A[k])), factor);
        }
        x[k * s] = d;
    };
    if constexpr (ASSOC2D) {
        c1p = C[0] / (B[0] + mx);
        d = div_(x[0], B[0] + mx);
    } else {
        const double factor = 1.0 / ((B[0] + mx) + my);
        c1p = C[0] * factor;
        d = mul_(x[0], factor);
    }
    x[0] = d;
    // the loads of kGU rows are issued together ahead of their steps, so
    // kGU memory latencies overlap instead of one per step
    int k = 1;
    for (; k + kGU <= n - 1; k += kGU) {
        E r[kGU];
#pragma unroll
        for (int u = 0; u < kGU; u++) r[u] = x[(long)(k + u) * s];
#pragma unroll
        for (int u = 0; u < kGU; u++) fwd(k + u, r[u]);
    }
    for (; k < n - 1; k++) fwd(k, x[(long)k * s]);
    if constexpr (ASSOC2D) {
        const double frac = (B[n - 1] + mx) - A[n - 1] * c1p;
        d = frac != 0.0 ? div_(sub_(x[(n - 1) * s], mul_(d, A[n - 1])), frac) : zero_<E>();
    } else {
        const double factor = ((B[n - 1] + mx) + my) - A[n - 1] * c1p;
        d = factor != 0.0 ? div_(sub_(x[(n - 1) * s], mul_(d, A[n - 1])), factor) : zero_<E>();
    }
    x[(n - 1) * s] = d;
    E xn = d;
    k = n - 2;
    for (; k - kGU + 1 >= 0; k -= kGU) {
        E dk[kGU];
        double ck[kGU];
#pragma unroll
        for (int u = 0; u < kGU; u++) {
            dk[u] = x[(long)(k - u) * s];
            ck[u] = c1[(long)(k - u) * cs];
        }
#pragma unroll
        for (int u = 0; u < kGU; u++) {
            xn = sub_(dk[u], mul_(xn, ck[u]));
            x[(long)(k - u) * s] = xn;
        }
    }
    for (; k >= 0; k--) {
        xn = sub_(x[(long)k * s], mul_(xn, c1[(long)k * cs]));
        x[(long)k * s] = xn;
    }
}

template <int ASSOC2D> __global__ __launch_bounds__(kNB) void thomas_c1_kernel(ThomasGArgs a)
{
    const long t = (long)blockIdx.x * kNB + threadIdx.x;
    if (t >= a.ncol) return;
    const long o = t / a.ni, q = t % a.ni, s = a.s;
    double *c1 = a.c1 + o * a.os + q;
    const double mx = a.mwx[o], my = ASSOC2D ? 0.0 : a.mwy[q];
    const double *A = a.ta, *B = a.tb, *C = a.tc;
    double c1p;
    if constexpr (ASSOC2D) {
        c1p = C[0] / (B[0] + mx);
        c1[0] = c1p;
        for (int k = 1; k < a.n - 1; k++) {
            c1p = C[k] / ((B[k] - A[k] * c1p) + mx);
            c1[k * s] = c1p;
        }
    } else {
        double factor = 1.0 / ((B[0] + mx) + my);
        c1p = C[0] * factor;
        c1[0] = c1p;
        for (int k = 1; k < a.n - 1; k++) {
            factor = 1.0 / (((B[k] + mx) + my) - A[k] * c1p);
            c1p = C[k] * factor;
            c1[k * s] = c1p;
        }
    }
}

// 2-D solvers (ASSOC2D form, contiguous columns): a few hundred columns of up
// to 1024 cells, so one thread per column leaves the sweeps bound by memory
// latency, and the dependent chain of each step bounds any layout.  One wave
// per column stages the column, its stored c1 (thomas_c1_kernel) and the
// coefficients in LDS.  With c1 known the pivot (b_k + mx) - a_k c1_{k-1} is
// off the chain, which keeps one multiply, one subtract and one division per
// step; the real and imaginary parts of a complex column are independent
// chains, run by lanes 0 and 1.  Each lane reads its operands one step ahead
// of the step's store.  Same arithmetic as thomas_general_kernel<E, 1>.
// Dynamic LDS: E x[n], then double c1[n], A[n], B[n].
template <class E> __global__ __launch_bounds__(64) void thomas_column_kernel(ThomasGArgs a)
{
    extern __shared__ double smem[];
    constexpr int NC = sizeof(E) / sizeof(double);   // components per element
    const int n = a.n;
    const long o = blockIdx.x;
    const E *xg = static_cast<const E *>(a.x) + o * a.os;
    const double *c1g = a.c1 + o * a.os;
    double *x = smem;                 // element k, component l at x[NC * k + l]
    double *c1 = smem + (long)n * NC;
    double *A = c1 + n, *B = A + n;
    for (int k = threadIdx.x; k < n; k += 64) {
        reinterpret_cast<E *>(x)[k] = xg[k];
        c1[k] = c1g[k];
        A[k] = a.ta[k];
        B[k] = a.tb[k];
    }
    __syncthreads();
    const int l = threadIdx.x;
    if (l < NC) {
        const double mx = a.mwx[o];
        double d = x[l] / (B[0] + mx);
        x[l] = d;
        // operands of row k held one row ahead
        double Ak = A[1], Bk = B[1], cp = c1[0], xk = x[NC + l];
        for (int k = 1; k < n - 1; k++) {
            const double An = A[k + 1], Bn = B[k + 1], cn = c1[k], xn = x[NC * (k + 1) + l];
            d = (xk - d * Ak) / ((Bk + mx) - Ak * cp);
            x[NC * k + l] = d;
            Ak = An; Bk = Bn; cp = cn; xk = xn;
        }
        const double frac = (Bk + mx) - Ak * cp;   // row n-1 (n >= 2)
        d = frac != 0.0 ? (xk - d * Ak) / frac : 0.0;
        x[NC * (n - 1) + l] = d;
        double xr = d, dk = x[NC * (n - 2) + l], ck = c1[n - 2];
        for (int k = n - 2; k >= 0; k--) {
            const int kp = k > 0 ? k - 1 : 0;
            const double dn = x[NC * kp + l], cnx = c1[kp];
            xr = dk - xr * ck;
            x[NC * k + l] = xr;
            dk = dn; ck = cnx;
        }
    }
    __syncthreads();
    E *xo = static_cast<E *>(a.x) + o * a.os;
    for (int k = threadIdx.x; k < n; k += 64) xo[k] = reinterpret_cast<const E *>(x)[k];
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
#define FEN_NB_DISPATCH(n, MACRO)                                        \
    switch (n) {                                                         \
    case 8: MACRO(8); break;                                             \
    case 16: MACRO(16); break;                                           \
    case 32: MACRO(32); break;                                           \
    case 64: MACRO(64); break;                                           \
    case 128: MACRO(128); break;                                         \
    case 256: MACRO(256); break;                                         \
    case 512: MACRO(512); break;                                         \
    case 1024: MACRO(1024); break;                                       \
    default: throw std::runtime_error("unsupported transform length");   \
    }

static bool aligned16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

template <int N, int SRC, int MKL> static void launch_dct2_pairs(unsigned grid, hipStream_t s, const DctArgs &a)
{
    if constexpr (N <= 512) hipLaunchKernelGGL((dct2_pairs_kernel<N, SRC, MKL>), dim3(grid), dim3(kNB), 0, s, a);
    else throw std::logic_error("dct2_pairs_kernel takes N <= 512");
}

static unsigned row_blocks(long nrows, int N) { return (unsigned)((nrows + 2 * (kNB / (N / 8)) - 1) / (2 * (kNB / (N / 8)))); }

// x DCT-II of the rows of div(u*)*rho/dt (vel != nullptr) or of phi (direct)
// into the real spectrum A[kx*nrows + r], times post.
void launch_dct2_x(fen_ctx *c, const double *const *vel, const double *direct, long sx, long sy, int gl, double rho,
                   double dt, double post)
{
    DctArgs a{};
    if (vel) { a.u = vel[0]; a.v = vel[1]; a.w = vel[2]; }
    else { a.u = direct; a.sx = sx; a.sy = sy; a.gl = gl; }
    a.out = as<double>(c->d_spec);
    a.tw = as<double2>(c->d_twx);
    a.mk = c->d_mkx;
    a.b = c->box;
    a.id = 1.0 / c->delta;
    a.rho = rho;
    a.dt = dt;
    a.post = post;
    a.nrows = (long)c->ny * c->nz;
    // spectrum [k][kx][j]: element kx of row r = j + ny k at kx ny + j + k nx ny
    a.ok = c->ny;
    a.orr = 1;
    a.okz = (long)(c->nx - 1) * c->ny;
    a.d3 = c->dim == 3;
    const unsigned grid = row_blocks(a.nrows, c->nx);
    prof_begin(c, K_DIV_R2C);
    // the register-loading form: N <= 512, 16-byte aligned pairs (even gx, aligned fields)
    const bool pairs = vel && c->nx <= 512 && c->box.gx % 2 == 0 && aligned16(a.u) && aligned16(a.v) &&
                       (!a.d3 || aligned16(a.w));
#define L_D2(NN)                                                                              \
    if (pairs) launch_dct2_pairs<NN, 0, 0>(grid, c->stream, a);                               \
    else if (vel) hipLaunchKernelGGL((dct2_rows_kernel<NN, 0>), dim3(grid), dim3(kNB), 0, c->stream, a); \
    else hipLaunchKernelGGL((dct2_rows_kernel<NN, 1>), dim3(grid), dim3(kNB), 0, c->stream, a);
    FEN_NB_DISPATCH(c->nx, L_D2)
#undef L_D2
    prof_end(c, K_DIV_R2C, (vel ? (c->dim == 3 ? 3.0 : 2.0) : 1.0) * 8.0 * c->n_int + 8.0 * c->n_int);
}

bool dct3_writes_ghosts(const fen_ctx *c)
{
    if (c->box.dim != 3 || c->periodic[0]) return false;
    for (int ax = 0; ax < 3; ax++)
        for (int side = 0; side < 2; side++) {
            const int t = c->bc_type[FEN_PHI][2 * ax + side];
            if (c->periodic[ax] ? t != 0 : t != 2) return false;
        }
    return true;
}

// x DCT-III of A[kx*nrows + r] into the interior rows of phi, times post;
// with img (gl == 1 and dct3_writes_ghosts) also every ghost of phi.
void launch_dct3_x(fen_ctx *c, double *phi, int gl, long sx, long sy, double post, bool img)
{
    Dct3Args a{};
    a.in = as<double>(c->d_spec);
    a.out = phi;
    a.tw = as<double2>(c->d_twx);
    a.mk = c->d_mkx;
    a.nrows = (long)c->ny * c->nz;
    a.ik = c->ny;   // spectrum [k][kx][j], as the x DCT-II writes it
    a.ir = 1;
    a.ikz = (long)(c->nx - 1) * c->ny;
    a.sx = sx;
    a.sy = sy;
    a.gl = gl;
    a.ny = c->ny;
    a.nz = c->nz;
    if (img && (gl != 1 || !dct3_writes_ghosts(c))) throw std::runtime_error("DCT-III ghost writes need gl == 1");
    a.img = img ? 1 : 0;
    a.yper = c->periodic[1] ? 1 : 0;
    a.zper = c->periodic[2] ? 1 : 0;
    a.post = post;
    const unsigned grid = row_blocks(a.nrows, c->nx);
    prof_begin(c, K_C2R);
#define L_D3(NN) hipLaunchKernelGGL((dct3_rows_kernel<NN, 0>), dim3(grid), dim3(kNB), 0, c->stream, a);
    FEN_NB_DISPATCH(c->nx, L_D3)
#undef L_D3
    prof_end(c, K_C2R, 16.0 * c->n_int);
}

// y DCT-II / DCT-III in place on the real spectrum lines A[(kx*nz + k)*ny + j] (nnn).
void launch_dct_y(fen_ctx *c, int inverse, double post)
{
    const long nrows = (long)c->nx * c->nz;
    const unsigned grid = row_blocks(nrows, c->ny);
    const int kid = inverse ? K_Y_INV : K_Y_FWD;
    prof_begin(c, kid);
    if (!inverse) {
        DctArgs a{};
        a.u = as<double>(c->d_spec);
        a.out = as<double>(c->d_spec);
        a.tw = as<double2>(c->d_twy);
        a.mk = c->d_mky;
        a.post = post;
        a.nrows = nrows;
        a.ok = 1;
        a.orr = c->ny;
        // the register-loading form for N <= 512, with the DCT's (cos, sin) table in LDS
#define L_DY(NN)                                                                              \
    if (c->ny <= 512) launch_dct2_pairs<NN, 2, 1>(grid, c->stream, a);                        \
    else hipLaunchKernelGGL((dct2_rows_kernel<NN, 2>), dim3(grid), dim3(kNB), 0, c->stream, a);
        FEN_NB_DISPATCH(c->ny, L_DY)
#undef L_DY
    } else {
        Dct3Args a{};
        a.in = as<double>(c->d_spec);
        a.out = as<double>(c->d_spec);
        a.tw = as<double2>(c->d_twy);
        a.mk = c->d_mky;
        a.nrows = nrows;
        a.ik = 1;
        a.ir = c->ny;
        a.post = post;
#define L_DY3(NN) hipLaunchKernelGGL((dct3_rows_kernel<NN, 1>), dim3(grid), dim3(kNB), 0, c->stream, a);
        FEN_NB_DISPATCH(c->ny, L_DY3)
#undef L_DY3
    }
    prof_end(c, kid, 16.0 * c->n_int);
}

// y real FFT (forward: A -> B, times post) or its inverse (B -> A) (npn).
void launch_rfft_y(fen_ctx *c, int inverse, double post)
{
    RfftArgs a{};
    a.re = as<double>(c->d_spec);
    a.cx = c->d_spec2;
    a.tw = as<double2>(c->d_twy);
    a.nrows = (long)c->nx * c->nz;
    a.post = post;
    const unsigned grid = row_blocks(a.nrows, c->ny);
    const int kid = inverse ? K_Y_INV : K_Y_FWD;
    prof_begin(c, kid);
#define L_RF(NN)                                                                             \
    if (inverse) hipLaunchKernelGGL((irfft_rows_kernel<NN>), dim3(grid), dim3(kNB), 0, c->stream, a); \
    else hipLaunchKernelGGL((rfft_rows_kernel<NN>), dim3(grid), dim3(kNB), 0, c->stream, a);
    FEN_NB_DISPATCH(c->ny, L_RF)
#undef L_RF
    prof_end(c, kid, 16.0 * c->n_int);
}

// Column layout of the pn, nn, npn and nnn sweeps (and of ppn's, whose
// sweep kernel is thomas_kernel): element type and association form.
static ThomasGArgs thomas_layout(fen_ctx *c, bool *complex_out, bool *assoc2d_out)
{
    ThomasGArgs a{};
    a.c1 = c->d_c1;
    a.mwx = c->d_mwx;
    a.mwy = c->d_mwy;
    a.ta = c->d_ta;
    a.tb = c->d_tb;
    a.tc = c->d_tc;
    const int hy = c->ny / 2 + 1;
    bool complex_e = false, assoc2d = false;
    switch (c->solver) {
    case FEN_SOLVER_PN:   // complex [kx][j], columns along j
        a.x = c->d_spec; complex_e = true; assoc2d = true;
        a.ncol = c->hx; a.ni = 1; a.os = c->ny; a.s = 1; a.n = c->ny;
        break;
    case FEN_SOLVER_NN:   // real [kx][j]
        a.x = c->d_spec; assoc2d = true;
        a.ncol = c->nx; a.ni = 1; a.os = c->ny; a.s = 1; a.n = c->ny;
        break;
    case FEN_SOLVER_NPN:  // complex [k][kx][jy], columns along k: column kx hy + jy, one k row contiguous
        a.x = c->d_spec2; complex_e = true;
        a.ncol = (long)c->nx * hy; a.ni = hy; a.os = hy; a.s = (long)c->nx * hy; a.n = c->nz;
        break;
    case FEN_SOLVER_NNN:  // real [k][kx][j]
        a.x = c->d_spec;
        a.ncol = (long)c->nx * c->ny; a.ni = c->ny; a.os = c->ny; a.s = (long)c->nx * c->ny; a.n = c->nz;
        break;
    case FEN_SOLVER_PPN:  // complex [kx][k][j] (thomas_kernel)
        a.x = c->d_spec; complex_e = true;
        a.ncol = (long)c->hx * c->ny; a.ni = c->ny; a.os = (long)c->ny * c->nz; a.s = c->ny; a.n = c->nz;
        break;
    default: throw std::runtime_error("not a tridiagonal solver");
    }
    if (a.n < 2) throw std::runtime_error("the tridiagonal direction needs at least 2 cells");
    *complex_out = complex_e;
    *assoc2d_out = assoc2d;
    return a;
}

// c1 of every column, once per context (after the coefficients are uploaded).
void launch_thomas_c1(fen_ctx *c)
{
    bool complex_e = false, assoc2d = false;
    const ThomasGArgs a = thomas_layout(c, &complex_e, &assoc2d);
    const dim3 grid((unsigned)((a.ncol + kNB - 1) / kNB));
    if (assoc2d) hipLaunchKernelGGL(thomas_c1_kernel<1>, grid, dim3(kNB), 0, c->stream, a);
    else hipLaunchKernelGGL(thomas_c1_kernel<0>, grid, dim3(kNB), 0, c->stream, a);
}

// Thomas sweeps of the pn, nn, npn and nnn solvers.
void launch_thomas_neumann(fen_ctx *c)
{
    bool complex_e = false, assoc2d = false;
    if (c->solver == FEN_SOLVER_PPN) throw std::runtime_error("ppn sweeps use thomas_kernel");
    ThomasGArgs a = thomas_layout(c, &complex_e, &assoc2d);
    const dim3 grid((unsigned)((a.ncol + kNB - 1) / kNB));
    prof_begin(c, K_THOMAS);
    const size_t lds = (size_t)a.n * ((complex_e ? sizeof(double2) : sizeof(double)) + 3 * sizeof(double));
    if (assoc2d && a.s == 1 && a.os == a.n && lds <= 65536) {
        // contiguous columns that fit in LDS: one wave each
        if (complex_e)
            hipLaunchKernelGGL(thomas_column_kernel<double2>, dim3((unsigned)a.ncol), dim3(64), lds, c->stream, a);
        else
            hipLaunchKernelGGL(thomas_column_kernel<double>, dim3((unsigned)a.ncol), dim3(64), lds, c->stream, a);
    } else if (complex_e) {
        // real and imaginary parts on adjacent threads
        const dim3 g2((unsigned)((2 * a.ncol + kNB - 1) / kNB));
        if (assoc2d) hipLaunchKernelGGL((thomas_general_kernel<double, 1, 1>), g2, dim3(kNB), 0, c->stream, a);
        else hipLaunchKernelGGL((thomas_general_kernel<double, 0, 1>), g2, dim3(kNB), 0, c->stream, a);
    }
    else if (assoc2d) hipLaunchKernelGGL((thomas_general_kernel<double, 1>), grid, dim3(kNB), 0, c->stream, a);
    else hipLaunchKernelGGL((thomas_general_kernel<double, 0>), grid, dim3(kNB), 0, c->stream, a);
    // columns read and written in each sweep, c1 read in the back substitution (LDS kernel: once each)
    const double eb = complex_e ? 16.0 : 8.0, cells = (double)a.ncol * a.n;
    prof_end(c, K_THOMAS, (assoc2d && a.s == 1 && lds <= 65536 ? 2.0 * eb : 4.0 * eb + 8.0) * cells);
}

}  // namespace fen
