// fft_device.hpp -- LDS-resident FFT building blocks for gfx950.
//
// A line of N complex points (N a power of two, 8 <= N <= 1024) is
// transformed by T = N/8 threads with the Stockham autosort formulation:
// each stage loads radix-R groups x[j + r*N/R] into registers, applies the
// twiddles W_{Ns*R}^{r*(j mod Ns)}, does an R-point DFT in registers and
// stores to (j div Ns)*Ns*R + (j mod Ns) + r*Ns.  Stages use radix 8 and one
// final radix-4 or radix-2 stage when log2(N) is not a multiple of 3.
// Output is in natural order, unnormalised; SIGN = -1 is FFTW_FORWARD
// (exp(-2 pi i jk/N)) and SIGN = +1 is FFTW_BACKWARD.
//
// Twiddles come from a table W_N^m = exp(-2 pi i m/N), m < N, computed once
// on the host in long double and rounded to the element type.
#pragma once

#include <hip/hip_runtime.h>

namespace fen {

template <typename C> __device__ __forceinline__ C c_add(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C c_sub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C c_mul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C> __device__ __forceinline__ C c_conj(C a) { return C{a.x, -a.y}; }

// a * (-i) for SIGN = -1, a * (+i) for SIGN = +1
template <int SIGN, typename C> __device__ __forceinline__ C c_rot(C a)
{
    if constexpr (SIGN < 0) return C{a.y, -a.x};
    else return C{-a.y, a.x};
}

// a * W_8^m (m odd) with W_8 = exp(SIGN * 2 pi i / 8)
template <int SIGN, int M, typename C> __device__ __forceinline__ C c_w8(C a)
{
    using R = decltype(a.x);
    const R h = (R)0.70710678118654752440084436210484903928;
    constexpr int m = ((M % 8) + 8) % 8;
    // forward: W^1 = h(1 - i), W^3 = h(-1 - i); inverse: conjugates
    if constexpr ((m == 1 && SIGN < 0) || (m == 7 && SIGN > 0)) return C{h * (a.x + a.y), h * (a.y - a.x)};
    else if constexpr ((m == 3 && SIGN < 0) || (m == 5 && SIGN > 0)) return C{h * (a.y - a.x), -(h * (a.x + a.y))};
    else if constexpr ((m == 5 && SIGN < 0) || (m == 3 && SIGN > 0)) return C{-(h * (a.x + a.y)), h * (a.x - a.y)};
    else return C{h * (a.x - a.y), h * (a.x + a.y)};   // (m==7 fwd) or (m==1 inv)
}

// Two complex values carried together: two lines transformed by one thread
// share each LDS access (16 bytes for two float2) and each twiddle.  Every
// operation acts on each half as on a single value, so the two lines' results
// are bitwise those of single-line transforms.
template <typename C> struct CPair {
    C a, b;
};
template <typename C> __device__ __forceinline__ CPair<C> c_add(CPair<C> x, CPair<C> y)
{
    return CPair<C>{c_add(x.a, y.a), c_add(x.b, y.b)};
}
template <typename C> __device__ __forceinline__ CPair<C> c_sub(CPair<C> x, CPair<C> y)
{
    return CPair<C>{c_sub(x.a, y.a), c_sub(x.b, y.b)};
}
template <typename C> __device__ __forceinline__ CPair<C> c_mul(CPair<C> x, C w)
{
    return CPair<C>{c_mul(x.a, w), c_mul(x.b, w)};
}
template <int SIGN, typename C> __device__ __forceinline__ CPair<C> c_rot(CPair<C> x)
{
    return CPair<C>{c_rot<SIGN>(x.a), c_rot<SIGN>(x.b)};
}
template <int SIGN, int M, typename C> __device__ __forceinline__ CPair<C> c_w8(CPair<C> x)
{
    return CPair<C>{c_w8<SIGN, M>(x.a), c_w8<SIGN, M>(x.b)};
}

template <int SIGN, typename C> __device__ __forceinline__ void dft4(C &a0, C &a1, C &a2, C &a3)
{
    C t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
    C t2 = c_add(a1, a3), t3 = c_rot<SIGN>(c_sub(a1, a3));
    a0 = c_add(t0, t2); a2 = c_sub(t0, t2);
    a1 = c_add(t1, t3); a3 = c_sub(t1, t3);
}

template <int SIGN, typename C> __device__ __forceinline__ void dft8(C *v)
{
    C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    C o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4<SIGN>(e0, e1, e2, e3);
    dft4<SIGN>(o0, o1, o2, o3);
    o1 = c_w8<SIGN, 1>(o1);
    o2 = c_rot<SIGN>(o2);
    o3 = c_w8<SIGN, 3>(o3);
    v[0] = c_add(e0, o0); v[4] = c_sub(e0, o0);
    v[1] = c_add(e1, o1); v[5] = c_sub(e1, o1);
    v[2] = c_add(e2, o2); v[6] = c_sub(e2, o2);
    v[3] = c_add(e3, o3); v[7] = c_sub(e3, o3);
}

template <int R, int SIGN, typename C> __device__ __forceinline__ void dft_any(C *v)
{
    if constexpr (R == 8) dft8<SIGN>(v);
    else if constexpr (R == 4) dft4<SIGN>(v[0], v[1], v[2], v[3]);
    else { C a = v[0], b = v[1]; v[0] = c_add(a, b); v[1] = c_sub(a, b); }
}

// LDS index with one pad element every 16 to break power-of-two strides.
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
// Stride between lines in LDS: the padded length plus one element, so that
// the same position in consecutive lines falls in different banks (N + N/16
// complex values would be a multiple of 32 banks apart for N >= 32, and the
// staging loops that walk across lines would conflict).
template <int N> constexpr int lds_line() { return N + N / 16 + 1; }

// One Stockham stage and the recursion over the remaining ones.  `line`
// points at this thread's line in LDS (padded indexing), t in [0, N/8).
// All threads of the block must call this (it contains barriers).
template <int N, int NS, int SIGN, typename C>
__device__ __forceinline__ void fft_stages(C *line, int t, const C *tw)
{
    if constexpr (NS < N) {
        constexpr int REM = N / NS;
        constexpr int R = REM >= 8 ? 8 : REM;
        constexpr int T = N / 8;
        constexpr int NB = 8 / R;
        C v[8];
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int j = t + m * T;
#pragma unroll
            for (int r = 0; r < R; r++) v[m * R + r] = line[lds_pad(j + r * (N / R))];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int j = t + m * T;
            const int k = j & (NS - 1);
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    C w = tw[r * k * (N / (NS * R))];
                    if constexpr (SIGN > 0) w.y = -w.y;
                    v[m * R + r] = c_mul(v[m * R + r], w);
                }
            }
            dft_any<R, SIGN>(&v[m * R]);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) line[lds_pad(base + r * NS)] = v[m * R + r];
        }
        __syncthreads();
        fft_stages<N, NS * R, SIGN>(line, t, tw);
    }
}

template <int N, int SIGN, typename C>
__device__ __forceinline__ void fft_lds(C *line, int t, const C *tw)
{
    fft_stages<N, 1, SIGN>(line, t, tw);
}

// The same transform with the line's data in registers at both ends.  The
// first Stockham stage (radix 8, no twiddles) reads x[t + q N/8], q = 0..7,
// and the last stage writes X[t + q N/8]: the same eight positions.  So a
// thread can load its eight inputs straight from global memory into v[q] =
// x[t + q N/8], and receives v[q] = X[t + q N/8]; only the stages in between
// exchange through `line` in LDS (two round trips for N = 128..512 instead
// of four with LDS staging at both ends).  The arithmetic is fft_lds's, so
// the results are bitwise equal.  `line` is this thread's line in LDS
// (padded indexing); all threads of the block must call this (barriers),
// and on return every read of `line` is complete.  C may be a CPair of the
// twiddles' type TW (two lines per thread).
template <int N, int NS, int SIGN, typename C, typename TW>
__device__ __forceinline__ void fft_regs_stages(C (&v)[8], C *line, int t, const TW *tw)
{
    constexpr int REM = N / NS;
    constexpr int R = REM >= 8 ? 8 : REM;
    constexpr int T = N / 8;
    constexpr int NB = 8 / R;
    constexpr bool FIRST = NS == 1, LAST = NS * R == N;
    if constexpr (!FIRST) {
        // this stage's inputs x[j + r N/R], j = t + m T
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int j = t + m * T;
#pragma unroll
            for (int r = 0; r < R; r++) v[m * R + r] = line[lds_pad(j + r * (N / R))];
        }
        __syncthreads();   // every read of this stage precedes any write below (in place)
    }
#pragma unroll
    for (int m = 0; m < NB; m++) {
        const int j = t + m * T;
        const int k = j & (NS - 1);
        if constexpr (NS > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                TW w = tw[r * k * (N / (NS * R))];
                if constexpr (SIGN > 0) w.y = -w.y;
                v[m * R + r] = c_mul(v[m * R + r], w);
            }
        }
        dft_any<R, SIGN>(&v[m * R]);
    }
    if constexpr (LAST) {
        // outputs (j div NS) NS R + k + r NS with j < NS: position t + (m + NB r) N/8
        C out[8];
#pragma unroll
        for (int m = 0; m < NB; m++)
#pragma unroll
            for (int r = 0; r < R; r++) out[m + NB * r] = v[m * R + r];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = out[q];
    } else {
#pragma unroll
        for (int m = 0; m < NB; m++) {
            const int j = t + m * T;
            const int k = j & (NS - 1);
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) line[lds_pad(base + r * NS)] = v[m * R + r];
        }
        __syncthreads();
        fft_regs_stages<N, NS * R, SIGN>(v, line, t, tw);
    }
}

template <int N, int SIGN, typename C, typename TW>
__device__ __forceinline__ void fft_regs(C (&v)[8], C *line, int t, const TW *tw)
{
    fft_regs_stages<N, 1, SIGN>(v, line, t, tw);
    if constexpr (N > 8) __syncthreads();   // the last stage's LDS reads are done
}

// Copies the twiddle table W_N^m (m < N) into LDS.
template <int N, typename C>
__device__ __forceinline__ void load_twiddles(C *tw_lds, const C *tw_glob)
{
    for (int m = threadIdx.x; m < N; m += blockDim.x) tw_lds[m] = tw_glob[m];
}

}  // namespace fen
