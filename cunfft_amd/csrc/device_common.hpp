// device_common.hpp -- types and window arithmetic shared by the gfx950 kernels.
//
// Window and index rule follow the reference exactly:
//   phi(t) = (pi b)^-1/2 exp(-t^2 / b)                 src/cunfft_kernel.cu:162-166
//   u      = rint(n x) - signbit(x) - m  (half-even)   src/cunfft_kernel.cu:175-182
//   taps   l = u .. u + 2m + 1, wrapped with & (n-1)    src/cunfft_kernel.cu:528-531
// The split of n*x into integer cell and fraction is always done in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cunfft {

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx<T>::type;

// Per-transform geometry passed by value to every gridding kernel.
struct Geom {
    int d;
    int m;
    int W;                // 2m + 2 taps per axis
    int64_t N[3];
    int64_t n[3];
    int64_t M;
    double invb[3];       // 1 / b_t
    double nrm[3];        // (pi b_t)^-1/2
};

__device__ __forceinline__ int sign_bit(double x)
{
    return (int)((uint64_t)__double_as_longlong(x) >> 63);
}

// First window index u and the FP64 fraction c = n x - u (c in [m-1/2, m+1/2]
// up to the sign-bit shift).  rint rounds half to even like llrint.
__device__ __forceinline__ void window_start(double x, int64_t n, int m, int64_t &u,
                                             double &c)
{
    const double nx = (double)n * x;
    u = (int64_t)rint(nx) - sign_bit(x) - m;
    c = nx - (double)u;
}

// Index wrap onto [0, n) for any grid length n (the reference masks with
// n - 1, src/cunfft_kernel.cu:528-531, which needs a power of two).
// wrap_lo: a in [-n, n); wrap_hi: a in [0, 2n).  Window starts lie in
// [-n/2 - m - 1, n/2 - m + 1] and n >= 2m + 2, so one step suffices.
__device__ __forceinline__ int64_t wrap_lo(int64_t a, int64_t n) { return a < 0 ? a + n : a; }
__device__ __forceinline__ int64_t wrap_hi(int64_t a, int64_t n) { return a >= n ? a - n : a; }
__device__ __forceinline__ uint32_t wrap_hi32(uint32_t a, uint32_t n) { return a >= n ? a - n : a; }

template <typename T> __device__ __forceinline__ T win_exp(T v);
template <> __device__ __forceinline__ float win_exp<float>(float v) { return expf(v); }
template <> __device__ __forceinline__ double win_exp<double>(double v) { return exp(v); }

// w[l] = phi(c - l) for l = 0..W-1 in precision T; the offset c - l is formed
// in FP64 and rounded once.
template <typename T, int W>
__device__ __forceinline__ void window_weights(double c, double invb, double nrm, T (&w)[W])
{
#pragma unroll
    for (int l = 0; l < W; ++l) {
        const T s = (T)(c - (double)l);
        w[l] = win_exp<T>(-s * s * (T)invb) * (T)nrm;
    }
}

// Weights and their x-derivatives: d/dx phi(n x - l) = n * (-2 s / b) phi(s).
template <typename T, int W>
__device__ __forceinline__ void window_weights_dx(double c, double invb, double nrm, int64_t n,
                                                  T (&w)[W], T (&dw)[W])
{
#pragma unroll
    for (int l = 0; l < W; ++l) {
        const T s = (T)(c - (double)l);
        const T p = win_exp<T>(-s * s * (T)invb) * (T)nrm;
        w[l] = p;
        dw[l] = (T)n * ((T)-2.0 * s * (T)invb) * p;
    }
}

template <typename T> __device__ __forceinline__ cplx_t<T> cmake(T a, T b)
{
    cplx_t<T> r;
    r.x = a;
    r.y = b;
    return r;
}

}  // namespace cunfft
