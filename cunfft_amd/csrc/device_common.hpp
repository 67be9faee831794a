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

// Windows: the reference's Gaussian, and NFFT3's Kaiser-Bessel.
enum Window : int { WIN_GAUSSIAN = 0, WIN_KAISER_BESSEL = 1 };

// Per-transform geometry passed by value to every gridding kernel.
struct Geom {
    int d;
    int m;
    int W;                // 2m + 2 taps per axis
    int win;              // Window
    int64_t N[3];
    int64_t n[3];
    int64_t M;
    double invb[3];       // Gaussian: 1 / b_t
    double nrm[3];        // Gaussian: (pi b_t)^-1/2
    double kbb[3];        // Kaiser-Bessel: shape b_t = pi (2 - 1 / sigma_t)
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

// Kaiser-Bessel window in grid units (NFFT3's; the reference is
// Gaussian-only), shape b, support m, analytic continuation past |s| = m:
//   phi(s) = sinh(b r) / (pi r), r = sqrt(m^2 - s^2)        |s| < m
//          = sin(b r) / (pi r),  r = sqrt(s^2 - m^2)        |s| > m
// (b / pi at |s| = m).  Both are (1/pi) sum_j b^{2j+1} d^j / (2j+1)! in
// d = m^2 - s^2, whose d-derivative gives phi' near d = 0.
template <typename T> __device__ __forceinline__ T kb_phi(T s, T m, T b)
{
    const T inv_pi = (T)0.318309886183790671538;
    const T d = m * m - s * s;
    const T r = sqrt(fabs(d));
    if (r == (T)0) return b * inv_pi;
    return inv_pi * (d > (T)0 ? sinh(b * r) : sin(b * r)) / r;
}

template <typename T> __device__ __forceinline__ T kb_dphi(T s, T m, T b)
{
    const T inv_pi = (T)0.318309886183790671538;
    const T d = m * m - s * s;
    const T lim = sizeof(T) == 4 ? (T)1e-2 : (T)1e-4;
    if (fabs(d) < lim) {
        const T b2 = b * b;
        const T g = b * b2 * ((T)1 / (T)6 + b2 * d * ((T)2 / (T)120 + b2 * d * ((T)3 / (T)5040 +
                                                                          b2 * d * (T)4 / (T)362880)));
        return (T)-2 * s * g * inv_pi;
    }
    const T r = sqrt(fabs(d));
    const T br = b * r;
    if (d > (T)0) return -s * (br * cosh(br) - sinh(br)) * inv_pi / (r * r * r);
    return s * (br * cos(br) - sin(br)) * inv_pi / (r * r * r);
}

// Window weight of tap offset s along axis t.
template <typename T> __device__ __forceinline__ T window_weight(T s, const Geom &ge, int t)
{
    if (ge.win == WIN_KAISER_BESSEL) return kb_phi<T>(s, (T)ge.m, (T)ge.kbb[t]);
    return win_exp<T>(-s * s * (T)ge.invb[t]) * (T)ge.nrm[t];
}

// d/dx phi(n x - l) = n phi'(s).
template <typename T>
__device__ __forceinline__ T window_weight_dx(T s, T p, const Geom &ge, int t)
{
    if (ge.win == WIN_KAISER_BESSEL) return (T)ge.n[t] * kb_dphi<T>(s, (T)ge.m, (T)ge.kbb[t]);
    return (T)ge.n[t] * ((T)-2.0 * s * (T)ge.invb[t]) * p;
}

// w[l] = phi(c - l) for l = 0..W-1 along axis t in precision T; the offset
// c - l is formed in FP64 and rounded once.
template <typename T, int W>
__device__ __forceinline__ void window_weights(double c, const Geom &ge, int t, T (&w)[W])
{
#pragma unroll
    for (int l = 0; l < W; ++l) w[l] = window_weight<T>((T)(c - (double)l), ge, t);
}

// Weights and their x-derivatives.
template <typename T, int W>
__device__ __forceinline__ void window_weights_dx(double c, const Geom &ge, int t, T (&w)[W],
                                                  T (&dw)[W])
{
#pragma unroll
    for (int l = 0; l < W; ++l) {
        const T s = (T)(c - (double)l);
        w[l] = window_weight<T>(s, ge, t);
        dw[l] = window_weight_dx<T>(s, w[l], ge, t);
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
