// device_common.hpp -- types and window arithmetic shared by the gfx950 kernels.
//
// Window and index rule follow the reference exactly:
//   phi(t) = (pi b)^-1/2 exp(-t^2 / b)                 src/cunfft_kernel.cu:162-166
//   u      = rint(n x) - signbit(x) - m  (half-even)   src/cunfft_kernel.cu:175-182
//   taps   l = u .. u + 2m + 1, wrapped mod n           src/cunfft_kernel.cu:528-531
// The split of n*x into integer cell and fraction is always done in FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cunfft {

template <typename T> struct cplx;
template <> struct cplx<float> { using type = float2; };
template <> struct cplx<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx<T>::type;

// Windows: the reference's Gaussian, and NFFT3's Kaiser-Bessel and sinc power.
enum Window : int { WIN_GAUSSIAN = 0, WIN_KAISER_BESSEL = 1, WIN_SINC = 2 };

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
    double shp[3];        // Kaiser-Bessel: b_t = pi (2 - 1 / sigma_t); sinc: a_t (sinc_phi)
};

__host__ __device__ __forceinline__ int sign_bit(double x) { return __builtin_signbit(x) ? 1 : 0; }

// Index wrap onto [0, n) for any grid length n (the reference masks with
// n - 1, src/cunfft_kernel.cu:528-531, which needs a power of two).
// wrap_hi: a in [0, 2n) (a wrapped window start plus a tap offset < n).
__device__ __forceinline__ int64_t wrap_hi(int64_t a, int64_t n) { return a >= n ? a - n : a; }
__device__ __forceinline__ uint32_t wrap_hi32(uint32_t a, uint32_t n) { return a >= n ? a - n : a; }

// Floor modulo a mod n in [0, n) for any a; the division runs only outside
// [-n, n), i.e. for nodes outside [-1/2, 1/2).
__host__ __device__ __forceinline__ int64_t wrap_mod(int64_t a, int64_t n)
{
    if (a < -n || a >= n) a %= n;
    return a < 0 ? a + n : a;
}

// First window cell u mod n in [0, n) and the FP64 fraction c = n x - u for
// any node x.  For x in [-1/2, 1/2) this is the reference's rule exactly; for
// other finite x it is the reference's too, whose mask with n - 1
// (src/cunfft_kernel.cu:528-531) is u mod n for a power-of-two n.  A
// non-finite or huge n x (|rint(n x)| >= 2^52, where n x has no fraction
// left) is given the cell of x = 0 and keeps its non-finite fraction: its
// taps are NaN or zero, never an index outside the grid.  rint rounds half to
// even like the reference's llrint.
__host__ __device__ __forceinline__ int64_t window_cell(double x, int64_t n, int m, double &c)
{
    const double nx = (double)n * x;
    double r = rint(nx);
    r = fabs(r) < 0x1p52 ? r : 0.0;   // false for NaN as well
    const int64_t u = (int64_t)r - sign_bit(x) - m;
    c = nx - (double)u;
    return wrap_mod(u, n);
}

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

// Sinc-power window in grid units (NFFT3's SINC_POWER; the reference is
// Gaussian-only):
//   phi(s) = a sinc(pi a s)^(2m),   sinc(u) = sin(u) / u,   a = (2 sigma - 1) / (2 sigma m),
// whose Fourier transform is the centred cardinal B-spline M_2m(nu / a) of
// order 2m, zero for |nu| >= m a = 1 - 1/(2 sigma): the oversampled grid's
// first alias of the band edge.  The roll-off divides by M_2m(k / (n a)).
template <typename T> __device__ __forceinline__ T sin_pi(T v);
template <> __device__ __forceinline__ float sin_pi<float>(float v) { return sinpif(v); }
template <> __device__ __forceinline__ double sin_pi<double>(double v) { return sinpi(v); }
template <typename T> __device__ __forceinline__ T cos_pi(T v);
template <> __device__ __forceinline__ float cos_pi<float>(float v) { return cospif(v); }
template <> __device__ __forceinline__ double cos_pi<double>(double v) { return cospi(v); }

// v^k for k >= 0 by squaring.
template <typename T> __device__ __forceinline__ T ipow(T v, int k)
{
    T r = (T)1;
    for (; k > 0; k >>= 1) {
        if (k & 1) r *= v;
        v *= v;
    }
    return r;
}

template <typename T> __device__ __forceinline__ T sinc_pi(T t)   // sin(pi t) / (pi t)
{
    return t == (T)0 ? (T)1 : sin_pi<T>(t) / ((T)3.14159265358979323846 * t);
}

template <typename T> __device__ __forceinline__ T sinc_phi(T s, int m, T a)
{
    return a * ipow<T>(sinc_pi<T>(a * s), 2 * m);
}

// d phi / d s = a 2m sinc^(2m-1)(u) pi a g(u), u = pi a s, g(u) = d sinc / du
// = (u cos u - sin u) / u^2, by its series -u/3 + u^3/30 - u^5/840 + u^7/45360
// for |u| < 0.1 (the closed form cancels there).
template <typename T> __device__ __forceinline__ T sinc_dphi(T s, int m, T a)
{
    const T pi = (T)3.14159265358979323846;
    const T t = a * s, u = pi * t;
    T g;
    if (fabs(u) < (T)0.1) {
        const T u2 = u * u;
        g = u * ((T)(-1.0 / 3.0) + u2 * ((T)(1.0 / 30.0) + u2 * ((T)(-1.0 / 840.0) + u2 * (T)(1.0 / 45360.0))));
    } else {
        g = (u * cos_pi<T>(t) - sin_pi<T>(t)) / (u * u);
    }
    return a * (T)(2 * m) * ipow<T>(sinc_pi<T>(t), 2 * m - 1) * pi * a * g;
}

// Window weight of tap offset s along axis t.
template <typename T> __device__ __forceinline__ T window_weight(T s, const Geom &ge, int t)
{
    if (ge.win == WIN_KAISER_BESSEL) return kb_phi<T>(s, (T)ge.m, (T)ge.shp[t]);
    if (ge.win == WIN_SINC) return sinc_phi<T>(s, ge.m, (T)ge.shp[t]);
    return win_exp<T>(-s * s * (T)ge.invb[t]) * (T)ge.nrm[t];
}

// d/dx phi(n x - l) = n phi'(s).
template <typename T>
__device__ __forceinline__ T window_weight_dx(T s, T p, const Geom &ge, int t)
{
    if (ge.win == WIN_KAISER_BESSEL) return (T)ge.n[t] * kb_dphi<T>(s, (T)ge.m, (T)ge.shp[t]);
    if (ge.win == WIN_SINC) return (T)ge.n[t] * sinc_dphi<T>(s, ge.m, (T)ge.shp[t]);
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
