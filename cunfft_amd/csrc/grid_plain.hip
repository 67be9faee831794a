// grid_plain.hip -- node-parallel Gaussian gridding kernels for gfx950.
//
// One lane owns one node j.  Forward (gather, src/cunfft_kernel.cu:514-1093):
//   f_j = sum_{l in window(x_j)} g_l prod_t phi(n_t x_jt - l_t)
// Adjoint (scatter, src/cunfft_kernel.cu:608-1197):
//   g_l += f_j prod_t phi(n_t x_jt - l_t)   with hardware float/double atomics
// instead of the reference's CAS loops (src/cunfft_kernel.cu:97-130).
// The derivative gather (no reference counterpart) evaluates
//   d f_j / d x_jt = sum_l g_l n_t phi'(n_t x_jt - l_t) prod_{t'!=t} phi(..)
// for the same truncated window and contracts it with an upstream vector.
//
// The taps per axis W = 2m+2 are a template parameter for m = 1..8 (weights in
// registers, loops unrolled); larger m uses the WT = 0 kernels, which keep
// the weights in a private array indexed at run time.
#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include "internal.hpp"

namespace cunfft {

constexpr int kMaxW = 2 * 16 + 2;
constexpr int kBlock = 256;

// Weights for one node along every axis.  WS is the array extent, W the taps.
template <typename T, int D, int WS>
struct NodeWin {
    T w[D][WS];
    int64_t u[D];
};

template <typename T, int D, int WS>
__device__ __forceinline__ void load_window(const Geom &ge, const double *__restrict__ x,
                                            int64_t j, int W, NodeWin<T, D, WS> &nw)
{
#pragma unroll
    for (int t = 0; t < D; ++t) {
        double c;
        nw.u[t] = window_cell(x[j * D + t], ge.n[t], ge.m, c);
        if constexpr (WS <= 18) {
            window_weights<T, WS>(c, ge, t, nw.w[t]);
        } else {
            for (int l = 0; l < W; ++l) nw.w[t][l] = window_weight<T>((T)(c - (double)l), ge, t);
        }
    }
}

// ---------------------------------------------------------------- gather --
template <typename T, int D, int WT>
__global__ __launch_bounds__(kBlock) void k_gather_plain(const cplx_t<T> *__restrict__ g,
                                                         const double *__restrict__ x,
                                                         cplx_t<T> *__restrict__ f, Geom ge)
{
    constexpr int WS = WT ? WT : kMaxW;
    const int W = WT ? WT : ge.W;
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= ge.M) return;
    NodeWin<T, D, WS> nw;
    load_window<T, D, WS>(ge, x, j, W, nw);
    const int64_t n0 = ge.n[0];
    T sr = 0, si = 0;
    if constexpr (D == 1) {
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            const cplx_t<T> v = g[wrap_hi(nw.u[0] + a, n0)];
            sr += v.x * nw.w[0][a];
            si += v.y * nw.w[0][a];
        }
    } else if constexpr (D == 2) {
        const int64_t n1 = ge.n[1];
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            const cplx_t<T> *row = g + (wrap_hi(nw.u[0] + a, n0)) * n1;
            T r1 = 0, i1 = 0;
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                const cplx_t<T> v = row[wrap_hi(nw.u[1] + c, n1)];
                r1 += v.x * nw.w[1][c];
                i1 += v.y * nw.w[1][c];
            }
            sr += r1 * nw.w[0][a];
            si += i1 * nw.w[0][a];
        }
    } else {
        const int64_t n1 = ge.n[1], n2 = ge.n[2];
#pragma unroll 1
        for (int a = 0; a < W; ++a) {
            const cplx_t<T> *plane = g + (wrap_hi(nw.u[0] + a, n0)) * n1 * n2;
            T r1 = 0, i1 = 0;
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                const cplx_t<T> *row = plane + (wrap_hi(nw.u[1] + c, n1)) * n2;
                T r2 = 0, i2 = 0;
#pragma unroll
                for (int e = 0; e < WS; ++e) {
                    if (WT == 0 && e >= W) break;
                    const cplx_t<T> v = row[wrap_hi(nw.u[2] + e, n2)];
                    r2 += v.x * nw.w[2][e];
                    i2 += v.y * nw.w[2][e];
                }
                r1 += r2 * nw.w[1][c];
                i1 += i2 * nw.w[1][c];
            }
            // outer weight selected without dynamic register indexing
            T w0 = 0;
#pragma unroll
            for (int q = 0; q < WS; ++q)
                if (q == a) w0 = nw.w[0][q];
            sr += r1 * w0;
            si += i1 * w0;
        }
    }
    f[j] = cmake<T>(sr, si);
}

// --------------------------------------------------------------- scatter --
__device__ __forceinline__ void atomic_add_c(float2 *p, float re, float im)
{
    unsafeAtomicAdd(&p->x, re);
    unsafeAtomicAdd(&p->y, im);
}
__device__ __forceinline__ void atomic_add_c(double2 *p, double re, double im)
{
    unsafeAtomicAdd(&p->x, re);
    unsafeAtomicAdd(&p->y, im);
}

template <typename T, int D, int WT>
__global__ __launch_bounds__(kBlock) void k_scatter_atomic(const double *__restrict__ x,
                                                           const cplx_t<T> *__restrict__ f,
                                                           cplx_t<T> *__restrict__ g, Geom ge)
{
    constexpr int WS = WT ? WT : kMaxW;
    const int W = WT ? WT : ge.W;
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= ge.M) return;
    NodeWin<T, D, WS> nw;
    load_window<T, D, WS>(ge, x, j, W, nw);
    const cplx_t<T> fj = f[j];
    const int64_t n0 = ge.n[0];
    if constexpr (D == 1) {
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            atomic_add_c(g + (wrap_hi(nw.u[0] + a, n0)), fj.x * nw.w[0][a], fj.y * nw.w[0][a]);
        }
    } else if constexpr (D == 2) {
        const int64_t n1 = ge.n[1];
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            cplx_t<T> *row = g + (wrap_hi(nw.u[0] + a, n0)) * n1;
            const T fr = fj.x * nw.w[0][a], fi = fj.y * nw.w[0][a];
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                atomic_add_c(row + (wrap_hi(nw.u[1] + c, n1)), fr * nw.w[1][c], fi * nw.w[1][c]);
            }
        }
    } else {
        const int64_t n1 = ge.n[1], n2 = ge.n[2];
#pragma unroll 1
        for (int a = 0; a < W; ++a) {
            T w0 = 0;
#pragma unroll
            for (int q = 0; q < WS; ++q)
                if (q == a) w0 = nw.w[0][q];
            cplx_t<T> *plane = g + (wrap_hi(nw.u[0] + a, n0)) * n1 * n2;
            const T fr0 = fj.x * w0, fi0 = fj.y * w0;
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                cplx_t<T> *row = plane + (wrap_hi(nw.u[1] + c, n1)) * n2;
                const T fr = fr0 * nw.w[1][c], fi = fi0 * nw.w[1][c];
#pragma unroll
                for (int e = 0; e < WS; ++e) {
                    if (WT == 0 && e >= W) break;
                    atomic_add_c(row + (wrap_hi(nw.u[2] + e, n2)), fr * nw.w[2][e], fi * nw.w[2][e]);
                }
            }
        }
    }
}

// ------------------------------------------------------ derivative gather --
// gx[j*D+t] = Re(conj(v_j) * sum_l g_l dW_t(l)), dW_t = derivative on axis t.
template <typename T, int D, int WT>
__global__ __launch_bounds__(kBlock) void k_grad_x(const cplx_t<T> *__restrict__ g,
                                                   const double *__restrict__ x,
                                                   const cplx_t<T> *__restrict__ v,
                                                   T *__restrict__ gx, Geom ge)
{
    constexpr int WS = WT ? WT : kMaxW;
    const int W = WT ? WT : ge.W;
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= ge.M) return;
    T w[D][WS], dw[D][WS];
    int64_t u[D];
#pragma unroll
    for (int t = 0; t < D; ++t) {
        double c;
        u[t] = window_cell(x[j * D + t], ge.n[t], ge.m, c);
        if constexpr (WS <= 18) {
            window_weights_dx<T, WS>(c, ge, t, w[t], dw[t]);
        } else {
            for (int l = 0; l < W; ++l) {
                const T s = (T)(c - (double)l);
                w[t][l] = window_weight<T>(s, ge, t);
                dw[t][l] = window_weight_dx<T>(s, w[t][l], ge, t);
            }
        }
    }
    const cplx_t<T> vj = v[j];
    const int64_t n0 = ge.n[0];
    if constexpr (D == 1) {
        T sr = 0, si = 0;
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            const cplx_t<T> q = g[wrap_hi(u[0] + a, n0)];
            sr += q.x * dw[0][a];
            si += q.y * dw[0][a];
        }
        gx[j] = vj.x * sr + vj.y * si;
    } else if constexpr (D == 2) {
        const int64_t n1 = ge.n[1];
        T s0r = 0, s0i = 0, s1r = 0, s1i = 0;
#pragma unroll
        for (int a = 0; a < WS; ++a) {
            if (WT == 0 && a >= W) break;
            const cplx_t<T> *row = g + (wrap_hi(u[0] + a, n0)) * n1;
            T ar = 0, ai = 0, br = 0, bi = 0;  // sum g w1, sum g dw1
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                const cplx_t<T> q = row[wrap_hi(u[1] + c, n1)];
                ar += q.x * w[1][c];
                ai += q.y * w[1][c];
                br += q.x * dw[1][c];
                bi += q.y * dw[1][c];
            }
            s0r += ar * dw[0][a];
            s0i += ai * dw[0][a];
            s1r += br * w[0][a];
            s1i += bi * w[0][a];
        }
        gx[2 * j] = vj.x * s0r + vj.y * s0i;
        gx[2 * j + 1] = vj.x * s1r + vj.y * s1i;
    } else {
        const int64_t n1 = ge.n[1], n2 = ge.n[2];
        T s0r = 0, s0i = 0, s1r = 0, s1i = 0, s2r = 0, s2i = 0;
#pragma unroll 1
        for (int a = 0; a < W; ++a) {
            T w0 = 0, dw0 = 0;
#pragma unroll
            for (int q = 0; q < WS; ++q)
                if (q == a) {
                    w0 = w[0][q];
                    dw0 = dw[0][q];
                }
            const cplx_t<T> *plane = g + (wrap_hi(u[0] + a, n0)) * n1 * n2;
            T Ar = 0, Ai = 0, Br = 0, Bi = 0, Cr = 0, Ci = 0;
#pragma unroll
            for (int c = 0; c < WS; ++c) {
                if (WT == 0 && c >= W) break;
                const cplx_t<T> *row = plane + (wrap_hi(u[1] + c, n1)) * n2;
                T ar = 0, ai = 0, br = 0, bi = 0;  // sum g w2, sum g dw2
#pragma unroll
                for (int e = 0; e < WS; ++e) {
                    if (WT == 0 && e >= W) break;
                    const cplx_t<T> q = row[wrap_hi(u[2] + e, n2)];
                    ar += q.x * w[2][e];
                    ai += q.y * w[2][e];
                    br += q.x * dw[2][e];
                    bi += q.y * dw[2][e];
                }
                Ar += ar * w[1][c];   // -> d/dx0 after dw0
                Ai += ai * w[1][c];
                Br += br * w[1][c];   // -> d/dx2 after w0
                Bi += bi * w[1][c];
                Cr += ar * dw[1][c];  // -> d/dx1 after w0
                Ci += ai * dw[1][c];
            }
            s0r += Ar * dw0;
            s0i += Ai * dw0;
            s1r += Cr * w0;
            s1i += Ci * w0;
            s2r += Br * w0;
            s2i += Bi * w0;
        }
        gx[3 * j] = vj.x * s0r + vj.y * s0i;
        gx[3 * j + 1] = vj.x * s1r + vj.y * s1i;
        gx[3 * j + 2] = vj.x * s2r + vj.y * s2i;
    }
}

// ------------------------------------------------------------------ zero --
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zero(cplx_t<T> *__restrict__ p, int64_t n)
{
    const int64_t stride = (int64_t)gridDim.x * kBlock;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride)
        p[i] = cmake<T>(0, 0);
}

// ------------------------------------------------------------- dispatch --
static inline unsigned grid_for(int64_t M) { return (unsigned)((M + kBlock - 1) / kBlock); }

#define CUNFFT_W_CASES(MACRO, D) \
    switch (ge.W) {              \
    case 4: MACRO(D, 4); break;  \
    case 6: MACRO(D, 6); break;  \
    case 8: MACRO(D, 8); break;  \
    case 10: MACRO(D, 10); break; \
    case 12: MACRO(D, 12); break; \
    case 14: MACRO(D, 14); break; \
    case 16: MACRO(D, 16); break; \
    case 18: MACRO(D, 18); break; \
    default: MACRO(D, 0); break;  \
    }

#define CUNFFT_D_CASES(MACRO)                 \
    switch (ge.d) {                           \
    case 1: CUNFFT_W_CASES(MACRO, 1) break;   \
    case 2: CUNFFT_W_CASES(MACRO, 2) break;   \
    default: CUNFFT_W_CASES(MACRO, 3) break;  \
    }

template <typename T>
void launch_gather_plain(const Geom &ge, const cplx_t<T> *g, const double *x, cplx_t<T> *f,
                         hipStream_t s)
{
    if (ge.M == 0) return;
#define L_GATHER(D, W) k_gather_plain<T, D, W><<<grid_for(ge.M), kBlock, 0, s>>>(g, x, f, ge)
    CUNFFT_D_CASES(L_GATHER)
#undef L_GATHER
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
void launch_scatter_atomic(const Geom &ge, const double *x, const cplx_t<T> *f, cplx_t<T> *g,
                           hipStream_t s)
{
    if (ge.M == 0) return;
#define L_SCATTER(D, W) k_scatter_atomic<T, D, W><<<grid_for(ge.M), kBlock, 0, s>>>(x, f, g, ge)
    CUNFFT_D_CASES(L_SCATTER)
#undef L_SCATTER
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
void launch_grad_x(const Geom &ge, const cplx_t<T> *g, const double *x, const cplx_t<T> *v,
                   T *gx, hipStream_t s)
{
    if (ge.M == 0) return;
#define L_GRAD(D, W) k_grad_x<T, D, W><<<grid_for(ge.M), kBlock, 0, s>>>(g, x, v, gx, ge)
    CUNFFT_D_CASES(L_GRAD)
#undef L_GRAD
    CUNFFT_HIP(hipGetLastError());
}

template <typename T> void launch_zero(cplx_t<T> *p, int64_t n, hipStream_t s)
{
    if (n == 0) return;
    int64_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks > 8192) blocks = 8192;
    k_zero<T><<<(unsigned)blocks, kBlock, 0, s>>>(p, n);
    CUNFFT_HIP(hipGetLastError());
}

#define CUNFFT_INST(T)                                                                       \
    template void launch_gather_plain<T>(const Geom &, const cplx_t<T> *, const double *,   \
                                         cplx_t<T> *, hipStream_t);                          \
    template void launch_scatter_atomic<T>(const Geom &, const double *, const cplx_t<T> *, \
                                           cplx_t<T> *, hipStream_t);                        \
    template void launch_grad_x<T>(const Geom &, const cplx_t<T> *, const double *,         \
                                   const cplx_t<T> *, T *, hipStream_t);                     \
    template void launch_zero<T>(cplx_t<T> *, int64_t, hipStream_t);
CUNFFT_INST(float)
CUNFFT_INST(double)

}  // namespace cunfft
