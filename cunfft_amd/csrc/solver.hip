// solver.hip -- inverse NFFT: conjugate gradients on the normal equations.
//
// The reference's README names an inverse NFFT (README:1) but the code has
// only the forward and adjoint transforms; its CPU side links NFFT3, whose
// solver module provides it.  This is that solver's CGNR iteration for
//   min_f  sum_j w_j |y_j - (A f)_j|^2,   i.e.  A^H W A f = A^H W y,
// on the plan's own forward (A) and adjoint (A^H) transforms:
//   r = y - A f,  z = A^H W r,  p = z
//   repeat:  v = A p;  alpha = |z|^2 / |v|_W^2;  f += alpha p;  r -= alpha v
//            z' = A^H W r;  beta = |z'|^2 / |z|^2;  p = z' + beta p
// The scalars stay on the device: the norms are reduced in FP64 in a fixed
// order (deterministic), and the update kernels read alpha and beta from
// device memory, so an iteration queues without host round trips; the host
// reads the residuals every few iterations to stop.  On an inconsistent
// system |r| levels off while the normal-equation residual |z| keeps
// falling; past that point the recurrences lose orthogonality and grow, so
// the iteration also stops when |z| <= tol |z_0| or |z| exceeds ten times
// its smallest checked value.
#include <math.h>
#include <string.h>

#include <vector>

#include "internal.hpp"

namespace cunfft {

constexpr int kRedBlocks = 1024;   // partial sums per norm (fixed: same order every call)
constexpr int kRedThreads = 256;

// Device scalars: |v|_W^2, |y|_W^2, then the histories |r|_W^2 and |z|^2
// per iteration (max_iter + 1 entries each).
enum SolverSlot : int { SL_VV = 0, SL_YY = 1, SL_RR = 2 };

template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_norm2_part(const cplx_t<T> *__restrict__ x,
                                                            const T *__restrict__ w, int64_t n,
                                                            double *__restrict__ part)
{
    __shared__ double red[kRedThreads];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
         i += (int64_t)kRedBlocks * kRedThreads) {
        const double a = (double)x[i].x, b = (double)x[i].y;
        s += (w ? (double)w[i] : 1.0) * (a * a + b * b);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kRedThreads) void k_norm2_final(const double *__restrict__ part,
                                                             double *__restrict__ out)
{
    __shared__ double red[kRedThreads];
    double s = 0.0;
    for (int i = threadIdx.x; i < kRedBlocks; i += kRedThreads) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kRedThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// out = a - b
template <typename T>
__global__ void k_sub(cplx_t<T> *__restrict__ out, const cplx_t<T> *__restrict__ a,
                      const cplx_t<T> *__restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = cmake<T>(a[i].x - b[i].x, a[i].y - b[i].y);
}

// out = w .* x (w == NULL: copy)
template <typename T>
__global__ void k_wmul(cplx_t<T> *__restrict__ out, const T *__restrict__ w,
                       const cplx_t<T> *__restrict__ x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const T s = w ? w[i] : (T)1;
        out[i] = cmake<T>(s * x[i].x, s * x[i].y);
    }
}

// y += sgn * (num / den) x, the ratio read from the device scalars (0 if den == 0)
template <typename T>
__global__ void k_axpy_ratio(cplx_t<T> *__restrict__ y, const cplx_t<T> *__restrict__ x,
                             const double *__restrict__ scal, int num, int den, double sgn,
                             int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = scal[den];
    const T a = (T)(d > 0.0 ? sgn * scal[num] / d : 0.0);
    y[i] = cmake<T>(y[i].x + a * x[i].x, y[i].y + a * x[i].y);
}

// p = z + (num / den) p
template <typename T>
__global__ void k_xpby_ratio(cplx_t<T> *__restrict__ p, const cplx_t<T> *__restrict__ z,
                             const double *__restrict__ scal, int num, int den, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = scal[den];
    const T b = (T)(d > 0.0 ? scal[num] / d : 0.0);
    p[i] = cmake<T>(z[i].x + b * p[i].x, z[i].y + b * p[i].y);
}

static unsigned blocks_of(int64_t n) { return (unsigned)((n + 255) / 256); }

template <typename T>
static void norm2(const cplx_t<T> *x, const T *w, int64_t n, double *part, double *out,
                  hipStream_t s)
{
    k_norm2_part<T><<<kRedBlocks, kRedThreads, 0, s>>>(x, w, n, part);
    k_norm2_final<<<1, kRedThreads, 0, s>>>(part, out);
}

template <typename T>
void solve_cgnr(const SolverOps<T> &ops, const cplx_t<T> *y, const T *w, cplx_t<T> *f,
                int64_t M, int64_t N, int max_iter, double tol, double *res_hist, int *iters,
                hipStream_t s)
{
    using C = cplx_t<T>;
    if (max_iter < 0) throw Error(CUNFFT_ERR_INVALID_ARG, "max_iter must be >= 0");
    C *r = nullptr, *v = nullptr, *z = nullptr, *p = nullptr;
    C *fbest = nullptr;   // the iterate of the smallest checked |z|
    double *part = nullptr, *scal = nullptr;
    auto release = [&] {
        for (void *q : {(void *)r, (void *)v, (void *)z, (void *)p, (void *)fbest, (void *)part,
                        (void *)scal})
            if (q) (void)hipFree(q);
    };
    try {
        CUNFFT_HIP(hipMalloc((void **)&r, sizeof(C) * M));
        CUNFFT_HIP(hipMalloc((void **)&v, sizeof(C) * M));
        CUNFFT_HIP(hipMalloc((void **)&z, sizeof(C) * N));
        CUNFFT_HIP(hipMalloc((void **)&p, sizeof(C) * N));
        CUNFFT_HIP(hipMalloc((void **)&fbest, sizeof(C) * N));
        CUNFFT_HIP(hipMalloc((void **)&part, sizeof(double) * kRedBlocks));
        const int SL_ZH = SL_RR + max_iter + 1;   // |z|^2 history
        CUNFFT_HIP(hipMalloc((void **)&scal, sizeof(double) * (SL_ZH + max_iter + 1)));
        // r = y - A f, z = A^H W r, p = z
        ops.forward(f, v, s);
        k_sub<T><<<blocks_of(M), 256, 0, s>>>(r, y, v, M);
        k_wmul<T><<<blocks_of(M), 256, 0, s>>>(v, w, r, M);
        ops.adjoint(v, z, s);
        CUNFFT_HIP(hipMemcpyAsync(p, z, sizeof(C) * N, hipMemcpyDeviceToDevice, s));
        norm2<T>(y, w, M, part, scal + SL_YY, s);
        norm2<T>(r, w, M, part, scal + SL_RR, s);
        norm2<T>(z, (const T *)nullptr, N, part, scal + SL_ZH, s);
        CUNFFT_HIP(hipGetLastError());
        double yy = 0, rr = 0, zz0 = 0;
        CUNFFT_HIP(hipMemcpyAsync(&yy, scal + SL_YY, sizeof(double), hipMemcpyDeviceToHost, s));
        CUNFFT_HIP(hipMemcpyAsync(&rr, scal + SL_RR, sizeof(double), hipMemcpyDeviceToHost, s));
        CUNFFT_HIP(hipMemcpyAsync(&zz0, scal + SL_ZH, sizeof(double), hipMemcpyDeviceToHost, s));
        CUNFFT_HIP(hipStreamSynchronize(s));
        CUNFFT_HIP(hipMemcpyAsync(fbest, f, sizeof(C) * N, hipMemcpyDeviceToDevice, s));
        const double ynorm = yy > 0 ? sqrt(yy) : 1.0;
        const double z0 = sqrt(zz0);
        double zmin = z0;
        int done = 0;
        bool stop = sqrt(rr) <= tol * ynorm || z0 == 0.0;
        const int check = 8;   // iterations between residual reads
        while (!stop && done < max_iter) {
            const int zz = SL_ZH + done, zn = SL_ZH + done + 1;
            ops.forward(p, v, s);                                   // v = A p
            norm2<T>(v, w, M, part, scal + SL_VV, s);               // |v|_W^2
            k_axpy_ratio<T><<<blocks_of(N), 256, 0, s>>>(f, p, scal, zz, SL_VV, 1.0, N);
            k_axpy_ratio<T><<<blocks_of(M), 256, 0, s>>>(r, v, scal, zz, SL_VV, -1.0, M);
            norm2<T>(r, w, M, part, scal + SL_RR + done + 1, s);    // |r|_W^2 history
            k_wmul<T><<<blocks_of(M), 256, 0, s>>>(v, w, r, M);
            ops.adjoint(v, z, s);                                   // z' = A^H W r
            norm2<T>(z, (const T *)nullptr, N, part, scal + zn, s);
            k_xpby_ratio<T><<<blocks_of(N), 256, 0, s>>>(p, z, scal, zn, zz, N);
            CUNFFT_HIP(hipGetLastError());
            ++done;
            if (done % check == 0 || done == max_iter) {
                double zk = 0;
                CUNFFT_HIP(hipMemcpyAsync(&rr, scal + SL_RR + done, sizeof(double),
                                          hipMemcpyDeviceToHost, s));
                CUNFFT_HIP(hipMemcpyAsync(&zk, scal + SL_ZH + done, sizeof(double),
                                          hipMemcpyDeviceToHost, s));
                CUNFFT_HIP(hipStreamSynchronize(s));
                zk = sqrt(zk);
                const bool diverged = zk > 10.0 * zmin;
                stop = sqrt(rr) <= tol * ynorm || zk <= tol * z0 || diverged;
                if (diverged) {
                    // loss of orthogonality: return the best checked iterate, not
                    // the diverging one
                    CUNFFT_HIP(hipMemcpyAsync(f, fbest, sizeof(C) * N, hipMemcpyDeviceToDevice, s));
                } else if (zk < zmin) {
                    zmin = zk;
                    CUNFFT_HIP(hipMemcpyAsync(fbest, f, sizeof(C) * N, hipMemcpyDeviceToDevice, s));
                }
            }
        }
        std::vector<double> hist((size_t)done + 1);
        CUNFFT_HIP(hipMemcpyAsync(hist.data(), scal + SL_RR, sizeof(double) * (done + 1),
                                  hipMemcpyDeviceToHost, s));
        CUNFFT_HIP(hipStreamSynchronize(s));
        if (res_hist)
            for (int k = 0; k <= done; ++k) res_hist[k] = sqrt(hist[k] > 0 ? hist[k] : 0.0) / ynorm;
        if (iters) *iters = done;
    } catch (...) {
        (void)hipStreamSynchronize(s);
        release();
        throw;
    }
    release();
}

template void solve_cgnr<float>(const SolverOps<float> &, const float2 *, const float *, float2 *,
                                int64_t, int64_t, int, double, double *, int *, hipStream_t);
template void solve_cgnr<double>(const SolverOps<double> &, const double2 *, const double *,
                                 double2 *, int64_t, int64_t, int, double, double *, int *,
                                 hipStream_t);

}  // namespace cunfft
