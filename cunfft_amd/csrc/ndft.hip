// ndft.hip -- direct NDFT on the GPU in FP64 (the GPU oracle, "CUNDFT").
//
//   forward  f_j    = sum_k fhat_k exp(-2 pi i k.x_j)   src/cunfft_kernel.cu:1207-1285
//   adjoint  fhat_k = sum_j f_j    exp(+2 pi i k.x_j)   src/cunfft_kernel.cu:1289-1345
// k runs over the centred box, flat index row-major with the last axis
// fastest.  Unlike the reference (one thread per output with a serial sum),
// one 256-lane workgroup owns one output: lanes stride over the inner sum and
// a fixed-order tree reduction combines them, so results are reproducible and
// a few thousand sampled outputs still fill the chip.  Terms use sincospi on
// the FP64 phase; FP32 plans read FP32 inputs and round only the result.
#include <algorithm>

#include "internal.hpp"

namespace cunfft {

constexpr int kNdftThreads = 256;

struct NdftDims {
    int d;
    uint32_t N[3];
    uint32_t sN[3];  // flat stride of axis t in the coefficient array
};

__device__ __forceinline__ void block_sum2(double &a, double &b, double *sa, double *sb)
{
    const int tid = threadIdx.x;
    sa[tid] = a;
    sb[tid] = b;
    __syncthreads();
    for (int h = kNdftThreads / 2; h > 0; h >>= 1) {
        if (tid < h) {
            sa[tid] += sa[tid + h];
            sb[tid] += sb[tid + h];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

// An output index outside [0, M) (rows) or [0, N_total) (columns) gives NaN
// for that output; the check is uniform over the workgroup, before any barrier.
template <typename T> __device__ __forceinline__ cplx_t<T> nan_c()
{
    return cmake<T>((T)__builtin_nan(""), (T)__builtin_nan(""));
}

template <typename T>
__global__ __launch_bounds__(kNdftThreads) void k_ndft_rows(NdftDims nd, uint32_t Nt, int64_t M,
                                                            const double *__restrict__ x,
                                                            const cplx_t<T> *__restrict__ fhat,
                                                            const long long *__restrict__ rows,
                                                            int64_t r0,
                                                            cplx_t<T> *__restrict__ out)
{
    __shared__ double sa[kNdftThreads], sb[kNdftThreads];
    const int64_t r = r0 + blockIdx.x;
    const int64_t j = rows ? rows[r] : r;
    if (j < 0 || j >= M) {
        if (threadIdx.x == 0) out[r] = nan_c<T>();
        return;
    }
    double xj[3] = {0, 0, 0};
    for (int t = 0; t < nd.d; ++t) xj[t] = x[j * nd.d + t];
    double sr = 0, si = 0;
    for (uint32_t kl = threadIdx.x; kl < Nt; kl += kNdftThreads) {
        double ph = 0;
        for (int t = 0; t < nd.d; ++t) {
            const int k = (int)((kl / nd.sN[t]) % nd.N[t]) - (int)(nd.N[t] >> 1);
            ph += (double)k * xj[t];
        }
        double sn, cs;
        sincospi(-2.0 * ph, &sn, &cs);
        const cplx_t<T> a = fhat[kl];
        sr += (double)a.x * cs - (double)a.y * sn;
        si += (double)a.x * sn + (double)a.y * cs;
    }
    block_sum2(sr, si, sa, sb);
    if (threadIdx.x == 0) out[r] = cmake<T>((T)sr, (T)si);
}

template <typename T>
__global__ __launch_bounds__(kNdftThreads) void k_ndft_cols(NdftDims nd, int64_t M,
                                                            const double *__restrict__ x,
                                                            const cplx_t<T> *__restrict__ f,
                                                            const long long *__restrict__ cols,
                                                            int64_t r0, int64_t Nt,
                                                            cplx_t<T> *__restrict__ out)
{
    __shared__ double sa[kNdftThreads], sb[kNdftThreads];
    const int64_t r = r0 + blockIdx.x;
    const int64_t kc = cols ? cols[r] : r;
    if (kc < 0 || kc >= Nt) {
        if (threadIdx.x == 0) out[r] = nan_c<T>();
        return;
    }
    const uint32_t kl = (uint32_t)kc;
    double k[3] = {0, 0, 0};
    for (int t = 0; t < nd.d; ++t)
        k[t] = (double)((int)((kl / nd.sN[t]) % nd.N[t]) - (int)(nd.N[t] >> 1));
    double sr = 0, si = 0;
    for (int64_t j = threadIdx.x; j < M; j += kNdftThreads) {
        double ph = 0;
        for (int t = 0; t < nd.d; ++t) ph += k[t] * x[j * nd.d + t];
        double sn, cs;
        sincospi(2.0 * ph, &sn, &cs);
        const cplx_t<T> a = f[j];
        sr += (double)a.x * cs - (double)a.y * sn;
        si += (double)a.x * sn + (double)a.y * cs;
    }
    block_sum2(sr, si, sa, sb);
    if (threadIdx.x == 0) out[r] = cmake<T>((T)sr, (T)si);
}

static NdftDims make_dims(int d, const int64_t *N)
{
    NdftDims nd;
    nd.d = d;
    uint32_t s = 1;
    for (int t = d - 1; t >= 0; --t) {
        nd.N[t] = (uint32_t)N[t];
        nd.sN[t] = s;
        s *= (uint32_t)N[t];
    }
    for (int t = d; t < 3; ++t) {
        nd.N[t] = 1;
        nd.sN[t] = 1;
    }
    return nd;
}

// One workgroup per output, launched in pieces of kNdftMaxBlocks workgroups so
// the thread count of a launch stays below 2^32.
constexpr int64_t kNdftMaxBlocks = int64_t(1) << 22;

template <typename T>
void launch_ndft_rows(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *fhat,
                      const long long *rows, int64_t n_rows, cplx_t<T> *out, hipStream_t s)
{
    if (n_rows <= 0) return;
    NdftDims nd = make_dims(d, N);
    uint32_t Nt = 1;
    for (int t = 0; t < d; ++t) Nt *= (uint32_t)N[t];
    for (int64_t r0 = 0; r0 < n_rows; r0 += kNdftMaxBlocks) {
        const int64_t nb = std::min<int64_t>(kNdftMaxBlocks, n_rows - r0);
        k_ndft_rows<T><<<(unsigned)nb, kNdftThreads, 0, s>>>(nd, Nt, M, x, fhat, rows, r0, out);
        CUNFFT_HIP(hipGetLastError());
    }
}

template <typename T>
void launch_ndft_cols(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *f,
                      const long long *cols, int64_t n_cols, cplx_t<T> *out, hipStream_t s)
{
    if (n_cols <= 0) return;
    NdftDims nd = make_dims(d, N);
    int64_t Nt = 1;
    for (int t = 0; t < d; ++t) Nt *= N[t];
    for (int64_t r0 = 0; r0 < n_cols; r0 += kNdftMaxBlocks) {
        const int64_t nb = std::min<int64_t>(kNdftMaxBlocks, n_cols - r0);
        k_ndft_cols<T><<<(unsigned)nb, kNdftThreads, 0, s>>>(nd, M, x, f, cols, r0, Nt, out);
        CUNFFT_HIP(hipGetLastError());
    }
}

template void launch_ndft_rows<float>(int, const int64_t *, int64_t, const double *,
                                      const float2 *, const long long *, int64_t, float2 *,
                                      hipStream_t);
template void launch_ndft_rows<double>(int, const int64_t *, int64_t, const double *,
                                       const double2 *, const long long *, int64_t, double2 *,
                                       hipStream_t);
template void launch_ndft_cols<float>(int, const int64_t *, int64_t, const double *,
                                      const float2 *, const long long *, int64_t, float2 *,
                                      hipStream_t);
template void launch_ndft_cols<double>(int, const int64_t *, int64_t, const double *,
                                       const double2 *, const long long *, int64_t, double2 *,
                                       hipStream_t);

}  // namespace cunfft
