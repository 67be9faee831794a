// internal.hpp -- launch interface between the host plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "cunfft.h"
#include "device_common.hpp"

namespace cunfft {

// Error carried from the host code to the C ABI boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

void hip_check(hipError_t e, const char *what);
#define CUNFFT_HIP(call) ::cunfft::hip_check((call), #call)

// ----------------------------------------------------------------- FFT ---
// One Stockham pass along axis t of the row-major grid n_0 x .. x n_{d-1},
// viewed as [A][L][S] with L = n_t and S = prod_{t'>t} n_t'.  The lines of the
// pass run over the support (first and last N_t'/2 indices) of the axes
// before t and over all of the axes after t.  Roll-off/pad and crop are
// fused into the first forward and the last adjoint pass.
enum FftMode : int {
    FFT_IN_PRUNED = 1,   // only the support along the axis is read, rest is 0
    FFT_OUT_PRUNED = 2,  // only the support along the axis is written
    FFT_FROM_FHAT = 4,   // read fhat (centred order) times the roll-off factor
    FFT_TO_FHAT = 8,     // write fhat (centred order) times the roll-off factor
};

template <typename T> struct FftPass {
    cplx_t<T> *g;
    const cplx_t<T> *fhat_in;
    cplx_t<T> *fhat_out;
    const cplx_t<T> *tw;     // tw[q] = exp(-2 pi i q / L), q < L
    const double *roc;       // per-axis roll-off tables by pruned index p < N_t
    int64_t rocoff[3];
    int64_t n[3];
    int64_t N[3];
    int d, t;
    int L, logL;
    int64_t S;
    int64_t n_pa;            // number of lines over the axes before t
    int C, logC;             // lines per workgroup
    int ldsL;                // LDS row pitch in elements
    int sign;                // -1 forward, +1 adjoint
    int mode;
};

template <typename T>
void launch_fft_pass(const FftPass<T> &p, int64_t blocks, size_t lds_bytes, hipStream_t s);

// ------------------------------------------------------------ gridding ---
template <typename T>
void launch_gather_plain(const Geom &ge, const cplx_t<T> *g, const double *x, cplx_t<T> *f,
                         hipStream_t s);
template <typename T>
void launch_scatter_atomic(const Geom &ge, const double *x, const cplx_t<T> *f, cplx_t<T> *g,
                           hipStream_t s);
template <typename T>
void launch_grad_x(const Geom &ge, const cplx_t<T> *g, const double *x, const cplx_t<T> *v,
                   T *gx, hipStream_t s);
template <typename T> void launch_zero(cplx_t<T> *p, int64_t n, hipStream_t s);

// ---------------------------------------------------------------- NDFT ---
template <typename T>
void launch_ndft_rows(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *fhat,
                      const long long *rows, int64_t n_rows, cplx_t<T> *out, hipStream_t s);
template <typename T>
void launch_ndft_cols(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *f,
                      const long long *cols, int64_t n_cols, cplx_t<T> *out, hipStream_t s);

}  // namespace cunfft
