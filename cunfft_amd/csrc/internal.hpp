// internal.hpp -- launch interface between the host plan code and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <functional>
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

// Axis lengths one LDS-resident pass transforms: L = 2^a 3^b with a >= 1,
// L <= kFftMaxLine.  Longer axes (up to CUNFFT_MAX_AXIS) are split into two
// such factors, L = L1 L2, and run as two passes (fft_split).
constexpr int64_t kFftMaxLine = 4096;
inline bool fft_line_ok(int64_t n)
{
    if (n < 2 || n > kFftMaxLine || (n & 1)) return false;
    while (n % 3 == 0) n /= 3;
    return (n & (n - 1)) == 0;
}
// The factors of a split axis: L1 (the minor one, stride 1 within the axis)
// and L2 = L / L1, both valid lines, L1 the largest such factor up to
// sqrt(L) (balanced passes); false if there are none.
inline bool fft_split(int64_t n, int64_t &L1, int64_t &L2)
{
    L1 = L2 = 0;
    for (int64_t a = 2; a * a <= n; ++a) {
        if (n % a || !fft_line_ok(a) || !fft_line_ok(n / a)) continue;
        L1 = a;
        L2 = n / a;
    }
    return L1 > 0;
}
// FFT lengths this build runs: n = 2^a 3^b with a >= 1, n <= CUNFFT_MAX_AXIS.
inline bool fft_length_ok(int64_t n)
{
    if (fft_line_ok(n)) return true;
    if (n > CUNFFT_MAX_AXIS) return false;
    int64_t L1, L2;
    return fft_split(n, L1, L2);
}

// Geometry of one transform (window shape from sigma_t = n_t / N_t).
Geom make_geom(int d, const int64_t *N, const int64_t *n, int64_t M, int m,
               int win = WIN_GAUSSIAN);
// Device roll-off tables 1 / phihat_t(k / n_t) by pruned index (Gaussian:
// exp(fac_t k^2); Kaiser-Bessel: 1 / I0(m sqrt(b^2 - (2 pi k / n)^2))),
// offsets per axis.
double *make_rolloff(const Geom &ge, int64_t (&rocoff)[3]);

// ----------------------------------------------------------------- FFT ---
// Target of a cunfftHandle: twiddle tables tw[q] = exp(-2 pi i q / n_t) per
// axis, computed in extended precision on the host -- or, for an axis longer
// than kFftMaxLine, the tables of its two factors L1 (tw) and L2 (tw2).
struct FftTables {
    int prec;
    int d;
    int64_t n[3];
    void *tw[3];
    void *tw2[3];
};
FftTables *make_fft_tables(int prec, int d, const int64_t *n);
void free_fft_tables(void *h);

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

// A split axis of length Lf = L1 L2 (element i = i1 + L1 i2, output
// k = k2 + L2 k1) runs as two passes:
//   SPLIT_A  the L2-point FFTs over i2 for every i1 (stride L1 S), in place:
//            Y[i1 + L1 k2] = sum_i2 x[i1 + L1 i2] w_L2^(i2 k2);
//   SPLIT_B  per k2: the L1-point FFT over i1 of Y[k2 L1 + i1] w_Lf^(i1 k2),
//            written to X[k2 + L2 k1] of another grid (dst) in natural
//            order, or to fhat with the crop and roll-off.
// The pruning of the first forward / last adjoint pass and the roll-off use
// the full axis index.
enum FftSplit : int { SPLIT_NONE = 0, SPLIT_A = 1, SPLIT_B = 2 };

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
    int L, logL;             // logL = log2(L) for a power of two, else -1
    int n3;                  // radix-3 factors of L (L = 2^a 3^n3)
    int64_t S;
    int64_t n_pa;            // number of lines over the axes before t (times L2 for SPLIT_B)
    int C, logC;             // lines per workgroup
    int ldsL;                // LDS row pitch in elements
    int sign;                // -1 forward, +1 adjoint
    int mode;
    // split axes
    int split;               // FftSplit
    int64_t Lf;              // the axis length (L when not split)
    int64_t L1, L2;          // its factors (split passes)
    int64_t So;              // the axis stride (product of the later axes)
    cplx_t<T> *dst;          // SPLIT_B output grid (not for FFT_TO_FHAT)
};

template <typename T>
void launch_fft_pass(const FftPass<T> &p, int64_t blocks, size_t lds_bytes, hipStream_t s);

// Pass along axis t of one transform (forward: axes d-1 .. 0, the first from
// fhat with the roll-off; adjoint: axes 0 .. d-1, the last to fhat), with its
// launch shape.  tw is the twiddle table of length n_t.  Axes longer than
// kFftMaxLine need fft_axis_setup.
template <typename T> struct FftLaunch {
    FftPass<T> P;
    int64_t blocks;
    size_t lds;
};
template <typename T>
FftLaunch<T> fft_pass_setup(const Geom &ge, const cplx_t<T> *tw, const double *roc,
                            const int64_t *rocoff, int t, bool forward, const cplx_t<T> *fhat_in,
                            cplx_t<T> *fhat_out, cplx_t<T> *g);
// The passes of axis t (one, or two for a split axis): tw[0] is the twiddle
// table of the axis (n_t), or of L1 and tw[1] of L2 when it is split.  The
// axis's data is in *g; a split axis leaves it in g_other and swaps the two
// pointers (unless it writes fhat).  Returns the number of passes (<= 2).
template <typename T>
int fft_axis_setup(const Geom &ge, const cplx_t<T> *const *tw, const double *roc,
                   const int64_t *rocoff, int t, bool forward, const cplx_t<T> *fhat_in,
                   cplx_t<T> *fhat_out, cplx_t<T> **g, cplx_t<T> **g_other, FftLaunch<T> *out);

// Workgroup b of a batched launch works on block map[b].blk of job map[b].job.
struct BlockMap {
    uint32_t job;
    uint32_t blk;
};
// One launch of the passes of many transforms (one FftPass per job);
// max_elems = the largest C * L of the passes.
template <typename T>
void launch_fft_batch(const FftPass<T> *passes, const BlockMap *map, uint32_t nblk,
                      size_t lds_bytes, int max_elems, hipStream_t s);

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

// ------------------------------------------------------------ scratch ---
// Bytes of the device-resident scratch of the allocation-free adjoint.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
// Device address of the current device's scratch (hipGetSymbolAddress on the
// first call per device, kept after).
void *noalloc_scratch();

// -------------------------------------------------------------- solver ---
// The transforms the inverse NFFT iterates (device pointers, on stream s).
template <typename T> struct SolverOps {
    std::function<void(const cplx_t<T> *fhat, cplx_t<T> *f, hipStream_t s)> forward;
    std::function<void(const cplx_t<T> *f, cplx_t<T> *fhat, hipStream_t s)> adjoint;
};
// CGNR for min_f sum_j w_j |y_j - (A f)_j|^2 (w NULL: all 1); f is the
// initial guess and the result.  res_hist (may be NULL) receives
// |y - A f|_W / |y|_W for iterations 0 .. *iters.
template <typename T>
void solve_cgnr(const SolverOps<T> &ops, const cplx_t<T> *y, const T *w, cplx_t<T> *f,
                int64_t M, int64_t N, int max_iter, double tol, double *res_hist, int *iters,
                hipStream_t s);

// ---------------------------------------------------------------- NDFT ---
template <typename T>
void launch_ndft_rows(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *fhat,
                      const long long *rows, int64_t n_rows, cplx_t<T> *out, hipStream_t s);
template <typename T>
void launch_ndft_cols(int d, const int64_t *N, int64_t M, const double *x, const cplx_t<T> *f,
                      const long long *cols, int64_t n_cols, cplx_t<T> *out, hipStream_t s);

}  // namespace cunfft
