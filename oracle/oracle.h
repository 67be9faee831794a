/*
 * oracle.h -- CPU restatement of the CUNFFT NFFT / NDFT algorithm in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: tests/, the smoke
 * check in __graft_entry__.py and the cpu_baseline leg of bench.py may load
 * it; the product (cunfft_amd, libcunfft_amd.so) never links, calls or falls
 * back to it.
 *
 * Conventions restated (reference file:line):
 *   window      phi(t) = (pi b)^-1/2 exp(-t^2/b)          src/cunfft_kernel.cu:162-166
 *   shape       b = 2 sigma / (2 sigma - 1) * m / pi       src/cunfft.cu:346-350
 *   index rule  u = llrint(n x) - signbit(x) - m            src/cunfft_kernel.cu:175-182
 *               window l = u .. u+2m+1, wrapped mod n       src/cunfft_typedefs.h:38
 *   roll-off    ghat[k mod n] = fhat_k exp(sum_t fac_t k_t^2),
 *               fac_t = pi^2 b_t / n_t^2                    src/cunfft.cu:79-90,
 *                                                           src/cunfft_kernel.cu:240-412
 *   FFT         forward e^{-2 pi i k l / n}, adjoint e^{+}, unnormalised
 *                                                           src/cunfft_wrapper.cu:288-329
 *   layouts     fhat row-major over k_t+N_t/2 (last axis fastest),
 *               x interleaved x[d*j+t], g row-major n_0 x .. x n_{d-1}
 *
 * Complex arrays are interleaved (re, im) doubles.
 */
#ifndef CUNFFT_ORACLE_H
#define CUNFFT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* glibc-compatible drand48 stream with explicit 48-bit state (state 0 is
 * what an unseeded process starts from). Used by the reference's harness
 * input generators, src/cuda_util.cpp:151-212. */
double orc_drand48(uint64_t *state);

/* Example data exactly as getExampleData_uniDistr (src/cuda_util.cpp:151-160,
 * 204-212): fhat re/im = drand48() for k = 0..N_total-1, then
 * x[i] = drand48() - 0.5 for i = 0..d*M-1. */
void orc_example_data_unidistr(uint64_t *state, int64_t N_total, int64_t M, int d,
                               double *fhat, double *x);
/* getExampleDataAd_uniDistr (src/cuda_util.cpp:162-171): f then x. */
void orc_example_data_ad_unidistr(uint64_t *state, int64_t M, int d, double *f,
                                  double *x);

/* Window shape parameter b_t (src/cunfft.cu:346-350). */
double orc_shape_b(int64_t N, int64_t n, int m);
/* First window index (src/cunfft_kernel.cu:175-182). */
int64_t orc_window_start(int64_t n, double x, int m);
/* Gaussian window in grid units (src/cunfft_kernel.cu:162-166). */
double orc_phi(double t, double b);

/* Window of the transforms below: 0 Gaussian (the reference's, default),
 * 1 Kaiser-Bessel, 2 sinc power (NFFT3's, which the reference links for its
 * CPU side; no counterpart in the reference's GPU code).  Process-wide
 * setting. */
void orc_set_window(int kind);
int orc_get_window(void);
/* Kaiser-Bessel shape b = pi (2 - 1/sigma), window phi(s) in grid units,
 * d phi / d s, and I0 (power series). */
double orc_kb_shape(int64_t N, int64_t n);
double orc_phi_kb(double s, int m, double b);
double orc_dphi_kb(double s, int m, double b);
double orc_bessel_i0(double z);
/* Sinc-power window (window kind 2, NFFT3's SINC_POWER): shape
 * a = (2 sigma - 1) / (2 sigma m), phi(s) = a sinc(pi a s)^(2m), d phi / d s,
 * and the centred cardinal B-spline M_p of order p (its Fourier transform is
 * M_2m(nu / a)). */
double orc_sinc_shape(int64_t N, int64_t n, int m);
double orc_phi_sinc(double s, int m, double a);
double orc_dphi_sinc(double s, int m, double a);
double orc_bspline(int p, double x);

/* Forward roll-off + zero pad: fhat (N_total) -> ghat (n_total), zeros
 * outside the centred box (src/cunfft_wrapper.cu:32-75,
 * src/cunfft_kernel.cu:240-412). */
void orc_rolloff_pad(int d, const int64_t *N, const int64_t *n, int m,
                     const double *fhat, double *ghat);
/* Adjoint crop + roll-off: g (n_total) -> fhat (N_total)
 * (src/cunfft_wrapper.cu:77-121, src/cunfft_kernel.cu:264-467). */
void orc_crop_rolloff(int d, const int64_t *N, const int64_t *n, int m,
                      const double *g, double *fhat);
/* In-place unnormalised d-dim FFT, sign -1 forward / +1 adjoint
 * (cufftExecZ2Z at src/cunfft_wrapper.cu:288-329): radix 2 on power-of-two
 * axes, a direct DFT on the others. */
void orc_fft(int d, const int64_t *n, double *g, int sign);
/* Forward gridding (gather), f_j = sum_l g_l prod_t phi(n_t x_jt - l_t)
 * (src/cunfft_kernel.cu:514-1093). */
void orc_gather(int d, const int64_t *N, const int64_t *n, int m, const double *g,
                int64_t M, const double *x, double *f);
/* Adjoint gridding (scatter), g_l += f_j prod_t phi(...); g is zeroed first
 * (src/cunfft_wrapper.cu:178-247, src/cunfft_kernel.cu:608-1197). */
void orc_scatter(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                 const double *x, const double *f, double *g);
/* Complete transforms (src/cunfft.cu:196-224). g_work holds n_total complex. */
void orc_nfft_trafo(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                    const double *x, const double *fhat, double *f, double *g_work);
void orc_nfft_adjoint(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                      const double *x, const double *f, double *fhat, double *g_work);

/* Direct NDFT (src/cunfft_kernel.cu:1207-1345, Tests/Scripts/ndft.cpp:11-113):
 *   f_j = sum_k fhat_k e^{-2 pi i k.x_j},  fhat_k = sum_j f_j e^{+2 pi i k.x_j}.
 * rows/cols == NULL means all outputs; otherwise only the listed outputs are
 * computed (n_rows of them) and written densely to the output. nthreads <= 0
 * means all OpenMP threads. */
void orc_ndft_trafo(int d, const int64_t *N, int64_t M, const double *x,
                    const double *fhat, double *f, const int64_t *rows,
                    int64_t n_rows, int nthreads);
void orc_ndft_adjoint(int d, const int64_t *N, int64_t M, const double *x,
                      const double *f, double *fhat, const int64_t *cols,
                      int64_t n_cols, int nthreads);

/* Gradient of the NFFT forward map with respect to node coordinates:
 * dfdx[(j*d+t)*2 + {0,1}] = d f_j / d x_jt for the truncated, implemented
 * window (derivative of the gather at src/cunfft_kernel.cu:514-1093 with the
 * window of src/cunfft_kernel.cu:162-182 held fixed). g is the FFT'd grid of
 * the forward transform. */
void orc_gather_dx(int d, const int64_t *N, const int64_t *n, int m, const double *g,
                   int64_t M, const double *x, double *dfdx);
/* Analytic NDFT derivative d f_j / d x_jt = sum_k fhat_k (-2 pi i k_t)
 * e^{-2 pi i k.x_j} for the listed rows (all if rows == NULL). */
void orc_ndft_dx(int d, const int64_t *N, int64_t M, const double *x,
                 const double *fhat, double *dfdx, const int64_t *rows,
                 int64_t n_rows, int nthreads);

/* Error norms on interleaved complex vectors of n entries.
 * rel_linf = ||a-b||_inf / ||a||_inf   (src/cuda_util.cpp:237-258)
 * rel_l2   = ||a-b||_2 / ||a||_2 */
double orc_rel_linf(const double *a, const double *b, int64_t n);
double orc_rel_l2(const double *a, const double *b, int64_t n);

#ifdef __cplusplus
}
#endif
#endif
