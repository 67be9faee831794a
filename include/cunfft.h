/*
 * cunfft.h -- C ABI of the MI355X-native NFFT library (libcunfft_amd.so).
 *
 * Drop-in for the plan surface of CUNFFT 1.0, src/cunfft.h:35-261 and
 * src/cunfft_typedefs.h:36-178: same function names, argument order, void
 * returns and plan field names.  Two precision families are exported:
 *
 *   FP64  cunfft_plan,  cunfft_*, copyDataTo*, cundft_*   (reference names;
 *         the reference's default build, CMakeLists.txt:173)
 *   FP32  cunfftf_plan, cunfftf_*                          (suffixed family)
 *
 * A caller that defines CUNFFT_SINGLE_PRECISION before including this header
 * gets the reference names mapped onto the FP32 family, the way the
 * reference's CUNFFT_DOUBLE_PRECISION switch selects gpuComplex/dTyp
 * (src/cunfft_typedefs.h:59-69).  Node coordinates x are double in both
 * families, so the node-to-grid index split is always done in FP64.
 *
 * Differences from the reference, all deliberate:
 *  - m (cut-off) is a runtime plan parameter: cunfft_init uses
 *    CUNFFT_DEFAULT_CUT_OFF (6, as CMakeLists.txt:170), cunfft_initGuru_m
 *    takes it explicitly.
 *  - errors never exit(): every call records a status in plan->status,
 *    readable with cunfft_last_error(); after an error every call except
 *    finalize is a no-op until the plan is finalized.
 *  - the FFT handles are opaque (cunfftHandle) and the device state lives in
 *    the plan (plan->impl), so several plans may live at once; one plan is
 *    not re-entrant.
 *  - finalize does not reset the device.
 */
#ifndef CUNFFT_AMD_H_
#define CUNFFT_AMD_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ constants -- */

#define CUNFFT_DEFAULT_CUT_OFF 6
#define CUNFFT_MAX_CUT_OFF 16
#define CUNFFT_MAX_DIM 3
/** Largest oversampled axis length n_t (= 2^a 3^b, a >= 1).  Axes up to
 *  4096 run as one LDS-resident pass each; longer ones as two (n_t = L1 L2,
 *  both factors up to 4096). */
#define CUNFFT_MAX_AXIS 16777216

/* plan flags, same bits as src/cunfft_typedefs.h:46-53 */
#define CUFFT_INIT  (1U << 0)
#define CUFFT_BOTH  (1U << 1)
#define CUFFT_FORW  (1U << 2)
#define CUFFT_BACKW (1U << 3)
/* additions */
/** Record per-step device times in plan->CUNFFTTimes (MEASURED_TIMES at
 *  run time instead of compile time, src/timeMeasurement.h:38-48). */
#define CUNFFT_MEASURE_TIME   (1U << 8)
/** Do not allocate the pinned host arrays x, f_hat, f, g (device-pointer
 *  callers only; the copy functions then fail with CUNFFT_ERR_STATE). */
#define CUNFFT_NO_HOST_BUFFERS (1U << 9)
/** Adjoint gridding with one global atomic per tap instead of the binned
 *  tile scatter (the plain path). */
#define CUNFFT_ATOMIC_SCATTER (1U << 10)
/** Forward gridding over unsorted nodes (one node per lane, plain path). */
#define CUNFFT_PLAIN_GATHER   (1U << 11)
/** Tiled gridding with one node per lane instead of one lane per tap of the
 *  window row (d >= 2). */
#define CUNFFT_NODE_LANES     (1U << 12)
/** FP32 plans: compensated summation in the adjoint tiles -- Kahan sums in
 *  the plane-tiled accumulation (a second LDS tile holds the compensation
 *  terms) and Neumaier sums in the fold.  Costs LDS space and time; FP64
 *  plans ignore it.  Configurations whose doubled tile does not fit the
 *  LDS (FP32 3D 16x16x8, FP32 3D m >= 6) and the node-lane configuration
 *  are not offered; a plan with none left (FP32 3D m >= 6, or
 *  CUNFFT_NODE_LANES as well) is refused with CUNFFT_ERR_UNSUPPORTED.  1D
 *  plans accumulate their adjoint tiles in FP64 anyway. */
#define CUNFFT_COMPENSATED    (1U << 13)
/** Asynchronous copy boundary: copyDataToDevice[Ad], cunfft_transform and
 *  cunfft_adjoint return once their work is queued on the plan's stream
 *  (the reference returns after it, src/cunfft.cu:107-224); copyDataToHost[Ad]
 *  and the copy_*_ToHost calls still return with the data.  Several plans
 *  then overlap one plan's PCIe copies with another's kernels.  The caller
 *  must not change plan->x / f_hat / f until the plan's next copy to the
 *  host or cunfft_synchronize.  CUNFFT_MEASURE_TIME synchronizes anyway. */
#define CUNFFT_ASYNC_COPY     (1U << 14)
/** Kaiser-Bessel window instead of the reference's Gaussian (NFFT3's
 *  window, b = pi (2 - 1/sigma), roll-off 1 / I0(m sqrt(b^2 - (2 pi k/n)^2))):
 *  the same 2m+2 taps per axis reach ~1e-7 at m = 4 and ~1e-14 at m = 8
 *  against the NDFT, where the Gaussian reaches ~1e-4 and ~1e-8. */
#define CUNFFT_KAISER_BESSEL  (1U << 15)
/** Sinc-power window instead of the reference's Gaussian (NFFT3's
 *  SINC_POWER): phi(s) = a sinc(pi a s)^(2m) in grid units with
 *  a = (2 sigma - 1) / (2 sigma m), Fourier transform the centred cardinal
 *  B-spline of order 2m, roll-off 1 / M_2m(k / (n a)).  Needs n_t > N_t;
 *  excludes CUNFFT_KAISER_BESSEL. */
#define CUNFFT_SINC           (1U << 16)

/* status codes */
#define CUNFFT_SUCCESS          0
#define CUNFFT_ERR_INVALID_ARG  1
#define CUNFFT_ERR_UNSUPPORTED  2
#define CUNFFT_ERR_ALLOC        3
#define CUNFFT_ERR_HIP          4
#define CUNFFT_ERR_STATE        5

/* ---------------------------------------------------------------- types -- */

/** Index type: the reference's LARGE_INPUT build (CMakeLists.txt:172),
 *  src/cunfft_typedefs.h:71-87. */
typedef ptrdiff_t uint_t;
typedef ptrdiff_t sint_t;

/** Interleaved complex with .x = re, .y = im (as cuDoubleComplex). */
typedef struct { double x, y; } cunfftDoubleComplex;
typedef struct { float x, y; } cunfftFloatComplex;

/** Opaque FFT plan handle (replaces cufftHandle). */
typedef void *cunfftHandle;

/** Per-step times in seconds, src/timeMeasurement.h:22-36. time_ROC stays
 *  0 because roll-off, pad and crop are fused into the FFT passes and are
 *  counted in time_FFT. */
typedef struct {
    double runTime;
    double time_ROC;
    double time_CONV;
    double time_FFT;
    double time_COPY_IN;
    double time_COPY_OUT;
    double time_else;
} NFFTTimeSpec;

/** Per-axis transform constants, src/cunfft_typedefs.h:98-112. */
typedef struct {
    uint_t N[CUNFFT_MAX_DIM];
    uint_t n[CUNFFT_MAX_DIM];
    double b[CUNFFT_MAX_DIM];
    double bsqrt[CUNFFT_MAX_DIM];
    double fac[CUNFFT_MAX_DIM];
    double sigma[CUNFFT_MAX_DIM];
} kdata;

/* Plan fields of src/cunfft_typedefs.h:118-178 (same names, same meaning),
 * followed by the opaque device state and the status word. */
#define CUNFFT_PLAN_FIELDS(CT)                                                 \
    uint_t N_total;                                                            \
    uint_t M_total;                                                            \
    int d;                                                                     \
    uint_t *N;                                                                 \
    double *x;          /* pinned host nodes, x[d*j+t] in [-0.5,0.5) */        \
    CT *f_hat;          /* pinned host coefficients, N_total */                \
    CT *f;              /* pinned host samples, M_total */                     \
    double *sigma;                                                             \
    uint_t *n;                                                                 \
    uint_t n_total;                                                            \
    int m;                                                                     \
    double *b;                                                                 \
    CT *g;              /* pinned host grid, n_total (copy_g_ToHost) */        \
    CT *g_hat;          /* unused, kept for layout parity */                   \
    double *fg_exp_l_cpu; /* exp(-l^2/b_t), (2m+2) per axis */                 \
    CT *g_gpu;                                                                 \
    CT *fhat_gpu;                                                              \
    double *x_gpu;                                                             \
    CT *f_gpu;                                                                 \
    cunfftHandle fft_forward;                                                  \
    cunfftHandle fft_backward;                                                 \
    NFFTTimeSpec CUNFFTTimes;                                                  \
    unsigned flags;                                                            \
    void *impl;                                                                \
    int status;

typedef struct { CUNFFT_PLAN_FIELDS(cunfftDoubleComplex) } cunfft_plan;
typedef struct { CUNFFT_PLAN_FIELDS(cunfftFloatComplex) } cunfftf_plan;

/* ======================================================= FP64 family ==== */
/* Each declaration cites the reference declaration it replaces. */

/* src/cunfft.h:35 */
void warmUp_createPlan(uint_t *n);
/* src/cunfft.h:47 */
void copyDataToDevice(cunfft_plan *plan);
/* src/cunfft.h:57 */
void copyDataToDeviceAd(cunfft_plan *plan);
/* src/cunfft.h:66 */
void copyDataToHost(cunfft_plan *plan);
/* src/cunfft.h:75 */
void copyDataToHostAd(cunfft_plan *plan);
/* src/cunfft.h:83 */
void copy_f_hat_ToHost(cunfft_plan *plan, cunfftDoubleComplex *arr);
/* src/cunfft.h:91 */
void copy_f_ToHost(cunfft_plan *plan, cunfftDoubleComplex *arr);
/* src/cunfft.h:98 */
void copy_g_ToHost(cunfft_plan *plan);
/* src/cunfft.h:106 */
void copy_x_toHost(cunfft_plan *plan, double *arr);
/* src/cunfft.h:117 */
void cunfft_transform(cunfft_plan *plan);
/* src/cunfft.h:124 */
void cunfft_adjoint(cunfft_plan *plan);
/* src/cunfft.h:133 (cufftHandle -> cunfftHandle) */
void cunfft_initPlan(int dim, uint_t *n, cunfftHandle *h);
/* src/cunfft.h:139 */
void cunfft_freePlan(cunfft_plan *p);
/* src/cunfft.h:145 */
void cunfft_initDevicePara(cunfft_plan *plan);
/* src/cunfft.h:156 */
void cunfft_init(cunfft_plan *plan, int dim, uint_t *N, uint_t M_total);
/* src/cunfft.h:168 */
void cunfft_initGuru(cunfft_plan *plan, int dim, uint_t *N, uint_t M_total,
                     uint_t *n, unsigned flags);
/* src/cunfft.h:177 */
void cunfft_initHelp(cunfft_plan *plan);
/* src/cunfft.h:185 */
void cunfft_reinit(cunfft_plan *plan);
/* src/cunfft.h:194 */
void cunfft_reinitAd(cunfft_plan *plan);
/* src/cunfft.h:200 */
void cunfft_finalize(cunfft_plan *plan);
/* src/cunfft.h:206 */
void cunfft_finalizeDevice(cunfft_plan *plan);
/* src/cunfft.h:212 */
void resetCUNFFTTimes(NFFTTimeSpec *times);
/* src/cunfft.h:222 */
void cundft_transform(cunfft_plan *plan);
/* src/cunfft.h:228 */
void cundft_adjoint(cunfft_plan *plan);
/* src/cunfft.h:242; FILTER_SIZE = 2*CUNFFT_DEFAULT_CUT_OFF+2 entries */
void nfft_init_fg_exp_l(double *fg_exp_l, const double b, const int offset);
/* src/cunfft.h:256 */
void setProperties(uint_t *N, uint_t *n, double *b, double *sigma, int dim,
                   kdata *k_cpu);
/* src/cunfft.h:261: synchronizes the device and releases this library's
 * cached device state; the primary context is left alive. */
void resetDevice(void);

/* ---- additions (no counterpart in the reference) ---- */

/* cunfft_initGuru with a runtime cut-off m (the reference fixes
 * m = CUT_OFF at compile time, src/cunfft.cu:294,322). */
void cunfft_initGuru_m(cunfft_plan *plan, int dim, uint_t *N, uint_t M_total,
                       uint_t *n, int m, unsigned flags);
/* Status of the last call on this plan (CUNFFT_SUCCESS or CUNFFT_ERR_*). */
int cunfft_last_error(const cunfft_plan *plan);
/* FP32 variant of nfft_init_fg_exp_l with explicit cut-off. */
void nfft_init_fg_exp_l_m(double *fg_exp_l, const double b, const int offset, int m);
/* Text for a status code. */
const char *cunfft_error_string(int status);
/* Message of the last error recorded by this thread (empty if none). */
const char *cunfft_last_message(void);

/* Device-pointer entry points: operate on caller-owned device buffers on the
 * caller's HIP stream (NULL: the HIP default stream).  They return the
 * status; the node binning's buffers are allocated by the first transform of
 * a plan, nothing afterwards.  Buffers: fhat N_total complex, f M_total
 * complex, x d*M_total doubles. */
/* Replace the plan's nodes (device copy) and rebuild the node binning. */
int cunfft_set_nodes_dev(cunfft_plan *plan, const double *x_dev, void *stream);
/* f = A fhat (forward NFFT, src/cunfft.cu:196-207). */
int cunfft_transform_dev(cunfft_plan *plan, const cunfftDoubleComplex *fhat_dev,
                         cunfftDoubleComplex *f_dev, void *stream);
/* fhat = A^H f (adjoint NFFT, src/cunfft.cu:209-224). */
int cunfft_adjoint_dev(cunfft_plan *plan, const cunfftDoubleComplex *f_dev,
                       cunfftDoubleComplex *fhat_dev, void *stream);
/* fhat = A^H f like cunfft_adjoint_dev, without any device allocation (no
 * counterpart in the reference; for callers that cannot allocate: code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The node binning, sorted coefficients and partial tiles live in a
 * device-resident scratch array of 64 MiB that the library holds per GPU
 * until the process ends; a plan that needs more (roughly M_total d 8 plus
 * the tiles' slots, e.g. 2D 256^2 with 65536 nodes takes ~21 MiB) is refused
 * with CUNFFT_ERR_UNSUPPORTED.  The nodes are re-binned on every call.
 * Calls that overlap on one GPU (other streams, other threads, other plans)
 * share the array and race: serialize them.  Needs the plan's adjoint FFT
 * tables (CUFFT_INIT with CUFFT_BOTH or CUFFT_BACKW), else CUNFFT_ERR_STATE. */
int cunfft_adjoint_noalloc_dev(cunfft_plan *plan, const cunfftDoubleComplex *f_dev,
                               cunfftDoubleComplex *fhat_dev, void *stream);
/* Backward of f = A(x) fhat with respect to x: gx[j*d+t] =
 * Re(conj(v_j) * d f_j / d x_jt) for upstream v, the derivative-of-Gaussian
 * gather over the forward grid of fhat (weights recomputed, not stored). */
int cunfft_grad_x_dev(cunfft_plan *plan, const cunfftDoubleComplex *fhat_dev,
                      const cunfftDoubleComplex *v_dev, double *gx_dev, void *stream);
/* Backward of f = A(x) fhat (no counterpart in the reference; the gradient
 * entry point of a differentiable NFFT layer).  For an upstream gradient
 * grad_f in the conjugate-Wirtinger convention (for L = 0.5||f - y||^2 it is
 * f - y): grad_fhat = A^H grad_f (the adjoint NFFT) and grad_x[j*d+t] =
 * Re(conj(grad_f_j) * d f_j / d x_jt).  Either output may be NULL. */
int cunfft_backward_dev(cunfft_plan *plan, const cunfftDoubleComplex *fhat_dev,
                        const cunfftDoubleComplex *grad_f_dev, cunfftDoubleComplex *grad_fhat_dev,
                        double *grad_x_dev, void *stream);
/* Wait for all work queued on the plan's stream (CUNFFT_ASYNC_COPY plans). */
int cunfft_synchronize(cunfft_plan *plan);
/* Inverse NFFT (no counterpart in the reference's code; its README:1 names
 * it, and NFFT3's solver module, which the reference links for its CPU side,
 * provides it): conjugate gradients on the normal equations (CGNR) for
 *   min_fhat  sum_j w_j |y_j - (A fhat)_j|^2
 * with the plan's nodes.  y: M_total complex; w: M_total weights or NULL
 * (all 1, e.g. density compensation otherwise); fhat: the initial guess in,
 * the solution out.  Stops after max_iter iterations, once
 * |y - A fhat|_W <= tol |y|_W, or once the normal-equation residual
 * |A^H W (y - A fhat)| falls to tol times its start (the least-squares
 * solution of an inconsistent system) or grows tenfold over its smallest
 * value (loss of orthogonality; fhat is then the iterate of the smallest
 * |A^H W (y - A fhat)| seen, which is checked every 8 iterations).
 * res_hist (NULL or max_iter + 1 doubles)
 * receives the relative residual of each iteration, *iters (may be NULL)
 * the iterations done.  Allocates its work vectors and frees them; runs on
 * `stream` and synchronizes it. */
int cunfft_solve_dev(cunfft_plan *plan, const cunfftDoubleComplex *y_dev, const double *w_dev,
                     cunfftDoubleComplex *fhat_dev, int max_iter, double tol,
                     double *res_hist, int *iters, void *stream);
/* Gridding kernel configurations (no counterpart in the reference, whose
 * tile shape and kernels are fixed at compile time, src/cunfft_kernel.cu).
 * Configuration c < CUNFFT_TILE_SHAPES is the plane-tiled kernel pair (one
 * node per wavefront instruction, lanes over the rows of the node's window
 * planes) on tile shape c -- 2D: 32x32, 16x16, 64x64; 3D: 8x8x8, 8x8x16,
 * 16x16x8; configuration CUNFFT_TILE_SHAPES is the node-lane pair (64 nodes
 * per wavefront instruction; 32x32 / 8x8x8 tiles).  1D plans have the one
 * configuration 0.  The default is 0 (CUNFFT_TILE_SHAPES with the
 * CUNFFT_NODE_LANES flag). */
#define CUNFFT_TILE_SHAPES 3
#define CUNFFT_TUNE_CONFIGS (CUNFFT_TILE_SHAPES + 1)
#define CUNFFT_TUNE_FORWARD 1
#define CUNFFT_TUNE_ADJOINT 2
/* Autotuner: times every configuration that applies on the plan's current
 * nodes (set them first) in each direction of `directions`
 * (CUNFFT_TUNE_FORWARD and/or CUNFFT_TUNE_ADJOINT), best of `reps` (<= 0: 5),
 * and keeps the fastest per direction (the two directions may differ; each
 * then has its own node binning).  With both directions tuned and their
 * fastest configurations different, it also times a forward + adjoint step
 * with each of the two on both directions (one shared binning) and keeps
 * whichever of the three steps is fastest; the pair of different
 * configurations only if its step is more than 2% faster than the shared
 * one's (a shared binning takes half the memory and one sort per node set).
 * ms_out (may be NULL) receives
 * 2 * CUNFFT_TUNE_CONFIGS times in milliseconds, the forward times of the
 * configurations followed by the adjoint ones, -1 where a configuration does
 * not apply or a direction was not tuned.  Allocates and frees temporaries,
 * runs on `stream` and synchronizes it. */
int cunfft_autotune(cunfft_plan *plan, int directions, int reps, float *ms_out, void *stream);
/* The configuration of one direction (CUNFFT_TUNE_FORWARD or
 * CUNFFT_TUNE_ADJOINT); -1 for an uninitialised plan or another direction. */
int cunfft_get_config(const cunfft_plan *plan, int direction);
/* Select a configuration for the given directions; the binning is rebuilt by
 * the next transform.  Synchronizes the device.  A configuration that does
 * not exist (CUNFFT_ERR_INVALID_ARG) or whose LDS tile does not fit the
 * plan's window (CUNFFT_ERR_UNSUPPORTED) is refused, and the plan stays as
 * it was. */
int cunfft_set_config(cunfft_plan *plan, int directions, int config);
/* Direct NDFT on the GPU in FP64 for the listed outputs only (rows of f or
 * entries of fhat by flat index); rows == NULL means all outputs.  An index
 * outside [0, M_total) (rows) or [0, N_total) (columns) yields NaN for that
 * output; a negative count is CUNFFT_ERR_INVALID_ARG. */
int cundft_transform_rows_dev(cunfft_plan *plan, const cunfftDoubleComplex *fhat_dev,
                              const long long *rows_dev, long long n_rows,
                              cunfftDoubleComplex *out_dev, void *stream);
int cundft_adjoint_cols_dev(cunfft_plan *plan, const cunfftDoubleComplex *f_dev,
                            const long long *cols_dev, long long n_cols,
                            cunfftDoubleComplex *out_dev, void *stream);

/* ---- batches (no counterpart in the reference) ----
 * Many uneven problems and right-hand sides in one graph launch: P problems
 * of one dimension, cut-off m and precision, problem i with its own N_i, n_i
 * and M_i nodes, and R right-hand sides each.  Every stage of the transform
 * is one launch over all P*R (problem, right-hand side) pairs; the stage
 * launches are captured into a hipGraph once and replayed per call.  Each
 * problem is binned once per node set, for all its right-hand sides.  This
 * replaces a loop of cunfft_transform / cunfft_adjoint calls over single
 * plans (src/cunfft.cu:196-224). */
typedef struct cunfft_batch cunfft_batch;
/* precision 64 (FP64) or 32 (FP32); N and n hold dim entries per problem
 * (n == NULL: n_t = 2 * next_pow2(N_t) as cunfft_init, src/cunfft.cu:296-299);
 * M one per problem.  All device buffers are allocated here.  Every problem
 * must suit the tiled gridding kernels (n_t >= 256 in 1D, 32 in 2D, 8 in 3D;
 * m <= 8), else CUNFFT_ERR_UNSUPPORTED. */
int cunfft_batch_create(cunfft_batch **batch, int precision, int dim, int n_problems,
                        const uint_t *N, const uint_t *n, const uint_t *M, int m, int n_rhs);
void cunfft_batch_destroy(cunfft_batch *batch);
/* Nodes of one problem (d*M_i doubles, device); binned by the next execute. */
int cunfft_batch_set_nodes_dev(cunfft_batch *batch, int problem, const double *x_dev,
                               void *stream);
/* The batch's device buffers of one problem, R right-hand sides each, one
 * after the other: fhat_in (R*N_total, forward input), f_out (R*M, forward
 * output), f_in (R*M, adjoint input), fhat_out (R*N_total, adjoint output).
 * Complex of the batch's precision; any output pointer may be NULL. */
int cunfft_batch_buffers(cunfft_batch *batch, int problem, void **fhat_in, void **f_out,
                         void **f_in, void **fhat_out);
/* Forward (CUNFFT_TUNE_FORWARD: f_out = A fhat_in) and/or adjoint
 * (CUNFFT_TUNE_ADJOINT: fhat_out = A^H f_in) of every problem and right-hand
 * side, as one graph launch on `stream` (NULL: the default stream).  The
 * first call after new nodes or a new configuration bins the nodes and
 * synchronizes `stream` once. */
int cunfft_batch_execute(cunfft_batch *batch, int directions, void *stream);
/* Autotuner over the batch as a whole (cunfft_autotune for batches): times
 * each gridding configuration's graph per direction on nonzero inputs and
 * keeps the fastest; ms_out as for cunfft_autotune.  The input buffers are
 * saved before and restored after (temporary device copies); the output
 * buffers are overwritten.  Synchronizes `stream`. */
int cunfft_batch_autotune(cunfft_batch *batch, int directions, int reps, float *ms_out,
                          void *stream);
int cunfft_batch_get_config(const cunfft_batch *batch, int direction);
int cunfft_batch_set_config(cunfft_batch *batch, int directions, int config);

/* ======================================================= FP32 family ==== */

void cunfftf_copyDataToDevice(cunfftf_plan *plan);
void cunfftf_copyDataToDeviceAd(cunfftf_plan *plan);
void cunfftf_copyDataToHost(cunfftf_plan *plan);
void cunfftf_copyDataToHostAd(cunfftf_plan *plan);
void cunfftf_copy_f_hat_ToHost(cunfftf_plan *plan, cunfftFloatComplex *arr);
void cunfftf_copy_f_ToHost(cunfftf_plan *plan, cunfftFloatComplex *arr);
void cunfftf_copy_g_ToHost(cunfftf_plan *plan);
void cunfftf_copy_x_toHost(cunfftf_plan *plan, double *arr);
void cunfftf_transform(cunfftf_plan *plan);
void cunfftf_adjoint(cunfftf_plan *plan);
void cunfftf_initPlan(int dim, uint_t *n, cunfftHandle *h);
void cunfftf_freePlan(cunfftf_plan *p);
void cunfftf_initDevicePara(cunfftf_plan *plan);
void cunfftf_init(cunfftf_plan *plan, int dim, uint_t *N, uint_t M_total);
void cunfftf_initGuru(cunfftf_plan *plan, int dim, uint_t *N, uint_t M_total,
                      uint_t *n, unsigned flags);
void cunfftf_initHelp(cunfftf_plan *plan);
void cunfftf_reinit(cunfftf_plan *plan);
void cunfftf_reinitAd(cunfftf_plan *plan);
void cunfftf_finalize(cunfftf_plan *plan);
void cunfftf_finalizeDevice(cunfftf_plan *plan);
void cunfftf_ndft_transform(cunfftf_plan *plan);
void cunfftf_ndft_adjoint(cunfftf_plan *plan);
void cunfftf_initGuru_m(cunfftf_plan *plan, int dim, uint_t *N, uint_t M_total,
                        uint_t *n, int m, unsigned flags);
int cunfftf_last_error(const cunfftf_plan *plan);
int cunfftf_set_nodes_dev(cunfftf_plan *plan, const double *x_dev, void *stream);
int cunfftf_transform_dev(cunfftf_plan *plan, const cunfftFloatComplex *fhat_dev,
                          cunfftFloatComplex *f_dev, void *stream);
int cunfftf_adjoint_dev(cunfftf_plan *plan, const cunfftFloatComplex *f_dev,
                        cunfftFloatComplex *fhat_dev, void *stream);
int cunfftf_grad_x_dev(cunfftf_plan *plan, const cunfftFloatComplex *fhat_dev,
                       const cunfftFloatComplex *v_dev, float *gx_dev, void *stream);
int cunfftf_backward_dev(cunfftf_plan *plan, const cunfftFloatComplex *fhat_dev,
                         const cunfftFloatComplex *grad_f_dev, cunfftFloatComplex *grad_fhat_dev,
                         float *grad_x_dev, void *stream);
int cunfftf_synchronize(cunfftf_plan *plan);
int cunfftf_solve_dev(cunfftf_plan *plan, const cunfftFloatComplex *y_dev, const float *w_dev,
                      cunfftFloatComplex *fhat_dev, int max_iter, double tol,
                      double *res_hist, int *iters, void *stream);
int cunfftf_autotune(cunfftf_plan *plan, int directions, int reps, float *ms_out, void *stream);
int cunfftf_get_config(const cunfftf_plan *plan, int direction);
int cunfftf_set_config(cunfftf_plan *plan, int directions, int config);
int cunfftf_ndft_transform_rows_dev(cunfftf_plan *plan, const cunfftFloatComplex *fhat_dev,
                                    const long long *rows_dev, long long n_rows,
                                    cunfftFloatComplex *out_dev, void *stream);
int cunfftf_ndft_adjoint_cols_dev(cunfftf_plan *plan, const cunfftFloatComplex *f_dev,
                                  const long long *cols_dev, long long n_cols,
                                  cunfftFloatComplex *out_dev, void *stream);

/* --------------------------------------------- reference-name mapping ---- */

#ifdef CUNFFT_SINGLE_PRECISION
typedef cunfftFloatComplex gpuComplex;
typedef float dTyp;
#define cunfft_plan cunfftf_plan
#define copyDataToDevice cunfftf_copyDataToDevice
#define copyDataToDeviceAd cunfftf_copyDataToDeviceAd
#define copyDataToHost cunfftf_copyDataToHost
#define copyDataToHostAd cunfftf_copyDataToHostAd
#define copy_f_hat_ToHost cunfftf_copy_f_hat_ToHost
#define copy_f_ToHost cunfftf_copy_f_ToHost
#define copy_g_ToHost cunfftf_copy_g_ToHost
#define copy_x_toHost cunfftf_copy_x_toHost
#define cunfft_transform cunfftf_transform
#define cunfft_adjoint cunfftf_adjoint
#define cunfft_initPlan cunfftf_initPlan
#define cunfft_freePlan cunfftf_freePlan
#define cunfft_initDevicePara cunfftf_initDevicePara
#define cunfft_init cunfftf_init
#define cunfft_initGuru cunfftf_initGuru
#define cunfft_initHelp cunfftf_initHelp
#define cunfft_reinit cunfftf_reinit
#define cunfft_reinitAd cunfftf_reinitAd
#define cunfft_finalize cunfftf_finalize
#define cunfft_finalizeDevice cunfftf_finalizeDevice
#define cundft_transform cunfftf_ndft_transform
#define cundft_adjoint cunfftf_ndft_adjoint
#define cunfft_initGuru_m cunfftf_initGuru_m
#define cunfft_last_error cunfftf_last_error
#define cunfft_set_nodes_dev cunfftf_set_nodes_dev
#define cunfft_transform_dev cunfftf_transform_dev
#define cunfft_adjoint_dev cunfftf_adjoint_dev
#define cunfft_grad_x_dev cunfftf_grad_x_dev
#define cunfft_backward_dev cunfftf_backward_dev
#define cunfft_synchronize cunfftf_synchronize
#define cunfft_solve_dev cunfftf_solve_dev
#define cunfft_autotune cunfftf_autotune
#define cunfft_get_config cunfftf_get_config
#define cunfft_set_config cunfftf_set_config
#define cundft_transform_rows_dev cunfftf_ndft_transform_rows_dev
#define cundft_adjoint_cols_dev cunfftf_ndft_adjoint_cols_dev
#else
typedef cunfftDoubleComplex gpuComplex;
typedef double dTyp;
#endif

#ifdef __cplusplus
}
#endif
#endif /* CUNFFT_AMD_H_ */
