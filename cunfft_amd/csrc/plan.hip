// plan.hip -- plan lifecycle, step orchestration and the C ABI of cunfft.h.
//
// Mirrors the reference plan surface (src/cunfft.h:35-261, src/cunfft.cu:60-444)
// with per-plan device state, one HIP stream per plan, runtime m and
// precision, status codes instead of exit(), and the additive device-pointer
// entry points.  Step order:
//   cunfft_transform  (src/cunfft.cu:196-207): FFT passes with the roll-off/pad
//                     fused into the first pass, then the gather.
//   cunfft_adjoint    (src/cunfft.cu:209-224): scatter, then FFT passes with the
//                     crop/roll-off fused into the last pass.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "cunfft.h"
#include "internal.hpp"
#include "bin.hpp"

namespace cunfft {

static_assert(kTileVariants == CUNFFT_TILE_SHAPES, "tile shapes of the C ABI and the kernels");

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        const int code = (e == hipErrorOutOfMemory) ? CUNFFT_ERR_ALLOC : CUNFFT_ERR_HIP;
        (void)hipGetLastError();
        throw Error(code, std::string(what) + ": " + hipGetErrorString(e));
    }
}

static thread_local std::string g_last_message;

template <typename T> struct Traits;
template <> struct Traits<double> {
    using plan = cunfft_plan;
    using hc = cunfftDoubleComplex;
};
template <> struct Traits<float> {
    using plan = cunfftf_plan;
    using hc = cunfftFloatComplex;
};

// ------------------------------------------------------------ FFT tables --
FftTables *make_fft_tables(int prec, int d, const int64_t *n)
{
    FftTables *ft = new FftTables();
    ft->prec = prec;
    ft->d = d;
    for (int t = 0; t < 3; ++t) {
        ft->n[t] = t < d ? n[t] : 1;
        ft->tw[t] = nullptr;
        ft->tw2[t] = nullptr;
    }
    try {
        // one table per line length: the axis, or its two factors when it is split
        for (int k = 0; k < 2 * d; ++k) {
            const int t = k / 2;
            int64_t L = n[t], L1 = 0, L2 = 0;
            if (!fft_line_ok(L)) {
                if (!fft_split(L, L1, L2)) throw Error(CUNFFT_ERR_UNSUPPORTED, "FFT length");
                L = k % 2 == 0 ? L1 : L2;
            } else if (k % 2 == 1) {
                continue;
            }
            void *&dst = k % 2 == 0 ? ft->tw[t] : ft->tw2[t];
            const size_t es = prec == 64 ? sizeof(double2) : sizeof(float2);
            std::vector<unsigned char> host((size_t)L * es);
            for (int64_t q = 0; q < L; ++q) {
                // angle reduced to the first octant symmetry-free form in long double
                const long double ang = -2.0L * 3.141592653589793238462643383279502884L *
                                        (long double)q / (long double)L;
                const long double c = cosl(ang), s = sinl(ang);
                if (prec == 64) {
                    double2 v;
                    v.x = (double)c;
                    v.y = (double)s;
                    memcpy(host.data() + q * es, &v, es);
                } else {
                    float2 v;
                    v.x = (float)c;
                    v.y = (float)s;
                    memcpy(host.data() + q * es, &v, es);
                }
            }
            CUNFFT_HIP(hipMalloc(&dst, host.size()));
            CUNFFT_HIP(hipMemcpy(dst, host.data(), host.size(), hipMemcpyHostToDevice));
        }
    } catch (...) {
        for (int t = 0; t < 3; ++t) {
            if (ft->tw[t]) (void)hipFree(ft->tw[t]);
            if (ft->tw2[t]) (void)hipFree(ft->tw2[t]);
        }
        delete ft;
        throw;
    }
    return ft;
}

void free_fft_tables(void *h)
{
    FftTables *ft = static_cast<FftTables *>(h);
    if (!ft) return;
    for (int t = 0; t < 3; ++t) {
        if (ft->tw[t]) (void)hipFree(ft->tw[t]);
        if (ft->tw2[t]) (void)hipFree(ft->tw2[t]);
    }
    delete ft;
}

// ----------------------------------------------------------------- impl --
template <typename T> struct Impl {
    hipStream_t stream = nullptr;
    Geom geom{};
    double *roc = nullptr;  // device roll-off tables, N_t entries per axis
    int64_t rocoff[3] = {0, 0, 0};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // Node binning of the forward [0] and adjoint [1] gridding kernels; the
    // adjoint shares [0] while both directions use one configuration.
    Binning bins[2];
    bool dirty[2] = {true, true};
    int config[2] = {0, 0};  // gridding kernel configuration per direction
    bool compensated = false;  // CUNFFT_COMPENSATED on an FP32 plan
    // a second grid for the axes longer than one FFT pass (their second pass
    // writes the other buffer), and the buffer the last transform left its
    // grid in (copy_g_ToHost)
    cplx_t<T> *g2 = nullptr;
    cplx_t<T> *g_last = nullptr;
    int slot(int dir) const { return config[1] == config[0] ? 0 : dir; }
    void nodes_changed() { dirty[0] = dirty[1] = true; }
};

template <typename T> static Impl<T> *impl_of(typename Traits<T>::plan *p)
{
    if (!p->impl) throw Error(CUNFFT_ERR_STATE, "plan is not initialised or was finalised");
    return static_cast<Impl<T> *>(p->impl);
}

static int64_t next_pow2(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}


static int ilog2(int64_t v)
{
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

template <typename T> static void validate(const typename Traits<T>::plan *p)
{
    if (p->d < 1 || p->d > 3) throw Error(CUNFFT_ERR_INVALID_ARG, "d must be 1, 2 or 3");
    if (p->m < 1 || p->m > CUNFFT_MAX_CUT_OFF)
        throw Error(CUNFFT_ERR_INVALID_ARG, "cut-off m must be in [1, 16]");
    if (p->M_total < 1) throw Error(CUNFFT_ERR_INVALID_ARG, "M_total must be >= 1");
    if (!p->N || !p->n) throw Error(CUNFFT_ERR_INVALID_ARG, "N and n must not be NULL");
    for (int t = 0; t < p->d; ++t) {
        if (p->N[t] < 2 || (p->N[t] & 1))
            throw Error(CUNFFT_ERR_INVALID_ARG, "N_t must be even and >= 2");
        if (!fft_length_ok(p->n[t]))
            throw Error(CUNFFT_ERR_UNSUPPORTED, "n_t must be 2^a 3^b with a >= 1 (and <= 2^24)");
        if (p->n[t] < p->N[t]) throw Error(CUNFFT_ERR_INVALID_ARG, "n_t must be >= N_t");
        if (p->n[t] < 2 * p->m + 2)
            throw Error(CUNFFT_ERR_INVALID_ARG, "n_t must be >= 2m+2 (window wider than grid)");
        // the sinc window's roll-off 1 / M_2m(k / (n a)) is infinite at the band
        // edge when sigma_t = 1
        if ((p->flags & CUNFFT_SINC) && p->n[t] <= p->N[t])
            throw Error(CUNFFT_ERR_INVALID_ARG, "CUNFFT_SINC needs n_t > N_t");
    }
    if ((p->flags & CUNFFT_SINC) && (p->flags & CUNFFT_KAISER_BESSEL))
        throw Error(CUNFFT_ERR_INVALID_ARG, "CUNFFT_SINC and CUNFFT_KAISER_BESSEL exclude each other");
}

template <typename T> static void zero_fields(typename Traits<T>::plan *p)
{
    memset((void *)p, 0, sizeof(*p));
}

template <typename T> struct Impl;

template <typename T> static void free_device(typename Traits<T>::plan *p)
{
    if (p->impl) {
        Impl<T> *im = static_cast<Impl<T> *>(p->impl);
        if (im->g2) (void)hipFree(im->g2);
        im->g2 = im->g_last = nullptr;
    }
    if (p->g_gpu) (void)hipFree(p->g_gpu);
    if (p->fhat_gpu) (void)hipFree(p->fhat_gpu);
    if (p->x_gpu) (void)hipFree(p->x_gpu);
    if (p->f_gpu) (void)hipFree(p->f_gpu);
    p->g_gpu = nullptr;
    p->fhat_gpu = nullptr;
    p->x_gpu = nullptr;
    p->f_gpu = nullptr;
}

// src/cunfft.cu:262-280: allocate and zero the device arrays.
template <typename T> static void init_device_para(typename Traits<T>::plan *p)
{
    using C2 = typename Traits<T>::hc;
    free_device<T>(p);
    CUNFFT_HIP(hipMalloc((void **)&p->g_gpu, sizeof(C2) * p->n_total));
    CUNFFT_HIP(hipMalloc((void **)&p->fhat_gpu, sizeof(C2) * p->N_total));
    CUNFFT_HIP(hipMalloc((void **)&p->x_gpu, sizeof(double) * p->M_total * p->d));
    CUNFFT_HIP(hipMalloc((void **)&p->f_gpu, sizeof(C2) * p->M_total));
    CUNFFT_HIP(hipMemset(p->g_gpu, 0, sizeof(C2) * p->n_total));
    CUNFFT_HIP(hipMemset(p->fhat_gpu, 0, sizeof(C2) * p->N_total));
    CUNFFT_HIP(hipMemset(p->x_gpu, 0, sizeof(double) * p->M_total * p->d));
    CUNFFT_HIP(hipMemset(p->f_gpu, 0, sizeof(C2) * p->M_total));
    if (p->impl) {
        Impl<T> *im = static_cast<Impl<T> *>(p->impl);
        im->nodes_changed();
        bool split = false;
        for (int t = 0; t < p->d; ++t) split = split || !fft_line_ok(p->n[t]);
        if (split) CUNFFT_HIP(hipMalloc((void **)&im->g2, sizeof(C2) * p->n_total));
        im->g_last = reinterpret_cast<cplx_t<T> *>(p->g_gpu);
    }
}

// Window shape b_t = 2 sigma_t / (2 sigma_t - 1) * m / pi with sigma_t = n_t / N_t
// (src/cunfft.cu:351-357) and the gridding geometry of one transform.
Geom make_geom(int d, const int64_t *N, const int64_t *n, int64_t M, int m, int win)
{
    Geom ge{};
    ge.d = d;
    ge.m = m;
    ge.W = 2 * m + 2;
    ge.M = M;
    ge.win = win;
    for (int t = 0; t < 3; ++t) {
        ge.N[t] = t < d ? N[t] : 1;
        ge.n[t] = t < d ? n[t] : 1;
        if (t < d) {
            const double sigma = (double)n[t] / (double)N[t];
            const double b = (2.0 * sigma) / (2.0 * sigma - 1.0) * ((double)m / M_PI);
            ge.invb[t] = 1.0 / b;
            ge.nrm[t] = 1.0 / sqrt(M_PI * b);
            // Kaiser-Bessel shape of NFFT3: b = pi (2 - 1 / sigma); sinc power:
            // a = (2 sigma - 1) / (2 sigma m)
            ge.shp[t] = win == WIN_SINC ? (2.0 * sigma - 1.0) / (2.0 * sigma * (double)m)
                                        : M_PI * (2.0 - 1.0 / sigma);
        } else {
            ge.invb[t] = ge.nrm[t] = ge.shp[t] = 0.0;
        }
    }
    return ge;
}

// Modified Bessel I0 by its power series in long double (z <= m pi 2 < 101).
static double bessel_i0(double z)
{
    const long double q = (long double)z * z / 4.0L;
    long double term = 1.0L, sum = 1.0L;
    for (int j = 1; j < 1000; ++j) {
        term *= q / ((long double)j * j);
        sum += term;
        if (term < 1e-22L * sum) break;
    }
    return (double)sum;
}

// Centred cardinal B-spline M_p(x) of order p (support [-p/2, p/2], unit
// integral), the Fourier transform of the sinc-power window: de Boor's
// recursion on the knots 0..p at t = x + p/2, over the p splines of order 1
// around the knot interval of t only.
static double bspline_centred(int p, double x)
{
    const long double t = (long double)x + 0.5L * p;
    if (t <= 0.0L || t >= (long double)p) return 0.0;
    const int j = (int)floorl(t);
    long double B[2 * CUNFFT_MAX_CUT_OFF + 2] = {};   // B[i] = B_k^(i)(t), i <= j
    B[j] = 1.0L;
    for (int k = 2; k <= p; ++k)
        for (int i = j - k + 1 < 0 ? 0 : j - k + 1; i <= j; ++i) {
            const long double right = i + 1 <= p - 1 ? B[i + 1] : 0.0L;
            B[i] = ((t - i) * B[i] + (i + k - t) * right) / (k - 1);
        }
    return (double)B[0];
}

// Device roll-off tables by pruned index q < N_t: exp(fac_t k^2) with
// k = q or q - N_t and fac_t = pi^2 b_t / n_t^2 (src/cunfft.cu:79-90), or
// 1 / phihat_t(k / n_t) of the other windows.
double *make_rolloff(const Geom &ge, int64_t (&rocoff)[3])
{
    std::vector<double> roc;
    for (int t = 0; t < 3; ++t) rocoff[t] = 0;
    for (int t = 0; t < ge.d; ++t) {
        rocoff[t] = (int64_t)roc.size();
        const double sigma = (double)ge.n[t] / (double)ge.N[t];
        const double b = (2.0 * sigma) / (2.0 * sigma - 1.0) * ((double)ge.m / M_PI);
        const double fac = M_PI * M_PI * b / ((double)ge.n[t] * (double)ge.n[t]);
        const int64_t Nn = ge.N[t];
        for (int64_t q = 0; q < Nn; ++q) {
            const int64_t k = q < Nn / 2 ? q : q - Nn;
            if (ge.win == WIN_KAISER_BESSEL) {
                const double w = 2.0 * M_PI * (double)k / (double)ge.n[t];
                const double z = (double)ge.m * sqrt(ge.shp[t] * ge.shp[t] - w * w);
                roc.push_back(1.0 / bessel_i0(z));
            } else if (ge.win == WIN_SINC) {
                const double nu = (double)k / ((double)ge.n[t] * ge.shp[t]);
                roc.push_back(1.0 / bspline_centred(2 * ge.m, nu));
            } else {
                roc.push_back(exp(fac * (double)k * (double)k));
            }
        }
    }
    double *dev = nullptr;
    CUNFFT_HIP(hipMalloc((void **)&dev, sizeof(double) * roc.size()));
    const hipError_t e =
        hipMemcpy(dev, roc.data(), sizeof(double) * roc.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) (void)hipFree(dev);
    CUNFFT_HIP(e);
    return dev;
}

// src/cunfft.cu:60-76: exp(-l^2 / b) for l = 0 .. 2m+1 (evaluated directly).
static void fg_exp_l(double *out, double b, int offset, int m)
{
    for (int l = 0; l < 2 * m + 2; ++l) out[offset + l] = exp(-(double)l * (double)l / b);
}

// Gridding configuration c: c < kTileVariants is the plane-tiled kernel pair
// with tile shape c, c == kTileVariants the node-lane kernels.
static bool config_exists(int d, int c) { return c >= 0 && c <= kTileVariants && (d >= 2 || c == 0); }

// False when the tiled kernels of configuration c do not apply; a
// compensated FP32 plan of d >= 2 also refuses the node-lane configuration,
// which has no compensated variant.
static bool plan_bins(Binning &b, const Geom &ge, int elem_bytes, int c, bool comp)
{
    const bool lanes = c == kTileVariants;
    b.compensated = comp && elem_bytes == 4;
    if (b.compensated && lanes && ge.d >= 2) {
        b.usable = 0;
        return false;
    }
    return b.plan(ge, elem_bytes, lanes ? 1 : 0, lanes ? 0 : c);
}

// Forward configuration cf, adjoint ca: re-plans the binnings (their buffers
// are re-allocated by the next transform); false when the tiled kernels of
// either do not apply (that direction then uses the untiled kernels).
template <typename T> static bool plan_configs(Impl<T> *im, int cf, int ca)
{
    im->bins[0].release();
    im->bins[1].release();
    im->nodes_changed();
    im->config[0] = cf;
    im->config[1] = ca;
    bool ok = plan_bins(im->bins[0], im->geom, (int)sizeof(T), cf, im->compensated);
    if (ca != cf) ok = plan_bins(im->bins[1], im->geom, (int)sizeof(T), ca, im->compensated) && ok;
    return ok;
}

// src/cunfft.cu:334-378 (cunfft_initHelp) on the fields set by init/initGuru.
template <typename T> static void init_help(typename Traits<T>::plan *p)
{
    using C2 = typename Traits<T>::hc;
    validate<T>(p);
    const int d = p->d;
    p->N_total = 1;
    p->n_total = 1;
    for (int t = 0; t < d; ++t) {
        p->N_total *= p->N[t];
        p->n_total *= p->n[t];
    }
    p->sigma = (double *)calloc(d, sizeof(double));
    p->b = (double *)calloc(d, sizeof(double));
    p->fg_exp_l_cpu = (double *)calloc((size_t)(2 * p->m + 2) * CUNFFT_MAX_DIM, sizeof(double));
    if (!p->sigma || !p->b || !p->fg_exp_l_cpu) throw Error(CUNFFT_ERR_ALLOC, "host allocation");
    for (int t = 0; t < d; ++t) {
        p->sigma[t] = (double)p->n[t] / (double)p->N[t];
        p->b[t] = (2.0 * p->sigma[t]) / (2.0 * p->sigma[t] - 1.0) * ((double)p->m / M_PI);
        // per-axis table (the reference reuses axis 0's table for every axis,
        // src/cunfft_kernel.cu:504)
        fg_exp_l(p->fg_exp_l_cpu, p->b[t], t * (2 * p->m + 2), p->m);
    }
    memset(&p->CUNFFTTimes, 0, sizeof(p->CUNFFTTimes));
    p->x = nullptr;
    p->f_hat = nullptr;
    p->f = nullptr;
    p->g = nullptr;
    p->g_hat = nullptr;
    p->g_gpu = p->fhat_gpu = p->f_gpu = nullptr;
    p->x_gpu = nullptr;
    p->fft_forward = p->fft_backward = nullptr;
    p->impl = nullptr;
    if (!(p->flags & CUNFFT_NO_HOST_BUFFERS)) {
        CUNFFT_HIP(hipHostMalloc((void **)&p->x, sizeof(double) * p->M_total * d, hipHostMallocDefault));
        CUNFFT_HIP(hipHostMalloc((void **)&p->f_hat, sizeof(C2) * p->N_total, hipHostMallocDefault));
        CUNFFT_HIP(hipHostMalloc((void **)&p->f, sizeof(C2) * p->M_total, hipHostMallocDefault));
        CUNFFT_HIP(hipHostMalloc((void **)&p->g, sizeof(C2) * p->n_total, hipHostMallocDefault));
        memset(p->x, 0, sizeof(double) * p->M_total * d);
        memset(p->f_hat, 0, sizeof(C2) * p->N_total);
        memset(p->f, 0, sizeof(C2) * p->M_total);
        memset(p->g, 0, sizeof(C2) * p->n_total);
    }
    Impl<T> *im = new Impl<T>();
    p->impl = im;
    CUNFFT_HIP(hipStreamCreateWithFlags(&im->stream, hipStreamNonBlocking));
    for (int k = 0; k < 4; ++k) CUNFFT_HIP(hipEventCreate(&im->ev[k]));

    int64_t NN[3], nn3[3];
    for (int t = 0; t < d; ++t) {
        NN[t] = p->N[t];
        nn3[t] = p->n[t];
    }
    im->geom = make_geom(d, NN, nn3, p->M_total, p->m,
                         (p->flags & CUNFFT_KAISER_BESSEL) ? WIN_KAISER_BESSEL
                         : (p->flags & CUNFFT_SINC)        ? WIN_SINC
                                                           : WIN_GAUSSIAN);
    const int c0 = (p->flags & CUNFFT_NODE_LANES) && d >= 2 ? kTileVariants : 0;
    im->compensated = (p->flags & CUNFFT_COMPENSATED) && sizeof(T) == 4;
    bool tiled = plan_configs<T>(im, c0, c0);
    if (im->compensated && (p->flags & CUNFFT_ATOMIC_SCATTER))
        throw Error(CUNFFT_ERR_UNSUPPORTED,
                    "CUNFFT_COMPENSATED needs the tiled adjoint; CUNFFT_ATOMIC_SCATTER adds with "
                    "global atomics");
    if (im->compensated) {
        // the compensated sums live in the tiled adjoint kernels only (the
        // untiled fallback adds with global atomics): take the first tile
        // shape that holds the doubled tile, or refuse the plan
        for (int c = 1; c < kTileVariants && !tiled && c0 == 0 && d >= 2; ++c)
            tiled = plan_configs<T>(im, c, c);
        if (!tiled)
            throw Error(CUNFFT_ERR_UNSUPPORTED,
                        "CUNFFT_COMPENSATED: no compensated tiled gridding configuration fits this "
                        "plan's window and grid (or CUNFFT_NODE_LANES was asked for)");
    }
    im->roc = make_rolloff(im->geom, im->rocoff);
    init_device_para<T>(p);
    int64_t nn[3];
    for (int t = 0; t < d; ++t) nn[t] = p->n[t];
    const int prec = sizeof(T) == 8 ? 64 : 32;
    if (sizeof(T) == 8) {
        // set-up of the allocation-free adjoint's scratch on this device (its
        // address is taken once per device; a failure is reported by that
        // entry point, not here)
        try {
            (void)noalloc_scratch();
        } catch (const Error &) {
        }
    }
    if (p->flags & CUFFT_INIT) {
        if (p->flags & (CUFFT_BOTH | CUFFT_FORW)) p->fft_forward = make_fft_tables(prec, d, nn);
        if (p->flags & (CUFFT_BOTH | CUFFT_BACKW)) p->fft_backward = make_fft_tables(prec, d, nn);
    }
}

template <typename T> static void finalize(typename Traits<T>::plan *p)
{
    Impl<T> *im = static_cast<Impl<T> *>(p->impl);
    if (im && im->stream) (void)hipStreamSynchronize(im->stream);
    free_fft_tables(p->fft_forward);
    free_fft_tables(p->fft_backward);
    p->fft_forward = p->fft_backward = nullptr;
    if (p->g) (void)hipHostFree(p->g);
    if (p->f) (void)hipHostFree(p->f);
    if (p->x) (void)hipHostFree(p->x);
    if (p->f_hat) (void)hipHostFree(p->f_hat);
    p->g = p->f = p->f_hat = nullptr;
    p->x = nullptr;
    free(p->sigma);
    free(p->n);
    free(p->N);
    free(p->b);
    free(p->fg_exp_l_cpu);
    p->sigma = p->b = p->fg_exp_l_cpu = nullptr;
    p->n = p->N = nullptr;
    free_device<T>(p);
    if (im) {
        im->bins[0].release();
        im->bins[1].release();
        if (im->roc) (void)hipFree(im->roc);
        for (int k = 0; k < 4; ++k)
            if (im->ev[k]) (void)hipEventDestroy(im->ev[k]);
        if (im->stream) (void)hipStreamDestroy(im->stream);
        delete im;
    }
    p->impl = nullptr;
    p->status = CUNFFT_SUCCESS;
}

// ------------------------------------------------------------ FFT passes --
// The d axes' passes on grid g (and the plan's second grid for split axes);
// returns the buffer holding the grid after them.
template <typename T>
static cplx_t<T> *fft_passes(typename Traits<T>::plan *p, Impl<T> *im, FftTables *ft, bool forward,
                             const cplx_t<T> *fhat_in, cplx_t<T> *fhat_out, cplx_t<T> *g,
                             hipStream_t s)
{
    const int d = p->d;
    cplx_t<T> *cur = g;
    cplx_t<T> *other = im->g2 == g ? reinterpret_cast<cplx_t<T> *>(p->g_gpu) : im->g2;
    for (int k = 0; k < d; ++k) {
        const int t = forward ? d - 1 - k : k;
        const cplx_t<T> *tw[2] = {static_cast<const cplx_t<T> *>(ft->tw[t]),
                                  static_cast<const cplx_t<T> *>(ft->tw2[t])};
        FftLaunch<T> L[2];
        const int np = fft_axis_setup<T>(im->geom, tw, im->roc, im->rocoff, t, forward, fhat_in,
                                         fhat_out, &cur, &other, L);
        for (int i = 0; i < np; ++i) launch_fft_pass<T>(L[i].P, L[i].blocks, L[i].lds, s);
    }
    im->g_last = cur;
    return cur;
}

template <typename T> static FftTables *fft_tables(typename Traits<T>::plan *p, bool forward)
{
    cunfftHandle &h = forward ? p->fft_forward : p->fft_backward;
    if (!h) {
        int64_t nn[3];
        for (int t = 0; t < p->d; ++t) nn[t] = p->n[t];
        h = make_fft_tables(sizeof(T) == 8 ? 64 : 32, p->d, nn);
    }
    return static_cast<FftTables *>(h);
}

// Binning of direction dir (0 forward, 1 adjoint), built on s if the nodes
// changed since.
template <typename T>
static const Binning &ensure_bins(Impl<T> *im, int dir, const double *x, hipStream_t s)
{
    const int k = im->slot(dir);
    if (im->dirty[k]) {
        im->bins[k].build(im->geom, x, s);
        im->dirty[k] = false;
    }
    return im->bins[k];
}

// x-gradient gx[j d + t] = Re(conj(v_j) d f_j / d x_jt) over the forward grid g:
// the binned plane-tiled kernel on the forward configuration's binning, else
// the node-parallel kernel (1D, node lanes, the plain gather, or no fitting
// tile).
template <typename T>
static void grad_x(typename Traits<T>::plan *p, Impl<T> *im, const cplx_t<T> *g, const cplx_t<T> *v,
                   T *gx, hipStream_t s)
{
    if (!(p->flags & CUNFFT_PLAIN_GATHER)) {
        const Binning &b = ensure_bins<T>(im, 0, p->x_gpu, s);
        if (b.usable && b.built && b.planes() && launch_tile_grad<T>(im->geom, b, g, v, gx, s)) return;
    }
    launch_grad_x<T>(im->geom, g, p->x_gpu, v, gx, s);
}

// f = A fhat.  Returns the event-timed (fft, conv) seconds when timing.
template <typename T>
static void run_forward(typename Traits<T>::plan *p, Impl<T> *im, const cplx_t<T> *fhat,
                        cplx_t<T> *f, hipStream_t s, bool timing, double *t_fft, double *t_conv)
{
    const double *x = p->x_gpu;
    if (timing) CUNFFT_HIP(hipEventRecord(im->ev[0], s));
    const cplx_t<T> *g = fft_passes<T>(p, im, fft_tables<T>(p, true), true, fhat, nullptr,
                                       reinterpret_cast<cplx_t<T> *>(p->g_gpu), s);
    if (timing) CUNFFT_HIP(hipEventRecord(im->ev[1], s));
    if (p->flags & CUNFFT_PLAIN_GATHER) {
        launch_gather_plain<T>(im->geom, g, x, f, s);
    } else {
        launch_gather_binned<T>(im->geom, ensure_bins<T>(im, 0, x, s), g, x, f, s);
    }
    if (timing) {
        CUNFFT_HIP(hipEventRecord(im->ev[2], s));
        CUNFFT_HIP(hipEventSynchronize(im->ev[2]));
        float a = 0, b = 0;
        CUNFFT_HIP(hipEventElapsedTime(&a, im->ev[0], im->ev[1]));
        CUNFFT_HIP(hipEventElapsedTime(&b, im->ev[1], im->ev[2]));
        *t_fft = a * 1e-3;
        *t_conv = b * 1e-3;
    }
}

// fhat = A^H f.  Timing: time_CONV = adjoint gridding (tile accumulation and
// fold), time_else = the fold alone, time_FFT = the FFT passes.
template <typename T>
static void run_adjoint(typename Traits<T>::plan *p, Impl<T> *im, const cplx_t<T> *f,
                        cplx_t<T> *fhat, hipStream_t s, bool timing, double *t_fft,
                        double *t_conv, double *t_fold)
{
    cplx_t<T> *g = reinterpret_cast<cplx_t<T> *>(p->g_gpu);
    const double *x = p->x_gpu;
    bool folded = false;
    if (p->flags & CUNFFT_ATOMIC_SCATTER) {
        if (timing) CUNFFT_HIP(hipEventRecord(im->ev[0], s));
        launch_zero<T>(g, p->n_total, s);
        launch_scatter_atomic<T>(im->geom, x, f, g, s);
    } else {
        const Binning &b = ensure_bins<T>(im, 1, x, s);
        if (timing) CUNFFT_HIP(hipEventRecord(im->ev[0], s));
        folded = launch_scatter_binned<T>(im->geom, b, x, f, g, s, timing ? im->ev[3] : nullptr);
    }
    if (timing && !folded) CUNFFT_HIP(hipEventRecord(im->ev[3], s));
    if (timing) CUNFFT_HIP(hipEventRecord(im->ev[1], s));
    fft_passes<T>(p, im, fft_tables<T>(p, false), false, nullptr, fhat, g, s);
    if (timing) {
        CUNFFT_HIP(hipEventRecord(im->ev[2], s));
        CUNFFT_HIP(hipEventSynchronize(im->ev[2]));
        float a = 0, b = 0, c = 0;
        CUNFFT_HIP(hipEventElapsedTime(&a, im->ev[0], im->ev[1]));
        CUNFFT_HIP(hipEventElapsedTime(&b, im->ev[1], im->ev[2]));
        CUNFFT_HIP(hipEventElapsedTime(&c, im->ev[3], im->ev[1]));
        *t_conv = a * 1e-3;
        *t_fft = b * 1e-3;
        *t_fold = c * 1e-3;
    }
}

// ------------------------------------------------------------ autotuner --

// Times every configuration that applies on the plan's current nodes (best of
// reps forward and adjoint passes over constant data, event-timed on s) and keeps
// the fastest per tuned direction, then, with both directions tuned to
// different configurations, the fastest forward + adjoint step of that pair and
// of either configuration on both.  The binning of each candidate is built
// before its timed passes: it is paid once per node set.  ms_out: forward
// times of the configurations, then adjoint times (-1: not timed).
template <typename T>
static void autotune(typename Traits<T>::plan *p, Impl<T> *im, int dirs, int reps, float *ms_out,
                     hipStream_t s)
{
    constexpr int NC = kTileVariants + 1;
    if (dirs & ~3 || !dirs) throw Error(CUNFFT_ERR_INVALID_ARG, "directions must be 1, 2 or 3");
    if (reps <= 0) reps = 5;
    void *fh = nullptr, *f = nullptr, *fh2 = nullptr;   // fh2: the adjoint output of a step
    float ms[2][NC];
    int best[2] = {im->config[0], im->config[1]};
    auto finish = [&] {
        (void)hipStreamSynchronize(s);
        if (fh) (void)hipFree(fh);
        if (f) (void)hipFree(f);
        if (fh2) (void)hipFree(fh2);
        plan_configs<T>(im, best[0], best[1]);
    };
    try {
        CUNFFT_HIP(hipMalloc(&fh, sizeof(cplx_t<T>) * p->N_total));
        CUNFFT_HIP(hipMalloc(&f, sizeof(cplx_t<T>) * p->M_total));
        // nonzero data (words 0x3ff00000: 1.875f, or 1.0000002 as a double):
        // the node-lane scatter skips its global atomics for zero partial sums
        CUNFFT_HIP(hipMemsetD32Async((hipDeviceptr_t)fh, 0x3ff00000,
                                     sizeof(cplx_t<T>) / 4 * p->N_total, s));
        CUNFFT_HIP(hipMemsetD32Async((hipDeviceptr_t)f, 0x3ff00000,
                                     sizeof(cplx_t<T>) / 4 * p->M_total, s));
        float best_ms[2] = {-1.0f, -1.0f};
        for (int c = 0; c < NC; ++c) {
            ms[0][c] = ms[1][c] = -1.0f;
            if (!config_exists(p->d, c) || !plan_configs<T>(im, c, c)) continue;
            for (int dir = 0; dir < 2; ++dir) {
                if (!(dirs & (1 << dir))) continue;
                double a, b, e;
                auto pass = [&] {
                    if (dir == 0)
                        run_forward<T>(p, im, static_cast<const cplx_t<T> *>(fh),
                                       static_cast<cplx_t<T> *>(f), s, false, &a, &b);
                    else
                        run_adjoint<T>(p, im, static_cast<const cplx_t<T> *>(f),
                                       static_cast<cplx_t<T> *>(fh), s, false, &a, &b, &e);
                };
                pass();   // builds the binning, warms up
                float t_min = 0;
                for (int r = 0; r < reps; ++r) {
                    CUNFFT_HIP(hipEventRecord(im->ev[0], s));
                    pass();
                    CUNFFT_HIP(hipEventRecord(im->ev[1], s));
                    CUNFFT_HIP(hipEventSynchronize(im->ev[1]));
                    float t = 0;
                    CUNFFT_HIP(hipEventElapsedTime(&t, im->ev[0], im->ev[1]));
                    t_min = r == 0 ? t : std::min(t_min, t);
                }
                ms[dir][c] = t_min;
                if (best_ms[dir] < 0 || t_min < best_ms[dir]) {
                    best_ms[dir] = t_min;
                    best[dir] = c;
                }
            }
        }
        // Both directions tuned to different configurations: a forward and an
        // adjoint on one configuration share one binning (one build per node set,
        // the same node data read by both passes), so the step of each
        // direction's best on both directions is timed beside the step of the
        // pair, and the fastest step is kept, the pair only when it is more than
        // 2% faster than the faster shared one.
        if (dirs == 3 && best[0] != best[1]) {
            CUNFFT_HIP(hipMalloc(&fh2, sizeof(cplx_t<T>) * p->N_total));
            const int pairs[3][2] = {{best[0], best[1]}, {best[1], best[1]}, {best[0], best[0]}};
            int pick = 0;
            float pick_ms = -1.0f;
            for (int k = 0; k < 3; ++k) {
                const int cf = pairs[k][0], ca = pairs[k][1];
                if (ms[0][cf] < 0 || ms[1][ca] < 0 || !plan_configs<T>(im, cf, ca)) continue;
                double a, b, e;
                auto step = [&] {
                    run_forward<T>(p, im, static_cast<const cplx_t<T> *>(fh),
                                   static_cast<cplx_t<T> *>(f), s, false, &a, &b);
                    run_adjoint<T>(p, im, static_cast<const cplx_t<T> *>(f),
                                   static_cast<cplx_t<T> *>(fh2), s, false, &a, &b, &e);
                };
                step();
                float t_min = 0;
                for (int r = 0; r < reps; ++r) {
                    CUNFFT_HIP(hipEventRecord(im->ev[0], s));
                    step();
                    CUNFFT_HIP(hipEventRecord(im->ev[1], s));
                    CUNFFT_HIP(hipEventSynchronize(im->ev[1]));
                    float t = 0;
                    CUNFFT_HIP(hipEventElapsedTime(&t, im->ev[0], im->ev[1]));
                    t_min = r == 0 ? t : std::min(t_min, t);
                }
                // a shared configuration (k > 0) also halves the binning memory and
                // sorts the nodes once when they change: it is kept unless the
                // pair's step is more than 2% faster
                const float scaled = k > 0 ? t_min : t_min * 1.02f;
                if (pick_ms < 0 || scaled < pick_ms) {
                    pick_ms = scaled;
                    pick = k;
                }
            }
            best[0] = pairs[pick][0];
            best[1] = pairs[pick][1];
        }
        CUNFFT_HIP(hipStreamSynchronize(s));
    } catch (...) {
        finish();
        throw;
    }
    finish();
    if (ms_out)
        for (int dir = 0; dir < 2; ++dir)
            for (int c = 0; c < NC; ++c) ms_out[dir * NC + c] = ms[dir][c];
}

// A copy on the plan stream; it returns after the copy unless `async` (and
// not timing).
template <typename T>
static double timed_copy(Impl<T> *im, bool timing, void *dst, const void *src, size_t bytes,
                         hipMemcpyKind kind, bool async = false)
{
    if (timing) CUNFFT_HIP(hipEventRecord(im->ev[0], im->stream));
    CUNFFT_HIP(hipMemcpyAsync(dst, src, bytes, kind, im->stream));
    if (timing) CUNFFT_HIP(hipEventRecord(im->ev[1], im->stream));
    if (!async || timing) CUNFFT_HIP(hipStreamSynchronize(im->stream));
    if (!timing) return 0.0;
    float ms = 0;
    CUNFFT_HIP(hipEventElapsedTime(&ms, im->ev[0], im->ev[1]));
    return ms * 1e-3;
}

template <typename T> static void need_host(const typename Traits<T>::plan *p)
{
    if (!p->x || !p->f || !p->f_hat || !p->g)
        throw Error(CUNFFT_ERR_STATE, "plan was created with CUNFFT_NO_HOST_BUFFERS");
}

template <typename T> static void need_device(const typename Traits<T>::plan *p)
{
    if (!p->g_gpu || !p->fhat_gpu || !p->x_gpu || !p->f_gpu)
        throw Error(CUNFFT_ERR_STATE, "device arrays were released (cunfft_finalizeDevice)");
}

// ------------------------------------------------------- guarded calls --
static void report(const char *what)
{
    g_last_message = what;
    if (getenv("CUNFFT_VERBOSE")) fprintf(stderr, "cunfft: %s\n", what);
}

void report_error(const char *what) { report(what); }

template <typename P, typename F> static int guard(P *p, F &&fn)
{
    if (!p) return CUNFFT_ERR_INVALID_ARG;
    if (p->status != CUNFFT_SUCCESS) return p->status;
    try {
        fn();
    } catch (const Error &e) {
        p->status = e.code;
        report(e.what());
    } catch (const std::bad_alloc &) {
        p->status = CUNFFT_ERR_ALLOC;
        report("host allocation failed");
    } catch (const std::exception &e) {
        p->status = CUNFFT_ERR_HIP;
        report(e.what());
    }
    return p->status;
}

template <typename T>
static void api_init_guru(typename Traits<T>::plan *p, int dim, const uint_t *N, uint_t M,
                          const uint_t *n, int m, unsigned flags, bool default_n)
{
    if (!p) return;
    zero_fields<T>(p);
    guard(p, [&] {
        if (dim < 1 || dim > 3) throw Error(CUNFFT_ERR_INVALID_ARG, "d must be 1, 2 or 3");
        if (!N) throw Error(CUNFFT_ERR_INVALID_ARG, "N is NULL");
        if (!default_n && !n) throw Error(CUNFFT_ERR_INVALID_ARG, "n is NULL");
        p->d = dim;
        p->N = (uint_t *)malloc(sizeof(uint_t) * dim);
        p->n = (uint_t *)malloc(sizeof(uint_t) * dim);
        if (!p->N || !p->n) throw Error(CUNFFT_ERR_ALLOC, "host allocation");
        for (int t = 0; t < dim; ++t) {
            p->N[t] = N[t];
            // src/cunfft.cu:296-299: n_t = 2 * next_pow2(N_t)
            p->n[t] = default_n ? 2 * next_pow2(N[t]) : n[t];
        }
        p->M_total = M;
        p->m = m;
        p->flags = flags;
        init_help<T>(p);
    });
}

}  // namespace cunfft

using namespace cunfft;

// =========================================================== C ABI ======
namespace {

template <typename T> void api_copy_in(typename Traits<T>::plan *p, bool adjoint)
{
    using C2 = typename Traits<T>::hc;
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_host<T>(p);
        need_device<T>(p);
        const bool timing = p->flags & CUNFFT_MEASURE_TIME;
        const bool async = p->flags & CUNFFT_ASYNC_COPY;
        double t = 0;
        if (adjoint)
            t += timed_copy(im, timing, p->f_gpu, p->f, sizeof(C2) * p->M_total, hipMemcpyHostToDevice,
                            async);
        else
            t += timed_copy(im, timing, p->fhat_gpu, p->f_hat, sizeof(C2) * p->N_total,
                            hipMemcpyHostToDevice, async);
        t += timed_copy(im, timing, p->x_gpu, p->x, sizeof(double) * p->M_total * p->d,
                        hipMemcpyHostToDevice, async);
        im->nodes_changed();
        p->CUNFFTTimes.time_COPY_IN = t;
    });
}

template <typename T> void api_copy_out(typename Traits<T>::plan *p, bool adjoint)
{
    using C2 = typename Traits<T>::hc;
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_host<T>(p);
        need_device<T>(p);
        const bool timing = p->flags & CUNFFT_MEASURE_TIME;
        double t;
        if (adjoint)
            t = timed_copy(im, timing, p->f_hat, p->fhat_gpu, sizeof(C2) * p->N_total,
                           hipMemcpyDeviceToHost);
        else
            t = timed_copy(im, timing, p->f, p->f_gpu, sizeof(C2) * p->M_total, hipMemcpyDeviceToHost);
        p->CUNFFTTimes.time_COPY_OUT = t;
    });
}

template <typename T>
void api_copy_dev_to(typename Traits<T>::plan *p, void *dst, const void *src, size_t bytes)
{
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        if (!dst) throw Error(CUNFFT_ERR_INVALID_ARG, "destination is NULL");
        need_device<T>(p);
        timed_copy(im, false, dst, src, bytes, hipMemcpyDeviceToHost);
    });
}

// The grid of the last transform (the buffer its FFT passes left it in: the
// plan's grid, or its second grid after an odd number of split axes).
template <typename T> void api_copy_g(typename Traits<T>::plan *p)
{
    if (!p) return;
    const void *src = p->impl && static_cast<Impl<T> *>(p->impl)->g_last
                          ? static_cast<const void *>(static_cast<Impl<T> *>(p->impl)->g_last)
                          : static_cast<const void *>(p->g_gpu);
    api_copy_dev_to<T>(p, p->g, src, sizeof(typename Traits<T>::hc) * p->n_total);
}

template <typename T> void api_transform(typename Traits<T>::plan *p)
{
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        const bool timing = p->flags & CUNFFT_MEASURE_TIME;
        double tf = 0, tc = 0;
        run_forward<T>(p, im, reinterpret_cast<const cplx_t<T> *>(p->fhat_gpu),
                       reinterpret_cast<cplx_t<T> *>(p->f_gpu), im->stream, timing, &tf, &tc);
        if (!(p->flags & CUNFFT_ASYNC_COPY)) CUNFFT_HIP(hipStreamSynchronize(im->stream));
        p->CUNFFTTimes.time_ROC = 0.0;
        p->CUNFFTTimes.time_FFT = tf;
        p->CUNFFTTimes.time_CONV = tc;
        p->CUNFFTTimes.runTime = tf + tc;
    });
}

template <typename T> void api_adjoint(typename Traits<T>::plan *p)
{
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        const bool timing = p->flags & CUNFFT_MEASURE_TIME;
        double tf = 0, tc = 0, tfold = 0;
        run_adjoint<T>(p, im, reinterpret_cast<const cplx_t<T> *>(p->f_gpu),
                       reinterpret_cast<cplx_t<T> *>(p->fhat_gpu), im->stream, timing, &tf, &tc,
                       &tfold);
        if (!(p->flags & CUNFFT_ASYNC_COPY)) CUNFFT_HIP(hipStreamSynchronize(im->stream));
        p->CUNFFTTimes.time_ROC = 0.0;
        p->CUNFFTTimes.time_FFT = tf;
        p->CUNFFTTimes.time_CONV = tc;
        p->CUNFFTTimes.time_else = tfold;
        p->CUNFFTTimes.runTime = tf + tc;
    });
}

template <typename T> void api_ndft(typename Traits<T>::plan *p, bool adjoint)
{
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        int64_t N[3];
        for (int t = 0; t < p->d; ++t) N[t] = p->N[t];
        if (p->N_total > 0x7fffffff) throw Error(CUNFFT_ERR_UNSUPPORTED, "NDFT needs N_total < 2^31");
        if (adjoint)
            launch_ndft_cols<T>(p->d, N, p->M_total, p->x_gpu,
                                reinterpret_cast<const cplx_t<T> *>(p->f_gpu), nullptr, p->N_total,
                                reinterpret_cast<cplx_t<T> *>(p->fhat_gpu), im->stream);
        else
            launch_ndft_rows<T>(p->d, N, p->M_total, p->x_gpu,
                                reinterpret_cast<const cplx_t<T> *>(p->fhat_gpu), nullptr,
                                p->M_total, reinterpret_cast<cplx_t<T> *>(p->f_gpu), im->stream);
        CUNFFT_HIP(hipStreamSynchronize(im->stream));
    });
}

template <typename T> void api_reinit(typename Traits<T>::plan *p, bool adjoint)
{
    using C2 = typename Traits<T>::hc;
    guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        memset(&p->CUNFFTTimes, 0, sizeof(p->CUNFFTTimes));
        CUNFFT_HIP(hipStreamSynchronize(im->stream));
        init_device_para<T>(p);
        if (p->g) memset(p->g, 0, sizeof(C2) * p->n_total);
        if (adjoint) {
            // src/cunfft.cu:247-257 re-allocates f_hat and g without freeing them
            // and leaves f_hat uninitialised; here both are zeroed in place.
            if (p->f_hat) memset(p->f_hat, 0, sizeof(C2) * p->N_total);
        } else {
            if (p->f) memset(p->f, 0, sizeof(C2) * p->M_total);
        }
    });
}

// Device-pointer entry points run on the caller's stream; NULL is the HIP
// default (legacy null) stream, as for HIP's own calls -- the stream torch
// reports as 0.  The plan's own stream serves only the host-buffer calls,
// which synchronize it before they return.
hipStream_t pick_stream(void *stream, hipStream_t /*own*/)
{
    return static_cast<hipStream_t>(stream);
}

template <typename T>
int api_set_nodes_dev(typename Traits<T>::plan *p, const double *x_dev, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!x_dev) throw Error(CUNFFT_ERR_INVALID_ARG, "x is NULL");
        CUNFFT_HIP(hipMemcpyAsync(p->x_gpu, x_dev, sizeof(double) * p->M_total * p->d,
                                  hipMemcpyDeviceToDevice, pick_stream(stream, im->stream)));
        im->nodes_changed();
    });
}

template <typename T>
int api_transform_dev(typename Traits<T>::plan *p, const void *fhat, void *f, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!fhat || !f) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        double a, b;
        run_forward<T>(p, im, static_cast<const cplx_t<T> *>(fhat), static_cast<cplx_t<T> *>(f),
                       pick_stream(stream, im->stream), false, &a, &b);
    });
}

template <typename T>
int api_adjoint_dev(typename Traits<T>::plan *p, const void *f, void *fhat, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!fhat || !f) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        double a, b, c;
        run_adjoint<T>(p, im, static_cast<const cplx_t<T> *>(f), static_cast<cplx_t<T> *>(fhat),
                       pick_stream(stream, im->stream), false, &a, &b, &c);
    });
}

// The adjoint of api_adjoint_dev with its node binning, sorted coefficients
// and partial-tile slots in the device-resident scratch (scratch.hip) instead
// of the plan's lazily allocated binning: no allocation.  The binning is
// rebuilt on every call (the scratch keeps whatever the last call on this
// device left); every buffer the kernels read is written first in the call.
template <typename T>
int api_adjoint_noalloc_dev(typename Traits<T>::plan *p, const void *f, void *fhat, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!fhat || !f) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        FftTables *ft = static_cast<FftTables *>(p->fft_backward);
        if (!ft)
            throw Error(CUNFFT_ERR_STATE, "the allocation-free adjoint needs the plan's adjoint FFT "
                                          "tables (CUFFT_INIT with CUFFT_BOTH or CUFFT_BACKW)");
        hipStream_t s = pick_stream(stream, im->stream);
        cplx_t<T> *g = reinterpret_cast<cplx_t<T> *>(p->g_gpu);
        const cplx_t<T> *fin = static_cast<const cplx_t<T> *>(f);
        Binning b;   // describes the scratch; owns nothing
        if (!(p->flags & CUNFFT_ATOMIC_SCATTER) &&
            plan_bins(b, im->geom, (int)sizeof(T), im->config[1], im->compensated)) {
            const size_t need = b.bytes_for(p->M_total);
            if (need > kNoAllocScratchBytes) {
                char msg[200];
                snprintf(msg, sizeof msg,
                         "the allocation-free adjoint needs %.1f MiB of scratch for this plan; it "
                         "holds %.0f MiB",
                         need / 1048576.0, kNoAllocScratchBytes / 1048576.0);
                throw Error(CUNFFT_ERR_UNSUPPORTED, msg);
            }
            if (!b.bind(noalloc_scratch(), kNoAllocScratchBytes, p->M_total))
                throw Error(CUNFFT_ERR_STATE, "scratch binding");
            b.build(im->geom, p->x_gpu, s);
        }
        // an unusable binning makes launch_scatter_binned clear the grid and add
        // with atomics: no scratch either
        launch_scatter_binned<T>(im->geom, b, p->x_gpu, fin, g, s, nullptr);
        fft_passes<T>(p, im, ft, false, nullptr, static_cast<cplx_t<T> *>(fhat), g, s);
    });
}

template <typename T>
int api_grad_x_dev(typename Traits<T>::plan *p, const void *fhat, const void *v, T *gx, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!fhat || !v || !gx) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        hipStream_t s = pick_stream(stream, im->stream);
        const cplx_t<T> *g = fft_passes<T>(p, im, fft_tables<T>(p, true), true,
                                           static_cast<const cplx_t<T> *>(fhat), nullptr,
                                           reinterpret_cast<cplx_t<T> *>(p->g_gpu), s);
        grad_x<T>(p, im, g, static_cast<const cplx_t<T> *>(v), gx, s);
    });
}

// Backward of f = A(x) fhat for an upstream gradient gf (complex gradient in
// the conjugate-Wirtinger convention of torch): gfhat = A^H gf, and
// gx[j*d+t] = Re(conj(gf_j) * d f_j / d x_jt).  Either output may be NULL.
// The x-gradient reads the forward grid of fhat, so it runs first; the
// adjoint then reuses the grid buffer.
template <typename T>
int api_backward_dev(typename Traits<T>::plan *p, const void *fhat, const void *gf, void *gfhat,
                     T *gx, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!gf) throw Error(CUNFFT_ERR_INVALID_ARG, "upstream gradient is NULL");
        if (gx && !fhat) throw Error(CUNFFT_ERR_INVALID_ARG, "the x-gradient needs fhat");
        hipStream_t s = pick_stream(stream, im->stream);
        if (gx) {
            const cplx_t<T> *g = fft_passes<T>(p, im, fft_tables<T>(p, true), true,
                                               static_cast<const cplx_t<T> *>(fhat), nullptr,
                                               reinterpret_cast<cplx_t<T> *>(p->g_gpu), s);
            grad_x<T>(p, im, g, static_cast<const cplx_t<T> *>(gf), gx, s);
        }
        if (gfhat) {
            double a, b, c;
            run_adjoint<T>(p, im, static_cast<const cplx_t<T> *>(gf),
                           static_cast<cplx_t<T> *>(gfhat), s, false, &a, &b, &c);
        }
    });
}

template <typename T>
int api_ndft_dev(typename Traits<T>::plan *p, bool adjoint, const void *in, const long long *idx,
                 long long n_idx, void *out, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!in || !out) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        if (idx && n_idx < 0) throw Error(CUNFFT_ERR_INVALID_ARG, "negative output count");
        if (p->N_total > 0x7fffffff) throw Error(CUNFFT_ERR_UNSUPPORTED, "NDFT needs N_total < 2^31");
        int64_t N[3];
        for (int t = 0; t < p->d; ++t) N[t] = p->N[t];
        const int64_t nout = idx ? n_idx : (adjoint ? p->N_total : p->M_total);
        hipStream_t s = pick_stream(stream, im->stream);
        if (adjoint)
            launch_ndft_cols<T>(p->d, N, p->M_total, p->x_gpu, static_cast<const cplx_t<T> *>(in),
                                idx, nout, static_cast<cplx_t<T> *>(out), s);
        else
            launch_ndft_rows<T>(p->d, N, p->M_total, p->x_gpu, static_cast<const cplx_t<T> *>(in),
                                idx, nout, static_cast<cplx_t<T> *>(out), s);
    });
}

template <typename T>
int api_autotune(typename Traits<T>::plan *p, int dirs, int reps, float *ms_out, void *stream)
{
    if (p && p->status == CUNFFT_SUCCESS && (dirs & ~3 || !dirs)) {
        report("directions must be 1, 2 or 3");
        return CUNFFT_ERR_INVALID_ARG;   // refused; the plan stays usable
    }
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        autotune<T>(p, im, dirs, reps, ms_out, pick_stream(stream, im->stream));
    });
}

template <typename T> int api_get_config(const typename Traits<T>::plan *p, int direction)
{
    if (!p || !p->impl || (direction != 1 && direction != 2)) return -1;
    return static_cast<const Impl<T> *>(p->impl)->config[direction - 1];
}

// A configuration that does not exist or does not fit is refused with its
// status and leaves the plan as it was (the status is not recorded on it).
template <typename T> int api_set_config(typename Traits<T>::plan *p, int dirs, int c)
{
    if (!p) return CUNFFT_ERR_INVALID_ARG;
    if (p->status != CUNFFT_SUCCESS) return p->status;
    if (!p->impl) return CUNFFT_ERR_STATE;
    Impl<T> *im = static_cast<Impl<T> *>(p->impl);
    if (dirs & ~3 || !dirs) {
        report("directions must be 1, 2 or 3");
        return CUNFFT_ERR_INVALID_ARG;
    }
    if (!config_exists(p->d, c)) {
        report("no such gridding configuration for this plan");
        return CUNFFT_ERR_INVALID_ARG;
    }
    Binning probe;
    if (!plan_bins(probe, im->geom, (int)sizeof(T), c, im->compensated)) {
        report(im->compensated && c == kTileVariants
                   ? "CUNFFT_COMPENSATED plans have no node-lane configuration"
                   : "the tiled kernels of this configuration do not fit the plan's window and grid");
        return CUNFFT_ERR_UNSUPPORTED;
    }
    return guard(p, [&] {
        CUNFFT_HIP(hipDeviceSynchronize());
        plan_configs<T>(im, (dirs & 1) ? c : im->config[0], (dirs & 2) ? c : im->config[1]);
    });
}

template <typename T>
int api_solve_dev(typename Traits<T>::plan *p, const void *y, const T *w, void *fhat,
                  int max_iter, double tol, double *res_hist, int *iters, void *stream)
{
    return guard(p, [&] {
        Impl<T> *im = impl_of<T>(p);
        need_device<T>(p);
        if (!y || !fhat) throw Error(CUNFFT_ERR_INVALID_ARG, "NULL device pointer");
        SolverOps<T> ops;
        ops.forward = [&](const cplx_t<T> *in, cplx_t<T> *out, hipStream_t s) {
            double a, b;
            run_forward<T>(p, im, in, out, s, false, &a, &b);
        };
        ops.adjoint = [&](const cplx_t<T> *in, cplx_t<T> *out, hipStream_t s) {
            double a, b, c;
            run_adjoint<T>(p, im, in, out, s, false, &a, &b, &c);
        };
        solve_cgnr<T>(ops, static_cast<const cplx_t<T> *>(y), w, static_cast<cplx_t<T> *>(fhat),
                      p->M_total, p->N_total, max_iter, tol, res_hist, iters,
                      pick_stream(stream, im->stream));
    });
}

template <typename T> int api_synchronize(typename Traits<T>::plan *p)
{
    return guard(p, [&] { CUNFFT_HIP(hipStreamSynchronize(impl_of<T>(p)->stream)); });
}

template <typename T> void api_init_plan(int dim, uint_t *n, cunfftHandle *h)
{
    if (!h) return;
    *h = nullptr;
    if (!n || dim < 1 || dim > 3) return;
    try {
        int64_t nn[3];
        for (int t = 0; t < dim; ++t) {
            if (!fft_length_ok(n[t])) return;
            nn[t] = n[t];
        }
        *h = make_fft_tables(sizeof(T) == 8 ? 64 : 32, dim, nn);
    } catch (const std::exception &e) {
        report(e.what());
        *h = nullptr;
    }
}

template <typename T> void api_free_plan(typename Traits<T>::plan *p)
{
    if (!p) return;
    free_fft_tables(p->fft_forward);
    free_fft_tables(p->fft_backward);
    p->fft_forward = p->fft_backward = nullptr;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ FP64 family --
void warmUp_createPlan(uint_t *n)
{
    (void)n;
    (void)hipFree(nullptr);
}
void copyDataToDevice(cunfft_plan *p) { api_copy_in<double>(p, false); }
void copyDataToDeviceAd(cunfft_plan *p) { api_copy_in<double>(p, true); }
void copyDataToHost(cunfft_plan *p) { api_copy_out<double>(p, false); }
void copyDataToHostAd(cunfft_plan *p) { api_copy_out<double>(p, true); }
void copy_f_hat_ToHost(cunfft_plan *p, cunfftDoubleComplex *arr)
{
    if (p) api_copy_dev_to<double>(p, arr, p->fhat_gpu, sizeof(*arr) * p->N_total);
}
void copy_f_ToHost(cunfft_plan *p, cunfftDoubleComplex *arr)
{
    if (p) api_copy_dev_to<double>(p, arr, p->f_gpu, sizeof(*arr) * p->M_total);
}
void copy_g_ToHost(cunfft_plan *p) { api_copy_g<double>(p); }
void copy_x_toHost(cunfft_plan *p, double *arr)
{
    if (p) api_copy_dev_to<double>(p, arr, p->x_gpu, sizeof(double) * p->M_total * p->d);
}
void cunfft_transform(cunfft_plan *p) { api_transform<double>(p); }
void cunfft_adjoint(cunfft_plan *p) { api_adjoint<double>(p); }
void cunfft_initPlan(int dim, uint_t *n, cunfftHandle *h) { api_init_plan<double>(dim, n, h); }
void cunfft_freePlan(cunfft_plan *p) { api_free_plan<double>(p); }
void cunfft_initDevicePara(cunfft_plan *p)
{
    guard(p, [&] {
        impl_of<double>(p);
        init_device_para<double>(p);
    });
}
void cunfft_init(cunfft_plan *p, int dim, uint_t *N, uint_t M)
{
    api_init_guru<double>(p, dim, N, M, nullptr, CUNFFT_DEFAULT_CUT_OFF, CUFFT_INIT | CUFFT_BOTH, true);
}
void cunfft_initGuru(cunfft_plan *p, int dim, uint_t *N, uint_t M, uint_t *n, unsigned flags)
{
    api_init_guru<double>(p, dim, N, M, n, CUNFFT_DEFAULT_CUT_OFF, flags, false);
}
void cunfft_initGuru_m(cunfft_plan *p, int dim, uint_t *N, uint_t M, uint_t *n, int m,
                       unsigned flags)
{
    api_init_guru<double>(p, dim, N, M, n, m, flags, n == nullptr);
}
void cunfft_initHelp(cunfft_plan *p)
{
    if (!p) return;
    p->status = CUNFFT_SUCCESS;
    p->impl = nullptr;
    guard(p, [&] { init_help<double>(p); });
}
void cunfft_reinit(cunfft_plan *p) { api_reinit<double>(p, false); }
void cunfft_reinitAd(cunfft_plan *p) { api_reinit<double>(p, true); }
void cunfft_finalize(cunfft_plan *p)
{
    if (p) finalize<double>(p);
}
void cunfft_finalizeDevice(cunfft_plan *p)
{
    if (!p) return;
    Impl<double> *im = static_cast<Impl<double> *>(p->impl);
    if (im) (void)hipStreamSynchronize(im->stream);
    free_device<double>(p);
}
void resetCUNFFTTimes(NFFTTimeSpec *t)
{
    if (t) memset(t, 0, sizeof(*t));
}
void cundft_transform(cunfft_plan *p) { api_ndft<double>(p, false); }
void cundft_adjoint(cunfft_plan *p) { api_ndft<double>(p, true); }
void nfft_init_fg_exp_l(double *out, const double b, const int offset)
{
    if (out) fg_exp_l(out, b, offset, CUNFFT_DEFAULT_CUT_OFF);
}
void nfft_init_fg_exp_l_m(double *out, const double b, const int offset, int m)
{
    if (out && m >= 1 && m <= CUNFFT_MAX_CUT_OFF) fg_exp_l(out, b, offset, m);
}
// src/cunfft.cu:79-90
void setProperties(uint_t *N, uint_t *n, double *b, double *sigma, int dim, kdata *k)
{
    if (!N || !n || !b || !sigma || !k || dim < 1 || dim > 3) return;
    for (int t = 0; t < dim; ++t) {
        k->N[t] = N[t];
        k->n[t] = n[t];
        k->b[t] = b[t];
        k->bsqrt[t] = sqrt(b[t]);
        k->fac[t] = (M_PI * M_PI * b[t]) / ((double)n[t] * (double)n[t]);
        k->sigma[t] = sigma[t];
    }
}
void resetDevice(void) { (void)hipDeviceSynchronize(); }
int cunfft_last_error(const cunfft_plan *p) { return p ? p->status : CUNFFT_ERR_INVALID_ARG; }
const char *cunfft_error_string(int status)
{
    switch (status) {
    case CUNFFT_SUCCESS: return "success";
    case CUNFFT_ERR_INVALID_ARG: return "invalid argument";
    case CUNFFT_ERR_UNSUPPORTED: return "unsupported parameters";
    case CUNFFT_ERR_ALLOC: return "allocation failed";
    case CUNFFT_ERR_HIP: return "HIP runtime error";
    case CUNFFT_ERR_STATE: return "invalid plan state";
    default: return "unknown status";
    }
}
const char *cunfft_last_message(void) { return g_last_message.c_str(); }
int cunfft_set_nodes_dev(cunfft_plan *p, const double *x, void *s)
{
    return api_set_nodes_dev<double>(p, x, s);
}
int cunfft_transform_dev(cunfft_plan *p, const cunfftDoubleComplex *fhat, cunfftDoubleComplex *f,
                         void *s)
{
    return api_transform_dev<double>(p, fhat, f, s);
}
int cunfft_adjoint_dev(cunfft_plan *p, const cunfftDoubleComplex *f, cunfftDoubleComplex *fhat,
                       void *s)
{
    return api_adjoint_dev<double>(p, f, fhat, s);
}
int cunfft_adjoint_noalloc_dev(cunfft_plan *p, const cunfftDoubleComplex *f,
                               cunfftDoubleComplex *fhat, void *s)
{
    return api_adjoint_noalloc_dev<double>(p, f, fhat, s);
}
int cunfft_grad_x_dev(cunfft_plan *p, const cunfftDoubleComplex *fhat, const cunfftDoubleComplex *v,
                      double *gx, void *s)
{
    return api_grad_x_dev<double>(p, fhat, v, gx, s);
}

int cunfft_backward_dev(cunfft_plan *p, const cunfftDoubleComplex *fhat,
                        const cunfftDoubleComplex *grad_f, cunfftDoubleComplex *grad_fhat,
                        double *grad_x, void *s)
{
    return api_backward_dev<double>(p, fhat, grad_f, grad_fhat, grad_x, s);
}
int cunfft_autotune(cunfft_plan *p, int directions, int reps, float *ms, void *s)
{
    return api_autotune<double>(p, directions, reps, ms, s);
}
int cunfft_synchronize(cunfft_plan *p) { return api_synchronize<double>(p); }
int cunfft_solve_dev(cunfft_plan *p, const cunfftDoubleComplex *y, const double *w,
                     cunfftDoubleComplex *fhat, int max_iter, double tol, double *res_hist,
                     int *iters, void *s)
{
    return api_solve_dev<double>(p, y, w, fhat, max_iter, tol, res_hist, iters, s);
}
int cunfft_get_config(const cunfft_plan *p, int direction)
{
    return api_get_config<double>(p, direction);
}
int cunfft_set_config(cunfft_plan *p, int directions, int c)
{
    return api_set_config<double>(p, directions, c);
}
int cundft_transform_rows_dev(cunfft_plan *p, const cunfftDoubleComplex *fhat, const long long *rows,
                              long long n_rows, cunfftDoubleComplex *out, void *s)
{
    return api_ndft_dev<double>(p, false, fhat, rows, n_rows, out, s);
}
int cundft_adjoint_cols_dev(cunfft_plan *p, const cunfftDoubleComplex *f, const long long *cols,
                            long long n_cols, cunfftDoubleComplex *out, void *s)
{
    return api_ndft_dev<double>(p, true, f, cols, n_cols, out, s);
}

// ------------------------------------------------------------ FP32 family --
void cunfftf_copyDataToDevice(cunfftf_plan *p) { api_copy_in<float>(p, false); }
void cunfftf_copyDataToDeviceAd(cunfftf_plan *p) { api_copy_in<float>(p, true); }
void cunfftf_copyDataToHost(cunfftf_plan *p) { api_copy_out<float>(p, false); }
void cunfftf_copyDataToHostAd(cunfftf_plan *p) { api_copy_out<float>(p, true); }
void cunfftf_copy_f_hat_ToHost(cunfftf_plan *p, cunfftFloatComplex *arr)
{
    if (p) api_copy_dev_to<float>(p, arr, p->fhat_gpu, sizeof(*arr) * p->N_total);
}
void cunfftf_copy_f_ToHost(cunfftf_plan *p, cunfftFloatComplex *arr)
{
    if (p) api_copy_dev_to<float>(p, arr, p->f_gpu, sizeof(*arr) * p->M_total);
}
void cunfftf_copy_g_ToHost(cunfftf_plan *p) { api_copy_g<float>(p); }
void cunfftf_copy_x_toHost(cunfftf_plan *p, double *arr)
{
    if (p) api_copy_dev_to<float>(p, arr, p->x_gpu, sizeof(double) * p->M_total * p->d);
}
void cunfftf_transform(cunfftf_plan *p) { api_transform<float>(p); }
void cunfftf_adjoint(cunfftf_plan *p) { api_adjoint<float>(p); }
void cunfftf_initPlan(int dim, uint_t *n, cunfftHandle *h) { api_init_plan<float>(dim, n, h); }
void cunfftf_freePlan(cunfftf_plan *p) { api_free_plan<float>(p); }
void cunfftf_initDevicePara(cunfftf_plan *p)
{
    guard(p, [&] {
        impl_of<float>(p);
        init_device_para<float>(p);
    });
}
void cunfftf_init(cunfftf_plan *p, int dim, uint_t *N, uint_t M)
{
    api_init_guru<float>(p, dim, N, M, nullptr, CUNFFT_DEFAULT_CUT_OFF, CUFFT_INIT | CUFFT_BOTH, true);
}
void cunfftf_initGuru(cunfftf_plan *p, int dim, uint_t *N, uint_t M, uint_t *n, unsigned flags)
{
    api_init_guru<float>(p, dim, N, M, n, CUNFFT_DEFAULT_CUT_OFF, flags, false);
}
void cunfftf_initGuru_m(cunfftf_plan *p, int dim, uint_t *N, uint_t M, uint_t *n, int m,
                        unsigned flags)
{
    api_init_guru<float>(p, dim, N, M, n, m, flags, n == nullptr);
}
void cunfftf_initHelp(cunfftf_plan *p)
{
    if (!p) return;
    p->status = CUNFFT_SUCCESS;
    p->impl = nullptr;
    guard(p, [&] { init_help<float>(p); });
}
void cunfftf_reinit(cunfftf_plan *p) { api_reinit<float>(p, false); }
void cunfftf_reinitAd(cunfftf_plan *p) { api_reinit<float>(p, true); }
void cunfftf_finalize(cunfftf_plan *p)
{
    if (p) finalize<float>(p);
}
void cunfftf_finalizeDevice(cunfftf_plan *p)
{
    if (!p) return;
    Impl<float> *im = static_cast<Impl<float> *>(p->impl);
    if (im) (void)hipStreamSynchronize(im->stream);
    free_device<float>(p);
}
void cunfftf_ndft_transform(cunfftf_plan *p) { api_ndft<float>(p, false); }
void cunfftf_ndft_adjoint(cunfftf_plan *p) { api_ndft<float>(p, true); }
int cunfftf_last_error(const cunfftf_plan *p) { return p ? p->status : CUNFFT_ERR_INVALID_ARG; }
int cunfftf_set_nodes_dev(cunfftf_plan *p, const double *x, void *s)
{
    return api_set_nodes_dev<float>(p, x, s);
}
int cunfftf_transform_dev(cunfftf_plan *p, const cunfftFloatComplex *fhat, cunfftFloatComplex *f,
                          void *s)
{
    return api_transform_dev<float>(p, fhat, f, s);
}
int cunfftf_adjoint_dev(cunfftf_plan *p, const cunfftFloatComplex *f, cunfftFloatComplex *fhat,
                        void *s)
{
    return api_adjoint_dev<float>(p, f, fhat, s);
}
int cunfftf_grad_x_dev(cunfftf_plan *p, const cunfftFloatComplex *fhat, const cunfftFloatComplex *v,
                       float *gx, void *s)
{
    return api_grad_x_dev<float>(p, fhat, v, gx, s);
}

int cunfftf_backward_dev(cunfftf_plan *p, const cunfftFloatComplex *fhat,
                         const cunfftFloatComplex *grad_f, cunfftFloatComplex *grad_fhat,
                         float *grad_x, void *s)
{
    return api_backward_dev<float>(p, fhat, grad_f, grad_fhat, grad_x, s);
}
int cunfftf_autotune(cunfftf_plan *p, int directions, int reps, float *ms, void *s)
{
    return api_autotune<float>(p, directions, reps, ms, s);
}
int cunfftf_synchronize(cunfftf_plan *p) { return api_synchronize<float>(p); }
int cunfftf_solve_dev(cunfftf_plan *p, const cunfftFloatComplex *y, const float *w,
                      cunfftFloatComplex *fhat, int max_iter, double tol, double *res_hist,
                      int *iters, void *s)
{
    return api_solve_dev<float>(p, y, w, fhat, max_iter, tol, res_hist, iters, s);
}
int cunfftf_get_config(const cunfftf_plan *p, int direction)
{
    return api_get_config<float>(p, direction);
}
int cunfftf_set_config(cunfftf_plan *p, int directions, int c)
{
    return api_set_config<float>(p, directions, c);
}
int cunfftf_ndft_transform_rows_dev(cunfftf_plan *p, const cunfftFloatComplex *fhat,
                                    const long long *rows, long long n_rows, cunfftFloatComplex *out,
                                    void *s)
{
    return api_ndft_dev<float>(p, false, fhat, rows, n_rows, out, s);
}
int cunfftf_ndft_adjoint_cols_dev(cunfftf_plan *p, const cunfftFloatComplex *f, const long long *cols,
                                  long long n_cols, cunfftFloatComplex *out, void *s)
{
    return api_ndft_dev<float>(p, true, f, cols, n_cols, out, s);
}

}  // extern "C"
