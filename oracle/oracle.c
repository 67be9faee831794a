/*
 * oracle.c -- FP64 CPU restatement of the CUNFFT algorithm (test checker).
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker, never by the product path.
 * Each function names the reference file:line it restates; the header lists
 * the conventions.  Written from the mathematics of those lines, not copied.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

#define MAXD 3

/* ---------------------------------------------------------------- inputs -- */

/* glibc drand48: X <- (0x5DEECE66D X + 0xB) mod 2^48, value = X / 2^48. */
double orc_drand48(uint64_t *state)
{
    const uint64_t a = 0x5DEECE66DULL, c = 0xBULL, mask = (1ULL << 48) - 1ULL;
    *state = (a * (*state) + c) & mask;
    return ldexp((double)(*state), -48);
}

/* src/cuda_util.cpp:151-160 (getExampleData_uniDistr) with the coefficient
 * generator of src/cuda_util.cpp:204-212. */
void orc_example_data_unidistr(uint64_t *state, int64_t N_total, int64_t M, int d,
                               double *fhat, double *x)
{
    for (int64_t k = 0; k < N_total; ++k) {
        fhat[2 * k] = orc_drand48(state);
        fhat[2 * k + 1] = orc_drand48(state);
    }
    for (int64_t i = 0; i < M * (int64_t)d; ++i)
        x[i] = orc_drand48(state) - 0.5;
}

/* src/cuda_util.cpp:162-171 (getExampleDataAd_uniDistr). */
void orc_example_data_ad_unidistr(uint64_t *state, int64_t M, int d, double *f,
                                  double *x)
{
    for (int64_t j = 0; j < M; ++j) {
        f[2 * j] = orc_drand48(state);
        f[2 * j + 1] = orc_drand48(state);
    }
    for (int64_t i = 0; i < M * (int64_t)d; ++i)
        x[i] = orc_drand48(state) - 0.5;
}

/* ------------------------------------------------------- window, indices -- */

/* src/cunfft.cu:342-350: sigma = n/N, b = 2 sigma/(2 sigma-1) * m/pi. */
double orc_shape_b(int64_t N, int64_t n, int m)
{
    double sigma = (double)n / (double)N;
    return (2.0 * sigma) / (2.0 * sigma - 1.0) * ((double)m / M_PI);
}

/* src/cunfft_kernel.cu:175-182: llrint rounds half to even in the default
 * rounding mode; signbit(-0.0) is 1. */
int64_t orc_window_start(int64_t n, double x, int m)
{
    return (int64_t)llrint((double)n * x) - (signbit(x) ? 1 : 0) - m;
}

/* src/cunfft_kernel.cu:162-166: exp(-t^2 b' pi) sqrt(b') with b' = 1/(pi b). */
double orc_phi(double t, double b)
{
    return exp(-(t * t) / b) / sqrt(M_PI * b);
}

/* ------------------------------------------------------ Kaiser-Bessel -- */
/* The reference is Gaussian-only; the Kaiser-Bessel window restated here is
 * NFFT3's (the library the reference links for its CPU side), in grid units
 * s = n x - l with shape b = pi (2 - 1/sigma):
 *   phi(s) = sinh(b sqrt(m^2 - s^2)) / (pi sqrt(m^2 - s^2))   |s| < m
 *          = b / pi                                          |s| = m
 *          = sin(b sqrt(s^2 - m^2)) / (pi sqrt(s^2 - m^2))    |s| > m
 * and its Fourier transform at xi = k/n, phihat = I0(m sqrt(b^2 - (2 pi k/n)^2)).
 * Both branches are one power series in d = m^2 - s^2,
 *   phi = (1/pi) sum_j b^{2j+1} d^j / (2j+1)!,
 * used near d = 0. */
static int g_window = 0;   /* 0 Gaussian, 1 Kaiser-Bessel, 2 sinc power */

void orc_set_window(int kind) { g_window = kind; }
int orc_get_window(void) { return g_window; }

double orc_kb_shape(int64_t N, int64_t n) { return M_PI * (2.0 - (double)N / (double)n); }

double orc_phi_kb(double s, int m, double b)
{
    const double d = (double)m * m - s * s;
    if (fabs(d) < 1e-8) return (b + b * b * b * d / 6.0) / M_PI;
    const double r = sqrt(fabs(d));
    return (d > 0 ? sinh(b * r) : sin(b * r)) / (M_PI * r);
}

double orc_dphi_kb(double s, int m, double b)
{
    const double d = (double)m * m - s * s;
    if (fabs(d) < 1e-4) {
        const double b2 = b * b;
        /* d/dd of the series, times dd/ds = -2 s */
        const double g = b * b2 / 6.0 + 2.0 * b * b2 * b2 * d / 120.0 +
                         3.0 * b * b2 * b2 * b2 * d * d / 5040.0 +
                         4.0 * b * b2 * b2 * b2 * b2 * d * d * d / 362880.0;
        return -2.0 * s * g / M_PI;
    }
    const double r = sqrt(fabs(d));
    const double br = b * r;
    if (d > 0) return -s * (br * cosh(br) - sinh(br)) / (M_PI * r * r * r);
    return s * (br * cos(br) - sin(br)) / (M_PI * r * r * r);
}

/* ---------------------------------------------------------- sinc power -- */
/* NFFT3's SINC_POWER window (the reference is Gaussian-only) in grid units
 * s = n x - l:  phi(s) = a sinc(pi a s)^(2m), sinc(u) = sin(u)/u, with
 * a = (2 sigma - 1) / (2 sigma m).  Its Fourier transform (frequency nu in
 * cycles per cell) is the centred cardinal B-spline M_2m(nu / a) of order 2m,
 * which vanishes for |nu| >= m a = 1 - 1/(2 sigma), the first alias of the
 * band edge of the oversampled grid. */
double orc_sinc_shape(int64_t N, int64_t n, int m)
{
    const double sigma = (double)n / (double)N;
    return (2.0 * sigma - 1.0) / (2.0 * sigma * (double)m);
}

static double sinc_of(double u) { return u == 0.0 ? 1.0 : sin(u) / u; }

double orc_phi_sinc(double s, int m, double a)
{
    return a * pow(sinc_of(M_PI * a * s), 2 * m);
}

/* d sinc / du = (u cos u - sin u) / u^2, by its Taylor series near 0. */
static double dsinc_of(double u)
{
    if (fabs(u) < 0.1) {
        const double u2 = u * u;
        return u * (-1.0 / 3.0 + u2 * (1.0 / 30.0 + u2 * (-1.0 / 840.0 + u2 / 45360.0)));
    }
    return (u * cos(u) - sin(u)) / (u * u);
}

double orc_dphi_sinc(double s, int m, double a)
{
    const double u = M_PI * a * s;
    return a * 2.0 * m * pow(sinc_of(u), 2 * m - 1) * dsinc_of(u) * M_PI * a;
}

/* Centred cardinal B-spline of order p: the p-fold convolution of the unit
 * box, support [-p/2, p/2], by its explicit sum
 *   M_p(x) = 1/(p-1)! sum_{j=0..p} (-1)^j C(p, j) (x + p/2 - j)_+^(p-1)
 * in long double, at -|x| (M_p is even).  The alternating sum cancels:
 * long double keeps ~1e-13 up to p = 16 (m = 8), less beyond. */
double orc_bspline(int p, double x)
{
    /* M_p is even: the left half has the fewest (and smallest) terms */
    const long double t = 0.5L * p - fabsl((long double)x);
    if (t <= 0.0L || t >= (long double)p) return 0.0;
    long double sum = 0.0L, binom = 1.0L, fact = 1.0L;
    for (int k = 2; k < p; ++k) fact *= k;
    for (int j = 0; j <= p; ++j) {
        const long double v = t - j;
        if (v > 0.0L) sum += ((j & 1) ? -binom : binom) * powl(v, p - 1);
        binom = binom * (p - j) / (j + 1);
    }
    return (double)(sum / fact);
}

/* Modified Bessel I0 by its power series in long double. */
double orc_bessel_i0(double z)
{
    long double q = (long double)z * z / 4.0L, term = 1.0L, sum = 1.0L;
    for (int j = 1; j < 500; ++j) {
        term *= q / ((long double)j * j);
        sum += term;
        if (term < 1e-22L * sum) break;
    }
    return (double)sum;
}

/* Roll-off factor 1 / phihat(k / n) of the current window for one axis. */
static double rolloff_factor(int64_t N, int64_t n, int m, int64_t k)
{
    if (g_window == 1) {
        const double b = orc_kb_shape(N, n), w = 2.0 * M_PI * (double)k / (double)n;
        return 1.0 / orc_bessel_i0((double)m * sqrt(b * b - w * w));
    }
    if (g_window == 2) {
        const double a = orc_sinc_shape(N, n, m);
        return 1.0 / orc_bspline(2 * m, (double)k / ((double)n * a));
    }
    const double b = orc_shape_b(N, n, m);
    return exp(M_PI * M_PI * b / ((double)n * (double)n) * (double)k * (double)k);
}

static int64_t pmod(int64_t a, int64_t n)
{
    int64_t r = a % n;
    return r < 0 ? r + n : r;
}

static int64_t prod(const int64_t *v, int d)
{
    int64_t p = 1;
    for (int t = 0; t < d; ++t) p *= v[t];
    return p;
}

/* ------------------------------------------------------------- roll-off -- */

/* Forward D: src/cunfft_wrapper.cu:32-75 (memset, then kernelRollOf_{1,2,3}d,
 * src/cunfft_kernel.cu:240-254, 286-312, 355-412): every centred k lands at
 * (k mod n) multiplied by exp(sum_t fac_t k_t^2); the rest of the grid is 0. */
void orc_rolloff_pad(int d, const int64_t *N, const int64_t *n, int m,
                     const double *fhat, double *ghat)
{
    int64_t Nt = prod(N, d), nt = prod(n, d);
    memset(ghat, 0, sizeof(double) * 2 * (size_t)nt);
    for (int64_t kl = 0; kl < Nt; ++kl) {
        int64_t rem = kl, gidx = 0, gstride = 1;
        double s = 1.0;
        int64_t k[MAXD];
        for (int t = d - 1; t >= 0; --t) {
            k[t] = rem % N[t] - N[t] / 2;
            rem /= N[t];
        }
        for (int t = d - 1; t >= 0; --t) {
            gidx += pmod(k[t], n[t]) * gstride;
            gstride *= n[t];
            s *= rolloff_factor(N[t], n[t], m, k[t]);
        }
        ghat[2 * gidx] = fhat[2 * kl] * s;
        ghat[2 * gidx + 1] = fhat[2 * kl + 1] * s;
    }
}

/* Adjoint D^T: src/cunfft_wrapper.cu:77-121 and kernelRollOf_adjoint_{1,2,3}d
 * (src/cunfft_kernel.cu:264-275, 318-343, 418-467). */
void orc_crop_rolloff(int d, const int64_t *N, const int64_t *n, int m,
                      const double *g, double *fhat)
{
    int64_t Nt = prod(N, d);
    for (int64_t kl = 0; kl < Nt; ++kl) {
        int64_t rem = kl, gidx = 0, gstride = 1;
        double s = 1.0;
        int64_t k[MAXD];
        for (int t = d - 1; t >= 0; --t) {
            k[t] = rem % N[t] - N[t] / 2;
            rem /= N[t];
        }
        for (int t = d - 1; t >= 0; --t) {
            gidx += pmod(k[t], n[t]) * gstride;
            gstride *= n[t];
            s *= rolloff_factor(N[t], n[t], m, k[t]);
        }
        fhat[2 * kl] = g[2 * gidx] * s;
        fhat[2 * kl + 1] = g[2 * gidx + 1] * s;
    }
}

/* ------------------------------------------------------------------ FFT -- */

/* Iterative radix-2 FFT of one contiguous line (length L, power of two),
 * unnormalised, kernel e^{sign 2 pi i j k / L}; twiddles from sin/cos of the
 * exact angle, one call per twiddle. */
static void fft_line(double *a, int64_t L, int sign, const double *tw)
{
    for (int64_t i = 1, j = 0; i < L; ++i) {
        int64_t bit = L >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double tr = a[2 * i], ti = a[2 * i + 1];
            a[2 * i] = a[2 * j]; a[2 * i + 1] = a[2 * j + 1];
            a[2 * j] = tr; a[2 * j + 1] = ti;
        }
    }
    for (int64_t len = 2; len <= L; len <<= 1) {
        int64_t half = len >> 1, step = L / len;
        for (int64_t s = 0; s < L; s += len) {
            for (int64_t q = 0; q < half; ++q) {
                double wr = tw[2 * (q * step)], wi = sign * tw[2 * (q * step) + 1];
                double *p0 = a + 2 * (s + q), *p1 = a + 2 * (s + q + half);
                double xr = p1[0] * wr - p1[1] * wi, xi = p1[0] * wi + p1[1] * wr;
                p1[0] = p0[0] - xr; p1[1] = p0[1] - xi;
                p0[0] += xr; p0[1] += xi;
            }
        }
    }
}

/* Direct DFT of one line of any length L (the oracle's path for lengths that
 * are not powers of two): out_k = sum_q a_q e^{sign 2 pi i q k / L}, the
 * exponent reduced mod L before the table lookup. */
static void dft_line(double *a, int64_t L, int sign, const double *tw, double *tmp)
{
    for (int64_t k = 0; k < L; ++k) {
        double sr = 0.0, si = 0.0;
        for (int64_t q = 0; q < L; ++q) {
            const int64_t e = (q * k) % L;
            const double wr = tw[2 * e], wi = sign * tw[2 * e + 1];
            sr += a[2 * q] * wr - a[2 * q + 1] * wi;
            si += a[2 * q] * wi + a[2 * q + 1] * wr;
        }
        tmp[2 * k] = sr;
        tmp[2 * k + 1] = si;
    }
    for (int64_t k = 0; k < 2 * L; ++k) a[k] = tmp[k];
}

/* cufftExecZ2Z(..., CUFFT_FORWARD / CUFFT_INVERSE), src/cunfft_wrapper.cu:288-329.
 * Power-of-two lengths use the radix-2 FFT, others the direct DFT. */
void orc_fft(int d, const int64_t *n, double *g, int sign)
{
    int64_t nt = prod(n, d);
    for (int t = 0; t < d; ++t) {
        int64_t L = n[t], stride = 1;
        for (int s = t + 1; s < d; ++s) stride *= n[s];
        int64_t lines = nt / L;
        double *tw = (double *)malloc(sizeof(double) * 2 * (size_t)L);
        for (int64_t q = 0; q < L; ++q) {
            double ang = 2.0 * M_PI * (double)q / (double)L;
            tw[2 * q] = cos(ang);
            tw[2 * q + 1] = sin(ang);
        }
        /* sign convention: forward uses e^{-...}; fft_line uses e^{sign ...} */
#pragma omp parallel
        {
            double *buf = (double *)malloc(sizeof(double) * 4 * (size_t)L);
            const int pow2 = (L & (L - 1)) == 0;
#pragma omp for schedule(static)
            for (int64_t li = 0; li < lines; ++li) {
                int64_t outer = li / stride, inner = li % stride;
                int64_t base = outer * L * stride + inner;
                for (int64_t q = 0; q < L; ++q) {
                    buf[2 * q] = g[2 * (base + q * stride)];
                    buf[2 * q + 1] = g[2 * (base + q * stride) + 1];
                }
                if (pow2)
                    fft_line(buf, L, sign, tw);
                else
                    dft_line(buf, L, sign, tw, buf + 2 * L);
                for (int64_t q = 0; q < L; ++q) {
                    g[2 * (base + q * stride)] = buf[2 * q];
                    g[2 * (base + q * stride) + 1] = buf[2 * q + 1];
                }
            }
            free(buf);
        }
        free(tw);
    }
}

/* --------------------------------------------------------------- gridding -- */

struct node_win {
    int64_t idx[MAXD][64];
    double w[MAXD][64];
    double dw[MAXD][64];
};

static void node_window(int d, const int64_t *N, const int64_t *n, int m,
                        const double *xj, struct node_win *nw, int want_dw)
{
    int W = 2 * m + 2;
    for (int t = 0; t < d; ++t) {
        double b = orc_shape_b(N[t], n[t], m);
        double bk = orc_kb_shape(N[t], n[t]);
        double as = orc_sinc_shape(N[t], n[t], m);
        int64_t u = orc_window_start(n[t], xj[t], m);
        double nx = (double)n[t] * xj[t];
        for (int l = 0; l < W; ++l) {
            double s = nx - (double)(u + l);
            nw->idx[t][l] = pmod(u + l, n[t]);
            if (g_window == 1) {
                nw->w[t][l] = orc_phi_kb(s, m, bk);
                if (want_dw) nw->dw[t][l] = (double)n[t] * orc_dphi_kb(s, m, bk);
            } else if (g_window == 2) {
                nw->w[t][l] = orc_phi_sinc(s, m, as);
                if (want_dw) nw->dw[t][l] = (double)n[t] * orc_dphi_sinc(s, m, as);
            } else {
                double p = orc_phi(s, b);
                nw->w[t][l] = p;
                /* d/dx phi(n x - l) = n * (-2 s / b) phi(s) */
                if (want_dw) nw->dw[t][l] = (double)n[t] * (-2.0 * s / b) * p;
            }
        }
    }
}

/* Forward B: src/cunfft_kernel.cu:514-1093 (kernelConvWithGauss_{1,2,3}d);
 * the FG_PSI recurrence of :491-505 is an algebraic rewrite of the same
 * window values, evaluated here directly. */
void orc_gather(int d, const int64_t *N, const int64_t *n, int m, const double *g,
                int64_t M, const double *x, double *f)
{
    int W = 2 * m + 2;
    int64_t s1 = d == 3 ? n[2] : (d == 2 ? n[1] : 1);
    int64_t s0 = d == 3 ? n[1] * n[2] : s1;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        struct node_win nw;
        node_window(d, N, n, m, x + d * j, &nw, 0);
        double sr = 0.0, si = 0.0;
        if (d == 1) {
            for (int a = 0; a < W; ++a) {
                const double *p = g + 2 * nw.idx[0][a];
                sr += p[0] * nw.w[0][a];
                si += p[1] * nw.w[0][a];
            }
        } else if (d == 2) {
            for (int a = 0; a < W; ++a) {
                double r1 = 0, i1 = 0;
                for (int c = 0; c < W; ++c) {
                    const double *p = g + 2 * (nw.idx[0][a] * s1 + nw.idx[1][c]);
                    r1 += p[0] * nw.w[1][c];
                    i1 += p[1] * nw.w[1][c];
                }
                sr += r1 * nw.w[0][a];
                si += i1 * nw.w[0][a];
            }
        } else {
            for (int a = 0; a < W; ++a) {
                double r1 = 0, i1 = 0;
                for (int c = 0; c < W; ++c) {
                    double r2 = 0, i2 = 0;
                    for (int e = 0; e < W; ++e) {
                        const double *p =
                            g + 2 * (nw.idx[0][a] * s0 + nw.idx[1][c] * s1 + nw.idx[2][e]);
                        r2 += p[0] * nw.w[2][e];
                        i2 += p[1] * nw.w[2][e];
                    }
                    r1 += r2 * nw.w[1][c];
                    i1 += i2 * nw.w[1][c];
                }
                sr += r1 * nw.w[0][a];
                si += i1 * nw.w[0][a];
            }
        }
        f[2 * j] = sr;
        f[2 * j + 1] = si;
    }
}

/* Adjoint B^T: src/cunfft_wrapper.cu:178-247 (memset g, then
 * kernelConvWithGauss_adjoint_{1,2,3}d, src/cunfft_kernel.cu:608-1197).
 * Serial over nodes so the summation order is fixed. */
void orc_scatter(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                 const double *x, const double *f, double *g)
{
    int W = 2 * m + 2;
    int64_t nt = prod(n, d);
    int64_t s1 = d == 3 ? n[2] : (d == 2 ? n[1] : 1);
    int64_t s0 = d == 3 ? n[1] * n[2] : s1;
    memset(g, 0, sizeof(double) * 2 * (size_t)nt);
    struct node_win nw;
    for (int64_t j = 0; j < M; ++j) {
        node_window(d, N, n, m, x + d * j, &nw, 0);
        double fr = f[2 * j], fi = f[2 * j + 1];
        if (d == 1) {
            for (int a = 0; a < W; ++a) {
                double *p = g + 2 * nw.idx[0][a];
                p[0] += fr * nw.w[0][a];
                p[1] += fi * nw.w[0][a];
            }
        } else if (d == 2) {
            for (int a = 0; a < W; ++a)
                for (int c = 0; c < W; ++c) {
                    double w = nw.w[0][a] * nw.w[1][c];
                    double *p = g + 2 * (nw.idx[0][a] * s1 + nw.idx[1][c]);
                    p[0] += fr * w;
                    p[1] += fi * w;
                }
        } else {
            for (int a = 0; a < W; ++a)
                for (int c = 0; c < W; ++c) {
                    double w01 = nw.w[0][a] * nw.w[1][c];
                    for (int e = 0; e < W; ++e) {
                        double w = w01 * nw.w[2][e];
                        double *p =
                            g + 2 * (nw.idx[0][a] * s0 + nw.idx[1][c] * s1 + nw.idx[2][e]);
                        p[0] += fr * w;
                        p[1] += fi * w;
                    }
                }
        }
    }
}

/* cunfft_transform: ROC -> FFT(-) -> gather, src/cunfft.cu:196-207. */
void orc_nfft_trafo(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                    const double *x, const double *fhat, double *f, double *g_work)
{
    orc_rolloff_pad(d, N, n, m, fhat, g_work);
    orc_fft(d, n, g_work, -1);
    orc_gather(d, N, n, m, g_work, M, x, f);
}

/* cunfft_adjoint: scatter -> FFT(+) -> crop/ROC, src/cunfft.cu:209-224. */
void orc_nfft_adjoint(int d, const int64_t *N, const int64_t *n, int m, int64_t M,
                      const double *x, const double *f, double *fhat, double *g_work)
{
    orc_scatter(d, N, n, m, M, x, f, g_work);
    orc_fft(d, n, g_work, +1);
    orc_crop_rolloff(d, N, n, m, g_work, fhat);
}

/* ----------------------------------------------------------------- NDFT -- */

static void set_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
    else omp_set_num_threads(omp_get_num_procs());
#else
    (void)nthreads;
#endif
}

static void k_of(int d, const int64_t *N, int64_t kl, int64_t *k)
{
    for (int t = d - 1; t >= 0; --t) {
        k[t] = kl % N[t] - N[t] / 2;
        kl /= N[t];
    }
}

/* kernelNDFT_{1,2,3}d / NDFT_forElem, src/cunfft_kernel.cu:1207-1285;
 * NFFT3 trafo_direct as wrapped by Tests/Scripts/ndft.cpp:11-58. */
void orc_ndft_trafo(int d, const int64_t *N, int64_t M, const double *x,
                    const double *fhat, double *f, const int64_t *rows,
                    int64_t n_rows, int nthreads)
{
    int64_t Nt = prod(N, d);
    int64_t nout = rows ? n_rows : M;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < nout; ++r) {
        int64_t j = rows ? rows[r] : r;
        double sr = 0.0, si = 0.0;
        int64_t k[MAXD];
        for (int64_t kl = 0; kl < Nt; ++kl) {
            k_of(d, N, kl, k);
            double ph = 0.0;
            for (int t = 0; t < d; ++t) ph += (double)k[t] * x[d * j + t];
            double sn, cs;
            sincos(-2.0 * M_PI * ph, &sn, &cs);
            sr += fhat[2 * kl] * cs - fhat[2 * kl + 1] * sn;
            si += fhat[2 * kl] * sn + fhat[2 * kl + 1] * cs;
        }
        f[2 * r] = sr;
        f[2 * r + 1] = si;
    }
}

/* kernelNDFT_adjoint_{1,2,3}d / NDFT_adjoint_forElem,
 * src/cunfft_kernel.cu:1289-1345; NFFT3 adjoint_direct as wrapped by
 * Tests/Scripts/ndft.cpp:60-113. */
void orc_ndft_adjoint(int d, const int64_t *N, int64_t M, const double *x,
                      const double *f, double *fhat, const int64_t *cols,
                      int64_t n_cols, int nthreads)
{
    int64_t Nt = prod(N, d);
    int64_t nout = cols ? n_cols : Nt;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < nout; ++r) {
        int64_t kl = cols ? cols[r] : r;
        int64_t k[MAXD];
        k_of(d, N, kl, k);
        double sr = 0.0, si = 0.0;
        for (int64_t j = 0; j < M; ++j) {
            double ph = 0.0;
            for (int t = 0; t < d; ++t) ph += (double)k[t] * x[d * j + t];
            double sn, cs;
            sincos(2.0 * M_PI * ph, &sn, &cs);
            sr += f[2 * j] * cs - f[2 * j + 1] * sn;
            si += f[2 * j] * sn + f[2 * j + 1] * cs;
        }
        fhat[2 * r] = sr;
        fhat[2 * r + 1] = si;
    }
}

/* ------------------------------------------------------------ gradients -- */

/* Derivative of the gather (src/cunfft_kernel.cu:514-1093) with respect to
 * x_jt, window index set held fixed; phi'(s) = -(2 s / b) phi(s). */
void orc_gather_dx(int d, const int64_t *N, const int64_t *n, int m, const double *g,
                   int64_t M, const double *x, double *dfdx)
{
    int W = 2 * m + 2;
    int64_t s1 = d == 3 ? n[2] : (d == 2 ? n[1] : 1);
    int64_t s0 = d == 3 ? n[1] * n[2] : s1;
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < M; ++j) {
        struct node_win nw;
        node_window(d, N, n, m, x + d * j, &nw, 1);
        for (int tt = 0; tt < d; ++tt) {
            double sr = 0.0, si = 0.0;
            if (d == 1) {
                for (int a = 0; a < W; ++a) {
                    const double *p = g + 2 * nw.idx[0][a];
                    sr += p[0] * nw.dw[0][a];
                    si += p[1] * nw.dw[0][a];
                }
            } else if (d == 2) {
                for (int a = 0; a < W; ++a)
                    for (int c = 0; c < W; ++c) {
                        double w = (tt == 0 ? nw.dw[0][a] : nw.w[0][a]) *
                                   (tt == 1 ? nw.dw[1][c] : nw.w[1][c]);
                        const double *p = g + 2 * (nw.idx[0][a] * s1 + nw.idx[1][c]);
                        sr += p[0] * w;
                        si += p[1] * w;
                    }
            } else {
                for (int a = 0; a < W; ++a)
                    for (int c = 0; c < W; ++c)
                        for (int e = 0; e < W; ++e) {
                            double w = (tt == 0 ? nw.dw[0][a] : nw.w[0][a]) *
                                       (tt == 1 ? nw.dw[1][c] : nw.w[1][c]) *
                                       (tt == 2 ? nw.dw[2][e] : nw.w[2][e]);
                            const double *p = g + 2 * (nw.idx[0][a] * s0 +
                                                        nw.idx[1][c] * s1 + nw.idx[2][e]);
                            sr += p[0] * w;
                            si += p[1] * w;
                        }
            }
            dfdx[2 * (j * d + tt)] = sr;
            dfdx[2 * (j * d + tt) + 1] = si;
        }
    }
}

/* d/dx_jt of the direct sum of src/cunfft_kernel.cu:1207-1285. */
void orc_ndft_dx(int d, const int64_t *N, int64_t M, const double *x,
                 const double *fhat, double *dfdx, const int64_t *rows,
                 int64_t n_rows, int nthreads)
{
    int64_t Nt = prod(N, d);
    int64_t nout = rows ? n_rows : M;
    set_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t r = 0; r < nout; ++r) {
        int64_t j = rows ? rows[r] : r;
        double sr[MAXD] = {0, 0, 0}, si[MAXD] = {0, 0, 0};
        int64_t k[MAXD];
        for (int64_t kl = 0; kl < Nt; ++kl) {
            k_of(d, N, kl, k);
            double ph = 0.0;
            for (int t = 0; t < d; ++t) ph += (double)k[t] * x[d * j + t];
            double sn, cs;
            sincos(-2.0 * M_PI * ph, &sn, &cs);
            /* term = fhat e^{-i theta}; derivative factor -2 pi i k_t */
            double tr = fhat[2 * kl] * cs - fhat[2 * kl + 1] * sn;
            double ti = fhat[2 * kl] * sn + fhat[2 * kl + 1] * cs;
            for (int t = 0; t < d; ++t) {
                double c = -2.0 * M_PI * (double)k[t];
                /* (tr + i ti) * (i c) = -c ti + i c tr */
                sr[t] += -c * ti;
                si[t] += c * tr;
            }
        }
        for (int t = 0; t < d; ++t) {
            dfdx[2 * (r * d + t)] = sr[t];
            dfdx[2 * (r * d + t) + 1] = si[t];
        }
    }
}

/* ---------------------------------------------------------------- norms -- */

/* src/cuda_util.cpp:237-258 (l_infty, compute_error_l_infty). */
double orc_rel_linf(const double *a, const double *b, int64_t n)
{
    double num = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double dr = a[2 * i] - b[2 * i], di = a[2 * i + 1] - b[2 * i + 1];
        double e = hypot(dr, di), v = hypot(a[2 * i], a[2 * i + 1]);
        if (e > num) num = e;
        if (v > den) den = v;
    }
    return num / den;
}

double orc_rel_l2(const double *a, const double *b, int64_t n)
{
    long double num = 0.0L, den = 0.0L;
    for (int64_t i = 0; i < n; ++i) {
        long double dr = a[2 * i] - b[2 * i], di = a[2 * i + 1] - b[2 * i + 1];
        num += dr * dr + di * di;
        den += (long double)a[2 * i] * a[2 * i] + (long double)a[2 * i + 1] * a[2 * i + 1];
    }
    return (double)sqrtl(num / den);
}
