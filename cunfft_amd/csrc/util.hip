// util.hip -- the reference's harness functions (include/cunfft_util.h):
// example data, printing, the rel-l_inf norm and event timers
// (src/cuda_util.cpp:130-258, src/cunfft_util.h:69-308,
// src/timeMeasurement.h:107-130).  Host code only.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "cunfft.h"
#include "cunfft_util.h"

namespace {

template <typename C> void show_coeff(const C *x, int n, const char *text)
{
    if (text) printf("\n %s\n", text);
    if (!x) return;
    for (int k = 0; k < n; ++k) {
        if (k % 4 == 0) printf("%6d.\t", k);
        printf("%+.2E %+.2Ei, ", (double)x[k].x, (double)x[k].y);
        if (k % 4 == 3) printf("\n");
    }
    if (n % 4 != 0) printf("\n");
    printf("\n");
}

// src/cuda_util.cpp:130-139: re = sin(10 k / N) exp(-(k / N^0.8)^2), im = 0
// over the centred k = -N/2 .. N/2 - 1.
template <typename C> void coeff_clustered(C *in, uint_t N)
{
    if (!in) return;
    for (uint_t k = 0; k < N; ++k) {
        const double kk = (double)k - (double)(N / 2);
        in[k].x = (decltype(in[k].x))(sin(10.0 * kk / (double)N) *
                                      exp(-pow(kk / pow((double)N, 0.8), 2.0)));
        in[k].y = 0;
    }
}

template <typename C> void coeff_unidistr(C *in, uint_t N)
{
    if (!in) return;
    for (uint_t k = 0; k < N; ++k) {
        in[k].x = (decltype(in[k].x))drand48();
        in[k].y = (decltype(in[k].y))drand48();
    }
}

template <typename C> void coeff_null(C *in, uint_t N)
{
    if (!in) return;
    for (uint_t k = 0; k < N; ++k) in[k].x = in[k].y = 0;
}

template <typename P> void nodes_unidistr(P *p)
{
    if (!p->x) return;
    for (uint_t i = 0; i < p->M_total * p->d; ++i) p->x[i] = drand48() - 0.5;
}

// src/cuda_util.cpp:141-149: x_i = j^(1/100) over all M d coordinates, j
// stepping by 1 / (M d) from 0 (accumulated, as the reference does, so the
// nodes agree to the last bit).
template <typename P> void nodes_clustered(P *p)
{
    if (!p->x) return;
    const uint_t n = p->M_total * p->d;
    double j = 0.0;
    for (uint_t i = 0; i < n; ++i, j += 1.0 / (double)n) p->x[i] = pow(j, 0.01);
}

template <typename C> double linf(const C *x, const C *y, int n)
{
    double m = 0.0;
    for (int k = 0; k < n; ++k) {
        const double dr = (double)x[k].x - (y ? (double)y[k].x : 0.0);
        const double di = (double)x[k].y - (y ? (double)y[k].y : 0.0);
        const double a = hypot(dr, di);
        m = a > m ? a : m;
    }
    return m;
}

template <typename C> double rel_linf(const C *x, const C *y, int n)
{
    return linf(x, y, n) / linf(x, (const C *)nullptr, n);
}

}  // namespace

extern "C" {

GPUTimer getTimeGPU(void)
{
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    (void)hipEventRecord(e, nullptr);
    return (GPUTimer)e;
}

double elapsedGPUTime(GPUTimer start, GPUTimer stop)
{
    float ms = 0.0f;
    hipEvent_t a = (hipEvent_t)start, b = (hipEvent_t)stop;
    if (a && b) {
        (void)hipEventSynchronize(b);
        (void)hipEventElapsedTime(&ms, a, b);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    return (double)ms * 1e-3;
}

void showTimes(NFFTTimeSpec *t, int tfac)
{
    if (!t) return;
    const double f = 1000.0 / (tfac > 0 ? tfac : 1);
    printf("\tRuntime:\t %.2E ms\n", t->runTime * f);
    printf("\t  roll-off:\t %.2E ms\n", t->time_ROC * f);
    printf("\t  FFT:\t\t %.2E ms\n", t->time_FFT * f);
    printf("\t  convolution:\t %.2E ms\n", t->time_CONV * f);
    printf("\t  copy in:\t %.2E ms\n", t->time_COPY_IN * f);
    printf("\t  copy out:\t %.2E ms\n", t->time_COPY_OUT * f);
    printf("\t  else:\t\t %.2E ms\n", t->time_else * f);
}

uint_t prod_int(uint_t *vec, int d)
{
    uint_t p = 1;
    for (int t = 0; t < d; ++t) p *= vec[t];
    return p;
}

uint_t next_power_of_2(uint_t N)
{
    uint_t p = 1;
    while (p < N) p <<= 1;
    return p;
}

void showCoeff_double(const double *x, int n, const char *text)
{
    if (text) printf("\n %s\n", text);
    if (!x) return;
    for (int k = 0; k < n; ++k) {
        if (k % 8 == 0) printf("%6d.\t", k);
        printf("%+.2E, ", x[k]);
        if (k % 8 == 7) printf("\n");
    }
    printf("\n");
}

// ---------------------------------------------------------------- FP64 --
void showCoeff_cuComplex(const cunfftDoubleComplex *x, int n, const char *text) { show_coeff(x, n, text); }
void createFourierCoeff_clustered(cunfftDoubleComplex *in, uint_t N) { coeff_clustered(in, N); }
void createFourierCoeff_uniDistr(cunfftDoubleComplex *in, uint_t N) { coeff_unidistr(in, N); }
void createFourierCoeff_null(cunfftDoubleComplex *in, uint_t N) { coeff_null(in, N); }
void getExampleData_clustered(cunfft_plan *p)
{
    if (!p) return;
    coeff_clustered(p->f_hat, p->N_total);
    nodes_clustered(p);
}
void getExampleData_uniDistr(cunfft_plan *p)
{
    if (!p) return;
    coeff_unidistr(p->f_hat, p->N_total);
    nodes_unidistr(p);
}
void getExampleDataAd_uniDistr(cunfft_plan *p)
{
    if (!p) return;
    coeff_unidistr(p->f, p->M_total);
    nodes_unidistr(p);
}
void getExampleData_null(cunfft_plan *p)
{
    if (!p) return;
    coeff_null(p->f_hat, p->N_total);
    nodes_unidistr(p);
}
void setExampleData(cunfft_plan *p, cunfftDoubleComplex *f, double *x)
{
    if (!p || !p->x || !p->f) return;
    for (uint_t i = 0; i < p->M_total * p->d; ++i) p->x[i] = x[i];
    for (uint_t i = 0; i < p->M_total; ++i) p->f[i] = f[i];
}
double l_infty(cunfftDoubleComplex *x, cunfftDoubleComplex *y, int n) { return linf(x, y, n); }
double compute_error_l_infty(cunfftDoubleComplex *x, cunfftDoubleComplex *y, int n)
{
    return rel_linf(x, y, n);
}

// ---------------------------------------------------------------- FP32 --
void cunfftf_showCoeff_cuComplex(const cunfftFloatComplex *x, int n, const char *text)
{
    show_coeff(x, n, text);
}
void cunfftf_createFourierCoeff_clustered(cunfftFloatComplex *in, uint_t N) { coeff_clustered(in, N); }
void cunfftf_createFourierCoeff_uniDistr(cunfftFloatComplex *in, uint_t N) { coeff_unidistr(in, N); }
void cunfftf_createFourierCoeff_null(cunfftFloatComplex *in, uint_t N) { coeff_null(in, N); }
void cunfftf_getExampleData_clustered(cunfftf_plan *p)
{
    if (!p) return;
    coeff_clustered(p->f_hat, p->N_total);
    nodes_clustered(p);
}
void cunfftf_getExampleData_uniDistr(cunfftf_plan *p)
{
    if (!p) return;
    coeff_unidistr(p->f_hat, p->N_total);
    nodes_unidistr(p);
}
void cunfftf_getExampleDataAd_uniDistr(cunfftf_plan *p)
{
    if (!p) return;
    coeff_unidistr(p->f, p->M_total);
    nodes_unidistr(p);
}
void cunfftf_getExampleData_null(cunfftf_plan *p)
{
    if (!p) return;
    coeff_null(p->f_hat, p->N_total);
    nodes_unidistr(p);
}
void cunfftf_setExampleData(cunfftf_plan *p, cunfftFloatComplex *f, double *x)
{
    if (!p || !p->x || !p->f) return;
    for (uint_t i = 0; i < p->M_total * p->d; ++i) p->x[i] = x[i];
    for (uint_t i = 0; i < p->M_total; ++i) p->f[i] = f[i];
}
double cunfftf_l_infty(cunfftFloatComplex *x, cunfftFloatComplex *y, int n) { return linf(x, y, n); }
double cunfftf_compute_error_l_infty(cunfftFloatComplex *x, cunfftFloatComplex *y, int n)
{
    return rel_linf(x, y, n);
}

}  // extern "C"
