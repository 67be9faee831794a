/*
 * cunfft_util.h -- the harness functions of CUNFFT 1.0 that its callers link
 * against (src/cunfft_util.h:69-308, src/timeMeasurement.h:16-130): example
 * data from the C library's drand48, coefficient and time printing, the
 * rel-l_inf error norm, and event timers.  The reference's tutorial
 * (src/simpleTest.cpp:18-148) compiles and runs against cunfft.h plus this
 * header and libcunfft_amd.so.
 *
 * Same names and argument order; gpuComplex / dTyp / cunfft_plan follow the
 * caller's CUNFFT_SINGLE_PRECISION switch as in cunfft.h (FP32 names carry the
 * cunfftf_ prefix in the library).  Node coordinates are double in both
 * precisions (cunfft.h).  GPUTimer is an opaque event handle instead of a
 * cudaEvent_t.
 */
#ifndef CUNFFT_AMD_UTIL_H_
#define CUNFFT_AMD_UTIL_H_

#include <time.h>

#include "cunfft.h"

#ifdef __cplusplus
extern "C" {
#endif

/* src/timeMeasurement.h:16-19 */
typedef void *GPUTimer;
typedef struct timespec CPUTimer;

/* src/timeMeasurement.h:107-115: an event recorded on the default stream. */
GPUTimer getTimeGPU(void);
/* src/timeMeasurement.h:124-130: seconds between two events (waits for stop,
 * releases both). */
double elapsedGPUTime(GPUTimer start, GPUTimer stop);

/* src/timeMeasurement.h:84-90, 118-122 */
static inline CPUTimer getTimeCPU(void)
{
    CPUTimer t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t;
}
static inline double elapsedCPUTime(CPUTimer start, CPUTimer stop)
{
    return (double)(stop.tv_sec - start.tv_sec) + (double)(stop.tv_nsec - start.tv_nsec) * 1e-9;
}

/* src/cunfft_util.h:94: prints the members of times divided by tfac. */
void showTimes(NFFTTimeSpec *times, int tfac);
/* src/cunfft_util.h:221, 229 */
uint_t prod_int(uint_t *vec, int d);
uint_t next_power_of_2(uint_t N);

/* ----- FP64 (the reference's default build) ----- */
/* src/cunfft_util.h:187, 197: the first n entries, with a caption. */
void showCoeff_double(const double *x, int n, const char *text);
void showCoeff_cuComplex(const cunfftDoubleComplex *x, int n, const char *text);
/* src/cunfft_util.h:240-286, src/cuda_util.cpp:130-212: example inputs.
 * uniDistr: re, im = drand48() per coefficient, then x = drand48() - 0.5
 * per coordinate (the unseeded C library stream, as the reference). */
void createFourierCoeff_clustered(cunfftDoubleComplex *in, uint_t N);
void createFourierCoeff_uniDistr(cunfftDoubleComplex *in, uint_t N);
void createFourierCoeff_null(cunfftDoubleComplex *in, uint_t N);
void getExampleData_clustered(cunfft_plan *plan);
void getExampleData_uniDistr(cunfft_plan *plan);
void getExampleDataAd_uniDistr(cunfft_plan *plan);
void getExampleData_null(cunfft_plan *plan);
void setExampleData(cunfft_plan *plan, cunfftDoubleComplex *f, double *x);
/* src/cunfft_util.h:298, 308, src/cuda_util.cpp:214-258:
 * l_infty = max_i |x_i - y_i|; compute_error_l_infty = that / max_i |x_i|. */
double l_infty(cunfftDoubleComplex *x, cunfftDoubleComplex *y, int n);
double compute_error_l_infty(cunfftDoubleComplex *x, cunfftDoubleComplex *y, int n);

/* ----- FP32 ----- */
void cunfftf_showCoeff_cuComplex(const cunfftFloatComplex *x, int n, const char *text);
void cunfftf_createFourierCoeff_clustered(cunfftFloatComplex *in, uint_t N);
void cunfftf_createFourierCoeff_uniDistr(cunfftFloatComplex *in, uint_t N);
void cunfftf_createFourierCoeff_null(cunfftFloatComplex *in, uint_t N);
void cunfftf_getExampleData_clustered(cunfftf_plan *plan);
void cunfftf_getExampleData_uniDistr(cunfftf_plan *plan);
void cunfftf_getExampleDataAd_uniDistr(cunfftf_plan *plan);
void cunfftf_getExampleData_null(cunfftf_plan *plan);
void cunfftf_setExampleData(cunfftf_plan *plan, cunfftFloatComplex *f, double *x);
double cunfftf_l_infty(cunfftFloatComplex *x, cunfftFloatComplex *y, int n);
double cunfftf_compute_error_l_infty(cunfftFloatComplex *x, cunfftFloatComplex *y, int n);

#ifdef CUNFFT_SINGLE_PRECISION
#define showCoeff_cuComplex cunfftf_showCoeff_cuComplex
#define createFourierCoeff_clustered cunfftf_createFourierCoeff_clustered
#define createFourierCoeff_uniDistr cunfftf_createFourierCoeff_uniDistr
#define createFourierCoeff_null cunfftf_createFourierCoeff_null
#define getExampleData_clustered cunfftf_getExampleData_clustered
#define getExampleData_uniDistr cunfftf_getExampleData_uniDistr
#define getExampleDataAd_uniDistr cunfftf_getExampleDataAd_uniDistr
#define getExampleData_null cunfftf_getExampleData_null
#define setExampleData cunfftf_setExampleData
#define l_infty cunfftf_l_infty
#define compute_error_l_infty cunfftf_compute_error_l_infty
#endif

#ifdef __cplusplus
}
#endif
#endif /* CUNFFT_AMD_UTIL_H_ */
