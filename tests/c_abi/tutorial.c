/*
 * tutorial.c -- a C caller of libcunfft_amd.so in the reference's tutorial
 * sequence (src/simpleTest.cpp:18-148, cunfftDoc.pdf p.9): cunfft_init,
 * getExampleData_uniDistr, then CUNDFT, CUNFFT (after cunfft_reinit), the
 * CUNDFT adjoint and the CUNFFT adjoint (each after cunfft_reinitAd), for a
 * 1D and a 2D plan.  Built with gcc as plain C, once for the FP64 names and
 * once with -DCUNFFT_SINGLE_PRECISION.  Besides printing as the reference
 * does, it writes every input and result to <outdir>/<d>d_<name>.bin (raw
 * doubles) so a test can check them against the oracle.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cunfft.h"
#include "cunfft_util.h"

static const char *g_dir = ".";

static void dump(int d, const char *name, const void *data, size_t n_doubles_or_floats, size_t elem)
{
    char path[1024];
    snprintf(path, sizeof path, "%s/%dd_%s.bin", g_dir, d, name);
    FILE *f = fopen(path, "wb");
    if (!f) {
        perror(path);
        exit(2);
    }
    fwrite(data, elem, n_doubles_or_floats, f);
    fclose(f);
}

static void check(cunfft_plan *p, const char *what)
{
    if (cunfft_last_error(p) != CUNFFT_SUCCESS) {
        fprintf(stderr, "%s failed: %s (%s)\n", what, cunfft_error_string(cunfft_last_error(p)),
                cunfft_last_message());
        exit(1);
    }
}

static void run(int d, uint_t *N, uint_t M)
{
    cunfft_plan p;
    GPUTimer t;
    double rt;
    resetDevice();
    cunfft_init(&p, d, N, M);
    check(&p, "cunfft_init");
    getExampleData_uniDistr(&p);
    dump(d, "x", p.x, (size_t)(p.M_total * p.d), sizeof(double));
    dump(d, "fhat_in", p.f_hat, (size_t)(2 * p.N_total), sizeof(dTyp));

    printf("CUNDFT(%dD):\n", p.d);
    t = getTimeGPU();
    copyDataToDevice(&p);
    cundft_transform(&p);
    copyDataToHost(&p);
    rt = elapsedGPUTime(t, getTimeGPU());
    check(&p, "cundft_transform");
    showCoeff_cuComplex(p.f, 8, "vector f (first few entries)");
    p.CUNFFTTimes.runTime = rt;
    showTimes(&p.CUNFFTTimes, 1);
    dump(d, "f_ndft", p.f, (size_t)(2 * p.M_total), sizeof(dTyp));

    printf("CUNFFT(%dD):\n", p.d);
    cunfft_reinit(&p);
    copyDataToDevice(&p);
    cunfft_transform(&p);
    copyDataToHost(&p);
    check(&p, "cunfft_transform");
    showCoeff_cuComplex(p.f, 8, "vector f (first few entries)");
    dump(d, "f_nfft", p.f, (size_t)(2 * p.M_total), sizeof(dTyp));

    /* the adjoints take the forward result f as input, as in the tutorial */
    printf("CUNDFT(%dD) adjoint:\n", p.d);
    cunfft_reinitAd(&p);
    copyDataToDeviceAd(&p);
    cundft_adjoint(&p);
    copyDataToHostAd(&p);
    check(&p, "cundft_adjoint");
    showCoeff_cuComplex(p.f_hat, 8, "vector f_hat (first few entries)");
    dump(d, "fhat_ndft", p.f_hat, (size_t)(2 * p.N_total), sizeof(dTyp));

    printf("CUNFFT(%dD) adjoint:\n", p.d);
    cunfft_reinitAd(&p);
    copyDataToDeviceAd(&p);
    cunfft_adjoint(&p);
    copyDataToHostAd(&p);
    check(&p, "cunfft_adjoint");
    showCoeff_cuComplex(p.f_hat, 8, "vector f_hat (first few entries)");
    showTimes(&p.CUNFFTTimes, 1);
    dump(d, "fhat_nfft", p.f_hat, (size_t)(2 * p.N_total), sizeof(dTyp));
    {
        /* the reference's error norm between the two adjoints (showing only) */
        gpuComplex *a = (gpuComplex *)malloc(sizeof(gpuComplex) * (size_t)p.N_total);
        FILE *f;
        char path[1024];
        snprintf(path, sizeof path, "%s/%dd_fhat_ndft.bin", g_dir, d);
        f = fopen(path, "rb");
        if (f && fread(a, sizeof(gpuComplex), (size_t)p.N_total, f) == (size_t)p.N_total)
            printf("rel l_inf CUNFFT vs CUNDFT adjoint: %.3e\n",
                   compute_error_l_infty(a, p.f_hat, (int)p.N_total));
        if (f) fclose(f);
        free(a);
    }
    cunfft_finalize(&p);
}

int main(int argc, char **argv)
{
    uint_t N1[3] = {32, 1, 1}, N2[3] = {32, 32, 1};
    if (argc > 1) g_dir = argv[1];
    run(1, N1, N1[0]);
    run(2, N2, N2[0] * N2[1]);
    printf("TUTORIAL_OK\n");
    return 0;
}
