/*
 * layout.c -- prints sizeof and the field offsets of cunfft_plan and
 * cunfftf_plan as a C compiler lays them out from include/cunfft.h, so a
 * test can compare them with the ctypes mirror in cunfft_amd/__init__.py.
 * No library call.
 */
#include <stddef.h>
#include <stdio.h>

#include "cunfft.h"

#define F(T, name) printf("%s %s %zu\n", #T, #name, offsetof(T, name))
#define ALL(T)                                                                         \
    printf("%s sizeof %zu\n", #T, sizeof(T));                                         \
    F(T, N_total); F(T, M_total); F(T, d); F(T, N); F(T, x); F(T, f_hat); F(T, f);   \
    F(T, sigma); F(T, n); F(T, n_total); F(T, m); F(T, b); F(T, g); F(T, g_hat);     \
    F(T, fg_exp_l_cpu); F(T, g_gpu); F(T, fhat_gpu); F(T, x_gpu); F(T, f_gpu);       \
    F(T, fft_forward); F(T, fft_backward); F(T, CUNFFTTimes); F(T, flags); F(T, impl); \
    F(T, status)

int main(void)
{
    ALL(cunfft_plan);
    ALL(cunfftf_plan);
    printf("NFFTTimeSpec sizeof %zu\n", sizeof(NFFTTimeSpec));
    printf("kdata sizeof %zu\n", sizeof(kdata));
    return 0;
}
