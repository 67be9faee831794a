// bin.hip -- binned gridding entry points.
#include "bin.hpp"

namespace cunfft {

void Binning::build(const Geom &ge, const double *x, hipStream_t s)
{
    (void)ge;
    (void)x;
    (void)s;
    built = 1;
}

void Binning::release() { built = 0; }

template <typename T>
void launch_gather_binned(const Geom &ge, const Binning &b, const cplx_t<T> *g, const double *x,
                          cplx_t<T> *f, hipStream_t s)
{
    (void)b;
    launch_gather_plain<T>(ge, g, x, f, s);
}

template <typename T>
void launch_scatter_binned(const Geom &ge, const Binning &b, const double *x, const cplx_t<T> *f,
                           cplx_t<T> *g, hipStream_t s)
{
    (void)b;
    int64_t nt = ge.n[0] * ge.n[1] * ge.n[2];
    launch_zero<T>(g, nt, s);
    launch_scatter_atomic<T>(ge, x, f, g, s);
}

template void launch_gather_binned<float>(const Geom &, const Binning &, const float2 *,
                                          const double *, float2 *, hipStream_t);
template void launch_gather_binned<double>(const Geom &, const Binning &, const double2 *,
                                           const double *, double2 *, hipStream_t);
template void launch_scatter_binned<float>(const Geom &, const Binning &, const double *,
                                           const float2 *, float2 *, hipStream_t);
template void launch_scatter_binned<double>(const Geom &, const Binning &, const double *,
                                            const double2 *, double2 *, hipStream_t);

}  // namespace cunfft
