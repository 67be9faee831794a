// bin.hpp -- node binning state shared by the tiled gridding kernels.
#pragma once
#include "internal.hpp"

namespace cunfft {

// Nodes sorted by the grid tile that holds their window start.  Built once per
// set of nodes on the plan stream and reused by every forward and adjoint
// call until the nodes change.
struct Binning {
    int built = 0;
    void build(const Geom &ge, const double *x, hipStream_t s);
    void release();
};

template <typename T>
void launch_gather_binned(const Geom &ge, const Binning &b, const cplx_t<T> *g, const double *x,
                          cplx_t<T> *f, hipStream_t s);
template <typename T>
void launch_scatter_binned(const Geom &ge, const Binning &b, const double *x, const cplx_t<T> *f,
                           cplx_t<T> *g, hipStream_t s);

}  // namespace cunfft
