// grid_tile_f64.hip -- FP64 instances of the plane-tiled gridding kernels
// (grid_tile.hpp); one file per precision so the two compile in parallel.
#include "grid_tile.hpp"

namespace cunfft {

template bool tile_kernels_fit<double>(const Geom &, int, bool);
template int64_t tile_slot_cells<double>(const Geom &, int);
template void launch_tile_gather<double>(const Geom &, const Binning &, const double2 *, double2 *,
                                       hipStream_t);
template void launch_tile_spread<double>(const Geom &, const Binning &, const double2 *, double2 *,
                                       hipStream_t, hipEvent_t);
template WinConst<double> win_const<double>(const Geom &);
template bool launch_tile_grad<double>(const Geom &, const Binning &, const double2 *, const double2 *, double *,
                                     hipStream_t);
template void launch_tile_batch<double>(BatchStage, const GridJob<double> *, const BlockMap *, uint32_t,
                                      int, int, int, hipStream_t);

}  // namespace cunfft
