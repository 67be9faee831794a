// grid_tile_f32.hip -- FP32 instances of the plane-tiled gridding kernels
// (grid_tile.hpp); one file per precision so the two compile in parallel.
#include "grid_tile.hpp"

namespace cunfft {

template bool tile_kernels_fit<float>(const Geom &, int, bool);
template int64_t tile_slot_cells<float>(const Geom &, int);
template void launch_tile_gather<float>(const Geom &, const Binning &, const float2 *, float2 *,
                                       hipStream_t);
template void launch_tile_spread<float>(const Geom &, const Binning &, const float2 *, float2 *,
                                       hipStream_t, hipEvent_t);
template WinConst<float> win_const<float>(const Geom &);
template bool launch_tile_grad<float>(const Geom &, const Binning &, const float2 *, const float2 *, float *,
                                     hipStream_t);
template void launch_tile_batch<float>(BatchStage, const GridJob<float> *, const BlockMap *, uint32_t,
                                      int, int, int, hipStream_t);

}  // namespace cunfft
