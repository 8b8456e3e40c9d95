// vll_fast_h6.hip — the FAST kernels for half-window size 6 (one
// translation unit per size, so the three compile in parallel).
#include "vll_fast_kernels.hpp"

namespace smx {

template hipError_t fast_launch<6>(int, bool, dim3, size_t, hipStream_t, const VllParams *, const CellArgs &,
                                   const FastArgs &);
template hipError_t fast_launch_passes<6>(bool, dim3, dim3, size_t, hipStream_t, const VllParams *,
                                          const CellArgs &, const FastArgs &, const FastArgs &);
template hipError_t fast_launch_batch<6>(int, bool, int, size_t, hipStream_t, const BatchArgs &);
template int fast_batch_blocks_per_cu<6>(int, bool, size_t);

}  // namespace smx
