// scan.hip — exclusive prefix sums over per-cell counts (CSR offsets of the
// voxel cube and of the height-run work list), via hipCUB.
#include <hip/hip_runtime.h>
#include <hipcub/device/device_scan.hpp>
#include <stdint.h>

namespace smx {

// Exclusive sum of n int64 values. Call with tmp == nullptr to size tmp_bytes.
hipError_t scan_i64(const int64_t *in, int64_t *out, long n, void *tmp, size_t &tmp_bytes, hipStream_t s)
{
    return hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, in, out, (int)n, s);
}

}  // namespace smx
