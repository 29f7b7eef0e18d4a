// order.hip -- the voxel order of the candidate rescore: voxels sorted by
// their best low-precision candidate (a radix sort of (key, voxel) pairs),
// so that the waves working at one time rescore voxels whose candidate rows
// are largely the same atoms and meet them in the L2 and MALL instead of
// fetching each voxel's rows from HBM.
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

namespace mrf {

size_t voxel_sort_tmp_bytes(size_t V)
{
    size_t bytes = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr,
                                           (uint32_t *)nullptr, (const uint32_t *)nullptr,
                                           (uint32_t *)nullptr, (int)V, 0, 32,
                                           (hipStream_t)0) != hipSuccess)
        return 0;
    return bytes;
}

int sort_voxels(void *tmp, size_t tmp_bytes, const uint32_t *keys_in, uint32_t *keys_out,
                const uint32_t *vox_in, uint32_t *vox_out, size_t V, int bits, hipStream_t s)
{
    size_t bytes = tmp_bytes;
    const hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vox_in,
                                                            vox_out, (int)V, 0, bits, s);
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "voxel sort: %s", hipGetErrorString(e));
}

}  // namespace mrf
