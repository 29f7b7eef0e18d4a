// lowp.hip -- BF16 copies of complex64 rows (dictionary atoms or signals) for
// the low-precision matcher: hi = bf16(x) (round to nearest even),
// lo = bf16(x - hi).  Each complex value becomes one 32-bit word (re in the
// low half, im in the high); rows are padded with zero words to the stride.
#include "kernels.h"
#include "snapmrf.h"

#include <hip/hip_runtime.h>

namespace mrf {

__device__ __forceinline__ uint32_t bf16_rne(float f)
{
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (u >> 16) | ((u & 0xffffu) ? 0x40u : 0u);
    return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ float bf16_to_f(uint32_t h) { return __uint_as_float(h << 16); }

namespace {

// One work-group per row (grid-stride over rows), one word per thread.
__global__ void lowp_kernel(const float2 *src, const int32_t *perm, size_t rows, int T, int ld,
                            uint32_t *hi, uint32_t *lo)
{
    for (size_t r = blockIdx.x; r < rows; r += gridDim.x) {
        const float2 *row = src + (size_t)(perm ? perm[r] : (int32_t)r) * T;
        uint32_t *h = hi + r * ld;
        uint32_t *l = lo ? lo + r * ld : nullptr;
        for (int t = threadIdx.x; t < ld; t += blockDim.x) {
            uint32_t hw = 0u, lw = 0u;
            if (t < T) {
                const float2 v = row[t];
                const uint32_t hr = bf16_rne(v.x), hm = bf16_rne(v.y);
                hw = hr | (hm << 16);
                lw = bf16_rne(v.x - bf16_to_f(hr)) | (bf16_rne(v.y - bf16_to_f(hm)) << 16);
            }
            h[t] = hw;
            if (l) l[t] = lw;
        }
    }
}

}  // namespace

int launch_lowp_prepare(const float2 *src, const int32_t *perm, size_t rows, int T, int ld,
                        uint32_t *hi, uint32_t *lo, hipStream_t s)
{
    if (rows == 0 || T == 0) return 0;
    if (ld < T || (ld & 31)) return mrf_fail(MRF_E_ARG, "lowp: bad row stride %d for T=%d", ld, T);
    const size_t blocks = rows < 65536 ? rows : 65536;
    lowp_kernel<<<dim3((unsigned)blocks), 256, 0, s>>>(src, perm, rows, T, ld, hi, lo);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "lowp launch: %s", hipGetErrorString(e));
}

}  // namespace mrf
