// lowp.hip -- BF16 copies of a complex64 dictionary for the low-precision
// matcher: hi = bf16(x) (round to nearest even), lo = bf16(x - hi).  Each
// complex value becomes one 32-bit word (re in the low half, im in the high).
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

__global__ void lowp_kernel(const float2 *atoms, size_t n, uint32_t *hi, uint32_t *lo)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const float2 v = atoms[i];
        const uint32_t hr = bf16_rne(v.x), hm = bf16_rne(v.y);
        hi[i] = hr | (hm << 16);
        if (lo) {
            const uint32_t lr = bf16_rne(v.x - bf16_to_f(hr));
            const uint32_t lm = bf16_rne(v.y - bf16_to_f(hm));
            lo[i] = lr | (lm << 16);
        }
    }
}

}  // namespace

int launch_lowp_prepare(const float2 *atoms, size_t n, int T, uint32_t *hi, uint32_t *lo,
                        hipStream_t s)
{
    const size_t total = n * (size_t)T;
    if (total == 0) return 0;
    size_t blocks = (total + 255) / 256;
    if (blocks > 65536 * 4) blocks = 65536 * 4;
    lowp_kernel<<<dim3((unsigned)blocks), 256, 0, s>>>(atoms, total, hi, lo);
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? 0 : mrf_fail(MRF_E_GPU, "lowp launch: %s", hipGetErrorString(e));
}

}  // namespace mrf
