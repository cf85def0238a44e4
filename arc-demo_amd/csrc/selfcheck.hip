/*
 * selfcheck.hip -- exhaustive on-device check of the engine's correctly
 * rounded reciprocal (xpbd_core.h recipf): every one of the 2^32 binary32 bit
 * patterns, recipf(x) against the IEEE quotient 1.0f / x, compared bit for
 * bit (NaN against NaN by class).  The fast path of recipf is taken only
 * where this check proves it exact.
 */
#include <hip/hip_runtime.h>
#include "xpbd_core.h"

namespace xpbd {

/* thread t checks the 256 patterns whose top 24 bits are t */
__global__ __launch_bounds__(256) void k_check_recip(unsigned long long *mismatches)
{
    const uint32_t hi = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long bad = 0;
    for (uint32_t lo = 0; lo < 256u; ++lo) {
        const float x = __uint_as_float((hi << 8) | lo);
        const float a = recipf(x), b = 1.0f / x;
        const bool same = __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
        bad += same ? 0ull : 1ull;
    }
    if (bad)
        atomicAdd(mismatches, bad);
}

} /* namespace xpbd */

extern "C" int arcx_check_recip(unsigned long long *mismatches)
{
    unsigned long long *d = nullptr;
    hipError_t e = hipMalloc((void **)&d, sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipMemset(d, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(xpbd::k_check_recip, dim3((1u << 24) / 256), dim3(256), 0, 0, d);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (d)
        (void)hipFree(d);
    return e == hipSuccess ? 0 : -1;
}
