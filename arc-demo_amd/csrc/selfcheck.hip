/*
 * selfcheck.hip -- exhaustive on-device check of the engine's correctly
 * rounded reciprocal (xpbd_core.h recipf): every one of the 2^32 binary32 bit
 * patterns, recipf(x) against the IEEE quotient 1.0f / x, compared bit for
 * bit (NaN against NaN by class).  The fast path of recipf is taken only
 * where this check proves it exact.  The same kind of check for the
 * three-instruction quotient by a per-simulation divisor (xpbd_core.h qdiv):
 * every numerator of its domain (zero, 2^-96 <= |a| < 2^100) against a / d,
 * run once per divisor and device before a simulation enables that divisor's
 * fast path.
 */
#include <hip/hip_runtime.h>
#include <map>
#include <mutex>
#include <utility>
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

/* thread t checks the 256 patterns whose top 24 bits are t */
__global__ __launch_bounds__(256) void k_check_quotient(Divisor D, unsigned long long *mismatches)
{
    const uint32_t hi = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long bad = 0;
    for (uint32_t lo = 0; lo < 256u; ++lo) {
        const float a = __uint_as_float((hi << 8) | lo);
        if (!qdiv_domain(a))
            continue;
        const float f = qdiv(a, D), q = a / D.d;
        bad += __float_as_uint(f) == __float_as_uint(q) ? 0ull : 1ull;
    }
    if (bad)
        atomicAdd(mismatches, bad);
}

/* mismatches of qdiv's fast path for divisor d over all numerators, on the
 * current device; -1 on a HIP error */
static long long quotient_mismatches(float d)
{
    unsigned long long *dm = nullptr, h = 0;
    hipError_t e = hipMalloc((void **)&dm, sizeof(unsigned long long));
    if (e == hipSuccess)
        e = hipMemset(dm, 0, sizeof(unsigned long long));
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_check_quotient, dim3((1u << 24) / 256), dim3(256), 0, 0, Divisor{ d, 1.0f / d, 1 }, dm);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpy(&h, dm, sizeof(unsigned long long), hipMemcpyDeviceToHost);
    if (dm)
        (void)hipFree(dm);
    return e == hipSuccess ? (long long)h : -1;
}

bool quotient_checked(float d)
{
    static std::mutex mu;
    static std::map<std::pair<int, uint32_t>, bool> done;
    int dev = 0;
    if (!(fabsf(d) > 0.0f && fabsf(d) < INFINITY) || hipGetDevice(&dev) != hipSuccess)
        return false;
    uint32_t bits;
    __builtin_memcpy(&bits, &d, 4);
    const auto key = std::make_pair(dev, bits);
    {
        std::lock_guard<std::mutex> g(mu);
        const auto it = done.find(key);
        if (it != done.end())
            return it->second;
    }
    const bool ok = quotient_mismatches(d) == 0;
    std::lock_guard<std::mutex> g(mu);
    done[key] = ok;
    return ok;
}

} /* namespace xpbd */

extern "C" int arcx_check_quotient(float d, unsigned long long *mismatches)
{
    const long long m = xpbd::quotient_mismatches(d);
    if (m < 0)
        return -1;
    *mismatches = (unsigned long long)m;
    return 0;
}

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
