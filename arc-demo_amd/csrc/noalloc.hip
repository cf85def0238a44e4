/*
 * noalloc.hip -- the static device scratch of arcx_step_device_noalloc.
 *
 * The step kernels' scratch (per-substep contact lists, written by the
 * position pass and read back by the velocity pass of the same launch, and
 * the persistent distribution's work counter) normally lives in buffers each
 * sim allocates.  Callers that cannot allocate (code under graph capture,
 * tight serving loops, harnesses that forbid hipMalloc) step through this
 * one array instead: a file-scope device array, one copy per GPU, held from
 * the library's load until the process ends.
 *
 * Host code never passes `scratch` itself to a kernel: in host code the name
 * stands for a host-side shadow, not a device address.  Its device address is
 * looked up once per GPU, with that GPU current, by noalloc_scratch_init
 * (called from arc_sim_create) and kept in a per-device table; kernels get
 * that address as an argument.
 */
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <mutex>

#include "noalloc.h"

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES] = {};
} /* namespace */

hipError_t noalloc_scratch_init(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_ptr[device])
        return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return e;
    }
    g_ptr[device] = p;
    return hipSuccess;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= MAX_DEVICES)
        return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_ptr[device];
}
