// Allocation-free execution: dedisp_amd_execute_noalloc.
//
// A TDD (or legacy) execute on device pointers whose device buffers (the
// channel-major transposed rows and the delay tables) come from one file-scope
// device array instead of hipMalloc, for callers that must not allocate
// device memory (code under HIP graph capture, tight serving loops). The array
// is DEDISP_AMD_NOALLOC_SCRATCH_BYTES per device, held from the library's load
// until the process ends. Its device address is looked up once per device
// with hipGetSymbolAddress (in host code the array's name is not a device
// address); every call on that device shares it, so calls that overlap on one
// device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>

#include "dedisp_amd.h"
#include "internal.hpp"

namespace {

static_assert(DEDISP_AMD_NOALLOC_SCRATCH_BYTES % 256 == 0, "scratch size must keep 256-byte alignment");

static __device__ __align__(256) unsigned char scratch[DEDISP_AMD_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    void* ptr = nullptr;               // device address of `scratch` on this device
    unsigned long long tag = 0;        // tables the scratch holds (0: none)
};

std::mutex g_mutex;
DeviceScratch g_dev[kMaxDevices];

// The device address of this device's copy of `scratch`, looked up on the
// first call on the device (which must be current).
DeviceScratch& device_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) DD_THROW(DEDISP_INVALID_DEVICE_INDEX);
    std::lock_guard<std::mutex> lock(g_mutex);
    DeviceScratch& d = g_dev[device];
    if (!d.ptr) {
        void* p = nullptr;
        DD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) DD_THROW(DEDISP_INTERNAL_GPU_ERROR);
        d.ptr = p;
        d.tag = 0;
    }
    return d;
}

}  // namespace

extern "C" dedisp_error dedisp_amd_execute_noalloc(dedisp_plan plan, dedisp_size nsamps, const dedisp_byte* in,
                                                   dedisp_size in_nbits, dedisp_byte* out, dedisp_size out_nbits,
                                                   unsigned flags)
{
    if (!plan || !plan->ptr) return DEDISP_INVALID_PLAN;
    {
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess && e != hipErrorNoDevice) return DEDISP_PRIOR_GPU_ERROR;
    }
    auto* t = dynamic_cast<dedisp::TDDPlan*>(plan->ptr.get());
    if (!t) return DEDISP_INVALID_PLAN;
    try {
        const int device = t->get_device();
        DD_HIP(hipSetDevice(device));
        DeviceScratch& d = device_scratch(device);
        t->execute_noalloc(nsamps, in, in_nbits, out, out_nbits, flags, d.ptr, DEDISP_AMD_NOALLOC_SCRATCH_BYTES,
                           &d.tag);
    } catch (const dedisp::Error& e) {
        return e.code();
    } catch (const std::bad_alloc&) {
        return DEDISP_MEM_ALLOC_FAILED;
    } catch (...) {
        return DEDISP_UNKNOWN_ERROR;
    }
    return DEDISP_NO_ERROR;
}
