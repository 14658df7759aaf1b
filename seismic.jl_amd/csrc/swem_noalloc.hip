// swem_noalloc.hip -- the fixed device scratch of the allocation-free migration entry point
// (swem_plan_migrate_noalloc).
//
// The scratch is a file-scope device array: every GPU the library runs on holds its own copy,
// for the life of the process.  Host code never passes the array's name to a kernel or a copy
// (on the host it names a stand-in, not device memory): the per-device address comes from
// hipGetSymbolAddress, taken once per device with that device current, and only that address
// is used.
#include <mutex>

#include "swem_internal.hpp"

namespace swem {

static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {
std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_scratch(int device, void** ptr, size_t* bytes)
{
    if (device < 0 || device >= kMaxDevices) return set_error(SWEM_EINVAL, "device %d out of range", device);
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[device]) {
        int cur = -1;
        SWEM_HIP(hipGetDevice(&cur));
        if (cur != device) SWEM_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p)
            return set_error(SWEM_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_ptr[device] = p;
    }
    *ptr = g_ptr[device];
    *bytes = kNoallocScratchBytes;
    return SWEM_OK;
}

}  // namespace swem
