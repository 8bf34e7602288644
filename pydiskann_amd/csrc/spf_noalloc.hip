// Fixed device scratch of the no-allocation search entry point
// (spf_search_batch_device_noalloc): a file-scope device array, one copy per
// GPU the library runs on, held until the process ends.  Host code never uses
// the array's name as an address: the device address is looked up once per
// device with hipGetSymbolAddress (that device current) and kept per device.
#include <mutex>

#include "spf_internal.h"

static __device__ __align__(256) unsigned char scratch[spf::kNoAllocScratchBytes];

namespace spf {

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_addr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch(int device, void **ptr) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (!g_addr[device]) {   // the caller has made `device` current
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_addr[device] = p;
    }
    *ptr = g_addr[device];
    return hipSuccess;
}

}  // namespace spf
