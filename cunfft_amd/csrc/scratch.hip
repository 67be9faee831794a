// scratch.hip -- the device-resident scratch of the allocation-free adjoint
// (cunfft_adjoint_noalloc_dev).
//
// One file-scope device array of kNoAllocScratchBytes holds that entry
// point's node binning and adjoint partial sums, so the entry point calls no
// allocator.  Each GPU the library runs on holds its own copy, at its own
// address, until the process ends.  In host code the array's name is only a
// stand-in: its device address is taken once per device with
// hipGetSymbolAddress (at plan set-up, the device current) and kept, and
// only that address is handed to kernels.
#include <mutex>

#include "internal.hpp"

namespace cunfft {

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxDevices] = {};
}  // namespace

void *noalloc_scratch()
{
    int dev = 0;
    CUNFFT_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        throw Error(CUNFFT_ERR_UNSUPPORTED, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[dev]) {
        void *ptr = nullptr;
        CUNFFT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        g_scratch_addr[dev] = ptr;
    }
    return g_scratch_addr[dev];
}

}  // namespace cunfft
