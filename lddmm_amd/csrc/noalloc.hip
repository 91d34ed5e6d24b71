// noalloc.hip -- the fixed per-GPU scratch of the entry points that allocate
// no device memory (lddmm_multiply_kernel_cross_noalloc).
//
// `scratch` is a file-scope device array: every GPU the library runs on holds
// its own copy, at its own address, from module load to process end.  Host
// code never passes the array's name to a kernel (in host code it names a
// host-side stand-in); it asks the runtime for the device address once per
// device, with that device current, and uses that pointer from then on.
// Every call on one GPU shares that GPU's array, so calls that overlap on one
// GPU (two streams, two threads) race on it.
#include <mutex>
#include <vector>

#include "kernels.hpp"

namespace lddmm {

namespace {

__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

std::mutex g_mu;
std::vector<double *> g_ptr;   // device address of `scratch`, by device id

}  // namespace

double *noalloc_scratch()
{
    int dev = 0;
    LDDMM_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_ptr.size() <= dev) g_ptr.resize(dev + 1, nullptr);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        LDDMM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(kErrHip, "no device address for the no-allocation scratch");
        g_ptr[dev] = static_cast<double *>(p);
    }
    return g_ptr[dev];
}

}  // namespace lddmm
