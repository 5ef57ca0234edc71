// Device scratch of the allocation-free entry point
// (beagleExtCalculateRootLogLikelihoodsNoAlloc): the workgroup sums and the
// result slot of the root log-likelihood live in one fixed array per GPU
// instead of an instance buffer grown on demand, so the call itself
// allocates nothing.
//
// In host code `scratch` names a host-side stand-in, not device memory; the
// device address is looked up once per device with hipGetSymbolAddress (with
// that device current, when an instance is created on it) and only that
// address is handed to kernels and copies.  Each GPU holds its own copy of
// the array, at its own address, until the process ends.  Every call on a
// GPU uses that GPU's array: calls that overlap on one GPU (different
// instances on different streams) race.
#include "kernels.h"

#include <mutex>

namespace hipbeagle {

static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
double* gScratch[kMaxDevices] = {};
}  // namespace

double* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return nullptr;
        }
        gScratch[device] = static_cast<double*>(p);
    }
    return gScratch[device];
}

}  // namespace hipbeagle
