// The FP64 interface heat-flux check without any device allocation, for
// callers that cannot allocate: code under HIP graph capture, tight serving
// or time-stepping loops, harnesses that forbid hipMalloc.
//
// The reduction partials and the two results live in `scratch`, a
// file-scope device array each GPU holds once for the life of the process.
// Host code never passes the array's name to a kernel or copy: it takes the
// array's device address once per device with hipGetSymbolAddress and uses
// that.  Every call on one GPU shares the array and overwrites what the last
// call left (every partial the sum reads is written first, so nothing needs
// clearing); calls that overlap in time on one GPU -- two contexts on
// different streams, or two threads -- race on it and must be serialised by
// the caller.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "batch.hpp"
#include "noalloc.hpp"

namespace mrfg {

namespace {

static __device__ __align__(256) double scratch[kFluxScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex gMu;
double *gPtr[kMaxDevices] = {};

}  // namespace

double *fluxScratch(int device)
{
    if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range for the flux scratch");
    std::lock_guard<std::mutex> lk(gMu);
    if (!gPtr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
        if (cur != device) throw std::runtime_error("flux scratch: the device is not current");
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("HIP error: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        gPtr[device] = (double *)p;
    }
    return gPtr[device];
}

}  // namespace mrfg
