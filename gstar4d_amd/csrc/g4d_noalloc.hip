// gStar4D for MI355X — allocation-free triangulation entry point.
//
// g4d_triangulate_device_noalloc runs the same pipeline as g4d_triangulate_device, but every
// intermediate buffer is carved from `scratch`, a file-scope device array of
// G4D_NOALLOC_SCRATCH_BYTES bytes that each GPU holds for the life of the process; the call
// makes no hipMalloc or hipFree.  Its device address is looked up once per device at context
// creation (g4d_create -> noalloc_setup) with hipGetSymbolAddress; host code only ever passes
// that address, never the array's name.  Every call on a device shares the device's array, so
// calls that overlap on one GPU race; the pipeline clears every counter and table it reads
// before use, so nothing depends on what the previous call left behind.
#include <hip/hip_runtime.h>

#include <mutex>

#include "gstar4d.h"
#include "g4d_noalloc.h"

namespace g4d {

static __device__ __align__(256) char scratch[G4D_NOALLOC_SCRATCH_BYTES];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[NOALLOC_MAX_DEVICES];

hipError_t noalloc_setup(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice; // the caller makes `device` current
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!ptr) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = ptr;
    return hipSuccess;
}

void *noalloc_scratch(int device)
{
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

} // namespace g4d
