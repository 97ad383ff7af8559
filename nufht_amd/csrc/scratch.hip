// The allocation-free entry point's scratch: a file-scope static device
// array that each GPU the library runs on holds until the process ends.
// It lives in its own translation unit (its own code object), so that the
// array is placed on a GPU only when nufht_batch_device_noalloc first runs
// there, not when the engine's kernels first load.  Host code reaches it
// only through the per-device address hipGetSymbolAddress returns, looked up
// once per device with that device current.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"

constexpr long long kScratchDoubles = 1LL << 24;   // 128 MiB per GPU
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kScratchMaxDevices];

// Device address of the current device's scratch array and its length in
// doubles; 0, or -550 / -510 with the error text set.
extern "C" int nufht_scratch_array(double **out, long long *count)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        nufht_set_error("hipGetDevice failed: %s", hipGetErrorString(e));
        return -510;
    }
    if (dev < 0 || dev >= kScratchMaxDevices) {
        nufht_set_error("device %d is beyond the %d the scratch table covers", dev,
                        kScratchMaxDevices);
        return -550;
    }
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            nufht_set_error("hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return -510;
        }
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    *count = kScratchDoubles;
    return 0;
}
