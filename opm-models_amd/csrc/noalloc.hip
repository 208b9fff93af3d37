// noalloc.hip -- the scalar product for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc): device inputs and output, the caller's stream, and the
// exact accumulator in a file-scope device array instead of an allocation.
//
// `scratch` holds the exact sum's device accumulators and ticket (DotScratch,
// 1.3 KiB) per GPU for the life of the process.  In host code its name is
// only a stand-in; the device address of each GPU's copy is looked up once per
// device with hipGetSymbolAddress and used for every launch.  All calls on one
// GPU share it: calls that overlap on one GPU (two streams, two threads) race
// and give wrong sums.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>

#include "omodels_internal.h"

namespace omodels {

namespace {

static __device__ __align__(256) DotScratch scratch[1];

constexpr int kMaxDevices = 64;
DotScratch* g_scratchPtr[kMaxDevices] = {nullptr};
std::mutex g_scratchMutex;

// the device address of this GPU's `scratch`, looked up on the first call
DotScratch* scratchOnCurrentDevice()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) throw std::runtime_error("no HIP device available");
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lk(g_scratchMutex);
    if (!g_scratchPtr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratchPtr[dev] = static_cast<DotScratch*>(p);
    }
    return g_scratchPtr[dev];
}

// clears the accumulators and the ticket (a kernel, not hipMemsetAsync, so
// that a caller's graph capture records an ordinary kernel node); the last
// workgroup of every product leaves them clear, this undoes an interrupted one
__global__ void k_clear_scratch(DotScratch* sc)
{
    for (int t = (int)threadIdx.x; t < kAccLimbs; t += (int)blockDim.x) {
        sc->acc[0].limb[t] = 0ull;
        sc->acc[1].limb[t] = 0ull;
    }
    if (threadIdx.x == 0) { sc->acc[0].flags = 0u; sc->acc[1].flags = 0u; sc->ticket = 0u; }
}

}  // namespace

void scalarProductNoAlloc(long m, const double* x, const double* y, double* out, hipStream_t st)
{
    if (m < 0 || m >= (1L << 28))
        throw std::invalid_argument("scalar product without allocation: 0 <= m < 2^28 (the accumulator's limit)");
    if (!out || (m > 0 && (!x || !y))) throw std::invalid_argument("scalar product without allocation: null pointer");
    DotScratch* sc = scratchOnCurrentDevice();
    hipLaunchKernelGGL(k_clear_scratch, dim3(1), dim3(64), 0, st, sc);
    launchScalarProduct(m, x, y, sc, out, st);
    OM_HIP(hipGetLastError());
}

}  // namespace omodels
