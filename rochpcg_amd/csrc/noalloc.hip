// noalloc.hip -- ComputeDotProduct for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).
//
// hpcgx_dot reduces through the block partials a problem allocates at
// creation.  hpcgx_dot_noalloc runs the same two kernels (k_dot's block
// partials, then the one-block fold, launch_dot in ops.hip) with that scratch
// in a file-scope device array instead: no problem handle, no hipMalloc, no
// hipFree, no allocating library call.
//
// The array holds RED_BLOCKS partials and the result.  The dot's grid is
// capped at RED_BLOCKS blocks (ops.hip red_grid), whatever n is, so every
// length fits it; the entry point still checks the block count against the
// capacity and refuses a call that would need more.  Every block writes its
// own partial before the fold reads it, so nothing has to be cleared between
// calls.
//
// Host code never passes the array's name to a kernel or a copy (on the host
// it is a stand-in, not a device address): the device address comes from
// hipGetSymbolAddress, once per device on the first call with that device
// current, and is kept per device.  Each GPU holds its own copy of the array
// until the process ends, and every call on a GPU shares it: calls that
// overlap on one GPU (two host threads, or two streams whose work is in flight
// at once) race on it.
#include "ops.hpp"

#include <mutex>

namespace hpcgx
{

constexpr int64_t NOALLOC_PARTIALS = RED_BLOCKS;        // block partials of one reduction
constexpr int64_t NOALLOC_SLOTS = NOALLOC_PARTIALS + 32; // + the result, padded to 256 bytes
constexpr int NOALLOC_MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_SLOTS];

static std::mutex g_scratch_mu;
static double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// device address of this device's copy of `scratch`, looked up on first use
static int scratch_of_current_device(double** out)
{
    int dev = -1;
    HPCGX_HIP(hipGetDevice(&dev));
    if(dev < 0 || dev >= NOALLOC_MAX_DEVICES)
    {
        set_error("hpcgx_dot_noalloc: device ordinal " + std::to_string(dev) + " beyond the " +
                  std::to_string(NOALLOC_MAX_DEVICES) + " devices it keeps scratch addresses for");
        return -2;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if(!g_scratch_ptr[dev])
    {
        double* ptr = nullptr;
        HPCGX_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
        if(!ptr)
        {
            set_error("hpcgx_dot_noalloc: hipGetSymbolAddress returned a null address for the scratch array");
            return -1;
        }
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

int dot_noalloc(int64_t n, const double* x, const double* y, double* result, hipStream_t s)
{
    if(n < 0)
    {
        set_error("hpcgx_dot_noalloc: negative length");
        return -2;
    }
    const int64_t parts = reduction_blocks(n);
    if(parts > NOALLOC_PARTIALS)
    {
        set_error("hpcgx_dot_noalloc: the reduction needs " + std::to_string(parts) +
                  " block partials, beyond the scratch's " + std::to_string(NOALLOC_PARTIALS));
        return -2;
    }
    double* ptr = nullptr;
    HPCGX_TRY(scratch_of_current_device(&ptr));
    RedArgs ra;
    ra.partial = ptr;
    ra.op = OP_STORE;
    ra.out = ptr + NOALLOC_PARTIALS;
    ra.cap = NOALLOC_PARTIALS;
    HPCGX_TRY(launch_dot(n, x, true, y, true, ra, s, nullptr));
    HPCGX_HIP(hipMemcpyAsync(result, ra.out, sizeof(double), hipMemcpyDeviceToHost, s));
    HPCGX_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // namespace hpcgx
