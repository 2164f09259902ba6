// The fixed device scratch behind tk_coarse_no_alloc.
//
// This file is compiled as an ordinary HIP translation unit (host and device
// parts), so the HIP runtime registers `scratch` and allocates one copy of it on
// each device the process uses, for the life of the process.  The kernels live in
// the separately loaded code object and never name the array: the host takes its
// device address once per device with hipGetSymbolAddress and passes that address
// to the kernels like any other buffer.  In host code `scratch` itself is only a
// host-side stand-in for the symbol; its address is not a device address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "tk_internal.h"

namespace {

static __device__ __align__(256) unsigned char scratch[TK_NO_ALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_lock;
void *g_address[kMaxDevices] = {};

}  // namespace

namespace tk {

tk_status scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) {
        set_error("device %d is outside the %d devices the no-alloc scratch supports", device, kMaxDevices);
        return TK_ERR_INVALID;
    }
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_address[device]) {
        int current = -1;
        hipError_t e = hipGetDevice(&current);
        if (e != hipSuccess || current != device) {
            set_error("no-alloc scratch: device %d is not current (%s)", device, hipGetErrorString(e));
            return TK_ERR_HIP;
        }
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("hipGetSymbolAddress of the no-alloc scratch failed on device %d: %s", device,
                      hipGetErrorString(e));
            return TK_ERR_HIP;
        }
        g_address[device] = p;
    }
    *out = g_address[device];
    return TK_OK;
}

}  // namespace tk
