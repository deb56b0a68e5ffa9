// Device scratch of the no-allocation entry point (nmfg_run_device_noalloc):
// one statically sized array per device, resolved to a device address once
// per device when a context is created on it.
#include "nmfg_internal.h"

static __device__ __align__(256) unsigned char scratch[nmfg::NOALLOC_SCRATCH_BYTES];

namespace nmfg {

// Address of this device's copy of `scratch`; call with the device current.
// The array's name is only a host-side stand-in for the symbol: every use
// of its memory goes through the address returned here.
hipError_t noalloc_scratch_address(void **out) {
    *out = nullptr;
    return hipGetSymbolAddress(out, HIP_SYMBOL(scratch));
}

}  // namespace nmfg
