// Pass scratch of the allocation-free single-row entry point
// (dcx_klist_noalloc): a file-scope device array, so the entry point needs
// no device allocation.  Every device the library runs on holds its own
// copy until the process ends; its address on the current device comes
// from hipGetSymbolAddress (the array's name is not a device address in
// host code).
#include <hip/hip_runtime.h>

#include "dcx_internal.hpp"

namespace dcx {

static __device__ __align__(256) uint32_t scratch[(size_t)kScratchArrays * kNoAllocMaxVertices];

hipError_t noalloc_scratch_address(void** ptr) {
    *ptr = nullptr;
    return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

}  // namespace dcx
