// gdp_devmem.h - device memory of the library's problems, servers and
// temporaries, with a process-wide cache of freed blocks.
//
// A drop-in call creates and destroys a problem every time, and hipMalloc and
// hipFree of its arrays (and of every call's temporaries) cost milliseconds.
// Freed blocks are therefore kept per device and handed out again for
// requests of up to 1.25 times less; a release waits for the device first,
// as hipFree does, so a block is never reused while a kernel may still touch
// it.  The cache holds at most 4 GiB per device; an allocation that fails
// empties it and tries again.  Blocks still cached when the process ends are
// reclaimed with the process's device memory.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gdp {

hipError_t cached_malloc(void **p, size_t bytes);
template <typename T>
hipError_t cached_malloc(T **p, size_t bytes) {
    return cached_malloc(reinterpret_cast<void **>(p), bytes);
}
// Waits for the current device, then keeps the block for reuse (or frees it
// when the cache is full).  A null pointer is ignored.
hipError_t cached_free(void *p);
// Bytes of free blocks kept for the current device.
size_t cached_bytes();

}  // namespace gdp
