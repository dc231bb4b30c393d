// Device and pinned host memory of the library. Every buffer the library
// allocates goes through these two functions, which count the allocations,
// so that a caller (and a test) can check that no call after a workspace's
// creation allocates (fpr_alloc_count).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace fpr {

hipError_t dev_alloc_bytes(void **p, size_t bytes);
hipError_t pinned_alloc_bytes(void **p, size_t bytes);
int64_t alloc_count();

template <class P>
inline hipError_t dev_alloc(P **p, size_t bytes) {
    return dev_alloc_bytes((void **)p, bytes);
}
template <class P>
inline hipError_t pinned_alloc(P **p, size_t bytes) {
    return pinned_alloc_bytes((void **)p, bytes);
}

}  // namespace fpr
