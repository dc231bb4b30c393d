// The counted allocations of alloc.hpp.
//
// SPDX-License-Identifier: MIT
#include "alloc.hpp"

#include <atomic>

namespace fpr {

static std::atomic<int64_t> g_allocs{0};

hipError_t dev_alloc_bytes(void **p, size_t bytes) {
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    return hipMalloc(p, bytes);
}

hipError_t pinned_alloc_bytes(void **p, size_t bytes) {
    g_allocs.fetch_add(1, std::memory_order_relaxed);
    return hipHostMalloc(p, bytes, 0);
}

int64_t alloc_count() { return g_allocs.load(std::memory_order_relaxed); }

}  // namespace fpr
