// The allocation-free solve (static_solve.cpp): one solve of a uniform batch
// on contiguous device arrays with every device buffer carved from a fixed
// device array of kStaticScratchBytes.
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "tv_common.hpp"
#include "tv_kernels.hpp"

namespace fpr {

constexpr size_t kStaticScratchBytes = size_t(64) << 20;  // per GPU, held until the process ends
constexpr int kStaticMaxImages = 256;
constexpr int kStaticMaxDevices = 16;

struct StaticSolve {
    int tsize = 4;
    Problem P{};             // k, iso, sw, w, sched, wmax_norm, alpha_max (tol, max_iter below)
    KernelChoice kc{};
    int B = 0;               // images
    int64_t L[3] = {1, 1, 1};
    int64_t N = 0;
    const void *x = nullptr; // B contiguous images on the current device
    void *z = nullptr;       // B contiguous outputs
    bool stage_x = false;    // z overwrites x: copy x first
    const double *sw_host = nullptr;  // the operator's spatial weights (host doubles) or null
    double gamma = 0.0, tol = 1e-6;
    int max_iter = 1000;
    hipStream_t stream = nullptr;
};

struct StaticResult {
    double gap = 0.0;   // summed over the images, at the input's scale
    int iters = 0;      // the largest
    bool notconv = false, bad = false;
};

// The scratch bytes a solve of c needs (at most kStaticScratchBytes to run).
size_t static_bytes_needed(const StaticSolve &c);
// Runs c on the current device (blocks until done). A hipError_t, or -1
// when the schedule does not finish.
int static_solve(const StaticSolve &c, StaticResult *res);

}  // namespace fpr
