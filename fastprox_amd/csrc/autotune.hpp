// The step-kernel choice per workspace shape (autotune.cpp).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <cstdint>

#include "tv_kernels.hpp"

namespace fpr {

// The kernel for images of extents img (padded to 3) in a batch of B: the
// entry of the tuning table for (k, dtype, spatial weights) nearest to the
// image size, the built-in default without one.
KernelChoice default_kernel(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B);

// The tail kernel and the live-tile fraction that switches to it (Engine::
// set_tail); frac 0 when there is none.
KernelChoice default_tail(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B, double *frac);

// The step kernel of a serving class (2D: the batch kernel of a 512^2
// image; 3D: slabs of 8 slices), unless FASTPROX_ROCM_KERNEL /
// FASTPROX_ROCM_KERNEL3D name one.
KernelChoice serve_kernel(int k, int tsize, bool iso);

}  // namespace fpr
