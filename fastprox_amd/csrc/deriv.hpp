// Implicit-differentiation derivatives of image TV on the device (deriv.cpp).
//
// SPDX-License-Identifier: MIT
#pragma once

#include "../../include/fastprox_rocm.h"

namespace fpr {

struct Deriv;

fp_status jvp_impl(const fp_op *op, fp_workspace *ws, const fp_array *dx, double dgamma, const fp_array *dparams,
                   fp_array *dout);
fp_status vjp_impl(const fp_op *op, fp_workspace *ws, const fp_array *g, fp_array *gx, double *ggamma_host,
                   fp_array *ggamma_dev, fp_array *gparams);

}  // namespace fpr
