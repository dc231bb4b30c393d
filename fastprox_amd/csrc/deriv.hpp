// Implicit-differentiation derivatives of image TV on the device (deriv.hip):
// the structure at a solution (flat pixels or edges, their connected regions)
// and the JVP and VJP of a uniform batch of images.
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpr {

struct Deriv;

// A solved batch of B images of one shape, as the derivatives read it. Every
// pointer is device memory of the solve's dtype (tsize bytes) unless noted.
struct DerivSpec {
    bool iso = true;
    int k = 2;
    int64_t L[3] = {1, 1, 1};   // extents, padded in front with 1 (L[2] contiguous)
    int64_t B = 0;
    int tsize = 8;
    double w[3] = {1.0, 1.0, 1.0};  // the operator's axis weights (first k)
    const void *x = nullptr;    // the inputs, B contiguous images
    const void *z = nullptr;    // the outputs
    // The dual: iso, u at the input's scale (|u_p| <= gamma alpha_p); aniso,
    // the per-edge dual (|u_e| <= gamma beta_e). Image b's plane a at
    // u + b u_img + a u_plane.
    const void *u = nullptr;
    int64_t u_img = 0, u_plane = 0;
    // Spatial weights (iso alpha: one plane; aniso beta: k planes) or null;
    // sw_img = 0: one map shared by every image.
    const void *sw = nullptr;
    int64_t sw_img = 0, sw_plane = 0;
    const double *gamma = nullptr;  // host, B: the steps as given
    const double *gap = nullptr;    // host, B: the gaps at the input's scale (iso)
    const int *e = nullptr;         // host, B: the solves' scale exponents (iso), or null for 0
    // Aniso branch overrides (fp_workspace_set_branch), or side null: at an
    // edge (p, p + e_a) at a kink within side_tol (its jump and its 1D dual
    // within side_tol of 0 and of the bound), side[p] = -1 joins its two
    // sides and +1 keeps them apart with the sign of its dual; 0 keeps the
    // default rule. side: device int8, B contiguous images. q: the C-form
    // dual (k planes) the kink test sums along each fiber.
    const int8_t *side = nullptr;
    double side_tol = 0.0;
    const void *q = nullptr;
    int64_t q_img = 0, q_plane = 0;
};

// Allocates every buffer of the structure of B images of spec's shape (the
// shape fields: iso, k, L, B, tsize, whether spatial weights are given and
// whether they are shared); no later call on it allocates. An fp_status.
int deriv_reserve(const DerivSpec &spec, Deriv **out);
// Builds the structure of the solution spec (of the reserved shape) into the
// reserved buffers (one host synchronisation); an fp_status.
int deriv_build(Deriv *d, const DerivSpec &spec, hipStream_t s);
// deriv_reserve, then deriv_build.
int deriv_create(const DerivSpec &spec, hipStream_t s, Deriv **out);
void deriv_free(Deriv *d);

// The VJP of the cotangent g (B images): gx (B images); per-image gradients
// in gamma (device, B doubles) or null; the axis weights' gradient summed over
// the images (device, k doubles) or null; the spatial weights' gradient
// (device doubles shaped like the weights: per image, or summed over the
// images for a shared map) or null. tol 0: the default tolerance.
int deriv_vjp(Deriv *d, hipStream_t s, const void *g, void *gx, double *ggamma, double *gw, double *gsw, double tol,
              int max_iter);

// The JVP along dx (B images, or null for 0), per-image dgamma (device, B
// doubles, or null), dw (host, k doubles, or null) and dsw (device doubles
// shaped like the spatial weights, or null): dz (B images).
int deriv_jvp(Deriv *d, hipStream_t s, const void *dx, const double *dgamma, const double *dw, const double *dsw,
              void *dz, double tol, int max_iter);

// The last derivative's largest CG residual and iteration count over the
// images, and the number of regions of the structure. Waits for the stream.
void deriv_result(const Deriv *d, double *residual, int *iters, int64_t *regions);

// Per-image results of the last derivative (host arrays of B).
void deriv_image_result(const Deriv *d, double *residual, int *iters);

// The structure's regions (fp_workspace_regions): per entry of the B images
// (device, B N, C order), its region's number within its image in the order
// of the regions' first entries (int64 if lab64, else int32), or -1 in an
// image whose input holds a NaN; sign +1 (0 there). Either may be null.
// nregions (host, B): the images' region counts, -1 for a NaN image, or
// null. Waits for the stream when nregions is given. An fp_status.
int deriv_regions(Deriv *d, hipStream_t s, void *labels, bool lab64, int8_t *sign, int64_t *nregions);

// Kinks (fp_workspace_kinks) at the solution the structure was built from,
// with q its C-form dual (device, image b's plane a at q + b q_img + a
// q_plane): per entry the code into mask (device, B N int8; iso: 1 where
// |(W D z)_p| <= tol and |q_p| >= r_p - tol with r_p > 0; aniso: bit a where
// the edge (p, p + e_a) has |z_{p+e_a} - z_p| <= tol and its 1D dual, the
// running sum of q_a along the fiber, is within tol of its bound gamma w_a
// beta_e > 0), and the entries with a nonzero code into *count. Waits for the
// stream. An fp_status.
int deriv_kinks(Deriv *d, hipStream_t s, const void *q, int64_t q_img, int64_t q_plane, double tol, int8_t *mask,
                int64_t *count);

}  // namespace fpr
