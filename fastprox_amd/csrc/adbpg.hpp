// Anisotropic TV by the reference CPU library's method on the device
// (adbpg.hip): exact 1D TV steps on every fibre, the accelerated dual block
// proximal gradient method for two active axes, the three-block fast dual
// proximal gradient method for three, gap checks on their schedules, gradient
// restarts and the polish (src/ops/tv_aniso.cpp:259-616, 886-953;
// src/ops/tv1d.cpp).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpr {

struct AdSolver;

// A uniform batch of B images of one shape (extents L, padded in front with
// 1; L[2] contiguous), k = 2 or 3 image axes, axis weights w, dtype of tsize
// bytes. No edge weights.
struct AdSpec {
    int k = 2;
    int64_t L[3] = {1, 1, 1};
    int64_t B = 0;
    int tsize = 8;
    double w[3] = {1.0, 1.0, 1.0};
};

// One solve. Per image b, device arrays of the dtype (host arrays of B
// pointers): its input xs[b] and output zs[b] (contiguous), and its dual in C
// form (q_a = D_a^T p_a, z = x - sum_a q_a) at qs[b], plane a at qs[b] + a
// q_plane, read first on a warm start; usave, when not null, the per-edge
// dual u_a = p_a / w_a (|u_e| <= gamma), image b's plane a at usave + b u_img
// + a N, all at the input's scale.
struct AdCall {
    const void *const *xs = nullptr;
    void *const *zs = nullptr;
    void *const *qs = nullptr;
    int64_t q_plane = 0;
    void *usave = nullptr;
    int64_t u_img = 0;
    const double *gammas = nullptr;  // host, B
    bool warm = false;
    double tol = 1e-6;
    int max_iter = 1000;
};

// Per image of the last solve: the gap at the input's scale, iterations,
// whether it met its target, an infinite entry (status 1), the scale
// exponent of the solve, whether the polish certified it; the certified
// relative error of z (as the reference's driver.hpp:1186-1212 states it).
struct AdResult {
    double gap;
    double zerr;  // the certified relative error
    int iters;
    int converged;
    int status;
    int e;
    int polished;
};

// Allocates every device buffer of a solve of spec's shape; no later call
// allocates. An fp_status.
int ad_reserve(const AdSpec &spec, AdSolver **out);
void ad_free(AdSolver *s);
// Solves on stream st; waits for it. Results in ad_results. An fp_status.
int ad_solve(AdSolver *s, hipStream_t st, const AdCall &c);
const AdResult *ad_results(const AdSolver *s);
// The last solve's per-edge dual (as AdCall::usave) into usave, on stream st.
int ad_copy_edge_dual(AdSolver *s, hipStream_t st, void *usave, int64_t u_img);
// Launches and host synchronisations of the last solve.
void ad_stats(const AdSolver *s, int *launches, int *host_syncs, int *fallbacks);

struct AdDeleter {
    void operator()(AdSolver *s) const { ad_free(s); }
};

}  // namespace fpr
