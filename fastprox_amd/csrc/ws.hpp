// The opaque structs of the C ABI (fp_op, fp_workspace), shared by the
// entry points (api.cpp) and the derivatives (deriv.hip).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <mutex>
#include <vector>

#include "alloc.hpp"
#include "../../include/fastprox_rocm.h"
#include "deriv.hpp"
#include "adbpg.hpp"
#include "engine.hpp"

namespace fpr {
struct DerivDeleter {
    void operator()(Deriv *d) const { deriv_free(d); }
};
}  // namespace fpr

struct fp_op {
    bool iso = true;
    int k = 2;
    double w[3] = {1.0, 1.0, 1.0};
    std::vector<double> spatial;  // iso: alpha (N); aniso: beta (k N); empty for none
    int64_t sshape[3] = {1, 1, 1};
    double spatial_max = 1.0;
    // device copies of the spatial weights, per (device, dtype)
    mutable std::mutex mu;
    mutable std::vector<std::pair<int, void *>> dev_f32, dev_f64;
    ~fp_op() {
        for (auto &p : dev_f32) (void)hipFree(p.second);
        for (auto &p : dev_f64) (void)hipFree(p.second);
    }
    int64_t nparams() const { return k + (int64_t)spatial.size(); }
    double step_scale() const {
        double m = 0.0;
        for (int a = 0; a < 3; ++a) m = std::max(m, a < k ? w[a] : 0.0);
        return m * (spatial.empty() ? 1.0 : spatial_max);
    }
    // The spatial weights on device dev in dtype (4 or 8 bytes); null without.
    void *device_weights(int dev, int tsize, hipError_t &err) const {
        err = hipSuccess;
        if (spatial.empty()) return nullptr;
        std::lock_guard<std::mutex> lock(mu);
        auto &vec = tsize == 4 ? dev_f32 : dev_f64;
        for (auto &p : vec)
            if (p.first == dev) return p.second;
        void *ptr = nullptr;
        const size_t n = spatial.size();
        if ((err = fpr::dev_alloc(&ptr, n * (size_t)tsize))) return nullptr;
        if (tsize == 4) {
            std::vector<float> f(spatial.begin(), spatial.end());
            err = hipMemcpy(ptr, f.data(), n * 4, hipMemcpyHostToDevice);
        } else {
            err = hipMemcpy(ptr, spatial.data(), n * 8, hipMemcpyHostToDevice);
        }
        if (err) {
            (void)hipFree(ptr);
            return nullptr;
        }
        vec.emplace_back(dev, ptr);
        return ptr;
    }
};

struct fp_workspace {
    const fp_op *op = nullptr;
    int dtype = FP_F64;
    int ndim = 0;
    int64_t shape[FP_MAX_NDIM] = {};
    int64_t numel = 0;
    int k = 2;
    int64_t B = 0, N = 0;
    int64_t img[3] = {1, 1, 1};
    int device = 0;
    void *stream = nullptr;
    bool keep = false;
    size_t tsize = 8;
    fpr::KernelChoice kc;
    std::unique_ptr<fpr::Engine> eng;                       // iso, and aniso by dual FISTA
    std::unique_ptr<fpr::AdSolver, fpr::AdDeleter> ad;      // aniso by the reference's method
    int method = 0;                                         // aniso: FPR_METHOD_ADBPG or FPR_METHOD_FISTA
    void *dstage = nullptr, *zstage = nullptr, *q = nullptr;
    void *xsave = nullptr, *zsave = nullptr;  // keep_structure: the point (x, z) of the last kept call
    void *usave = nullptr;                    // keep_structure, aniso: the per-edge dual of that call
    std::vector<int> exps;                    // keep_structure: the scale exponents of that call
    std::vector<char> host_stage;             // host arrays, packed
    std::vector<double> gammas;
    std::vector<fpr::ImagePtrs> ptrs;
    // results of the last call
    bool dual_valid = false, warm = false, seeded = false;
    // A solve enqueued on a capturing stream: its per-image results are read
    // (after a wait for the stream) at the first query of them.
    bool pending = false;
    bool info_valid = false;
    fp_status info_status = FP_OK;
    fp_info info{0.0, 0, 0};
    std::vector<double> gaps, zerrs, smooth;
    std::vector<int> iters;
    std::vector<char> exact;
    // derivatives
    bool valid = false;      // a kept point exists
    bool steps_on = false;
    double gamma = 0.0;
    double deriv_tol = 0.0;
    int deriv_max_iter = 10000;
    double deriv_res = 0.0;
    int deriv_its = 0;
    // keep_structure: the derivative structure's buffers, reserved at the
    // workspace's creation and rebuilt into at the first derivative or query
    // after each kept call (deriv_built)
    std::unique_ptr<fpr::Deriv, fpr::DerivDeleter> deriv;
    bool deriv_built = false;
    double *dscr = nullptr;  // device doubles for the derivatives' small arrays
    size_t dscr_n = 0;
    void *qbuf = nullptr;    // keep_structure: the queries' device output (9 numel bytes)
    // fp_workspace_set_branch (aniso): the side array (device int8, numel,
    // C order) and its kink tolerance, until the next prox call
    int8_t *branch = nullptr;
    bool branch_on = false;
    double branch_tol = 0.0;
    ~fp_workspace() {
        for (void *p : {dstage, zstage, q, xsave, zsave, usave, (void *)dscr, (void *)branch, qbuf})
            if (p) (void)hipFree(p);
    }
};
