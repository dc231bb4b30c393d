// The C ABI of libfastprox_rocm (include/fastprox_rocm.h): fastprox's image-TV
// entry points over the HIP engine.
//
// Argument rules follow the reference's boundary (src/api/api.cpp): every
// argument is checked before any work (a rejected call leaves the workspace as
// it was, api.cpp:480-511), no exception crosses the boundary (guard,
// api.cpp:28-38), and a workspace owns every buffer a solve needs
// (build_workspace, api.cpp:223-291), so fp_prox allocates nothing.
//
// SPDX-License-Identifier: MIT
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <vector>

#include "alloc.hpp"
#include "../../include/fastprox_rocm.h"
#include "autotune.hpp"
#include "deriv.hpp"
#include "engine.hpp"
#include "ws.hpp"
#include "serve.hpp"
#include "static_solve.hpp"

#ifndef FASTPROX_ROCM_BUILD
#define FASTPROX_ROCM_BUILD "1.0.6"
#endif

// ---------------------------------------------------------------------------
// Operators and workspaces
// ---------------------------------------------------------------------------


namespace {

using namespace fpr;

thread_local char g_last_error[256] = "";

// Records why the current call fails (fpr_last_error) and returns s.
fp_status fail(fp_status s, const char *why) {
    std::snprintf(g_last_error, sizeof g_last_error, "%s", why);
    return s;
}

constexpr double kMaxScaledStep = 0x1p960;  // the step range (op.hpp:48)
constexpr int64_t kMaxElems = (int64_t(1) << 60) - 1;

// The status for the exception being handled: out of memory (a failed
// allocation, or a container asked to outgrow its maximum size) or an
// internal error. No exception leaves an entry point (the reference's rule,
// src/api/api.cpp:28-38).
fp_status status_of_exception() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        return FP_ENOMEM;
    } catch (const std::length_error &) {
        return FP_ENOMEM;
    } catch (...) {
        return FP_EINTERNAL;
    }
}

template <class F>
fp_status guard(F &&f) noexcept {
    try {
        return f();
    } catch (...) {
        return status_of_exception();
    }
}

fp_status hip_status(int e) {
    if (e == 0) return FP_OK;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return FP_ENOMEM;
    if (e == -1) return FP_EINTERNAL;
    return FP_EDEVICE;
}

// The int a C caller stored in an enum field or argument, read from its
// bytes: a C enum may hold any int, while a C++ read of an enum object out of
// its enumeration's range is undefined.
template <class E>
int enum_bits(const E &e) noexcept {
    static_assert(sizeof(E) == sizeof(int32_t), "the C enums of the ABI are 32-bit");
    const unsigned char *b = reinterpret_cast<const unsigned char *>(&e);
    uint32_t v = 0;
    for (int i = (int)sizeof(E) - 1; i >= 0; --i) v = (v << 8) | b[i];  // little endian
    return (int)v;
}


struct Arr {
    char *data = nullptr;
    int dtype = FP_F64;
    int ndim = 0;
    int device = FP_DEVICE_CPU;
    int64_t shape[FP_MAX_NDIM] = {};
    int64_t strides[FP_MAX_NDIM] = {};
    int64_t numel() const {
        int64_t n = 1;
        for (int d = 0; d < ndim; ++d) n *= shape[d];
        return n;
    }
    size_t itemsize() const { return dtype == FP_F32 || dtype == FP_I32 ? 4 : (dtype == FP_I8 ? 1 : 8); }
    bool contiguous() const {
        int64_t s = 1;
        for (int d = ndim - 1; d >= 0; --d) {
            if (shape[d] != 1 && strides[d] != s) return false;
            s *= shape[d];
        }
        return true;
    }
    // [lo, hi] byte range of the entries (numel > 0).
    void span(intptr_t &lo, intptr_t &hi) const {
        int64_t a = 0, b = 0;
        for (int d = 0; d < ndim; ++d) {
            if (shape[d] < 2) continue;
            const int64_t e = (shape[d] - 1) * strides[d];
            if (e < 0) a += e;
            else b += e;
        }
        lo = (intptr_t)data + (intptr_t)(a * (int64_t)itemsize());
        hi = (intptr_t)data + (intptr_t)((b + 1) * (int64_t)itemsize()) - 1;
    }
};

bool dtype_allowed(int dt, bool floating) {
    return floating ? (dt == FP_F32 || dt == FP_F64) : (dt == FP_I8 || dt == FP_I32 || dt == FP_I64);
}

// An fp_array descriptor checked against the boundary's rules and converted
// (the rules of the reference's descriptor check, src/api/api.cpp:60-108,
// with FP_DEVICE_ROCM accepted besides the host):
//   the device is the host or a ROCm device (else FP_EINVAL); the dtype is a
//   float one (floating) or an integer one (FP_EDTYPE); 0 <= ndim <=
//   FP_MAX_NDIM (FP_ESHAPE), a 0-d array being one entry of shape [1]; shape
//   is given when ndim > 0 (FP_EINVAL); no extent is negative, the product of
//   the extents above 1 and the offset span sum_d (L_d - 1) |s_d| stay within
//   kMaxElems, as does every stride along an extent above 1 (FP_ESHAPE);
//   NULL strides mean C order; data is given when there are entries
//   (FP_EINVAL).
// The bounds are checked in 128-bit arithmetic, one axis at a time from the
// last, so that no product can wrap before it is compared.
fp_status to_arr(const fp_array *a, Arr &out, bool floating = true) noexcept {
    if (!a) return FP_EINVAL;
    const int dev = enum_bits(a->device);
    if (dev != FP_DEVICE_CPU && dev != FP_DEVICE_ROCM) return FP_EINVAL;
    const int dt = enum_bits(a->dtype);
    if (!dtype_allowed(dt, floating)) return FP_EDTYPE;
    if (a->ndim < 0 || a->ndim > FP_MAX_NDIM) return FP_ESHAPE;
    if (a->ndim > 0 && !a->shape) return FP_EINVAL;
    Arr r;
    r.data = static_cast<char *>(a->data);
    r.dtype = dt;
    r.device = dev;
    if (a->ndim == 0) {
        r.ndim = 1;
        r.shape[0] = 1;
        r.strides[0] = 1;
    } else {
        r.ndim = a->ndim;
        const __int128 cap = kMaxElems;
        __int128 entries = 1, span = 0, c_stride = 1;
        for (int d = a->ndim - 1; d >= 0; --d) {
            const int64_t n = a->shape[d];
            if (n < 0) return FP_ESHAPE;
            const int64_t st = a->strides ? a->strides[d] : (int64_t)c_stride;
            r.shape[d] = n;
            r.strides[d] = st;
            if (n > 1) {
                const __int128 mag = st < 0 ? -(__int128)st : (__int128)st;
                entries *= n;
                span += (__int128)(n - 1) * mag;
                if (entries > cap || mag > cap || span > cap) return FP_ESHAPE;
            }
            c_stride *= n;  // at most the entries so far, or 0
        }
    }
    if (!r.data && r.numel() > 0) return FP_EINVAL;
    out = r;
    return FP_OK;
}

// Copies any float array (host or device, any strides) into host doubles in C order.
fp_status read_doubles(const Arr &x, std::vector<double> &v) {
    const int64_t n = x.numel();
    v.resize((size_t)n);
    if (n == 0) return FP_OK;
    intptr_t lo, hi;
    x.span(lo, hi);
    const size_t bytes = (size_t)(hi - lo + 1);
    std::vector<char> host;
    const char *base = nullptr;
    if (x.device == FP_DEVICE_ROCM) {
        host.resize(bytes);
        if (hipMemcpy(host.data(), (const void *)lo, bytes, hipMemcpyDeviceToHost)) return FP_EDEVICE;
        base = host.data() + ((intptr_t)x.data - lo);
    } else {
        base = x.data;
    }
    const int64_t item = (int64_t)x.itemsize();
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, off = 0;
        for (int d = x.ndim - 1; d >= 0; --d) {
            const int64_t c = r % x.shape[d];
            r /= x.shape[d];
            off += c * x.strides[d];
        }
        const char *p = base + off * item;
        v[(size_t)i] = x.dtype == FP_F32 ? (double)*(const float *)p : *(const double *)p;
    }
    return FP_OK;
}

fp_status read_vec(const fp_array *a, std::vector<double> &v) {
    v.clear();
    if (!a) return FP_OK;
    Arr x;
    if (fp_status s = to_arr(a, x)) return s;
    if (a->ndim > 1) return FP_ESHAPE;
    if (x.shape[0] == 0) return FP_ESHAPE;
    return read_doubles(x, v);
}

bool finite_nonneg(double v) { return std::isfinite(v) && v >= 0.0; }

fp_status make_tv(bool iso, int ndim, const fp_array *axis_weights, const fp_array *spatial, fp_op **out) {
    if (!out) return FP_EINVAL;
    *out = nullptr;
    if (ndim != 2 && ndim != 3) return FP_EPARAM;
    std::vector<double> w;
    if (fp_status s = read_vec(axis_weights, w)) return s;
    double ww[3] = {1.0, 1.0, 1.0};
    if (w.size() == 1) {
        for (int a = 0; a < ndim; ++a) ww[a] = w[0];
    } else if (!w.empty()) {
        if ((int)w.size() != ndim) return FP_ESHAPE;
        for (int a = 0; a < ndim; ++a) ww[a] = w[(size_t)a];
    }
    for (int a = 0; a < ndim; ++a)
        if (!finite_nonneg(ww[a])) return FP_EPARAM;
    std::unique_ptr<fp_op> op(new fp_op());
    op->iso = iso;
    op->k = ndim;
    for (int a = 0; a < 3; ++a) op->w[a] = a < ndim ? ww[a] : 0.0;
    if (spatial) {
        Arr s;
        if (fp_status st = to_arr(spatial, s)) return st;
        const int want = iso ? ndim : ndim + 1;
        if (spatial->ndim != want) return FP_ESHAPE;
        if (!iso && s.shape[0] != ndim) return FP_ESHAPE;
        for (int a = 0; a < ndim; ++a) op->sshape[a] = s.shape[iso ? a : a + 1];
        double n = 1.0;
        for (int a = 0; a < ndim; ++a) n *= (double)op->sshape[a];
        if (!(n < 2147483647.0)) return FP_ESHAPE;
        if (fp_status st = read_doubles(s, op->spatial)) return st;
        if (op->spatial.empty()) return FP_ESHAPE;
        const int64_t N = (int64_t)n;
        double mx = 0.0;
        for (size_t i = 0; i < op->spatial.size(); ++i) {
            double &b = op->spatial[i];
            if (!finite_nonneg(b)) return FP_EPARAM;
            if (!iso) {  // the weight at the last entry along an axis belongs to no edge (tv_aniso.cpp:1447-1451)
                const int a = (int)((int64_t)i / N);
                const int64_t p = (int64_t)i % N;
                int64_t stride = 1;
                for (int c = ndim - 1; c > a; --c) stride *= op->sshape[c];
                if ((p / stride) % op->sshape[a] == op->sshape[a] - 1) b = 0.0;
            }
            mx = std::max(mx, b);
        }
        op->spatial_max = mx;
    }
    *out = op.release();
    return FP_OK;
}

}  // namespace


namespace {

fp_status check_output(const Arr *in, const Arr &out) {
    if (out.numel() == 0) return FP_OK;
    for (int d = 0; d < out.ndim; ++d)
        if (out.shape[d] > 1 && out.strides[d] == 0) return FP_EINVAL;
    if (!in || in->numel() == 0 || in->device != out.device) return FP_OK;
    bool same = in->data == out.data && in->ndim == out.ndim;
    for (int d = 0; same && d < in->ndim; ++d)
        same = in->shape[d] == out.shape[d] && (in->shape[d] < 2 || in->strides[d] == out.strides[d]);
    if (same) return FP_OK;
    intptr_t a0, a1, b0, b1;
    in->span(a0, a1);
    out.span(b0, b1);
    return (a1 < b0 || b1 < a0) ? FP_OK : FP_EINVAL;
}

bool same_shape(const Arr &a, const Arr &b) {
    if (a.ndim != b.ndim) return false;
    for (int d = 0; d < a.ndim; ++d)
        if (a.shape[d] != b.shape[d]) return false;
    return true;
}

fp_options defaults() {
    fp_options o;
    fp_options_init(&o);
    return o;
}

fp_status check_image(const fp_op &op, const int64_t *ishape) {
    double n = 1.0;
    for (int a = 0; a < op.k; ++a) n *= (double)ishape[a];
    if (!op.spatial.empty())
        for (int a = 0; a < op.k; ++a)
            if (ishape[a] != op.sshape[a]) return FP_ESHAPE;
    return n < 2147483647.0 ? FP_OK : FP_ESHAPE;
}

// The step range: gamma times the operator's largest weight (axis weight
// times spatial weight) at most 2^960, so that every rescaled radius is
// finite (the reference's range, op.hpp:48). Signs and finiteness are checked
// before, with the other arguments.
fp_status step_ok(const fp_op &op, double g) {
    const double largest = op.step_scale();
    const bool in_range = largest == 0.0 || g * largest <= kMaxScaledStep;
    return in_range ? FP_OK : FP_EPARAM;
}

fpr::StridedShape strided_of(const Arr &a) {
    fpr::StridedShape sh{};
    for (int d = 0; d < a.ndim; ++d) {
        sh.shape[d] = a.shape[d];
        sh.stride[d] = a.strides[d];
    }
    return sh;
}

// Host array -> contiguous host bytes of the array's dtype, C order.
void pack_host(const Arr &a, char *dst) {
    const int64_t n = a.numel(), item = (int64_t)a.itemsize();
    if (a.contiguous()) {
        std::memcpy(dst, a.data, (size_t)(n * item));
        return;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, off = 0;
        for (int d = a.ndim - 1; d >= 0; --d) {
            const int64_t c = r % a.shape[d];
            r /= a.shape[d];
            off += c * a.strides[d];
        }
        std::memcpy(dst + i * item, a.data + off * item, (size_t)item);
    }
}

void unpack_host(const char *src, const Arr &a) {
    const int64_t n = a.numel(), item = (int64_t)a.itemsize();
    if (a.contiguous()) {
        std::memcpy(a.data, src, (size_t)(n * item));
        return;
    }
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, off = 0;
        for (int d = a.ndim - 1; d >= 0; --d) {
            const int64_t c = r % a.shape[d];
            r /= a.shape[d];
            off += c * a.strides[d];
        }
        std::memcpy(a.data + off * item, src + i * item, (size_t)item);
    }
}

// Copies the array a (host or device, any strides, the workspace's dtype) into
// the contiguous device buffer dst.
fp_status stage_in(fp_workspace &ws, const Arr &a, void *dst, hipStream_t s) {
    const int64_t n = a.numel();
    const size_t bytes = (size_t)n * ws.tsize;
    if (a.device == FP_DEVICE_CPU) {
        ws.host_stage.resize(bytes);
        pack_host(a, ws.host_stage.data());
        if (hipMemcpyAsync(dst, ws.host_stage.data(), bytes, hipMemcpyHostToDevice, s)) return FP_EDEVICE;
        return hipStreamSynchronize(s) ? FP_EDEVICE : FP_OK;
    }
    if (a.contiguous()) return hipMemcpyAsync(dst, a.data, bytes, hipMemcpyDeviceToDevice, s) ? FP_EDEVICE : FP_OK;
    if (a.ndim > fpr::kMaxStrided) return FP_ESHAPE;
    const fpr::StridedShape sh = strided_of(a);
    if (ws.tsize == 4) fpr::launch_gather<float>((const float *)a.data, (float *)dst, n, a.ndim, sh, s);
    else fpr::launch_gather<double>((const double *)a.data, (double *)dst, n, a.ndim, sh, s);
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

fp_status stage_out(fp_workspace &ws, const void *src, const Arr &a, hipStream_t s) {
    const int64_t n = a.numel();
    const size_t bytes = (size_t)n * ws.tsize;
    if (a.device == FP_DEVICE_CPU) {
        ws.host_stage.resize(bytes);
        if (hipMemcpyAsync(ws.host_stage.data(), src, bytes, hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
        unpack_host(ws.host_stage.data(), a);
        return FP_OK;
    }
    if (a.contiguous()) return hipMemcpyAsync(a.data, src, bytes, hipMemcpyDeviceToDevice, s) ? FP_EDEVICE : FP_OK;
    if (a.ndim > fpr::kMaxStrided) return FP_ESHAPE;
    const fpr::StridedShape sh = strided_of(a);
    if (ws.tsize == 4) fpr::launch_scatter<float>((const float *)src, (float *)a.data, n, a.ndim, sh, s);
    else fpr::launch_scatter<double>((const double *)src, (double *)a.data, n, a.ndim, sh, s);
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

fp_status build_engine(fp_workspace &w) {
    const fp_op &op = *w.op;
    fpr::Problem P{};
    P.k = op.k;
    P.iso = op.iso ? 1 : 0;
    P.sw = op.spatial.empty() ? 0 : 1;
    P.sw_per_image = 0;
    for (int a = 0; a < 3; ++a) P.w[a] = op.w[a];
    P.sched = op.iso ? fpr::kSchedC : fpr::kSchedDevice;
    P.wmax_norm = op.iso ? 1 : 0;
    P.alpha_max = op.spatial.empty() ? 1.0 : op.spatial_max;
    std::vector<fpr::ImageSpec> specs((size_t)w.B);
    for (int64_t b = 0; b < w.B; ++b) {
        fpr::ImageSpec &s = specs[(size_t)b];
        for (int a = 0; a < 3; ++a) s.L[a] = w.img[a];
        s.q_cstride = w.B * w.N;
        s.sw_cstride = w.N;
    }
    w.ptrs.assign((size_t)w.B, fpr::ImagePtrs{});
    w.eng.reset(new fpr::Engine());
    if (int e = w.eng->setup((int)w.tsize, P, w.kc, specs)) return hip_status(e);
    double frac = 0.0;
    const fpr::KernelChoice tk = fpr::default_tail(op.k, (int)w.tsize, w.img, op.iso, !op.spatial.empty(), w.B, &frac);
    if (frac > 0.0 && w.kc.stream && !w.kc.generic && tk.iters() >= w.kc.iters())
        return hip_status(w.eng->set_tail(tk, frac));
    return FP_OK;
}

// The ADBPG solver of an anisotropic workspace (its buffers for ws's shape).
fp_status build_ad(fp_workspace &w) {
    const fp_op &op = *w.op;
    fpr::AdSpec sp;
    sp.k = op.k;
    for (int a = 0; a < 3; ++a) sp.L[a] = w.img[a];
    sp.B = w.B;
    sp.tsize = (int)w.tsize;
    for (int a = 0; a < 3; ++a) sp.w[a] = op.w[a];
    fpr::AdSolver *a = nullptr;
    if (int r = fpr::ad_reserve(sp, &a)) return (fp_status)r;
    w.ad.reset(a);
    w.ptrs.assign((size_t)w.B, fpr::ImagePtrs{});
    return FP_OK;
}

// The default anisotropic method: workspaces (fp_prox, the reference CPU
// library's call) take the reference's method (ADBPG), batches (the device
// path's batched solve) dual FISTA; FASTPROX_ROCM_ANISO=adbpg or fista sets
// both. Edge-weighted operators take FISTA.
int default_method(const fp_op &op, bool batch) {
    if (op.iso || !op.spatial.empty()) return FPR_METHOD_FISTA;
    if (const char *e = std::getenv("FASTPROX_ROCM_ANISO")) {
        if (std::strcmp(e, "fista") == 0) return FPR_METHOD_FISTA;
        if (std::strcmp(e, "adbpg") == 0) return FPR_METHOD_ADBPG;
    }
    return batch ? FPR_METHOD_FISTA : FPR_METHOD_ADBPG;
}

// kind 0: the streaming kernel (tt iterations, segments of ry rows); 1: the
// generic one-iteration kernel; 2: the block kernel (tt, ry).
fpr::KernelChoice generic_kernel() {
    fpr::KernelChoice kc;
    kc.generic = true;
    kc.stream = false;
    kc.TT = 1;
    kc.RY = 4;
    return kc;
}

fp_status kernel_choice(int kind, int tt, int ry, int k, int tsize, bool sw, fpr::KernelChoice &kc) {
    kc = fpr::KernelChoice{};
    if (kind == 1) {
        kc.generic = true;
        kc.stream = false;
        kc.TT = 1;
        kc.RY = 4;
        return FP_OK;
    }
    if (kind == 3) {
        kc.generic = false;
        kc.stream = false;
        kc.s3 = true;
        kc.TT = tt < 1 ? 1 : tt;
        kc.RY = ry;
        if (k != 3) return fail(FP_EINVAL, "the 3D streaming kernel is 3D only");
        if (!fpr::kernel_available_s3(ry, kc.TT, tsize, sw))
            return fail(FP_EINVAL, "3D streaming kernel: slices per tile out of range, or several iterations per "
                                   "launch asked of float64 or pixel-weighted volumes");
        return FP_OK;
    }
    if (kind != 0 && kind != 2)
        return fail(FP_EINVAL, "kernel kind must be 0 (stream), 1 (generic), 2 (block) or 3 (3D stream)");
    kc.stream = kind == 0;
    kc.TT = tt;
    kc.RY = ry;
    if (k != 2) return fail(FP_EINVAL, "the tiled kernels are 2D only");
    if (!fpr::kernel_available(tsize, kc.stream, tt, ry)) return fail(FP_EINVAL, "no such kernel built");
    return FP_OK;
}

fp_status alloc(void **p, size_t bytes) {
    if (fpr::dev_alloc(p, std::max<size_t>(bytes, 16))) return FP_ENOMEM;
    return FP_OK;
}

// keep_structure: every buffer the derivatives and the structure queries of a
// kept call use, so that fp_prox, fp_prox_steps, fp_jvp, fp_vjp, the
// queries and fp_workspace_set_branch allocate no device memory: the saved
// point, the derivative structure's buffers (fpr::deriv_reserve), the small
// per-image arrays of the derivatives, the queries' output and the branch
// overrides (the reference sizes these here too, api.cpp:265-289).
fp_status reserve_kept(fp_workspace &w) {
    const fp_op *op = w.op;
    const size_t bytes = (size_t)w.numel * w.tsize;
    if (fp_status s = alloc(&w.xsave, bytes)) return s;
    if (fp_status s = alloc(&w.zsave, bytes)) return s;
    if (!op->iso)
        if (fp_status s = alloc(&w.usave, bytes * (size_t)op->k)) return s;
    w.exps.assign((size_t)w.B, 0);
    fpr::DerivSpec sp;
    sp.iso = op->iso;
    sp.k = op->k;
    for (int a = 0; a < 3; ++a) sp.L[a] = w.img[a];
    sp.B = w.B;
    sp.tsize = (int)w.tsize;
    for (int a = 0; a < op->k; ++a) sp.w[a] = op->w[a];
    hipError_t err;
    sp.sw = op->device_weights(w.device, (int)w.tsize, err);
    if (err) return hip_status(err);
    sp.sw_img = 0;
    fpr::Deriv *d = nullptr;
    if (int r = fpr::deriv_reserve(sp, &d)) return (fp_status)r;
    w.deriv.reset(d);
    w.dscr_n = (size_t)(w.B + op->k + (int64_t)op->spatial.size());
    if (fp_status s = alloc((void **)&w.dscr, std::max<size_t>(w.dscr_n, 16) * sizeof(double))) return s;
    if (fp_status s = alloc(&w.qbuf, 9 * (size_t)w.numel)) return s;
    if (!op->iso)
        if (fp_status s = alloc((void **)&w.branch, (size_t)w.numel)) return s;
    return FP_OK;
}

fp_status create_impl(const fp_op *op, const fp_array *like, const fp_options *opt, fp_workspace **out) {
    if (!out) return FP_EINVAL;
    *out = nullptr;
    if (!op || !like) return FP_EINVAL;
    Arr x;
    if (fp_status s = to_arr(like, x)) return s;
    const fp_options o = opt ? *opt : defaults();
    if (x.ndim < op->k) return FP_ESHAPE;
    std::unique_ptr<fp_workspace> w(new fp_workspace());
    w->op = op;
    w->dtype = x.dtype;
    w->ndim = x.ndim;
    for (int d = 0; d < x.ndim; ++d) w->shape[d] = x.shape[d];
    w->numel = x.numel();
    w->k = op->k;
    w->tsize = x.itemsize();
    w->keep = o.keep_structure != 0;
    int64_t N = 1;
    for (int a = 0; a < op->k; ++a) {
        w->img[3 - op->k + a] = x.shape[x.ndim - op->k + a];
        N *= x.shape[x.ndim - op->k + a];
    }
    w->N = N;
    w->B = N > 0 ? w->numel / N : 0;
    if (hipGetDevice(&w->device)) return FP_EDEVICE;
    const int64_t B = w->B;
    w->gaps.assign((size_t)B, 0.0);
    w->zerrs.assign((size_t)B, 0.0);
    w->smooth.assign((size_t)B, 0.0);
    w->iters.assign((size_t)B, 0);
    w->exact.assign((size_t)B, 1);
    w->gammas.assign((size_t)B, 0.0);
    if (w->numel == 0) {
        *out = w.release();
        return FP_OK;
    }
    if (fp_status s = check_image(*op, x.shape + (x.ndim - op->k))) return s;
    w->kc = fpr::default_kernel(op->k, (int)w->tsize, w->img, op->iso, !op->spatial.empty(), B);
    if (w->kc.stream && !fpr::stream_fits((int)w->tsize, w->N)) w->kc = generic_kernel();
    if (w->kc.s3t() && !fpr::stream_fits((int)w->tsize, w->N)) w->kc.TT = 1;  // the one-iteration 3D kernel
    w->method = default_method(*op, false);
    if (w->method == FPR_METHOD_ADBPG) {
        if (fp_status s = build_ad(*w)) return s;
    } else if (fp_status s = build_engine(*w)) {
        return s;
    }
    const size_t bytes = (size_t)w->numel * w->tsize;
    if (fp_status s = alloc(&w->dstage, bytes)) return s;
    if (fp_status s = alloc(&w->zstage, bytes)) return s;
    if (fp_status s = alloc(&w->q, bytes * (size_t)op->k)) return s;
    hipError_t err;
    (void)op->device_weights(w->device, (int)w->tsize, err);
    if (err) return hip_status(err);
    if (w->keep)
        if (fp_status s = reserve_kept(*w)) return s;
    // host arrays are packed through host_stage: sized here for x's dtype
    w->host_stage.reserve(bytes * (size_t)op->k);
    *out = w.release();
    return FP_OK;
}

struct DevGuard {  // the workspace's device current for the call
    int dev, was;
    explicit DevGuard(int d) : dev(d), was(d) {
        (void)hipGetDevice(&was);
        if (was != dev) (void)hipSetDevice(dev);
    }
    ~DevGuard() {
        if (was != dev) (void)hipSetDevice(was);
    }
};

// The per-image results of the engine's last solve into ws (driver.hpp:
// 1216-1236, 1331-1341): gaps at the input's scale, iterations, certified
// errors, fp_info; FP_EINVAL for an infinite entry.
fp_status record_results(const fp_op *op, fp_workspace *ws, fp_info *info, fp_status *status) {
    bool bad = false, notconv = false;
    double gsum = 0.0;
    int itmax = 0;
    if (ws->ad) {
        const fpr::AdResult *R = fpr::ad_results(ws->ad.get());
        for (int64_t b = 0; b < ws->B; ++b) {
            const fpr::AdResult &t = R[b];
            if (t.status) bad = true;
            ws->gaps[(size_t)b] = t.gap;
            ws->iters[(size_t)b] = t.iters;
            ws->zerrs[(size_t)b] = t.zerr;
            if (!t.converged) notconv = true;
            gsum += t.gap;
            itmax = std::max(itmax, t.iters);
        }
    }
    const fpr::ImgState *S = ws->ad ? nullptr : ws->eng->state();
    for (int64_t b = 0; b < ws->B && S; ++b) {
        const fpr::ImgState &t = S[b];
        if (t.status) bad = true;
        double g = t.gp;
        if (op->iso && g == g) g = g < 0.0 ? 0.0 : g;
        if (op->iso && t.e) g = std::ldexp(g, 2 * t.e);
        ws->gaps[(size_t)b] = g;
        ws->iters[(size_t)b] = t.it;
        ws->zerrs[(size_t)b] = t.zerr;
        if (!t.converged) notconv = true;
        gsum += g;
        itmax = std::max(itmax, t.it);
    }
    if (bad) {
        ws->dual_valid = false;
        return FP_EINVAL;
    }
    ws->dual_valid = true;
    *status = notconv ? FP_ENOTCONV : FP_OK;
    ws->info = fp_info{gsum, itmax, 0};
    ws->info_status = *status;
    ws->info_valid = true;
    if (info) *info = ws->info;
    return FP_OK;
}

// The kept point's extras after xsave and zsave hold it: the per-edge dual
// (aniso), the scale exponents and the smoothing of the derivative system
// (tv_iso.cpp:566-568: mu = 0.1 w_max sqrt(2 gap / N) at the solved scale when
// the structure is read off the dual, two or more active axes; back at the
// input's scale). Waits for the stream.
fp_status record_kept(const fp_op *op, fp_workspace *ws, hipStream_t s) {
    const int k = op->k;
    // ADBPG wrote the per-edge dual during the solve
    if (!op->iso && !ws->ad && ws->eng->copy_dual_iterate(s, ws->usave, (int64_t)k * ws->N)) return FP_EDEVICE;
    if (hipStreamSynchronize(s)) return FP_EDEVICE;
    const fpr::ImgState *S = ws->ad ? nullptr : ws->eng->state();
    const fpr::AdResult *R = ws->ad ? fpr::ad_results(ws->ad.get()) : nullptr;
    for (int64_t b = 0; b < ws->B; ++b) {
        const int te = R ? R[b].e : S[b].e;
        ws->exps[(size_t)b] = te;
        double mu = 0.0;
        if (op->iso) {
            double wmax = 0.0, wL = 0.0;
            int na = 0;
            for (int a = 0; a < k; ++a)
                if (ws->img[3 - k + a] > 1) wL = std::max(wL, op->w[a]);
            for (int a = 0; a < k; ++a)
                if (ws->gammas[(size_t)b] > 0.0 && op->w[a] > 0.0 && ws->img[3 - k + a] > 1 &&
                    op->w[a] >= 1e-100 * wL) {
                    wmax = std::max(wmax, op->w[a]);
                    ++na;
                }
            const double gs = std::ldexp(ws->gaps[(size_t)b], -2 * te);
            if (na >= 2 && gs == gs) mu = 0.1 * wmax * std::sqrt(2.0 * std::max(gs, 0.0) / (double)ws->N);
        }
        ws->smooth[(size_t)b] = std::ldexp(mu, te);
        ws->exact[(size_t)b] = !(mu > 0.0);
    }
    ws->valid = true;
    return FP_OK;
}

// Whether stream s is being captured into a hipGraph.
bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive;
}

// The results of a solve enqueued under capture, once it has run: waits for
// the workspace's stream and records them.
fp_status settle(fp_workspace *ws) {
    if (!ws->pending) return FP_OK;
    DevGuard g(ws->device);
    if (hipStreamSynchronize((hipStream_t)ws->stream)) return FP_EDEVICE;
    ws->eng->collect();
    ws->pending = false;
    fp_status st = FP_OK;
    return record_results(ws->op, ws, nullptr, &st);
}

// One solve of ws's images from input d into z (device, contiguous), by the
// workspace's method; ws->ptrs hold the engine's per-image pointers. keep:
// the ADBPG solve also writes the per-edge dual the derivatives read.
fp_status run_solve(fp_workspace *ws, hipStream_t s, const void *d, void *z, bool warm, double tol, int max_iter,
                    bool keep) {
    if (ws->ad) {
        std::vector<const void *> xs((size_t)ws->B);
        std::vector<void *> zs((size_t)ws->B), qs((size_t)ws->B);
        for (int64_t b = 0; b < ws->B; ++b) {
            const size_t off = (size_t)(b * ws->N) * ws->tsize;
            xs[(size_t)b] = (const char *)d + off;
            zs[(size_t)b] = (char *)z + off;
            qs[(size_t)b] = (char *)ws->q + off;
        }
        fpr::AdCall c;
        c.xs = xs.data();
        c.zs = zs.data();
        c.qs = qs.data();
        c.q_plane = ws->B * ws->N;
        c.usave = keep ? ws->usave : nullptr;
        c.u_img = (int64_t)ws->k * ws->N;
        c.gammas = ws->gammas.data();
        c.warm = warm;
        c.tol = tol;
        c.max_iter = max_iter;
        return (fp_status)fpr::ad_solve(ws->ad.get(), s, c);
    }
    return hip_status(ws->eng->solve(s, ws->ptrs.data(), ws->gammas.data(), warm, tol, max_iter));
}

fp_status prox_impl(const fp_op *op, double step, const fp_array *xa, fp_array *oa, const fp_options *opt,
                    fp_workspace *ws, fp_info *info, const fp_array *sa) {
    if (!op || !xa || !oa) return FP_EINVAL;
    if (sa && ws && capturing((hipStream_t)ws->stream))
        return fail(FP_EINVAL, "fp_prox_steps reads its steps on the host: not under capture");
    if (!sa && (!std::isfinite(step) || step < 0)) return FP_EPARAM;
    const fp_options o = opt ? *opt : defaults();
    if (o.tie != 0 && o.tie != 1) return FP_EINVAL;
    Arr x, out;
    if (fp_status s = to_arr(xa, x)) return s;
    if (fp_status s = to_arr(oa, out)) return s;
    if (x.dtype != out.dtype) return FP_EDTYPE;
    if (!same_shape(x, out)) return FP_ESHAPE;
    if (fp_status s = check_output(&x, out)) return s;
    const int k = op->k;
    const bool keep = o.keep_structure != 0;
    if (keep && !ws) return FP_EINVAL;
    if (x.ndim < k) return FP_ESHAPE;
    std::vector<double> steps;
    Arr st;
    if (sa) {
        if (fp_status s = to_arr(sa, st)) return s;
        const int nb = x.ndim - k;
        if (sa->ndim != nb) return FP_ESHAPE;
        for (int d = 0; d < nb; ++d)
            if (st.shape[d] != x.shape[d]) return FP_ESHAPE;
        if (fp_status s = check_output(&st, out)) return s;
    }
    if (ws) {
        if (ws->op != op || ws->dtype != x.dtype || ws->ndim != x.ndim) return FP_EINVAL;
        for (int d = 0; d < x.ndim; ++d)
            if (ws->shape[d] != x.shape[d]) return FP_ESHAPE;
        if (keep && !ws->keep) return FP_EINVAL;
    }
    if (info) *info = fp_info{0.0, 0, 0};
    if (!(o.tol >= 0.0) || !std::isfinite(o.tol) || o.max_iter < 1) return FP_EPARAM;
    if (sa) {
        int64_t count = 1;
        for (int d = 0; d < sa->ndim; ++d) count *= st.shape[d];
        if (count > 0) {
            if (fp_status s = read_doubles(st, steps)) return s;
            for (double v : steps) {
                if (!std::isfinite(v) || v < 0) return FP_EPARAM;
                if (fp_status s = step_ok(*op, v)) return s;
            }
        }
    } else if (fp_status s = step_ok(*op, step)) {
        return s;
    }
    if (x.numel() != 0)
        if (fp_status s = check_image(*op, x.shape + (x.ndim - k))) return s;
    // Every argument is checked: from here the call replaces the saved point.
    std::unique_ptr<fp_workspace> local;
    if (!ws) {
        fp_workspace *w = nullptr;
        fp_options lo = o;
        lo.keep_structure = 0;
        if (fp_status s = create_impl(op, xa, &lo, &w)) return s;
        local.reset(w);
        ws = w;
    }
    ws->valid = false;
    ws->steps_on = false;
    ws->info_valid = false;
    ws->deriv_built = false;
    ws->branch_on = false;
    if (x.numel() == 0) {
        ws->dual_valid = true;
        ws->info = fp_info{0.0, 0, 0};
        ws->info_status = FP_OK;
        ws->info_valid = true;
        if (keep) {
            ws->valid = true;
            ws->gamma = step;
        }
        return FP_OK;
    }
    int cur_dev = 0;
    (void)hipGetDevice(&cur_dev);
    if (cur_dev != ws->device && hipSetDevice(ws->device)) return FP_EDEVICE;
    struct Restore {
        int dev, was;
        ~Restore() {
            if (dev != was) (void)hipSetDevice(was);
        }
    } restore{ws->device, cur_dev};
    hipStream_t s = (hipStream_t)ws->stream;
    const int64_t B = ws->B;
    for (int64_t b = 0; b < B; ++b) ws->gammas[(size_t)b] = sa ? steps[(size_t)b] : step;
    // Inputs: the caller's device array itself when contiguous and not
    // written by this call, else staged.
    const bool x_direct = x.device == FP_DEVICE_ROCM && x.contiguous() && x.data != out.data;
    const bool z_direct = out.device == FP_DEVICE_ROCM && out.contiguous();
    // On a capturing stream the solve is the static schedule (no host read
    // or wait): device arrays only, no host results, no kept point.
    ws->pending = false;
    const bool cap = capturing(s);
    if (cap) {
        if (ws->ad) return fail(FP_EINVAL, "ADBPG solves read the host between steps: not under capture");
        if (!x_direct || !z_direct) return fail(FP_EINVAL, "under capture x and out are contiguous device arrays");
        if (keep || info || sa) return fail(FP_EINVAL, "under capture: no host results, kept points or steps");
    }
    const void *d = x.data;
    if (!x_direct) {
        if (fp_status st2 = stage_in(*ws, x, ws->dstage, s)) return st2;
        d = ws->dstage;
    }
    void *z = z_direct ? (void *)out.data : ws->zstage;
    hipError_t err;
    const void *sw = op->device_weights(ws->device, (int)ws->tsize, err);
    if (err) return hip_status(err);
    const bool warm = ws->dual_valid && (ws->warm || ws->seeded);
    for (int64_t b = 0; b < B; ++b) {
        fpr::ImagePtrs &p = ws->ptrs[(size_t)b];
        p.d = (const char *)d + (size_t)(b * ws->N) * ws->tsize;
        p.z = (char *)z + (size_t)(b * ws->N) * ws->tsize;
        p.q = (char *)ws->q + (size_t)(b * ws->N) * ws->tsize;
        p.sw = sw;
    }
    ws->seeded = false;
    if (fp_status r = run_solve(ws, s, d, z, warm, o.tol, o.max_iter, keep)) {
        ws->dual_valid = false;
        return r;
    }
    if (cap) {  // the results are read at the first query after the graph has run
        ws->pending = true;
        ws->dual_valid = true;
        return FP_OK;
    }
    if (!z_direct)
        if (fp_status st2 = stage_out(*ws, z, out, s)) return st2;
    fp_status status = FP_OK;
    if (fp_status st2 = record_results(op, ws, info, &status)) return st2;
    if (keep) {
        const size_t bytes = (size_t)ws->numel * ws->tsize;
        if (hipMemcpyAsync(ws->xsave, d, bytes, hipMemcpyDeviceToDevice, s)) return FP_EDEVICE;
        if (hipMemcpyAsync(ws->zsave, z, bytes, hipMemcpyDeviceToDevice, s)) return FP_EDEVICE;
        if (fp_status st2 = record_kept(op, ws, s)) return st2;
        ws->steps_on = sa != nullptr;
        ws->gamma = step;
    }
    if (!z_direct || !x_direct)
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
    return status;
}

// fpr_prox_static: fp_prox without a workspace (the arguments checked as
// prox_impl checks them), on contiguous device arrays, with its device
// scratch in the fixed array of static_solve.cpp. The kernel is fixed (2D:
// the streaming kernel, 6 iterations per launch, 16-row segments; 3D: the
// streaming kernel with 16-slice slabs; the generic kernel for planes too
// large for the streaming kernel's offsets), so that nothing is read from
// the tuning table. Nothing here allocates.
fp_status static_prox_impl(const fp_op *op, double step, const fp_array *xa, fp_array *oa, const fp_options *opt,
                           void *stream, fp_info *info) {
    if (!op || !xa || !oa) return FP_EINVAL;
    if (!std::isfinite(step) || step < 0) return FP_EPARAM;
    const fp_options o = opt ? *opt : defaults();
    if (o.tie != 0 && o.tie != 1) return FP_EINVAL;
    Arr x, out;
    if (fp_status s = to_arr(xa, x)) return s;
    if (fp_status s = to_arr(oa, out)) return s;
    if (x.dtype != out.dtype) return FP_EDTYPE;
    if (!same_shape(x, out)) return FP_ESHAPE;
    if (fp_status s = check_output(&x, out)) return s;
    const int k = op->k;
    if (x.ndim < k) return FP_ESHAPE;
    if (info) *info = fp_info{0.0, 0, 0};
    if (!(o.tol >= 0.0) || !std::isfinite(o.tol) || o.max_iter < 1) return FP_EPARAM;
    if (fp_status s = step_ok(*op, step)) return s;
    if (x.numel() != 0)
        if (fp_status s = check_image(*op, x.shape + (x.ndim - k))) return s;
    if (o.keep_structure) return fail(FP_EINVAL, "fpr_prox_static keeps no structure (it has no workspace)");
    if (x.device != FP_DEVICE_ROCM || out.device != FP_DEVICE_ROCM || !x.contiguous() || !out.contiguous())
        return fail(FP_EINVAL, "fpr_prox_static takes contiguous device arrays");
    if (x.numel() == 0) return FP_OK;
    fpr::StaticSolve c;
    c.tsize = (int)x.itemsize();
    c.P.k = k;
    c.P.iso = op->iso ? 1 : 0;
    c.P.sw = op->spatial.empty() ? 0 : 1;
    c.P.sw_per_image = 0;
    for (int a = 0; a < 3; ++a) c.P.w[a] = op->w[a];
    c.P.sched = op->iso ? fpr::kSchedC : fpr::kSchedDevice;
    c.P.wmax_norm = op->iso ? 1 : 0;
    c.P.alpha_max = op->spatial.empty() ? 1.0 : op->spatial_max;
    int64_t N = 1;
    for (int a = 0; a < k; ++a) {
        c.L[3 - k + a] = x.shape[x.ndim - k + a];
        N *= c.L[3 - k + a];
    }
    c.N = N;
    const int64_t B = x.numel() / N;
    if (B > fpr::kStaticMaxImages)
        return fail(FP_ENOMEM, "fpr_prox_static: more images than its scratch holds (256 at most)");
    c.B = (int)B;
    if (k == 3) {
        c.kc.generic = false;
        c.kc.stream = false;
        c.kc.s3 = true;
        c.kc.TT = 1;
        c.kc.RY = 16;
    } else {
        c.kc.stream = true;
        c.kc.TT = 6;
        c.kc.RY = 16;
    }
    if ((c.kc.stream && !fpr::stream_fits(c.tsize, N)) || (k == 3 && !fpr::kernel_available_s3(16)))
        c.kc = generic_kernel();
    c.x = x.data;
    c.z = out.data;
    c.stage_x = x.data == out.data;
    c.sw_host = op->spatial.empty() ? nullptr : op->spatial.data();
    c.gamma = step;
    c.tol = o.tol;
    c.max_iter = o.max_iter;
    c.stream = (hipStream_t)stream;
    if (fpr::static_bytes_needed(c) > fpr::kStaticScratchBytes)
        return fail(FP_ENOMEM, "fpr_prox_static: the images need more than its 64 MiB of scratch");
    fpr::StaticResult R;
    if (int r = fpr::static_solve(c, &R)) return hip_status(r);
    if (R.bad) return FP_EINVAL;
    if (info) *info = fp_info{R.gap, R.iters, 0};
    return R.notconv ? FP_ENOTCONV : FP_OK;
}

// fp_workspace_refine (the reference's refine_impl, src/api/api.cpp:1185-1217):
// the kept call solved again from its saved input, warm-started from its
// dual, to tol within max_iter; the derivatives then read the refined point.
fp_status refine_impl(const fp_op *op, fp_workspace *ws, double tol, int max_iter, fp_array *oa, fp_info *info) {
    if (!op || !ws || ws->op != op || !ws->valid) return FP_EINVAL;
    if (!(tol >= 0.0) || !std::isfinite(tol) || max_iter < 1) return FP_EPARAM;
    Arr out;
    if (oa) {
        if (fp_status s = to_arr(oa, out)) return s;
        if (out.dtype != ws->dtype) return FP_EDTYPE;
        if (out.ndim != ws->ndim) return FP_ESHAPE;
        for (int d = 0; d < out.ndim; ++d)
            if (out.shape[d] != ws->shape[d]) return FP_ESHAPE;
        if (fp_status s = check_output(nullptr, out)) return s;
    }
    if (info) *info = fp_info{0.0, 0, 0};
    ws->info_valid = false;
    ws->deriv_built = false;
    ws->branch_on = false;
    if (ws->numel == 0) {
        ws->info = fp_info{0.0, 0, 0};
        ws->info_status = FP_OK;
        ws->info_valid = true;
        return FP_OK;
    }
    DevGuard guard(ws->device);
    hipStream_t s = (hipStream_t)ws->stream;
    hipError_t err;
    const void *sw = op->device_weights(ws->device, (int)ws->tsize, err);
    if (err) return hip_status(err);
    for (int64_t b = 0; b < ws->B; ++b) {
        fpr::ImagePtrs &p = ws->ptrs[(size_t)b];
        p.d = (const char *)ws->xsave + (size_t)(b * ws->N) * ws->tsize;
        p.z = (char *)ws->zsave + (size_t)(b * ws->N) * ws->tsize;
        p.q = (char *)ws->q + (size_t)(b * ws->N) * ws->tsize;
        p.sw = sw;
    }
    if (fp_status r = run_solve(ws, s, ws->xsave, ws->zsave, true, tol, max_iter, true)) {
        ws->dual_valid = false;
        ws->valid = false;
        return r;
    }
    fp_status status = FP_OK;
    if (fp_status st = record_results(op, ws, info, &status)) {
        ws->valid = false;
        return st;
    }
    if (fp_status st = record_kept(op, ws, s)) return st;
    if (oa) {
        if (out.device == FP_DEVICE_ROCM && out.contiguous()) {
            if (hipMemcpyAsync(out.data, ws->zsave, (size_t)ws->numel * ws->tsize, hipMemcpyDeviceToDevice, s))
                return FP_EDEVICE;
        } else if (fp_status st = stage_out(*ws, ws->zsave, out, s)) {
            return st;
        }
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
    }
    return status;
}

// The dual between the workspace (k planes of B N, x's dtype) and an array
// of shape (k, *x.shape) (api.cpp:1151-1180, 1235-1264).
fp_status dual_impl(fp_workspace *ws, const fp_array *aa, bool to_ws) {
    if (!ws || !aa || !ws->op) return FP_EINVAL;
    Arr a;
    if (fp_status s = to_arr(aa, a)) return s;
    if (aa->ndim != ws->ndim + 1 || a.shape[0] != ws->k) return FP_ESHAPE;
    for (int d = 0; d < ws->ndim; ++d)
        if (a.shape[d + 1] != ws->shape[d]) return FP_ESHAPE;
    if (!to_ws && !ws->dual_valid) return FP_EINVAL;
    if (!to_ws)
        if (fp_status s = check_output(nullptr, a)) return s;
    if (ws->numel == 0) {
        if (to_ws) ws->dual_valid = ws->seeded = true;
        return FP_OK;
    }
    hipStream_t s = (hipStream_t)ws->stream;
    const int64_t n = ws->numel * ws->k;
    // Through doubles on the host when dtypes differ or the array is on the
    // host; device to device otherwise.
    if (a.device == FP_DEVICE_ROCM && a.dtype == ws->dtype) {
        fp_status st = to_ws ? stage_in(*ws, a, ws->q, s) : stage_out(*ws, ws->q, a, s);
        if (st) return st;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
    } else {
        std::vector<double> v;
        if (to_ws) {
            if (fp_status st = read_doubles(a, v)) return st;
        } else {
            std::vector<char> raw((size_t)n * ws->tsize);
            if (hipMemcpy(raw.data(), ws->q, raw.size(), hipMemcpyDeviceToHost)) return FP_EDEVICE;
            v.resize((size_t)n);
            for (int64_t i = 0; i < n; ++i)
                v[(size_t)i] = ws->tsize == 4 ? (double)((const float *)raw.data())[i] : ((const double *)raw.data())[i];
            // write v into a (host or device) in a's dtype and strides
            std::vector<char> packed((size_t)n * a.itemsize());
            for (int64_t i = 0; i < n; ++i) {
                if (a.dtype == FP_F32) ((float *)packed.data())[i] = (float)v[(size_t)i];
                else ((double *)packed.data())[i] = v[(size_t)i];
            }
            if (a.device == FP_DEVICE_CPU) {
                unpack_host(packed.data(), a);
            } else {
                // the dual converted on the device straight into a's dtype and strides
                if (a.ndim > fpr::kMaxStrided) return FP_ESHAPE;
                const fpr::StridedShape sh = strided_of(a);
                if (ws->tsize == 4)
                    fpr::launch_convert_scatter<float, double>((const float *)ws->q, (double *)a.data, n, a.ndim, sh, s);
                else
                    fpr::launch_convert_scatter<double, float>((const double *)ws->q, (float *)a.data, n, a.ndim, sh, s);
                if (hipGetLastError() || hipStreamSynchronize(s)) return FP_EDEVICE;
            }
            return FP_OK;
        }
        std::vector<char> raw((size_t)n * ws->tsize);
        for (int64_t i = 0; i < n; ++i) {
            const double val = v[(size_t)i];
            if (ws->tsize == 4) ((float *)raw.data())[i] = (float)val;
            else ((double *)raw.data())[i] = val;
        }
        if (hipMemcpy(ws->q, raw.data(), raw.size(), hipMemcpyHostToDevice)) return FP_EDEVICE;
    }
    if (to_ws) {
        // A seed must not be infinite; NaN entries are stored as 0 (api.cpp:1241-1256).
        std::vector<char> raw((size_t)n * ws->tsize);
        if (hipMemcpy(raw.data(), ws->q, raw.size(), hipMemcpyDeviceToHost)) return FP_EDEVICE;
        bool ok = true, fix = false;
        for (int64_t i = 0; i < n; ++i) {
            double v = ws->tsize == 4 ? (double)((float *)raw.data())[i] : ((double *)raw.data())[i];
            if (std::isinf(v)) ok = false;
            if (v != v) {
                fix = true;
                if (ws->tsize == 4) ((float *)raw.data())[i] = 0.0f;
                else ((double *)raw.data())[i] = 0.0;
            }
        }
        if (!ok) {
            ws->dual_valid = false;
            return FP_EINVAL;
        }
        if (fix && hipMemcpy(ws->q, raw.data(), raw.size(), hipMemcpyHostToDevice)) return FP_EDEVICE;
        ws->dual_valid = true;
        ws->seeded = true;
    }
    return FP_OK;
}

int device_count() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

fp_status op_check(const fp_op *op, int *row_parallel_only) {
    if (!op) return FP_EINVAL;
    if (row_parallel_only) *row_parallel_only = 0;
    return FP_OK;
}

// ---------------------------------------------------------------------------
// Derivatives at a kept solve (the reference's jvp_impl / vjp_impl,
// src/api/api.cpp:584-760, over the device derivatives of deriv.hip)
// ---------------------------------------------------------------------------

fp_status check_like(const fp_workspace &ws, const Arr &a) {
    if (a.dtype != ws.dtype) return FP_EDTYPE;
    if (a.ndim != ws.ndim) return FP_ESHAPE;
    for (int d = 0; d < a.ndim; ++d)
        if (a.shape[d] != ws.shape[d]) return FP_ESHAPE;
    return FP_OK;
}

// Writes n doubles to a float array (host or device, any strides), in C order.
fp_status write_doubles(const Arr &x, const double *v) {
    const int64_t n = x.numel();
    if (n == 0) return FP_OK;
    intptr_t lo, hi;
    x.span(lo, hi);
    const size_t bytes = (size_t)(hi - lo + 1);
    std::vector<char> host;
    char *base = nullptr;
    if (x.device == FP_DEVICE_ROCM) {
        host.resize(bytes);
        if (hipMemcpy(host.data(), (const void *)lo, bytes, hipMemcpyDeviceToHost)) return FP_EDEVICE;
        base = host.data() + ((intptr_t)x.data - lo);
    } else {
        base = x.data;
    }
    const int64_t item = (int64_t)x.itemsize();
    for (int64_t i = 0; i < n; ++i) {
        int64_t r = i, off = 0;
        for (int d = x.ndim - 1; d >= 0; --d) {
            const int64_t c = r % x.shape[d];
            r /= x.shape[d];
            off += c * x.strides[d];
        }
        char *p = base + off * item;
        if (x.dtype == FP_F32) *(float *)p = (float)v[i];
        else *(double *)p = v[i];
    }
    if (x.device == FP_DEVICE_ROCM && hipMemcpy((void *)lo, host.data(), bytes, hipMemcpyHostToDevice))
        return FP_EDEVICE;
    return FP_OK;
}

// The derivatives' small device arrays (sized by reserve_kept).
fp_status scratch(fp_workspace *ws, size_t n, double **out) {
    if (!ws->dscr || ws->dscr_n < n) return FP_EINTERNAL;
    *out = ws->dscr;
    return FP_OK;
}

// The structure of the kept solve, built into its reserved buffers on first
// use after each kept call.
fp_status ensure_deriv(const fp_op *op, fp_workspace *ws, hipStream_t s) {
    if (ws->deriv_built) return FP_OK;
    if (!ws->deriv) return FP_EINVAL;
    fpr::DerivSpec sp;
    sp.iso = op->iso;
    sp.k = op->k;
    for (int a = 0; a < 3; ++a) sp.L[a] = ws->img[a];
    sp.B = ws->B;
    sp.tsize = (int)ws->tsize;
    for (int a = 0; a < op->k; ++a) sp.w[a] = op->w[a];
    sp.x = ws->xsave;
    sp.z = ws->zsave;
    if (op->iso) {
        sp.u = ws->q;  // (k, B, N): the solve's dual
        sp.u_img = ws->N;
        sp.u_plane = ws->B * ws->N;
    } else {
        sp.u = ws->usave;  // (B, k, N): the per-edge dual
        sp.u_img = (int64_t)op->k * ws->N;
        sp.u_plane = ws->N;
    }
    hipError_t err;
    sp.sw = op->device_weights(ws->device, (int)ws->tsize, err);
    if (err) return hip_status(err);
    sp.sw_img = 0;
    sp.sw_plane = ws->N;
    sp.gamma = ws->gammas.data();
    sp.gap = ws->gaps.data();
    sp.e = ws->exps.data();
    if (!op->iso && ws->branch_on) {
        sp.side = ws->branch;
        sp.side_tol = ws->branch_tol;
        sp.q = ws->q;  // (k, B, N): the C-form dual
        sp.q_img = ws->N;
        sp.q_plane = ws->B * ws->N;
    }
    if (int r = fpr::deriv_build(ws->deriv.get(), sp, s)) return (fp_status)r;
    ws->deriv_built = true;
    return FP_OK;
}

fp_status jvp_impl(const fp_op *op, fp_workspace *ws, const fp_array *dxa, double dgamma, const fp_array *dpa,
                   fp_array *oa) {
    if (!op || !ws || !oa) return FP_EINVAL;
    if (ws->op != op || !ws->valid) return FP_EINVAL;
    if (!std::isfinite(dgamma)) return FP_EPARAM;
    Arr dx, out;
    if (fp_status s = to_arr(oa, out)) return s;
    if (fp_status s = check_like(*ws, out)) return s;
    if (dxa) {
        if (fp_status s = to_arr(dxa, dx)) return s;
        if (fp_status s = check_like(*ws, dx)) return s;
    }
    if (fp_status s = check_output(dxa ? &dx : nullptr, out)) return s;
    const int64_t P = op->nparams();
    std::vector<double> dp;
    if (dpa && P > 0) {
        Arr t;
        if (fp_status s = to_arr(dpa, t)) return s;
        if (dpa->ndim > 1 || t.numel() != P) return FP_ESHAPE;
        if (fp_status s = read_doubles(t, dp)) return s;
    }
    if (ws->numel == 0) return FP_OK;
    DevGuard guard(ws->device);
    hipStream_t s = (hipStream_t)ws->stream;
    if (fp_status st = ensure_deriv(op, ws, s)) return st;
    const int64_t B = ws->B;
    const int64_t ns = (int64_t)op->spatial.size();
    const void *dxd = nullptr;
    if (dxa) {
        if (dx.device == FP_DEVICE_ROCM && dx.contiguous()) {
            dxd = dx.data;
        } else {
            if (fp_status st = stage_in(*ws, dx, ws->dstage, s)) return st;
            dxd = ws->dstage;
        }
    }
    const bool o_direct = out.device == FP_DEVICE_ROCM && out.contiguous();
    void *dz = o_direct ? (void *)out.data : ws->zstage;
    double *scr = nullptr;
    if (fp_status st = scratch(ws, (size_t)(B + ns), &scr)) return st;
    std::vector<double> up((size_t)(B + ns), dgamma);
    const bool sp_on = !dp.empty() && ns > 0;
    if (sp_on)
        for (int64_t i = 0; i < ns; ++i) up[(size_t)(B + i)] = dp[(size_t)(op->k + i)];
    if (hipMemcpyAsync(scr, up.data(), sizeof(double) * (size_t)(B + (sp_on ? ns : 0)), hipMemcpyHostToDevice, s))
        return FP_EDEVICE;
    double dw[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < op->k && !dp.empty(); ++a) dw[a] = dp[(size_t)a];
    int r = fpr::deriv_jvp(ws->deriv.get(), s, dxd, scr, dp.empty() ? nullptr : dw, sp_on ? scr + B : nullptr, dz,
                           ws->deriv_tol, ws->deriv_max_iter);
    if (r) return (fp_status)r;
    if (!o_direct)
        if (fp_status st = stage_out(*ws, dz, out, s)) return st;
    if (hipStreamSynchronize(s)) return FP_EDEVICE;
    fpr::deriv_result(ws->deriv.get(), &ws->deriv_res, &ws->deriv_its, nullptr);
    return FP_OK;
}

fp_status vjp_impl(const fp_op *op, fp_workspace *ws, const fp_array *ga, fp_array *gxa, double *ggamma,
                   const fp_array *gga, fp_array *gpa) {
    if (!op || !ws || !ga || !gxa) return FP_EINVAL;
    if (ws->op != op || !ws->valid) return FP_EINVAL;
    Arr g, gx, gp, gg;
    if (fp_status s = to_arr(ga, g)) return s;
    if (fp_status s = to_arr(gxa, gx)) return s;
    if (fp_status s = check_like(*ws, g)) return s;
    if (fp_status s = check_like(*ws, gx)) return s;
    if (fp_status s = check_output(&g, gx)) return s;
    const int64_t P = op->nparams();
    if (gpa) {  // validate before any work
        if (fp_status s = to_arr(gpa, gp)) return s;
        if (gpa->ndim > 1 || gp.numel() != P) return FP_ESHAPE;
    }
    if (gga) {
        if (fp_status s = to_arr(gga, gg)) return s;
        if (gga->ndim != 0) return FP_ESHAPE;
    }
    const int64_t B = ws->B;
    const int64_t ns = (int64_t)op->spatial.size();
    std::vector<double> h((size_t)(B + op->k + ns), 0.0);
    if (ws->numel > 0) {
        DevGuard guard(ws->device);
        hipStream_t s = (hipStream_t)ws->stream;
        if (fp_status st = ensure_deriv(op, ws, s)) return st;
        const void *gd = g.data;
        if (!(g.device == FP_DEVICE_ROCM && g.contiguous())) {
            if (fp_status st = stage_in(*ws, g, ws->dstage, s)) return st;
            gd = ws->dstage;
        }
        const bool o_direct = gx.device == FP_DEVICE_ROCM && gx.contiguous() && gx.data != gd;
        void *gxd = o_direct ? (void *)gx.data : ws->zstage;
        double *scr = nullptr;
        if (fp_status st = scratch(ws, h.size(), &scr)) return st;
        int r = fpr::deriv_vjp(ws->deriv.get(), s, gd, gxd, scr, scr + B, ns ? scr + B + op->k : nullptr,
                               ws->deriv_tol, ws->deriv_max_iter);
        if (r) return (fp_status)r;
        if (!o_direct)
            if (fp_status st = stage_out(*ws, gxd, gx, s)) return st;
        if (hipMemcpyAsync(h.data(), scr, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
        fpr::deriv_result(ws->deriv.get(), &ws->deriv_res, &ws->deriv_its, nullptr);
    }
    double tot = 0.0;  // the gamma gradient summed over the images
    for (int64_t b = 0; b < B; ++b) tot += h[(size_t)b];
    if (ggamma) *ggamma = tot;
    if (gga)
        if (fp_status st = write_doubles(gg, &tot)) return st;
    if (gpa && P > 0)
        if (fp_status st = write_doubles(gp, h.data() + B)) return st;
    return FP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Batches of uneven images
// ---------------------------------------------------------------------------

struct fpr_batch {
    const fp_op *op = nullptr;
    int64_t rep[3] = {1, 1, 1};  // a representative image for the tuning table
    int64_t maxN = 0;            // the largest image's pixel count
    int dtype = FP_F32;
    size_t tsize = 4;
    int weighted = 0;
    int device = 0;
    std::vector<fpr::ImageSpec> specs;
    fpr::KernelChoice kc;
    std::unique_ptr<fpr::Engine> eng;                   // dual FISTA (and iso)
    std::unique_ptr<fpr::AdSolver, fpr::AdDeleter> ad;  // aniso by the reference's method (uniform batches)
    int method = FPR_METHOD_FISTA;
    std::vector<fpr::ImagePtrs> ptrs;
    std::vector<double> gammas, gaps, zerrs;
    std::vector<int> iters, conv, exps;
    bool valid = false;
    // a solve enqueued on a capturing stream: results read at the first query
    bool pending = false;
    hipStream_t pending_stream = nullptr;
    bool uniform() const {
        for (size_t i = 1; i < specs.size(); ++i)
            for (int a = 0; a < 3; ++a)
                if (specs[i].L[a] != specs[0].L[a]) return false;
        return true;
    }
};

struct fpr_server {
    std::unique_ptr<fpr::Server> s;
    int device = 0;
    int tsize = 4;
};

struct fpr_deriv {
    std::unique_ptr<fpr::Deriv, fpr::DerivDeleter> d;
    int device = 0;
    int64_t B = 0;
    int k = 2;
    int64_t ns = 0;          // spatial entries (per image or shared), 0 without
    double *scr = nullptr;   // device: the axis weights' gradient
    ~fpr_deriv() {
        if (scr) (void)hipFree(scr);
    }
};

namespace {

// The ADBPG solver of a uniform, unweighted anisotropic batch.
fp_status batch_build_ad(fpr_batch &b) {
    const fp_op &op = *b.op;
    fpr::AdSpec sp;
    sp.k = op.k;
    for (int a = 0; a < 3; ++a) sp.L[a] = b.specs[0].L[a];
    sp.B = (int64_t)b.specs.size();
    sp.tsize = (int)b.tsize;
    for (int a = 0; a < 3; ++a) sp.w[a] = op.w[a];
    fpr::AdSolver *a = nullptr;
    if (int r = fpr::ad_reserve(sp, &a)) return (fp_status)r;
    b.ad.reset(a);
    return FP_OK;
}

// The engine's per-image results of the last solve into b.
fp_status batch_results(fpr_batch &b) {
    const fp_op &op = *b.op;
    const fpr::ImgState *S = b.eng->state();
    bool bad = false, notconv = false;
    for (size_t i = 0; i < b.specs.size(); ++i) {
        const fpr::ImgState &t = S[i];
        if (t.status) bad = true;
        double g = t.gp;
        if (op.iso && g == g) g = g < 0.0 ? 0.0 : g;
        if (op.iso && t.e) g = std::ldexp(g, 2 * t.e);
        b.gaps[i] = g;
        b.iters[i] = t.it;
        b.zerrs[i] = t.zerr;
        b.conv[i] = t.converged;
        b.exps[i] = t.e;
        if (!t.converged) notconv = true;
    }
    if (bad) {
        b.valid = false;
        return FP_EINVAL;
    }
    b.valid = true;
    return notconv ? FP_ENOTCONV : FP_OK;
}

// A captured solve's results, once it has run (waits for its stream).
fp_status settle_batch(const fpr_batch *cb) {
    fpr_batch *b = const_cast<fpr_batch *>(cb);
    if (!b->pending) return FP_OK;
    DevGuard g(b->device);
    if (hipStreamSynchronize(b->pending_stream)) return FP_EDEVICE;
    b->eng->collect();
    b->pending = false;
    const fp_status st = batch_results(*b);
    return st == FP_ENOTCONV ? FP_OK : st;
}

fp_status batch_build(fpr_batch &b) {
    const fp_op &op = *b.op;
    fpr::Problem P{};
    P.k = op.k;
    P.iso = op.iso ? 1 : 0;
    P.sw = b.weighted ? 1 : 0;
    P.sw_per_image = 1;
    for (int a = 0; a < 3; ++a) P.w[a] = op.w[a];
    P.sched = op.iso ? fpr::kSchedC : fpr::kSchedDevice;
    P.wmax_norm = op.iso ? 1 : 0;
    P.alpha_max = 1.0;  // per image, from its weights (the prologue)
    b.eng.reset(new fpr::Engine());
    if (int e = b.eng->setup((int)b.tsize, P, b.kc, b.specs)) return hip_status(e);
    // The table's tail kernel is measured on batches of one shape. In a batch
    // of different shapes the last live images are its largest and slowest,
    // which a short-segment tiling only slows down (the 256 uneven images with
    // pixel weights: 1,713 images/s with the table's tail, 2,121 without), so
    // such a batch gets no tail unless fpr_batch_set_tail asks for one.
    bool uniform = true;
    for (size_t i = 1; i < b.specs.size() && uniform; ++i)
        for (int a = 0; a < 3; ++a) uniform = uniform && b.specs[i].L[a] == b.specs[0].L[a];
    double frac = 0.0;
    const fpr::KernelChoice tk = fpr::default_tail(op.k, (int)b.tsize, b.rep, op.iso, b.weighted != 0, (int64_t)b.specs.size(), &frac);
    if (uniform && frac > 0.0 && b.kc.stream && !b.kc.generic && tk.iters() >= b.kc.iters())
        return hip_status(b.eng->set_tail(tk, frac));
    return FP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// Exported entry points
// ---------------------------------------------------------------------------

extern "C" {

FP_API void fp_options_init(fp_options *opt) {
    if (!opt) return;
    opt->axis = -1;
    opt->threads = 0;
    opt->tol = 1e-6;
    opt->max_iter = 1000;
    opt->tie = 0;
    opt->keep_structure = 0;
}

FP_API fp_status fp_op_tv_aniso(int ndim, const fp_array *axis_weights, fp_op **out) {
    return guard([&] { return make_tv(false, ndim, axis_weights, nullptr, out); });
}

FP_API fp_status fp_op_tv_aniso_weighted(int ndim, const fp_array *axis_weights, const fp_array *edge_weights,
                                         fp_op **out) {
    return guard([&] {
        if (!out) return FP_EINVAL;
        *out = nullptr;
        if (!edge_weights) return FP_EINVAL;
        return make_tv(false, ndim, axis_weights, edge_weights, out);
    });
}

FP_API fp_status fp_op_tv_iso(int ndim, const fp_array *axis_weights, fp_op **out) {
    return guard([&] { return make_tv(true, ndim, axis_weights, nullptr, out); });
}

FP_API fp_status fp_op_tv_iso_weighted(int ndim, const fp_array *axis_weights, const fp_array *pixel_weights,
                                       fp_op **out) {
    return guard([&] {
        if (!out) return FP_EINVAL;
        *out = nullptr;
        if (!pixel_weights) return FP_EINVAL;
        return make_tv(true, ndim, axis_weights, pixel_weights, out);
    });
}

FP_API void fp_op_free(fp_op *op) { delete op; }
FP_API int64_t fp_op_nparams(const fp_op *op) { return op ? op->nparams() : 0; }
FP_API const char *fp_op_name(const fp_op *op) { return op ? (op->iso ? "tv_iso" : "tv_aniso") : nullptr; }

FP_API fp_status fp_workspace_create(const fp_op *op, const fp_array *like, const fp_options *opt,
                                     fp_workspace **out) {
    return guard([&] { return create_impl(op, like, opt, out); });
}

FP_API void fp_workspace_free(fp_workspace *ws) { delete ws; }

FP_API fp_status fp_workspace_get_dual(const fp_workspace *ws, fp_array *out) {
    return guard([&] { return dual_impl(const_cast<fp_workspace *>(ws), out, false); });
}

FP_API fp_status fp_workspace_set_dual(fp_workspace *ws, const fp_array *dual) {
    return guard([&] { return dual_impl(ws, dual, true); });
}

FP_API fp_status fp_workspace_set_warm(fp_workspace *ws, int on) {
    if (!ws) return FP_EINVAL;
    ws->warm = on != 0;
    return FP_OK;
}

FP_API fp_status fp_workspace_set_deriv_tol(fp_workspace *ws, double tol, int max_iter) {
    if (!ws) return FP_EINVAL;
    if (!(tol >= 0.0) || !std::isfinite(tol) || max_iter < 1) return FP_EPARAM;
    ws->deriv_tol = tol;
    ws->deriv_max_iter = max_iter;
    return FP_OK;
}

FP_API fp_status fp_workspace_deriv_info(const fp_workspace *ws, double *residual, int *iters) {
    if (!ws) return FP_EINVAL;
    if (residual) *residual = ws->deriv_res;
    if (iters) *iters = ws->deriv_its;
    return FP_OK;
}

FP_API fp_status fp_prox(const fp_op *op, double gamma, const fp_array *x, fp_array *out, const fp_options *opt,
                         fp_workspace *ws, fp_info *info) {
    return guard([&] { return prox_impl(op, gamma, x, out, opt, ws, info, nullptr); });
}

FP_API fp_status fpr_prox_static(const fp_op *op, double gamma, const fp_array *x, fp_array *out,
                                 const fp_options *opt, void *stream, fp_info *info) {
    return guard([&] { return static_prox_impl(op, gamma, x, out, opt, stream, info); });
}

FP_API fp_status fp_prox_steps(const fp_op *op, const fp_array *gamma, const fp_array *x, fp_array *out,
                               const fp_options *opt, fp_workspace *ws) {
    return guard([&] {
        if (!gamma) return FP_EINVAL;
        return prox_impl(op, 0.0, x, out, opt, ws, nullptr, gamma);
    });
}

FP_API fp_status fp_jvp(const fp_op *op, const fp_workspace *ws, const fp_array *dx, double dgamma,
                        const fp_array *dparams, fp_array *dout) {
    return guard([&] { return jvp_impl(op, const_cast<fp_workspace *>(ws), dx, dgamma, dparams, dout); });
}

FP_API fp_status fp_vjp(const fp_op *op, const fp_workspace *ws, const fp_array *g, fp_array *gx, double *ggamma,
                        fp_array *gparams) {
    return guard([&] {
        return vjp_impl(op, const_cast<fp_workspace *>(ws), g, gx, ggamma, nullptr, gparams);
    });
}

FP_API fp_status fp_vjp_device(const fp_op *op, const fp_workspace *ws, const fp_array *g, fp_array *gx,
                               fp_array *ggamma, fp_array *gparams) {
    return guard([&] {
        return vjp_impl(op, const_cast<fp_workspace *>(ws), g, gx, nullptr, ggamma, gparams);
    });
}

FP_API int fp_device_available(fp_device dev) {
    const int d = enum_bits(dev);
    if (d == FP_DEVICE_CPU) return 1;
    return d == FP_DEVICE_ROCM && device_count() > 0 ? 1 : 0;
}

FP_API fp_status fp_op_device_check(const fp_op *op, fp_device dev, int *row_parallel_only) {
    if (!op) return FP_EINVAL;
    const int d = enum_bits(dev);
    if (d == FP_DEVICE_CPU) {
        if (row_parallel_only) *row_parallel_only = 0;
        return FP_OK;
    }
    if (d != FP_DEVICE_ROCM || device_count() == 0) return FP_EINVAL;
    return op_check(op, row_parallel_only);
}

FP_API fp_status fp_workspace_set_stream(fp_workspace *ws, void *stream) {
    if (!ws) return FP_EINVAL;
    ws->stream = stream;
    return FP_OK;
}

FP_API fp_status fp_workspace_get_stream(const fp_workspace *ws, void **stream) {
    if (!ws || !stream) return FP_EINVAL;
    *stream = ws->stream;
    return FP_OK;
}

FP_API fp_status fp_workspace_info(const fp_workspace *ws, fp_info *info) {
    if (!ws || !info) return FP_EINVAL;
    if (fp_status st = settle(const_cast<fp_workspace *>(ws))) return st;
    if (!ws->info_valid) return FP_EINVAL;
    *info = ws->info;
    return ws->info_status;
}

FP_API fp_status fp_workspace_refine(const fp_op *op, fp_workspace *ws, double tol, int max_iter, fp_array *out,
                                     fp_info *info) {
    return guard([&] { return refine_impl(op, ws, tol, max_iter, out, info); });
}

// ---- structure of the derivative and kinks (include/fastprox.h:440-579) ----

namespace {

// The queries read the point kept by the last prox call (src/api/api.cpp:
// 1031-1036 in the reference, which has them on the CPU only).
fp_status struct_ws(const fp_workspace *ws) {
    if (!ws || !ws->op || !ws->keep || !ws->valid) return FP_EINVAL;
    return FP_OK;
}

int64_t struct_units(const fp_workspace &ws) { return ws.numel == 0 ? 0 : ws.B; }

enum class OutKind { Int8, Label, Float };

// An output of the queries: x's shape, of the kind's dtype (Float: x's),
// writable (src/api/api.cpp:1058-1065).
fp_status struct_out(const fp_workspace &ws, const fp_array *a, Arr &out, OutKind kind) {
    if (fp_status s = to_arr(a, out, kind == OutKind::Float)) return s;
    if (kind == OutKind::Int8 && out.dtype != FP_I8) return FP_EDTYPE;
    if (kind == OutKind::Label && out.dtype != FP_I32 && out.dtype != FP_I64) return FP_EDTYPE;
    if (kind == OutKind::Float && out.dtype != ws.dtype) return FP_EDTYPE;
    if (out.ndim != ws.ndim) return FP_ESHAPE;
    for (int d = 0; d < out.ndim; ++d)
        if (out.shape[d] != ws.shape[d]) return FP_ESHAPE;
    if (out.device == FP_DEVICE_ROCM && !out.contiguous() && out.ndim > fpr::kMaxStrided) return FP_ESHAPE;
    return check_output(nullptr, out);
}


// The contiguous device entries src (C order, out's item size) into out
// (host or device, any strides).
fp_status put_struct(const void *src, const Arr &a, hipStream_t s) {
    const int64_t n = a.numel();
    if (n == 0) return FP_OK;
    const size_t bytes = (size_t)n * a.itemsize();
    if (a.device == FP_DEVICE_CPU) {
        std::vector<char> host(bytes);
        if (hipMemcpyAsync(host.data(), src, bytes, hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s))
            return FP_EDEVICE;
        unpack_host(host.data(), a);
        return FP_OK;
    }
    if (a.contiguous()) return hipMemcpyAsync(a.data, src, bytes, hipMemcpyDeviceToDevice, s) ? FP_EDEVICE : FP_OK;
    const fpr::StridedShape sh = strided_of(a);
    switch (a.itemsize()) {
        case 1: fpr::launch_scatter<int8_t>((const int8_t *)src, (int8_t *)a.data, n, a.ndim, sh, s); break;
        case 4: fpr::launch_scatter<int32_t>((const int32_t *)src, (int32_t *)a.data, n, a.ndim, sh, s); break;
        default: fpr::launch_scatter<int64_t>((const int64_t *)src, (int64_t *)a.data, n, a.ndim, sh, s); break;
    }
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

// Zeros into out (a state array of FREE entries, or a coefficient array):
// through the workspace's query buffer (9 numel bytes, at least the array's).
fp_status put_zeros(const fp_workspace &ws, const Arr &a, hipStream_t s) {
    const size_t bytes = (size_t)a.numel() * a.itemsize();
    if (!ws.qbuf || bytes > 9 * (size_t)ws.numel) return FP_EINTERNAL;
    fpr::launch_zero_bytes(ws.qbuf, bytes, s);
    if (hipGetLastError()) return FP_EDEVICE;
    return put_struct(ws.qbuf, a, s);
}

fp_status states_impl(const fp_workspace *ws, fp_array *sa, int64_t *nfixed) {
    if (fp_status s = struct_ws(ws)) return s;
    Arr st;
    if (sa)
        if (fp_status s = struct_out(*ws, sa, st, OutKind::Int8)) return s;
    if (sa && ws->numel) {
        DevGuard guard(ws->device);
        if (fp_status s = put_zeros(*ws, st, (hipStream_t)ws->stream)) return s;
    }
    if (nfixed) *nfixed = 0;  // every entry of an image operator is FREE
    return FP_OK;
}

fp_status regions_impl(const fp_op *op, fp_workspace *ws, fp_array *la, fp_array *ga, int64_t *nregions,
                       int64_t count) {
    if (fp_status s = struct_ws(ws)) return s;
    if (nregions && count != struct_units(*ws)) return FP_ESHAPE;
    Arr lab, sg;
    if (la)
        if (fp_status s = struct_out(*ws, la, lab, OutKind::Label)) return s;
    if (ga)
        if (fp_status s = struct_out(*ws, ga, sg, OutKind::Int8)) return s;
    if (ws->numel == 0) return FP_OK;
    DevGuard guard(ws->device);
    hipStream_t s = (hipStream_t)ws->stream;
    if (fp_status st = ensure_deriv(op, ws, s)) return st;
    // labels (up to 8 bytes each) then signs in the query buffer
    const size_t n = (size_t)ws->numel;
    char *lb = la ? (char *)ws->qbuf : nullptr;
    int8_t *sb = ga ? (int8_t *)ws->qbuf + 8 * n : nullptr;
    if (int r = fpr::deriv_regions(ws->deriv.get(), s, lb, la && lab.dtype == FP_I64, sb, nregions))
        return (fp_status)r;
    if (la)
        if (fp_status st = put_struct(lb, lab, s)) return st;
    if (ga)
        if (fp_status st = put_struct(sb, sg, s)) return st;
    return FP_OK;
}

fp_status constraint_impl(const fp_workspace *ws, fp_array *ca, int8_t *active, int64_t count) {
    if (fp_status s = struct_ws(ws)) return s;
    if (active && count != struct_units(*ws)) return FP_ESHAPE;
    Arr c;
    if (ca)
        if (fp_status s = struct_out(*ws, ca, c, OutKind::Float)) return s;
    if (ca && ws->numel) {
        DevGuard guard(ws->device);
        if (fp_status s = put_zeros(*ws, c, (hipStream_t)ws->stream)) return s;
    }
    if (active) std::fill(active, active + count, int8_t{0});  // TV has no linear constraint
    return FP_OK;
}

// fp_workspace_set_branch (src/api/api.cpp:993-1025): anisotropic TV only;
// every argument is checked before the workspace changes, so a failed call
// leaves an earlier override in place.
fp_status set_branch_impl(fp_workspace *ws, const fp_array *sa, double tol) {
    if (!ws || !ws->op) return FP_EINVAL;
    if (!sa) {
        if (ws->branch_on) {
            ws->branch_on = false;
            ws->deriv_built = false;
        }
        return FP_OK;
    }
    if (!(tol >= 0.0) || !std::isfinite(tol)) return FP_EPARAM;
    if (ws->op->iso) return FP_EINVAL;
    if (fp_status s = struct_ws(ws)) return s;
    Arr a;
    if (fp_status s = to_arr(sa, a, false)) return s;
    if (a.dtype != FP_I8) return FP_EDTYPE;
    if (a.ndim != ws->ndim) return FP_ESHAPE;
    for (int d = 0; d < a.ndim; ++d)
        if (a.shape[d] != ws->shape[d]) return FP_ESHAPE;
    if (ws->numel == 0) return FP_OK;
    DevGuard guard(ws->device);
    hipStream_t s = (hipStream_t)ws->stream;
    const size_t n = (size_t)ws->numel;
    std::vector<char> h(n);
    if (a.device == FP_DEVICE_CPU) {
        pack_host(a, h.data());
    } else {
        const void *src = a.data;
        if (!a.contiguous()) {
            if (a.ndim > fpr::kMaxStrided) return FP_ESHAPE;
            fpr::launch_gather<int8_t>((const int8_t *)a.data, (int8_t *)ws->qbuf, (int64_t)n, a.ndim, strided_of(a),
                                       s);
            src = ws->qbuf;
        }
        if (hipMemcpyAsync(h.data(), src, n, hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s)) return FP_EDEVICE;
    }
    for (size_t i = 0; i < n; ++i)
        if (h[i] < -1 || h[i] > 1) return FP_EPARAM;
    if (!ws->branch) return FP_EINTERNAL;  // reserved with the structure (aniso)
    if (hipMemcpyAsync(ws->branch, h.data(), n, hipMemcpyHostToDevice, s) || hipStreamSynchronize(s))
        return FP_EDEVICE;
    ws->branch_on = true;
    ws->branch_tol = tol;
    ws->deriv_built = false;  // rebuilt with the override on first use
    return FP_OK;
}

fp_status kinks_impl(const fp_op *op, fp_workspace *ws, double tol, fp_array *ma, int64_t *count) {
    if (!(tol >= 0.0) || !std::isfinite(tol)) return FP_EPARAM;
    if (fp_status s = struct_ws(ws)) return s;
    Arr m;
    if (ma)
        if (fp_status s = struct_out(*ws, ma, m, OutKind::Int8)) return s;
    int64_t c = 0;
    if (ws->numel) {
        DevGuard guard(ws->device);
        hipStream_t s = (hipStream_t)ws->stream;
        if (fp_status st = ensure_deriv(op, ws, s)) return st;
        int8_t *mb = (int8_t *)ws->qbuf;
        if (int r = fpr::deriv_kinks(ws->deriv.get(), s, ws->q, ws->N, ws->B * ws->N, tol, mb, &c))
            return (fp_status)r;
        if (ma)
            if (fp_status st = put_struct(mb, m, s)) return st;
    }
    if (count) *count = c;
    return FP_OK;
}

}  // namespace

FP_API fp_status fp_op_struct_kind(const fp_op *op, int *kind) {
    if (!op || !kind) return FP_EINVAL;
    *kind = op->iso ? FP_STRUCT_CURVED : FP_STRUCT_PROJECTOR;
    return FP_OK;
}

FP_API fp_status fp_workspace_units(const fp_workspace *ws, int64_t *count, int64_t *size) {
    if (fp_status s = struct_ws(ws)) return s;
    if (count) *count = struct_units(*ws);
    if (size) *size = ws->N;
    return FP_OK;
}

FP_API fp_status fp_workspace_states(const fp_workspace *ws, fp_array *state, int64_t *nfixed) {
    return guard([&] { return states_impl(ws, state, nfixed); });
}

FP_API fp_status fp_workspace_regions(const fp_workspace *ws, fp_array *labels, fp_array *sign, int64_t *nregions,
                                      int64_t count) {
    // The structure is built on first use: the workspace is logically const.
    return guard([&] {
        return regions_impl(ws ? ws->op : nullptr, const_cast<fp_workspace *>(ws), labels, sign, nregions, count);
    });
}

FP_API fp_status fp_workspace_constraint(const fp_workspace *ws, fp_array *coef, int8_t *active, int64_t count) {
    return guard([&] { return constraint_impl(ws, coef, active, count); });
}

FP_API fp_status fp_workspace_kinks(const fp_workspace *ws, double tol, fp_array *mask, int64_t *count) {
    return guard([&] { return kinks_impl(ws ? ws->op : nullptr, const_cast<fp_workspace *>(ws), tol, mask, count); });
}

FP_API fp_status fp_workspace_set_branch(fp_workspace *ws, const fp_array *side, double tol) {
    return guard([&] { return set_branch_impl(ws, side, tol); });
}

FP_API fp_status fp_workspace_deriv_exact(const fp_workspace *ws, int *exact, double *smoothing) {
    if (!ws || !ws->valid) return FP_EINVAL;
    int all = 1;
    double mu = 0.0;
    for (size_t i = 0; i < ws->exact.size(); ++i) {
        all = all && ws->exact[i];
        mu = std::max(mu, ws->smooth[i]);
    }
    if (exact) *exact = all;
    if (smoothing) *smoothing = mu;
    return FP_OK;
}

FP_API fp_status fp_workspace_image_info(const fp_workspace *ws, double *gaps, int *iters, double *zerrs,
                                         int64_t count) {
    if (!ws) return FP_EINVAL;
    if (fp_status st = settle(const_cast<fp_workspace *>(ws))) return st;
    if (!ws->info_valid) return FP_EINVAL;
    int64_t images = 1;
    for (int d = 0; d < ws->ndim - ws->k; ++d) images *= ws->shape[d];
    if (count != images) return FP_ESHAPE;
    const bool held = ws->numel != 0;
    for (int64_t r = 0; r < count; ++r) {
        if (gaps) gaps[r] = held ? ws->gaps[(size_t)r] : 0.0;
        if (iters) iters[r] = held ? ws->iters[(size_t)r] : 0;
        if (zerrs) zerrs[r] = held ? ws->zerrs[(size_t)r] : 0.0;
    }
    return FP_OK;
}

FP_API const char *fp_strerror(fp_status s) {
    switch (s) {
        case FP_OK: return "success";
        case FP_EINVAL: return "invalid argument";
        case FP_EDTYPE: return "unsupported or mismatched dtype";
        case FP_ESHAPE: return "mismatched shape or parameter length";
        case FP_ENOMEM: return "out of memory";
        case FP_ENOTCONV: return "iteration stopped before reaching the tolerance";
        case FP_EPARAM: return "invalid operator parameter or step";
        case FP_EINTERNAL: return "internal error";
        case FP_EDEVICE: return "device error";
    }
    return "unknown status";
}

FP_API int fp_abi_version(void) { return FP_ABI_VERSION; }
FP_API const char *fp_version_string(void) { return FASTPROX_ROCM_BUILD; }
FP_API int fp_default_threads(void) { return 1; }

static int plugin_device_count(void) { return device_count(); }
static fp_status plugin_op_check(const fp_op *op, int *rpo) { return op_check(op, rpo); }

FP_API const fpr_plugin_table *fastprox_rocm_plugin(void) {
    static const fpr_plugin_table table = {FPR_PLUGIN_ABI, FASTPROX_ROCM_BUILD, plugin_device_count,
                                           plugin_op_check};
    return &table;
}

FP_API const char *fpr_last_error(void) { return g_last_error; }

FP_API int64_t fpr_alloc_count(void) { return fpr::alloc_count(); }

FP_API fp_status fpr_workspace_set_kernel(fp_workspace *ws, int generic, int tt, int ry) {
    return guard([&] {
        g_last_error[0] = 0;
        if (!ws) return fail(FP_EINVAL, "null workspace");
        fpr::KernelChoice kc;
        if (fp_status st = kernel_choice(generic, tt, ry, ws->k, (int)ws->tsize, !ws->op->spatial.empty(), kc))
            return st;
        if ((kc.stream || kc.s3t()) && !fpr::stream_fits((int)ws->tsize, ws->N))
            return fail(FP_EINVAL, "images too large for the streaming kernel");
        ws->kc = kc;
        if (ws->numel == 0) return FP_OK;
        ws->dual_valid = false;
        // a step kernel is dual FISTA's: an anisotropic workspace takes that method
        ws->ad.reset();
        ws->method = FPR_METHOD_FISTA;
        return build_engine(*ws);
    });
}

FP_API fp_status fpr_workspace_set_method(fp_workspace *ws, int method) {
    return guard([&] {
        g_last_error[0] = 0;
        if (!ws || !ws->op) return fail(FP_EINVAL, "null workspace");
        if (ws->op->iso) return fail(FP_EINVAL, "isotropic TV has one method");
        if (method != FPR_METHOD_ADBPG && method != FPR_METHOD_FISTA) return fail(FP_EINVAL, "unknown method");
        if (method == FPR_METHOD_ADBPG && !ws->op->spatial.empty())
            return fail(FP_EINVAL, "edge-weighted anisotropic TV runs dual FISTA");
        if (method == ws->method) return FP_OK;
        ws->method = method;
        ws->valid = false;
        ws->deriv_built = false;
        if (ws->numel == 0) return FP_OK;
        DevGuard dg(ws->device);
        if (method == FPR_METHOD_ADBPG) {
            ws->eng.reset();
            return build_ad(*ws);
        }
        ws->ad.reset();
        return build_engine(*ws);
    });
}

FP_API fp_status fpr_workspace_get_method(const fp_workspace *ws, int *method) {
    if (!ws || !method) return FP_EINVAL;
    *method = ws->method;
    return FP_OK;
}

FP_API fp_status fpr_workspace_polished(const fp_workspace *ws, int *polished, int64_t count, int *fallbacks) {
    if (!ws || !ws->ad || count < 0 || count > ws->B) return FP_EINVAL;
    const fpr::AdResult *R = fpr::ad_results(ws->ad.get());
    for (int64_t b = 0; b < count && polished; ++b) polished[b] = R[b].polished;
    if (fallbacks) fpr::ad_stats(ws->ad.get(), nullptr, nullptr, fallbacks);
    return FP_OK;
}

FP_API fp_status fpr_workspace_get_kernel(const fp_workspace *ws, int *generic, int *tt, int *ry) {
    if (!ws) return FP_EINVAL;
    if (generic) *generic = ws->kc.s3 ? 3 : (ws->kc.generic ? 1 : (ws->kc.stream ? 0 : 2));
    if (tt) *tt = ws->kc.TT;
    if (ry) *ry = ws->kc.RY;
    return FP_OK;
}

FP_API fp_status fpr_batch_create(const fp_op *op, fp_dtype dtype, int64_t count, const int64_t *shapes,
                                  int weighted, fpr_batch **out) {
    return guard([&] {
        if (!out) return FP_EINVAL;
        *out = nullptr;
        if (!op || count < 0 || (count > 0 && !shapes)) return FP_EINVAL;
        const int dt = enum_bits(dtype);
        if (dt != FP_F32 && dt != FP_F64) return FP_EDTYPE;
        if (count > (int64_t(1) << 24)) return FP_ESHAPE;
        std::unique_ptr<fpr_batch> b(new fpr_batch());
        b->op = op;
        b->dtype = dt;
        b->tsize = dt == FP_F32 ? 4 : 8;
        b->weighted = weighted ? 1 : 0;
        if (hipGetDevice(&b->device)) return FP_EDEVICE;
        const int k = op->k;
        b->specs.resize((size_t)count);
        int64_t total = 0;
        for (int64_t i = 0; i < count; ++i) {
            fpr::ImageSpec &sp = b->specs[(size_t)i];
            sp.L[0] = sp.L[1] = sp.L[2] = 1;
            double n = 1.0;
            for (int a = 0; a < k; ++a) {
                const int64_t e = shapes[i * k + a];
                if (e < 1) return FP_ESHAPE;
                sp.L[3 - k + a] = e;
                n *= (double)e;
            }
            if (!(n < 2147483647.0)) return FP_ESHAPE;
            const int64_t N = (int64_t)n;
            sp.q_cstride = N;
            sp.sw_cstride = N;
            total += N;
            if (!b->weighted && !op->spatial.empty())
                for (int a = 0; a < k; ++a)
                    if (sp.L[3 - k + a] != op->sshape[a]) return FP_ESHAPE;
        }
        b->maxN = 0;
        for (const fpr::ImageSpec &sp : b->specs) b->maxN = std::max(b->maxN, sp.L[0] * sp.L[1] * sp.L[2]);
        // The tuning entry of a representative image: the batch's mean
        // pixel count, as a square (or cube).
        {
            const double mean = count ? (double)total / (double)count : 1.0;
            const double side = k == 3 ? std::cbrt(mean) : std::sqrt(mean);
            for (int a = 0; a < 3; ++a) b->rep[a] = (a < 3 - k) ? 1 : std::max<int64_t>(1, (int64_t)(side + 0.5));
        }
        b->kc = fpr::default_kernel(k, (int)b->tsize, b->rep, op->iso, b->weighted != 0, count);
        if (b->kc.stream && !fpr::stream_fits((int)b->tsize, b->maxN)) b->kc = generic_kernel();
        if (b->kc.s3t() && !fpr::stream_fits((int)b->tsize, b->maxN)) b->kc.TT = 1;
        b->ptrs.assign((size_t)count, fpr::ImagePtrs{});
        b->gammas.assign((size_t)count, 0.0);
        b->gaps.assign((size_t)count, 0.0);
        b->zerrs.assign((size_t)count, 0.0);
        b->iters.assign((size_t)count, 0);
        b->conv.assign((size_t)count, 1);
        b->exps.assign((size_t)count, 0);
        b->method = (count > 0 && !b->weighted && b->uniform()) ? default_method(*op, true) : FPR_METHOD_FISTA;
        if (b->method == FPR_METHOD_ADBPG) {
            if (fp_status s = batch_build_ad(*b)) return s;
        } else if (fp_status s = batch_build(*b)) {
            return s;
        }
        *out = b.release();
        return FP_OK;
    });
}

FP_API void fpr_batch_free(fpr_batch *b) { delete b; }

FP_API fp_status fpr_batch_solve(fpr_batch *b, const fpr_image *images, const fp_options *opt, int warm,
                                 void *stream) {
    return guard([&] {
        if (!b || (!images && !b->specs.empty())) return FP_EINVAL;
        const fp_options o = opt ? *opt : defaults();
        if (!(o.tol >= 0.0) || !std::isfinite(o.tol) || o.max_iter < 1) return FP_EPARAM;
        const fp_op &op = *b->op;
        const int64_t B = (int64_t)b->specs.size();
        hipError_t err;
        const void *opw = b->weighted ? nullptr : op.device_weights(b->device, (int)b->tsize, err);
        for (int64_t i = 0; i < B; ++i) {
            const fpr_image &im = images[i];
            if (!im.x || !im.z || !im.dual) return FP_EINVAL;
            if (b->weighted && !im.weights) return FP_EINVAL;
            if (!std::isfinite(im.gamma) || im.gamma < 0.0) return FP_EPARAM;
            if (!b->weighted)
                if (fp_status s = step_ok(op, im.gamma)) return s;
            fpr::ImagePtrs &p = b->ptrs[(size_t)i];
            p.d = im.x;
            p.z = im.z;
            p.q = im.dual;
            p.sw = b->weighted ? im.weights : opw;
            b->gammas[(size_t)i] = im.gamma;
        }
        b->valid = false;
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (cur != b->device && hipSetDevice(b->device)) return FP_EDEVICE;
        if (b->ad) {
            std::vector<const void *> xs((size_t)B);
            std::vector<void *> zs((size_t)B), qs((size_t)B);
            for (int64_t i = 0; i < B; ++i) {
                xs[(size_t)i] = images[i].x;
                zs[(size_t)i] = images[i].z;
                qs[(size_t)i] = images[i].dual;
            }
            fpr::AdCall c;
            c.xs = xs.data();
            c.zs = zs.data();
            c.qs = qs.data();
            c.q_plane = B ? b->specs[0].L[0] * b->specs[0].L[1] * b->specs[0].L[2] : 0;
            c.gammas = b->gammas.data();
            c.warm = warm != 0;
            c.tol = o.tol;
            c.max_iter = o.max_iter;
            const int r = B ? fpr::ad_solve(b->ad.get(), (hipStream_t)stream, c) : 0;
            if (cur != b->device) (void)hipSetDevice(cur);
            if (r) return (fp_status)r;
            const fpr::AdResult *R = fpr::ad_results(b->ad.get());
            bool bad = false, notconv = false;
            for (int64_t i = 0; i < B; ++i) {
                if (R[i].status) bad = true;
                b->gaps[(size_t)i] = R[i].gap;
                b->iters[(size_t)i] = R[i].iters;
                b->zerrs[(size_t)i] = R[i].zerr;
                b->conv[(size_t)i] = R[i].converged;
                b->exps[(size_t)i] = R[i].e;
                if (!R[i].converged) notconv = true;
            }
            if (bad) return FP_EINVAL;
            b->valid = true;
            return notconv ? FP_ENOTCONV : FP_OK;
        }
        b->pending = false;
        const bool cap = capturing((hipStream_t)stream);
        if (cap && b->ad) {
            if (cur != b->device) (void)hipSetDevice(cur);
            return fail(FP_EINVAL, "ADBPG solves read the host between steps: not under capture");
        }
        const int r = b->eng->solve((hipStream_t)stream, b->ptrs.data(), b->gammas.data(), warm != 0, o.tol,
                                    o.max_iter);
        if (cur != b->device) (void)hipSetDevice(cur);
        if (r) return hip_status(r);
        if (cap) {  // the results are read at the first query after the graph has run
            b->pending = true;
            b->pending_stream = (hipStream_t)stream;
            b->valid = true;
            return FP_OK;
        }
        return batch_results(*b);
    });
}

FP_API fp_status fpr_batch_image_info(const fpr_batch *b, double *gaps, int *iters, double *zerrs, int *converged,
                                      int64_t count) {
    if (!b) return FP_EINVAL;
    if (fp_status st = settle_batch(b)) return st;
    if (!b->valid) return FP_EINVAL;
    if (count != (int64_t)b->specs.size()) return FP_ESHAPE;
    for (int64_t i = 0; i < count; ++i) {
        if (gaps) gaps[i] = b->gaps[(size_t)i];
        if (iters) iters[i] = b->iters[(size_t)i];
        if (zerrs) zerrs[i] = b->zerrs[(size_t)i];
        if (converged) converged[i] = b->conv[(size_t)i];
    }
    return FP_OK;
}

FP_API fp_status fpr_batch_set_kernel(fpr_batch *b, int generic, int tt, int ry) {
    return guard([&] {
        if (!b) return FP_EINVAL;
        fpr::KernelChoice kc;
        if (fp_status st = kernel_choice(generic, tt, ry, b->op->k, (int)b->tsize, b->weighted != 0, kc)) return st;
        if ((kc.stream || kc.s3t()) && !fpr::stream_fits((int)b->tsize, b->maxN))
            return fail(FP_EINVAL, "images too large for the streaming kernel");
        b->kc = kc;
        b->valid = false;
        const bool timing = b->eng && b->eng->timing;
        // a step kernel is dual FISTA's: an anisotropic batch takes that method
        b->ad.reset();
        b->method = FPR_METHOD_FISTA;
        fp_status st = batch_build(*b);
        if (!st) b->eng->timing = timing;
        return st;
    });
}

FP_API fp_status fpr_batch_set_method(fpr_batch *b, int method) {
    return guard([&] {
        g_last_error[0] = 0;
        if (!b || !b->op) return fail(FP_EINVAL, "null batch");
        if (b->op->iso) return fail(FP_EINVAL, "isotropic TV has one method");
        if (method != FPR_METHOD_ADBPG && method != FPR_METHOD_FISTA) return fail(FP_EINVAL, "unknown method");
        if (method == b->method) return FP_OK;
        if (method == FPR_METHOD_ADBPG && (b->weighted || !b->op->spatial.empty() || !b->uniform() || b->specs.empty()))
            return fail(FP_EINVAL, "ADBPG takes uniform batches without edge weights");
        b->method = method;
        b->valid = false;
        DevGuard dg(b->device);
        if (method == FPR_METHOD_ADBPG) {
            b->eng.reset();
            return batch_build_ad(*b);
        }
        b->ad.reset();
        return batch_build(*b);
    });
}

FP_API fp_status fpr_batch_get_method(const fpr_batch *b, int *method) {
    if (!b || !method) return FP_EINVAL;
    *method = b->method;
    return FP_OK;
}

FP_API fp_status fpr_batch_get_kernel(const fpr_batch *b, int *generic, int *tt, int *ry) {
    if (!b) return FP_EINVAL;
    if (generic) *generic = b->kc.s3 ? 3 : (b->kc.generic ? 1 : (b->kc.stream ? 0 : 2));
    if (tt) *tt = b->kc.TT;
    if (ry) *ry = b->kc.RY;
    return FP_OK;
}

FP_API fp_status fpr_batch_set_tail(fpr_batch *b, int tt, int ry, double frac) {
    return guard([&] {
        if (!b || !b->eng) return FP_EINVAL;
        fpr::KernelChoice kc;
        kc.stream = true;
        kc.TT = tt;
        kc.RY = ry;
        // the tail runs at least the main kernel's iterations per launch
        if (frac > 0.0 && (!b->kc.stream || b->kc.generic || !fpr::kernel_available((int)b->tsize, true, tt, ry) ||
                           kc.iters() < b->kc.iters()))
            return FP_EINVAL;
        DevGuard dg(b->device);
        return hip_status(b->eng->set_tail(kc, frac));
    });
}

FP_API fp_status fpr_workspace_set_tail(fp_workspace *ws, int tt, int ry, double frac) {
    return guard([&] {
        if (!ws || !ws->eng) return FP_EINVAL;
        fpr::KernelChoice kc;
        kc.stream = true;
        kc.TT = tt;
        kc.RY = ry;
        if (frac > 0.0 && (!ws->kc.stream || ws->kc.generic || !fpr::kernel_available((int)ws->tsize, true, tt, ry) ||
                           kc.iters() < ws->kc.iters()))
            return FP_EINVAL;
        DevGuard dg(ws->device);
        return hip_status(ws->eng->set_tail(kc, frac));
    });
}

FP_API fp_status fpr_batch_scale_exps(const fpr_batch *b, int *exps, int64_t count) {
    if (!b || !exps) return FP_EINVAL;
    if (fp_status st = settle_batch(b)) return st;
    if (!b->valid) return FP_EINVAL;
    if (count != (int64_t)b->specs.size()) return FP_ESHAPE;
    for (int64_t i = 0; i < count; ++i) exps[i] = b->exps[(size_t)i];
    return FP_OK;
}

FP_API fp_status fpr_batch_copy_dual_iterate(const fpr_batch *b, void *dst, void *stream) {
    return guard([&] {
        if (!b || !dst || !b->valid || (!b->eng && !b->ad)) return FP_EINVAL;
        if (!b->uniform()) return FP_ESHAPE;
        const int64_t N = b->specs.empty() ? 0 : b->specs[0].L[0] * b->specs[0].L[1] * b->specs[0].L[2];
        DevGuard dg(b->device);
        if (b->ad) return (fp_status)fpr::ad_copy_edge_dual(b->ad.get(), (hipStream_t)stream, dst, (int64_t)b->op->k * N);
        return hip_status(b->eng->copy_dual_iterate((hipStream_t)stream, dst, (int64_t)b->op->k * N));
    });
}

FP_API fp_status fpr_deriv_create(const fp_op *op, fp_dtype dtype, int64_t count, const int64_t *img_shape,
                                  const void *x, const void *z, const void *dual, const void *spatial,
                                  int spatial_per_image, const double *gammas, const double *gaps,
                                  const int *scale_exps, void *stream, fpr_deriv **out) {
    return guard([&] {
        if (!out) return FP_EINVAL;
        *out = nullptr;
        if (!op || !img_shape || count <= 0 || !x || !z || !dual || !gammas || (op->iso && !gaps)) return FP_EINVAL;
        const int dt = enum_bits(dtype);
        if (dt != FP_F32 && dt != FP_F64) return FP_EDTYPE;
        if (!op->spatial.empty() && spatial) return FP_EINVAL;  // the op's own map, or the caller's
        std::unique_ptr<fpr_deriv> d(new fpr_deriv());
        fpr::DerivSpec sp;
        sp.iso = op->iso;
        sp.k = op->k;
        int64_t N = 1;
        for (int a = 0; a < op->k; ++a) {
            if (img_shape[a] < 1) return FP_ESHAPE;
            sp.L[3 - op->k + a] = img_shape[a];
            N *= img_shape[a];
        }
        sp.B = count;
        sp.tsize = dt == FP_F32 ? 4 : 8;
        for (int a = 0; a < op->k; ++a) sp.w[a] = op->w[a];
        sp.x = x;
        sp.z = z;
        sp.u = dual;
        sp.u_img = (int64_t)op->k * N;
        sp.u_plane = N;
        if (!op->spatial.empty()) {
            hipError_t err;
            int dev = 0;
            (void)hipGetDevice(&dev);
            sp.sw = op->device_weights(dev, sp.tsize, err);
            if (err) return hip_status(err);
            sp.sw_img = 0;
        } else {
            sp.sw = spatial;
            sp.sw_img = spatial && spatial_per_image ? (op->iso ? N : (int64_t)op->k * N) : 0;
        }
        sp.sw_plane = N;
        sp.gamma = gammas;
        sp.gap = gaps;
        sp.e = scale_exps;
        d->B = count;
        d->k = op->k;
        d->ns = sp.sw ? (op->iso ? N : (int64_t)op->k * N) * (sp.sw_img ? count : 1) : 0;
        (void)hipGetDevice(&d->device);
        if (fpr::dev_alloc((void **)&d->scr, 8 * (size_t)(count + 4))) return FP_ENOMEM;
        fpr::Deriv *dv = nullptr;
        if (int r = fpr::deriv_create(sp, (hipStream_t)stream, &dv)) return (fp_status)r;
        d->d.reset(dv);
        *out = d.release();
        return FP_OK;
    });
}

FP_API void fpr_deriv_free(fpr_deriv *d) { delete d; }

FP_API fp_status fpr_deriv_vjp(fpr_deriv *d, const void *g, void *gx, double *ggamma, double *gw, double *gspatial,
                               double tol, int max_iter, void *stream) {
    return guard([&] {
        if (!d || !g || !gx || !(tol >= 0.0) || max_iter < 1) return FP_EINVAL;
        if (gspatial && !d->ns) return FP_EINVAL;
        DevGuard dg(d->device);
        hipStream_t s = (hipStream_t)stream;
        if (int r = fpr::deriv_vjp(d->d.get(), s, g, gx, ggamma, d->scr, gspatial, tol, max_iter))
            return (fp_status)r;
        double h[3] = {0.0, 0.0, 0.0};
        if (hipMemcpyAsync(h, d->scr, 8 * (size_t)d->k, hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
        if (gw)
            for (int a = 0; a < d->k; ++a) gw[a] = h[a];
        return FP_OK;
    });
}

FP_API fp_status fpr_deriv_jvp(fpr_deriv *d, const void *dx, const double *dgamma, const double *dw,
                               const double *dspatial, void *dz, double tol, int max_iter, void *stream) {
    return guard([&] {
        if (!d || !dz || !(tol >= 0.0) || max_iter < 1) return FP_EINVAL;
        if (dspatial && !d->ns) return FP_EINVAL;
        DevGuard dg(d->device);
        hipStream_t s = (hipStream_t)stream;
        if (int r = fpr::deriv_jvp(d->d.get(), s, dx, dgamma, dw, dspatial, dz, tol, max_iter)) return (fp_status)r;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
        return FP_OK;
    });
}

FP_API fp_status fpr_deriv_info(const fpr_deriv *d, double *residual, int *iters, int64_t *regions) {
    if (!d) return FP_EINVAL;
    fpr::deriv_result(d->d.get(), residual, iters, regions);
    return FP_OK;
}

FP_API fp_status fpr_deriv_image_info(const fpr_deriv *d, double *residuals, int *iters, int64_t count) {
    if (!d) return FP_EINVAL;
    if (count != d->B) return FP_ESHAPE;
    fpr::deriv_image_result(d->d.get(), residuals, iters);
    return FP_OK;
}

FP_API fp_status fpr_server_create(const fp_op *op2d, const fp_op *op3d, fp_dtype dtype, size_t pool_bytes,
                                   int slots, fpr_server **out) {
    return guard([&] {
        if (!out) return FP_EINVAL;
        *out = nullptr;
        if ((!op2d && !op3d) || slots < 1 || pool_bytes < ((size_t)1 << 20)) return FP_EINVAL;
        if ((op2d && op2d->k != 2) || (op3d && op3d->k != 3)) return FP_EINVAL;
        if ((op2d && !op2d->spatial.empty()) || (op3d && !op3d->spatial.empty())) return FP_EINVAL;
        const int dt = enum_bits(dtype);
        if (dt != FP_F32 && dt != FP_F64) return FP_EDTYPE;
        const int tsize = dt == FP_F32 ? 4 : 8;
        fpr::Problem probs[2] = {};
        fpr::KernelChoice kcs[2];
        const fp_op *ops[2] = {op2d, op3d};
        for (int c = 0; c < 2; ++c) {
            const fp_op *op = ops[c];
            if (!op) continue;
            fpr::Problem &P = probs[c];
            P.k = op->k;
            P.iso = op->iso ? 1 : 0;
            P.sw = 0;
            P.sw_per_image = 0;
            for (int a = 0; a < 3; ++a) P.w[a] = op->w[a];
            P.sched = op->iso ? fpr::kSchedC : fpr::kSchedDevice;
            P.wmax_norm = op->iso ? 1 : 0;
            P.alpha_max = 1.0;
            kcs[c] = fpr::serve_kernel(op->k, tsize, op->iso);
        }
        std::unique_ptr<fpr_server> sv(new fpr_server());
        sv->tsize = tsize;
        if (hipGetDevice(&sv->device)) return FP_EDEVICE;
        sv->s.reset(new fpr::Server());
        if (int e = sv->s->setup(probs, kcs, tsize, slots, pool_bytes)) return hip_status(e);
        *out = sv.release();
        return FP_OK;
    });
}

FP_API void fpr_server_free(fpr_server *s) { delete s; }

FP_API fp_status fpr_server_submit(fpr_server *s, const fpr_request *r, double t_submit, int64_t *id) {
    return guard([&] {
        if (!s || !r) return FP_EINVAL;
        fpr::ServeRequest q;
        q.k = r->ndim;
        if (q.k != 2 && q.k != 3) return FP_EINVAL;
        q.has_xs = r->strides != nullptr;
        for (int a = 0; a < q.k; ++a) {
            q.L[3 - q.k + a] = r->shape[a];
            q.xs[3 - q.k + a] = q.has_xs ? r->strides[a] : 0;
        }
        q.caller = (hipStream_t)r->stream;
        q.x = r->x;
        q.z = r->z;
        q.gamma = r->gamma;
        q.tol = r->tol;
        q.max_iter = r->max_iter;
        q.keep_dual = r->keep_dual;
        q.warm_from = r->warm_from;
        q.t_submit = t_submit;
        DevGuard dg(s->device);
        return (fp_status)s->s->submit(q, id);
    });
}

FP_API fp_status fpr_server_step(fpr_server *s, int *busy, int *queued) {
    return guard([&] {
        if (!s) return FP_EINVAL;
        DevGuard dg(s->device);
        const int e = s->s->step(busy, queued);
        return e ? hip_status(e) : FP_OK;
    });
}

FP_API int64_t fpr_server_poll(fpr_server *s, fpr_result *out, int64_t max) {
    if (!s || !out || max <= 0) return 0;
    // results move through a small buffer on the stack: nothing allocates
    // and nothing can throw across the boundary
    fpr::ServeResult buf[64];
    int64_t n = 0;
    while (n < max) {
        const int64_t want = std::min<int64_t>(64, max - n);
        const int64_t got = s->s->poll(buf, want);
        for (int64_t i = 0; i < got; ++i) {
            const fpr::ServeResult &R = buf[i];
            out[n + i] = fpr_result{R.id, R.status, R.converged, R.iters, R.gap, R.zerr, R.t_submit, R.t_admit,
                                    R.t_done};
        }
        n += got;
        if (got < want) break;
    }
    return n;
}

FP_API double fpr_server_clock(const fpr_server *s) { return s ? s->s->now() : 0.0; }

FP_API fp_status fpr_server_release(fpr_server *s, int64_t id) {
    if (!s) return FP_EINVAL;
    return s->s->release_dual(id) ? FP_OK : FP_EINVAL;
}

FP_API fp_status fpr_server_stats(const fpr_server *s, int64_t *launches) {
    if (!s) return FP_EINVAL;
    if (launches) *launches = s->s->launches();
    return FP_OK;
}

FP_API fp_status fpr_server_get_kernel(const fpr_server *s, int k, int *kind, int *tt, int *ry) {
    if (!s || (k != 2 && k != 3)) return FP_EINVAL;
    const fpr::KernelChoice *kc = s->s->kernel(k - 2);
    if (!kc) return FP_EINVAL;
    if (kind) *kind = kc->s3 ? 3 : (kc->generic ? 1 : (kc->stream ? 0 : 2));
    if (tt) *tt = kc->TT;
    if (ry) *ry = kc->RY;
    return FP_OK;
}

FP_API fp_status fpr_batch_stats(const fpr_batch *b, int *launches, int *host_syncs) {
    if (!b) return FP_EINVAL;
    if (b->ad) {
        fpr::ad_stats(b->ad.get(), launches, host_syncs, nullptr);
        return FP_OK;
    }
    if (launches) *launches = b->eng ? b->eng->stats.launches : 0;
    if (host_syncs) *host_syncs = b->eng ? b->eng->stats.host_syncs : 0;
    return FP_OK;
}

FP_API fp_status fpr_batch_pixel_passes(const fpr_batch *b, double *pixel_passes) {
    if (!b || !pixel_passes) return FP_EINVAL;
    double n = 0.0;
    if (b->eng && b->valid) {
        const fpr::ImgState *S = b->eng->state();
        for (size_t i = 0; i < b->specs.size(); ++i) {
            const fpr::ImageSpec &sp = b->specs[i];
            n += (double)(sp.L[0] * sp.L[1] * sp.L[2]) * (double)S[i].passes;
        }
    }
    *pixel_passes = n;
    return FP_OK;
}

FP_API fp_status fpr_batch_set_timing(fpr_batch *b, int on) {
    if (!b || !b->eng) return FP_EINVAL;
    b->eng->timing = on != 0;
    return FP_OK;
}

FP_API fp_status fpr_batch_set_graph(fpr_batch *b, int on) {
    if (!b || !b->eng) return FP_EINVAL;
    b->eng->graphs = on != 0;
    return FP_OK;
}

FP_API fp_status fpr_batch_timing(const fpr_batch *b, double *step_ms, double *epilogue_ms, int *launches) {
    if (!b || !b->eng) return FP_EINVAL;
    if (step_ms) *step_ms = b->eng->stats.step_ms;
    if (epilogue_ms) *epilogue_ms = b->eng->stats.epi_ms;
    if (launches) *launches = b->eng->stats.timed;
    return FP_OK;
}

FP_API fp_status fpr_batch_busy(const fpr_batch *b, double *busy_ms) {
    if (!b || !b->eng || !busy_ms) return FP_EINVAL;
    *busy_ms = b->eng->stats.busy_ms;
    return FP_OK;
}

FP_API fp_status fpr_workspace_stats(const fp_workspace *ws, int *launches, int *host_syncs) {
    if (!ws) return FP_EINVAL;
    if (ws->ad) {
        fpr::ad_stats(ws->ad.get(), launches, host_syncs, nullptr);
        return FP_OK;
    }
    if (launches) *launches = ws->eng ? ws->eng->stats.launches : 0;
    if (host_syncs) *host_syncs = ws->eng ? ws->eng->stats.host_syncs : 0;
    return FP_OK;
}

}  // extern "C"

