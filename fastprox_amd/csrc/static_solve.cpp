// The allocation-free solve behind fpr_prox_static: fp_prox on device arrays
// without a workspace, whose device scratch (the dual iterates, the per-image
// descriptors and states, the tile lists, the partial sums, the dual output,
// the staged input and weights) is carved from one fixed array of this file,
// `scratch`, instead of from buffers the library allocates. For callers that
// cannot allocate device memory.
//
// Every GPU the library runs on holds its own copy of `scratch` (kScratchBytes)
// until the process ends. Its address on a device is read once, at the first
// call on that device, with hipGetSymbolAddress; host code then only uses
// that address (the array's name is a host-side stand-in). Every call on a
// GPU shares that GPU's copy: calls that overlap on one GPU race, so a caller
// serialises them (one stream, or its own lock). Every buffer carved from it
// is written before it is read in each call (the prologue fills the states
// and the counters are reset by a kernel), so nothing of an earlier call's
// contents is used.
//
// SPDX-License-Identifier: MIT
#include "static_solve.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "engine.hpp"

namespace fpr {

static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];

namespace {

struct DeviceSlot {
    bool ready = false;
    unsigned char *ptr = nullptr;  // this device's copy of scratch
    Engine *eng = nullptr;         // lends the solve loop its events (graphs and timing off)
    ImgDesc hdesc[kStaticMaxImages];
    ImgState hstate[kStaticMaxImages];
    double hgam[kStaticMaxImages];
    int hcount[2];
};

DeviceSlot g_dev[kStaticMaxDevices];
alignas(Engine) unsigned char g_eng_store[kStaticMaxDevices][sizeof(Engine)];
std::mutex g_setup_mu;

constexpr size_t kAlign = 256;
size_t round_up(size_t n) { return (n + kAlign - 1) / kAlign * kAlign; }

// The device's scratch address and the loop's events, once per device, with
// the device current.
int setup_device(int dev, DeviceSlot **out) {
    if (dev < 0 || dev >= kStaticMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_setup_mu);
    DeviceSlot &D = g_dev[dev];
    if (!D.ready) {
        void *p = nullptr;
        if (hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch))) return e;
        if (!p) return hipErrorInvalidSymbol;
        Engine *E = new (g_eng_store[dev]) Engine();  // static storage: nothing allocated
        E->graphs = false;
        E->timing = false;
        E->allow_split = false;
        if (int e = E->ensure_events(E->steps_per_sync, false)) return e;
        D.ptr = (unsigned char *)p;
        D.eng = E;
        D.ready = true;
    }
    *out = &D;
    return hipSuccess;
}

// The scratch layout of a solve, in bytes from the array's start.
struct Layout {
    size_t arena, desc, state, part, list0, list1, counts, gam, propart, q, xstage, swd, swt, end;
};

Layout layout_of(const StaticSolve &c, int tiles) {
    const size_t B = (size_t)c.B, N = (size_t)c.N, k = (size_t)c.P.k, ts = (size_t)c.tsize;
    Layout L{};
    size_t at = 0;
    auto take = [&](size_t bytes) {
        const size_t off = at;
        at += round_up(bytes ? bytes : 1);
        return off;
    };
    L.arena = take(4 * k * N * B * ts);
    L.desc = take(sizeof(ImgDesc) * B);
    L.state = take(sizeof(ImgState) * B);
    L.part = take(sizeof(double) * kSlots * (size_t)tiles);
    L.list0 = take(sizeof(Tile) * (size_t)tiles);
    L.list1 = take(sizeof(Tile) * (size_t)tiles);
    L.counts = take(64);
    L.gam = take(sizeof(double) * B);
    L.propart = take(sizeof(double) * 4 * B * (size_t)prologue_blocks((int64_t)N));
    L.q = take(k * B * N * ts);
    L.xstage = c.stage_x ? take(B * N * ts) : at;
    const size_t nsw = c.sw_host ? (c.P.iso ? N : k * N) : 0;
    L.swd = nsw ? take(nsw * sizeof(double)) : at;
    L.swt = (nsw && ts == 4) ? take(nsw * 4) : at;
    L.end = at;
    return L;
}

int tiles_of(const StaticSolve &c, const KernelChoice &kc) {
    const int64_t ntx = (c.L[2] + kc.tile_x() - 1) / kc.tile_x();
    const int64_t nty = (c.L[1] + kc.tile_y() - 1) / kc.tile_y();
    const int64_t t = ntx * nty * kc.tiles_z(c.L[0]);
    return (int)(t * c.B);
}

}  // namespace

size_t static_bytes_needed(const StaticSolve &c) {
    return layout_of(c, tiles_of(c, c.kc)).end;
}

int static_solve(const StaticSolve &c, StaticResult *res) {
    if (c.B < 1 || c.B > kStaticMaxImages) return hipErrorInvalidValue;
    int dev = 0;
    if (hipError_t e = hipGetDevice(&dev)) return e;
    DeviceSlot *D = nullptr;
    if (int e = setup_device(dev, &D)) return e;
    const KernelChoice &kc = c.kc;
    const int tiles = tiles_of(c, kc);
    const Layout Ly = layout_of(c, tiles);
    if (Ly.end > kStaticScratchBytes) return hipErrorInvalidValue;
    unsigned char *base = D->ptr;
    hipStream_t s = c.stream;
    const size_t N = (size_t)c.N, k = (size_t)c.P.k, ts = (size_t)c.tsize;
    const int B = c.B;
    // the input: the caller's array, or its copy when the output overwrites it
    const void *d = c.x;
    if (c.stage_x) {
        if (hipError_t e = hipMemcpyAsync(base + Ly.xstage, c.x, (size_t)B * N * ts, hipMemcpyDeviceToDevice, s))
            return e;
        d = base + Ly.xstage;
    }
    // the operator's spatial weights (one map for every image), in the dtype
    const void *sw = nullptr;
    if (c.sw_host) {
        const size_t nsw = c.P.iso ? N : k * N;
        if (hipError_t e = hipMemcpyAsync(base + Ly.swd, c.sw_host, nsw * sizeof(double), hipMemcpyHostToDevice, s))
            return e;
        if (ts == 4) {
            launch_convert<double, float>((const double *)(base + Ly.swd), (float *)(base + Ly.swt), (int64_t)nsw, s);
            sw = base + Ly.swt;
        } else {
            sw = base + Ly.swd;
        }
    }
    int tile_base = 0;
    for (int b = 0; b < B; ++b) {
        ImgDesc &I = D->hdesc[b];
        I = ImgDesc{};
        for (int a = 0; a < 3; ++a) I.L[a] = (int32_t)c.L[a];
        I.d = (const char *)d + (size_t)b * N * ts;
        I.z = (char *)c.z + (size_t)b * N * ts;
        I.q = base + Ly.q + (size_t)b * k * N * ts;
        I.sw = sw;
        I.st_off = (int64_t)b * 4 * (int64_t)k * (int64_t)N;
        I.q_cstride = (int64_t)N;
        I.sw_cstride = (int64_t)N;
        I.ntx = (int32_t)((c.L[2] + kc.tile_x() - 1) / kc.tile_x());
        I.nty = (int32_t)((c.L[1] + kc.tile_y() - 1) / kc.tile_y());
        I.ntz = (int32_t)kc.tiles_z(c.L[0]);
        I.ntiles = I.ntx * I.nty * I.ntz;
        I.tile_base = tile_base;
        tile_base += I.ntiles;
        D->hgam[b] = c.gamma;
    }
    hipError_t e = hipMemcpyAsync(base + Ly.desc, D->hdesc, sizeof(ImgDesc) * (size_t)B, hipMemcpyHostToDevice, s);
    if (!e) e = hipMemcpyAsync(base + Ly.gam, D->hgam, sizeof(double) * (size_t)B, hipMemcpyHostToDevice, s);
    if (e) return e;
    Problem P = c.P;
    P.TT = kc.iters();
    P.warm = 0;
    P.tol = c.tol;
    P.max_iter = c.max_iter;
    bool tv1d = false;  // one active axis: the exact 1D kernel (the engine's rule)
    if (P.iso && !P.sw) {
        int active = 0;
        for (int a = 0; a < P.k; ++a)
            if (c.L[3 - P.k + a] > 1 && P.w[a] > 0.0) ++active;
        tv1d = active == 1;
    }
    Tile *lists[2] = {(Tile *)(base + Ly.list0), (Tile *)(base + Ly.list1)};
    Engine &E = *D->eng;
    int r;
    if (ts == 4)
        r = run_solve_in<float>(E, P, kc, s, base + Ly.arena, (ImgDesc *)(base + Ly.desc),
                                (ImgState *)(base + Ly.state), (double *)(base + Ly.part), lists,
                                (int *)(base + Ly.counts), (double *)(base + Ly.gam), (double *)(base + Ly.propart),
                                B, tiles, (int64_t)N, tv1d, D->hcount, E.steps_per_sync, c.max_iter);
    else
        r = run_solve_in<double>(E, P, kc, s, base + Ly.arena, (ImgDesc *)(base + Ly.desc),
                                 (ImgState *)(base + Ly.state), (double *)(base + Ly.part), lists,
                                 (int *)(base + Ly.counts), (double *)(base + Ly.gam), (double *)(base + Ly.propart),
                                 B, tiles, (int64_t)N, tv1d, D->hcount, E.steps_per_sync, c.max_iter);
    if (r) return r;
    if ((e = hipMemcpyAsync(D->hstate, base + Ly.state, sizeof(ImgState) * (size_t)B, hipMemcpyDeviceToHost, s)))
        return e;
    if ((e = hipStreamSynchronize(s))) return e;
    // per-image results at the input's scale (the workspace path's rules)
    StaticResult R{};
    for (int b = 0; b < B; ++b) {
        const ImgState &t = D->hstate[b];
        if (t.status) R.bad = true;
        double g = t.gp;
        if (P.iso && g == g) g = g < 0.0 ? 0.0 : g;
        if (P.iso && t.e) g = std::ldexp(g, 2 * t.e);
        R.gap += g;
        R.iters = t.it > R.iters ? t.it : R.iters;
        if (!t.converged) R.notconv = true;
    }
    *res = R;
    return hipSuccess;
}

}  // namespace fpr
