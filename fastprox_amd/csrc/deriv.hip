// Derivatives of image TV at a solution, on the device.
//
// Isotropic TV (src/ops/tv_iso.cpp:551-631 save, :672-1067 Sys, cg, jvp_img,
// vjp_img): z - d + G^T u = 0 with G = W D. Pixels with the dual strictly
// inside the ball (|u_p| < r_p (1 - 1e-9)), or a gradient at the rounding
// level of the data (|g_p| <= w_max 64 2.2e-16 max|d|; a float32 dual: 64
// eps and max(1e-9, 64 eps) instead), are flat: z is
// constant across their forward edges, so dz lives on the connected regions
// of those edges (one unknown per region). On the others the dual moves on
// the sphere, du_p = H_p (G dz)_p with H_p = r_p / max(|g_p|, mu) (I - n n^T)
// and mu = 0.1 w_max sqrt(2 gap / N), which gives the SPD system
// M^T (I + G^T H G) M c = M^T rhs in region space, solved by conjugate
// gradients with the Jacobi preconditioner.
//
// Anisotropic TV, the device method's derivative
// (bindings/python/fastprox/_tv_torch.py:346-362, :528-585): flat edges
// (|u_e| < r_e (1 - eta), or a jump at the rounding level with r_e > 0) join
// their pixels; dz is the region mean of dx - sum_a dw_a D_a^T u_a -
// G^T (dr n), n = sign(u) on the other edges.
//
// Layout: a uniform batch of B images, pixel gp = b N + p. The structure:
//   labels: union-find over the joined edges with atomic hooks of the larger
//     root under the smaller (the component's smallest index ends as its
//     root whatever the order), full compression, an exclusive scan of the
//     roots (global region ids, ordered by image, then by first pixel);
//   members: a stable radix sort of (label, pixel), so each region's pixels
//     are contiguous and in index order; rstart marks each region's start.
// Every sum is in a fixed order (no floating-point atomics), so results
// repeat bit for bit run to run:
//   region sums: blocks of 4096 member positions (256 threads x 16), each
//     thread summing its runs in order, a block-wide segmented scan joining
//     runs across threads, and the pieces of regions that cross blocks added
//     in block order by the region's consumer;
//   per-image dot products in region space: blocks of 1024 regions of one
//     image, a fixed tree in each, the block partials added in order;
//   per-image pixel sums: blocks of 4096 pixels, the same way.
// CG runs every image of the batch in lockstep, each with its own scalars;
// an image that meets its stopping rule is frozen where the reference's
// per-image loop would stop (tv_iso.cpp:919-942), so iteration counts and
// residuals are per image. The host checks for a live image every 8
// iterations.
//
// SPDX-License-Identifier: MIT
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "alloc.hpp"
#include "../../include/fastprox_rocm.h"
#include "deriv.hpp"
#include "regions.hpp"

namespace fpr {
namespace {

constexpr double kDualEta = 1e-9;     // tv_iso.cpp:109
constexpr double kCurvRms = 0.1;      // tv_iso.cpp:113
constexpr double kDerivFloor = 1e-10; // tv_iso.cpp:115
constexpr int kThr = 256;
constexpr int kPerThr = 16;
constexpr int kChunk = kThr * kPerThr;  // member positions per region-sum block
constexpr int kRegPer = 4;
constexpr int kRB = kThr * kRegPer;     // regions per region-space block
constexpr int kPixBlk = kThr * 16;      // pixels per block of the per-image pixel sums
constexpr int kCheck = 8;               // CG iterations between host checks
constexpr int kMaxQ = 7;                // per-image pixel sums (2k + 1)
static_assert(kThr == kRsThr && kPerThr == kRsPer && kChunk == kRsChunk, "the region sums' blocks");

struct Geo {
    int k;
    int64_t L[3], S[3], N, BN;
};

struct DImg {
    double gamma_s;   // the step at the solved scale
    double sc;        // 2^e
    double gap_s;     // the gap at the solved scale
    double thr, rel, mu;
    double wsave[3];  // iso: the structure's weights (the solve's active axes)
    double wa[3];     // iso: the system's weights; aniso: w
    double floor_, eta;  // the flat test: the jumps' rounding floor (aniso), the dual margin
    int e, nan, from_dual, use_dual;
    int rbase, rcount, bbase, bcount;
};

struct CgState {
    double bn, rz, res, tol;
    int it, active, max_iter, pad_;
};

struct RBlock {
    int img, c0, c1, pad_;
};

// Pixel indices of the structure are below B N < 2^31 (checked when it is
// built), so coordinates and image indices come from 32-bit unsigned
// division, a short sequence on the GPU where 64-bit division is a long one.
__device__ __forceinline__ void coords(const Geo &G, int64_t p, int64_t c[3]) {
    const uint32_t q = (uint32_t)p, L2 = (uint32_t)G.L[2], L1 = (uint32_t)G.L[1];
    const uint32_t t = q / L2, u = t / L1;
    c[2] = q - t * L2;
    c[1] = t - u * L1;
    c[0] = u;
}
__device__ __forceinline__ int img_of(const Geo &G, int64_t gp) { return (int)((uint32_t)gp / (uint32_t)G.N); }

template <class T>
__device__ __forceinline__ double ld(const T *a, int64_t i) {
    return (double)a[i];
}

__device__ __forceinline__ double dmaxd(double a, double b) { return a < b ? b : a; }

// ---------------------------------------------------------------------------
// Fixed-order block reductions
// ---------------------------------------------------------------------------

__device__ double block_sum(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = kThr / 2; d > 0; d >>= 1) {
        if (t < d) sh[t] = sh[t] + sh[t + d];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double *sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int d = kThr / 2; d > 0; d >>= 1) {
        if (t < d) sh[t] = dmaxd(sh[t], sh[t + d]);
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Structure
// ---------------------------------------------------------------------------

// Per image: NaN, max |d_s| and |d_s|^2 (d_s = 2^-e x), then the flat
// threshold, the certified relative error and the curvature cap
// (tv_iso.cpp:557-569); aniso: the rounding floor of the jumps
// (_tv_torch.py:348-350).
// iso_floor_eps, iso_eta: the flat test's rounding floor (w_max 64
// iso_floor_eps max|d|) and dual margin: for a float64 dual the C kernels'
// 2.2e-16 and 1e-9 (tv_iso.cpp:109, 565); for a float32 dual the torch
// device path's 2^-23 and max(1e-9, 64 2^-23) (_tv_torch.py:192-195, 404-410),
// since a float32 dual sits on its sphere only to float32 rounding.
template <class T>
__global__ __launch_bounds__(kThr) void prep_kernel(Geo G, int iso, double eps, double iso_floor_eps, double iso_eta,
                                                    const T *x, DImg *di) {
    __shared__ double sh[kThr];
    const int b = blockIdx.x;
    DImg &D = di[b];
    const T *xb = x + (int64_t)b * G.N;
    double m = 0.0, ss = 0.0, nanv = 0.0;
    for (int64_t i = threadIdx.x; i < G.N; i += kThr) {
        const double v = ldexp(ld(xb, i), -D.e);
        if (v != v) nanv = 1.0;
        m = dmaxd(m, fabs(v));
    }
    m = block_max(m, sh);
    nanv = block_max(nanv, sh);
    for (int64_t i = threadIdx.x; i < G.N; i += kThr) {
        const double v = ldexp(ld(xb, i), -D.e);
        ss += v * v;
    }
    ss = block_sum(ss, sh);
    if (threadIdx.x == 0) {
        D.nan = nanv != 0.0;
        if (iso) {
            double wm = 0.0;
            for (int a = 0; a < G.k; ++a) wm = dmaxd(wm, D.wsave[a]);
            D.from_dual = D.from_dual && !D.nan;
            D.use_dual = D.use_dual && !D.nan;
            const double gap = D.nan ? 0.0 : D.gap_s;
            D.thr = D.from_dual ? wm * 64.0 * iso_floor_eps * m : 0.0;
            D.eta = iso_eta;
            D.rel = ss > 0.0 ? sqrt(2.0 * dmaxd(gap, 0.0) / ss) : 0.0;
            D.mu = D.from_dual ? kCurvRms * wm * sqrt(2.0 * dmaxd(gap, 0.0) / (double)G.N) : 0.0;
        } else {
            double wm = 0.0;
            for (int a = 0; a < G.k; ++a) wm = dmaxd(wm, D.wa[a]);
            D.floor_ = wm * 64.0 * eps * m;
            D.eta = dmaxd(1e-9, 64.0 * eps);
            D.rel = 0.0;
            D.mu = 0.0;
        }
    }
}

template <class T>
__device__ __forceinline__ double radius(const DImg &D, const T *sw, int64_t sw_img, int b, int64_t p) {
    if (!sw) return D.gamma_s;
    const double a = ld(sw, b * sw_img + p);
    if (a == 0.0) return 0.0;
    return D.gamma_s * a;
}

// Iso: gz = W D z_s (the system's weights), the flat test (the structure's
// weights), and the union-find parents.
template <class T>
__global__ __launch_bounds__(kThr) void flat_iso_kernel(Geo G, const T *x, const T *z, const T *u, int64_t u_img,
                                                        int64_t u_plane, const T *sw, int64_t sw_img, const DImg *di,
                                                        double *gz, uint8_t *flat, int *parent) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    const T *zsrc = D.nan ? x : z;
    int64_t c[3];
    coords(G, p, c);
    const double z0 = ldexp(ld(zsrc, gp), -D.e);
    double nn = 0.0;
    bool any = false;
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        const bool nx = c[ax] + 1 < G.L[ax];
        const double z1 = nx ? ldexp(ld(zsrc, gp + G.S[ax]), -D.e) : 0.0;
        gz[a * G.BN + gp] = (D.wa[a] != 0.0 && nx) ? D.wa[a] * (z1 - z0) : 0.0;
        if (D.wsave[a] != 0.0 && nx) {
            const double t = D.wsave[a] * (z1 - z0);
            nn += t * t;
            any = true;
        }
    }
    bool fl = !any || sqrt(nn) <= D.thr;
    if (D.from_dual && !fl) {
        double uu = 0.0;
        for (int a = 0; a < G.k; ++a) {
            const double ua = ldexp(ld(u, b * u_img + a * u_plane + p), -D.e);
            uu += ua * ua;
        }
        fl = sqrt(uu) < radius(D, sw, sw_img, b, p) * (1.0 - D.eta);
    }
    flat[gp] = (uint8_t)((fl ? 1 : 0) | (any ? 2 : 0));
    parent[gp] = (int)gp;
}

// Aniso: the flat edges (bit a of flat) and the signs n on the others.
template <class T>
__global__ __launch_bounds__(kThr) void flat_aniso_kernel(Geo G, const T *x, const T *z, const T *u, int64_t u_img,
                                                          int64_t u_plane, const T *sw, int64_t sw_img,
                                                          int64_t sw_plane, const DImg *di, int8_t *nsg, uint8_t *flat,
                                                          int *parent) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    int64_t c[3];
    coords(G, p, c);
    const double z0 = ld(z, gp);
    uint8_t bits = 0;
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        const bool nx = c[ax] + 1 < G.L[ax];
        const double be = sw ? ld(sw, b * sw_img + a * sw_plane + p) : 1.0;
        const double r = sw ? D.gamma_s * be : D.gamma_s;
        const double inside = r * (1.0 - D.eta);
        const double gj = nx ? D.wa[a] * (ld(z, gp + G.S[ax]) - z0) : 0.0;
        const double ua = ld(u, b * u_img + a * u_plane + p);
        const bool fl = fabs(ua) < inside || (fabs(gj) <= D.floor_ && inside > 0.0);
        double s = ua > 0.0 ? 1.0 : (ua < 0.0 ? -1.0 : 0.0);
        if (sw && be == 0.0) s = gj > 0.0 ? 1.0 : (gj < 0.0 ? -1.0 : 0.0);
        nsg[a * G.BN + gp] = (int8_t)(fl ? 0 : (int)s);
        if (fl && nx && D.wa[a] != 0.0) bits |= (uint8_t)(1u << a);
    }
    flat[gp] = bits;
    parent[gp] = (int)gp;
}

// Joins: iso, a flat pixel with every forward neighbour along the
// structure's active axes (tv_iso.cpp:598-605); aniso, the flat edges.
__global__ __launch_bounds__(kThr) void union_kernel(Geo G, int iso, const DImg *di, const uint8_t *flat, int *par) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const uint8_t f = flat[gp];
    if (iso && f != 3) return;  // flat and with a neighbour
    if (!iso && f == 0) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    int64_t c[3];
    coords(G, p, c);
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        const bool join = iso ? (D.wsave[a] != 0.0 && c[ax] + 1 < G.L[ax]) : ((f >> a) & 1);
        if (join) uf_unite(par, (int)gp, (int)(gp + G.S[ax]));
    }
}

__global__ __launch_bounds__(kThr) void compress_kernel(int64_t BN, int *par, int *isroot) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= BN) return;
    const int r = uf_find(par, (int)gp);
    isroot[gp] = r == (int)gp ? 1 : 0;
}

__global__ __launch_bounds__(kThr) void finish_roots_kernel(int64_t BN, int *par) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= BN) return;
    par[gp] = uf_find(par, (int)gp);
}

// label = the rank of the root; iota for the sort's values; per-image
// region bases (the first pixel of an image is a root) and the total.
__global__ __launch_bounds__(kThr) void label_kernel(Geo G, const int *par, const int *rank, const int *isroot,
                                                     int *label, int *iota, int *rbase) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    label[gp] = rank[par[gp]];
    iota[gp] = (int)gp;
    if (gp % G.N == 0) rbase[gp / G.N] = rank[gp];
    if (gp == G.BN - 1) rbase[G.BN / G.N] = rank[gp] + isroot[gp];
}

__global__ __launch_bounds__(kThr) void rstart_kernel(int64_t BN, const int *skey, int *rstart) {
    const int64_t m = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (m >= BN) return;
    if (m == 0 || skey[m] != skey[m - 1]) rstart[skey[m]] = (int)m;
    if (m == BN - 1) rstart[skey[m] + 1] = (int)BN;
}

__global__ __launch_bounds__(kThr) void inv_kernel(int nc, const int *rstart, double *inv) {
    const int C = blockIdx.x * kThr + threadIdx.x;
    if (C >= nc) return;
    inv[C] = 1.0 / (double)(rstart[C + 1] - rstart[C]);  // the count of tv_iso.cpp:617-618, exactly
}

// ---------------------------------------------------------------------------
// Region sums (rsum_kernel and region_value: regions.hpp)
// ---------------------------------------------------------------------------

template <class T>
struct FPix {  // a pixel array of the dtype
    const T *v;
    __device__ double operator()(int p, int) const { return ld(v, p); }
};

struct FPixD {  // a double pixel array
    const double *v;
    __device__ double operator()(int p, int) const { return v[p]; }
};

// Values already in member order (the apply's, apply_kernel).
struct FPos {
    const double *v;
    __device__ double operator()(int i, int) const { return v[i]; }
};

// ---------------------------------------------------------------------------
// Iso system pieces (per pixel)
// ---------------------------------------------------------------------------

// H_p y (tv_iso.cpp:810-828) in place on y (k values), for a non-flat pixel.
__device__ __forceinline__ void hess_apply(int k, const double *g, double h, double *y) {
    double nn = 0.0, dot = 0.0;
    for (int a = 0; a < k; ++a) nn += g[a] * g[a];
    const double nr = sqrt(nn);
    for (int a = 0; a < k; ++a) dot += g[a] * y[a];
    for (int a = 0; a < k; ++a) {
        const double na = g[a] / nr;
        y[a] = h * (y[a] - na * dot / nr);
    }
}

template <class T>
__device__ __forceinline__ double curv(const Geo &G, const DImg &D, const double *gzp, const T *sw, int64_t sw_img,
                                       int b, int64_t p, double *nr_out) {
    double nn = 0.0;
    for (int a = 0; a < G.k; ++a) nn += gzp[a] * gzp[a];
    const double nr = sqrt(nn);
    *nr_out = nr;
    return radius(D, sw, sw_img, b, p) / dmaxd(nr, D.mu);
}

// h = H_q G (M pv) at pixel gq (local index q, coordinates c) of image b
// (the apply's first half): 0 at a flat pixel.
template <class T>
__device__ __forceinline__ void h_at(const Geo &G, const DImg &D, const double *gz, const uint8_t *flat,
                                     const int *label, const double *pv, const T *sw, int64_t sw_img, int b,
                                     int64_t gq, int64_t q, const int64_t *c, double *y) {
    if (flat[gq] & 1) {
        for (int a = 0; a < G.k; ++a) y[a] = 0.0;
        return;
    }
    const double x0 = pv[label[gq]];
    double g[3];
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        y[a] = (D.wa[a] != 0.0 && c[ax] + 1 < G.L[ax]) ? D.wa[a] * (pv[label[gq + G.S[ax]]] - x0) : 0.0;
        g[a] = gz[a * G.BN + gq];
    }
    double nr;
    const double hs = curv(G, D, g, sw, sw_img, b, q, &nr);
    hess_apply(G.k, g, hs, y);
}

// h = H G (M pv) at every pixel of a live image (the apply's first half).
template <class T>
__global__ __launch_bounds__(kThr) void hgv_kernel(Geo G, const DImg *di, const CgState *cg, const double *gz,
                                                   const uint8_t *flat, const int *label, const double *pv,
                                                   const T *sw, int64_t sw_img, double *h) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    if (!cg[b].active) return;
    const int64_t p = gp - (int64_t)b * G.N;
    int64_t c[3];
    coords(G, p, c);
    double y[3];
    h_at(G, di[b], gz, flat, label, pv, sw, sw_img, b, gp, p, c, y);
    for (int a = 0; a < G.k; ++a) h[a * G.BN + gp] = y[a];
}

// The apply's second half in pixel order (tv_iso.cpp:881-887): the value of
// pixel p, pv_C - sum_a w_a (h_a[p] - h_a[p - e_a]) (0 for a frozen image),
// stored at p's position in the member order, so that the region sums read
// their members contiguously.
__global__ __launch_bounds__(kThr) void div_kernel(Geo G, const DImg *di, const CgState *cg, const int *label,
                                                   const int *pos, const double *pv, const double *h, double *vm) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    double v = 0.0;
    if (cg[b].active) {
        const DImg &D = di[b];
        int64_t c[3];
        coords(G, gp - (int64_t)b * G.N, c);
        double q = -pv[label[gp]];
        for (int a = 0; a < G.k; ++a) {
            if (D.wa[a] == 0.0) continue;
            const int ax = 3 - G.k + a;
            const double hp = h[a * G.BN + gp];
            q += c[ax] > 0 ? D.wa[a] * (hp - h[a * G.BN + gp - G.S[ax]]) : D.wa[a] * hp;
        }
        v = -q;
    }
    vm[pos[gp]] = v;
}

// 2D: both halves of the apply in one pass (h_at and div_kernel's value,
// the same expressions in the same order). A wave covers 64 columns, x0 - 1
// .. x0 + 62, of which lanes 1..62 are its own (lane 0 supplies h of the
// column before them, lane 63 the region value of the column after them),
// and walks rows y0 - 1 .. y0 + kApplyRows: at row y it loads the row's
// region values (one gather per pixel: the right neighbour's comes from the
// next lane, the lower neighbour's one row later), forms h of row y - 1,
// and the value of row y - 1 from h there, h of the column before (the
// previous lane) and h of the row above (the previous step).
constexpr int kApplyRows = 8;
constexpr int kApplyCols = 62;

__device__ __forceinline__ double lane_from_left(double v) {  // lane i: lane i - 1's v (lane 0: 0)
    return __shfl_up(v, 1, 64);
}
__device__ __forceinline__ double lane_from_right(double v) {  // lane i: lane i + 1's v (lane 63: its own)
    return __shfl_down(v, 1, 64);
}

template <class T>
__global__ __launch_bounds__(256) void apply2d_kernel(Geo G, const DImg *di, const CgState *cg, const double *gz,
                                                      const uint8_t *flat, const int *label, const int *pos,
                                                      const double *pv, const T *sw, int64_t sw_img, int ntx,
                                                      int nty, int B, double *vm, double *apart) {
    const int lane = threadIdx.x & 63;
    const int64_t tile = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t per = (int64_t)ntx * nty;
    const int b = (int)(tile / per);
    if (b >= B || !cg[b].active) return;
    const int r = (int)(tile - (int64_t)b * per);
    const int ty = r / ntx, tx = r - ty * ntx;
    const DImg &D = di[b];
    const int H = (int)G.L[1], W = (int)G.L[2];
    const int x = tx * kApplyCols - 1 + lane;
    const int y0 = ty * kApplyRows, yend = min(y0 + kApplyRows, H);
    const bool xin = x >= 0 && x < W;
    const bool own_col = xin && lane >= 1 && lane <= kApplyCols;
    const int64_t base = (int64_t)b * G.N;
    const double w0 = D.wa[0], w1 = D.wa[1];  // axis 0: rows, axis 1: columns
    // Row yr's values at this lane (region value, flat, g) are loaded one
    // step ahead and its label two steps ahead, so that neither the loads
    // nor the label -> region-value gather wait inside a step.
    auto row_in = [&](int yr) { return xin && yr >= 0 && yr < H; };
    auto gpix = [&](int yr) { return row_in(yr) ? base + (int64_t)yr * W + x : (int64_t)0; };
    int lab1 = row_in(y0) ? label[gpix(y0)] : 0;  // row y0's label
    double xv, g0, g1;
    bool fl;
    {
        const int yr = y0 - 1;
        const bool in = row_in(yr);
        const int64_t gp = gpix(yr);
        xv = in ? pv[label[gp]] : 0.0;
        fl = in ? (flat[gp] & 1) != 0 : true;
        g0 = in ? gz[gp] : 0.0;
        g1 = in ? gz[G.BN + gp] : 0.0;
    }
    // the previous row's (the row whose h is formed next)
    double pxv = 0.0, pg0 = 0.0, pg1 = 0.0;
    bool pfl = true;
    double ph0 = 0.0;  // h_0 of the row before that (the row above the value's row)
    double pq = 0.0;   // sum over the strip's pixels of pv_C(p) x value(p): its share of p . q
    for (int yr = y0 - 1; yr <= yend; ++yr) {
        // the next row's values and the row after's label, in flight during this step
        const bool in1 = row_in(yr + 1), in2 = row_in(yr + 2) && yr + 2 <= yend;
        const int64_t gp1 = gpix(yr + 1);
        const double xv1 = in1 ? pv[lab1] : 0.0;
        const bool fl1 = in1 ? (flat[gp1] & 1) != 0 : true;
        const double g01 = in1 ? gz[gp1] : 0.0, g11 = in1 ? gz[G.BN + gp1] : 0.0;
        const int lab2 = in2 ? label[gpix(yr + 2)] : 0;
        const double xr = lane_from_right(pxv);  // the right neighbour of row yr - 1
        const int y = yr - 1;                    // the row whose h is formed now
        if (y >= y0 - 1 && y >= 0) {
            const bool pin = xin && y < H;
            double h0 = 0.0, h1 = 0.0;
            if (pin && !pfl) {
                // h_at: y_a = w_a (pv[label[p + e_a]] - x0) where p + e_a is in the image
                double yy[2];
                yy[0] = (w0 != 0.0 && y + 1 < H) ? w0 * (xv - pxv) : 0.0;
                yy[1] = (w1 != 0.0 && x + 1 < W) ? w1 * (xr - pxv) : 0.0;
                double g[2] = {pg0, pg1};
                const int64_t q = (int64_t)y * W + x;
                double nr;
                const double hs = curv(G, D, g, sw, sw_img, b, q, &nr);
                hess_apply(2, g, hs, yy);
                h0 = yy[0];
                h1 = yy[1];
            }
            const double h1l = lane_from_left(h1);  // h_1 of the column before
            if (own_col && y >= y0 && y < yend) {
                double qv = -pxv;
                if (w0 != 0.0) qv += y > 0 ? w0 * (h0 - ph0) : w0 * h0;
                if (w1 != 0.0) qv += x > 0 ? w1 * (h1 - h1l) : w1 * h1;
                vm[pos[base + (int64_t)y * W + x]] = -qv;
                pq += pxv * -qv;
            }
            ph0 = h0;
        }
        pxv = xv;
        pfl = fl;
        pg0 = g0;
        pg1 = g1;
        xv = xv1;
        fl = fl1;
        g0 = g01;
        g1 = g11;
        lab1 = lab2;
    }
    // the strip's share of p . q, for the image's fixed-order sum over its strips
    for (int o = 32; o > 0; o >>= 1) pq += __shfl_xor(pq, o, 64);
    if (lane == 0) apart[tile] = pq;
}

// The member-order position of every pixel.
__global__ __launch_bounds__(kThr) void pos_kernel(int64_t BN, const int *members, int *pos) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i < BN) pos[members[i]] = (int)i;
}

// The Jacobi diagonal's terms moved onto pixels (tv_iso.cpp:833-870, in the
// gathered form of _tv_torch.py:446-485): own[q] = the term of q's region
// from q itself and from each backward neighbour p = q - e_a whose forward
// edge reaches q's region first.
template <class T>
__device__ double jac_term(const Geo &G, const DImg &D, const double *gz, const int *label, int64_t gp,
                           const int64_t *c, const T *sw, int64_t sw_img, int b, int64_t p, int C) {
    double g[3];
    for (int a = 0; a < G.k; ++a) g[a] = gz[a * G.BN + gp];
    double nr;
    const double hs = curv(G, D, g, sw, sw_img, b, p, &nr);
    const int l0 = label[gp];
    double hh = 0.0, nh = 0.0;
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        if (!(D.wa[a] != 0.0 && c[ax] + 1 < G.L[ax])) continue;
        const int la = label[gp + G.S[ax]];
        const double hv = D.wa[a] * ((la == C ? 1.0 : 0.0) - (l0 == C ? 1.0 : 0.0));
        hh += hv * hv;
        nh += hv * g[a] / nr;
    }
    return hs * (hh - nh * nh);
}

template <class T>
__global__ __launch_bounds__(kThr) void jacobi_own_kernel(Geo G, const DImg *di, const double *gz,
                                                          const uint8_t *flat, const int *label, const T *sw,
                                                          int64_t sw_img, double *own) {
    const int64_t gq = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gq >= G.BN) return;
    const int b = img_of(G, gq);
    const int64_t q = gq - (int64_t)b * G.N;
    const DImg &D = di[b];
    int64_t cq[3];
    coords(G, q, cq);
    const int C = label[gq];
    double s = 0.0;
    if (!(flat[gq] & 1)) s += jac_term(G, D, gz, label, gq, cq, sw, sw_img, b, q, C);
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        if (cq[ax] == 0 || D.wa[a] == 0.0) continue;
        const int64_t gp = gq - G.S[ax];
        if (flat[gp] & 1) continue;
        if (label[gp] == C) continue;
        int64_t cp[3] = {cq[0], cq[1], cq[2]};
        cp[ax] -= 1;
        bool seen = false;
        for (int a2 = 0; a2 < a; ++a2) {
            const int ax2 = 3 - G.k + a2;
            if (D.wa[a2] != 0.0 && cp[ax2] + 1 < G.L[ax2] && label[gp + G.S[ax2]] == C) seen = true;
        }
        if (seen) continue;
        s += jac_term(G, D, gz, label, gp, cp, sw, sw_img, b, q - G.S[ax], C);
    }
    own[gq] = s;
}

// ---------------------------------------------------------------------------
// Region-space kernels (blocks of kRB regions of one image)
// ---------------------------------------------------------------------------

// dinv = 1 / (|C| + the summed terms).
__global__ __launch_bounds__(kThr) void dinv_kernel(const RBlock *rb, const int *rstart, const double *osum,
                                                    const double *head, const double *tail, const double *inv,
                                                    double *dinv) {
    const RBlock B = rb[blockIdx.x];
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr)
        dinv[C] = 1.0 / (1.0 / inv[C] + region_value(C, osum, rstart, head, tail));
}

// Region means (aniso): out[C] = sum / |C|.
__global__ __launch_bounds__(kThr) void mean_kernel(const RBlock *rb, const int *rstart, const double *rs,
                                                    const double *head, const double *tail, const double *inv,
                                                    double *out) {
    const RBlock B = rb[blockIdx.x];
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr) out[C] = region_value(C, rs, rstart, head, tail) * inv[C];
}

// CG set-up (tv_iso.cpp:904-915): b finalised, x = 0, r = b, pv = r dinv;
// partials of b^2 inv and r pv.
__global__ __launch_bounds__(kThr) void cg_init_kernel(const RBlock *rb, const int *rstart, double *bv,
                                                       const double *head, const double *tail, const double *inv,
                                                       const double *dinv, double *x, double *r, double *pv,
                                                       double *part) {
    __shared__ double sh[kThr];
    const RBlock B = rb[blockIdx.x];
    double t0 = 0.0, t1 = 0.0;
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr) {
        const double v = region_value(C, bv, rstart, head, tail);
        bv[C] = v;
        t0 += v * v * inv[C];
        x[C] = 0.0;
        r[C] = v;
        pv[C] = v * dinv[C];
        t1 += v * pv[C];
    }
    t0 = block_sum(t0, sh);
    t1 = block_sum(t1, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = t0;
        part[3 * blockIdx.x + 1] = t1;
    }
}

__device__ __forceinline__ double image_total(const double *part, int j0, int n, int q) {
    double t = 0.0;
    for (int j = j0; j < j0 + n; ++j) t += part[3 * j + q];
    return t;
}

__global__ void cg_start_kernel(int B, const DImg *di, const double *part, double tol, int max_iter, CgState *cg) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const DImg &D = di[b];
    CgState S;
    S.bn = image_total(part, D.bbase, D.bcount, 0);
    S.rz = image_total(part, D.bbase, D.bcount, 1);
    S.tol = tol > 0.0 ? tol : dmaxd(kDerivFloor, D.rel);
    S.it = 0;
    S.max_iter = max_iter;
    S.pad_ = 0;
    if (S.bn == 0.0) {
        S.res = 0.0;
        S.active = 0;
    } else {
        S.res = 1.0;
        S.active = (S.res > S.tol && max_iter > 0) ? 1 : 0;
    }
    cg[b] = S;
}

// pq partials, with the apply's region sums finalised into qv.
__global__ __launch_bounds__(kThr) void cg_pq_kernel(const RBlock *rb, const CgState *cg, const int *rstart,
                                                     double *qv, const double *head, const double *tail,
                                                     const double *pv, double *part) {
    __shared__ double sh[kThr];
    const RBlock B = rb[blockIdx.x];
    if (!cg[B.img].active) return;
    double t = 0.0;
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr) {
        const double q = region_value(C, qv, rstart, head, tail);
        qv[C] = q;
        t += pv[C] * q;
    }
    t = block_sum(t, sh);
    if (threadIdx.x == 0) part[3 * blockIdx.x] = t;
}

// p . q of image b: the pq pass's region-block partials, or (apt > 0) the
// fused 2D apply's, summed per image by pqsum_kernel into pqv.
__device__ __forceinline__ double image_pq(const double *part, const DImg &D, const double *pqv, int apt, int b) {
    return apt == 0 ? image_total(part, D.bbase, D.bcount, 0) : pqv[b];
}

// The fused 2D apply's strip partials of p . q summed per image, one wave per
// image: each lane its strips in order, then a fixed butterfly.
__global__ __launch_bounds__(64) void pqsum_kernel(const CgState *cg, const double *apart, int apt, double *pqv) {
    const int b = blockIdx.x, lane = threadIdx.x;
    if (!cg[b].active) return;
    double t = 0.0;
    for (int j = lane; j < apt; j += 64) t += apart[(int64_t)b * apt + j];
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0) pqv[b] = t;
}

// alpha = rz / pq; x += alpha pv; r -= alpha qv; partials of r^2 dinv, r^2
// inv. With apt > 0 (no pq pass) the region sums of the apply are finalised
// here.
__global__ __launch_bounds__(kThr) void cg_update_kernel(const RBlock *rb, const DImg *di, const CgState *cg,
                                                         const double *pv, const double *qv, const double *dinv,
                                                         const double *inv, double *x, double *r, double *part,
                                                         const double *pqv, int apt, const int *rstart,
                                                         const double *head, const double *tail) {
    __shared__ double sh[kThr];
    const RBlock B = rb[blockIdx.x];
    const CgState S = cg[B.img];
    if (!S.active) return;
    const DImg &D = di[B.img];
    const double pq = image_pq(part, D, pqv, apt, B.img);
    if (!(pq > 0.0)) return;
    const double alpha = S.rz / pq;
    double t0 = 0.0, t1 = 0.0;
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr) {
        const double q = apt ? region_value(C, qv, rstart, head, tail) : qv[C];
        x[C] += alpha * pv[C];
        r[C] -= alpha * q;
        t0 += r[C] * r[C] * dinv[C];
        t1 += r[C] * r[C] * inv[C];
    }
    t0 = block_sum(t0, sh);
    t1 = block_sum(t1, sh);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 1] = t0;
        part[3 * blockIdx.x + 2] = t1;
    }
}

// beta = rzn / rz; pv = r dinv + beta pv; the image's first block writes the
// next state (tv_iso.cpp:935-941; a non-positive pq stops the image).
__global__ __launch_bounds__(kThr) void cg_dir_kernel(const RBlock *rb, const DImg *di, const CgState *cur,
                                                      CgState *nxt, const double *r, const double *dinv, double *pv,
                                                      const double *part, const double *pqv, int apt) {
    const RBlock B = rb[blockIdx.x];
    const DImg &D = di[B.img];
    const bool lead = (int)blockIdx.x == D.bbase;
    CgState S = cur[B.img];
    if (!S.active) {
        if (lead && threadIdx.x == 0) nxt[B.img] = S;
        return;
    }
    const double pq = image_pq(part, D, pqv, apt, B.img);
    if (!(pq > 0.0)) {
        if (lead && threadIdx.x == 0) {
            S.active = 0;
            nxt[B.img] = S;
        }
        return;
    }
    const double rzn = image_total(part, D.bbase, D.bcount, 1);
    const double rn = image_total(part, D.bbase, D.bcount, 2);
    const double beta = rzn / S.rz;
    for (int C = B.c0 + threadIdx.x; C < B.c1; C += kThr) pv[C] = r[C] * dinv[C] + beta * pv[C];
    if (lead && threadIdx.x == 0) {
        S.rz = rzn;
        S.it += 1;
        S.res = sqrt(rn / S.bn);
        S.active = (S.res > S.tol && S.it < S.max_iter) ? 1 : 0;
        nxt[B.img] = S;
    }
}

__global__ void count_active_kernel(int B, const CgState *cg, int *out) {
    __shared__ int sh;
    if (threadIdx.x == 0) sh = 0;
    __syncthreads();
    int n = 0;
    for (int b = threadIdx.x; b < B; b += blockDim.x) n += cg[b].active;
    if (n) atomicAdd(&sh, n);
    __syncthreads();
    if (threadIdx.x == 0) out[0] = sh;
}

// Region values back to the pixels, in the dtype.
template <class T>
__global__ __launch_bounds__(kThr) void spread_kernel(int64_t BN, const int *label, const double *xr, T *out) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= BN) return;
    out[gp] = (T)xr[label[gp]];
}

// ---------------------------------------------------------------------------
// JVP right-hand sides (per pixel)
// ---------------------------------------------------------------------------

// c_A at pixel q (tv_iso.cpp:969-985): H (dW D z) + dr n on non-flat pixels.
template <class T>
__device__ void jvp_c(const Geo &G, const DImg &D, const double *gz, const uint8_t *flat, int64_t gq, int b,
                      int64_t q, const double *dw, double dgam, const T *sw, int64_t sw_img, const double *dsw,
                      int64_t dsw_img, double *out) {
    for (int a = 0; a < G.k; ++a) out[a] = 0.0;
    if (flat[gq] & 1) return;
    double g[3];
    for (int a = 0; a < G.k; ++a) {
        g[a] = gz[a * G.BN + gq];
        const double wd = D.wa[a] != 0.0 ? g[a] / D.wa[a] : 0.0;
        out[a] = (dw && D.wa[a] != 0.0) ? D.sc * dw[a] * wd : 0.0;
    }
    double nr;
    const double hs = curv(G, D, g, sw, sw_img, b, q, &nr);
    hess_apply(G.k, g, hs, out);
    double dr = dgam;
    if (sw) dr = dgam * ld(sw, b * sw_img + q) + (dsw ? D.gamma_s * (D.sc * dsw[b * dsw_img + q]) : 0.0);
    for (int a = 0; a < G.k; ++a) out[a] += dr * g[a] / nr;
}

struct DwArg {
    double v[3];
    int on;
};

template <class T>
__global__ __launch_bounds__(kThr) void jvp_rhs_iso_kernel(Geo G, const DImg *di, const double *gz,
                                                           const uint8_t *flat, const T *dx, const T *u,
                                                           int64_t u_img, int64_t u_plane, DwArg dw,
                                                           const double *dgamma, const T *sw, int64_t sw_img,
                                                           const double *dsw, int64_t dsw_img, double *rhs) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    int64_t c[3];
    coords(G, p, c);
    double v = dx ? ld(dx, gp) : 0.0;
    const double *dwp = dw.on ? dw.v : nullptr;
    if (dwp)
        for (int a = 0; a < G.k; ++a) {
            const int ax = 3 - G.k + a;
            if (!(dw.v[a] != 0.0 && G.L[ax] > 1)) continue;
            const double cc = D.sc * dw.v[a];
            double up = D.use_dual ? ldexp(ld(u, b * u_img + a * u_plane + p), -D.e) : 0.0;
            if (c[ax] > 0) {
                const double um = D.use_dual ? ldexp(ld(u, b * u_img + a * u_plane + p - G.S[ax]), -D.e) : 0.0;
                v += cc * (up - um);
            } else {
                v += cc * up;
            }
        }
    const double dg = dgamma ? dgamma[b] : 0.0;
    double c0[3];
    jvp_c(G, D, gz, flat, gp, b, p, dwp, dg, sw, sw_img, dsw, dsw_img, c0);
    for (int a = 0; a < G.k; ++a) {
        if (D.wa[a] == 0.0) continue;
        const int ax = 3 - G.k + a;
        if (c[ax] > 0) {
            double cm[3];
            jvp_c(G, D, gz, flat, gp - G.S[ax], b, p - G.S[ax], dwp, dg, sw, sw_img, dsw, dsw_img, cm);
            v += D.wa[a] * (c0[a] - cm[a]);
        } else {
            v += D.wa[a] * c0[a];
        }
    }
    rhs[gp] = v;
}

// Aniso: r = dx - sum_a dw_a D_a^T u_a - G^T (dr n) (_tv_torch.py:528-547).
template <class T>
__global__ __launch_bounds__(kThr) void jvp_rhs_aniso_kernel(Geo G, const DImg *di, const int8_t *nsg, const T *dx,
                                                             const T *u, int64_t u_img, int64_t u_plane, DwArg dw,
                                                             const double *dgamma, const T *sw, int64_t sw_img,
                                                             int64_t sw_plane, const double *dsw, int64_t dsw_img,
                                                             double *rhs) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    int64_t c[3];
    coords(G, p, c);
    const double dg = dgamma ? dgamma[b] : 0.0;
    auto edge = [&](int a, int64_t q) {  // dw_a u_a[q] + w_a dr_q n_q on the edge (q, q + e_a)
        double t = dw.on ? dw.v[a] * ld(u, b * u_img + a * u_plane + q) : 0.0;
        double dr = dg;
        if (sw) dr = dg * ld(sw, b * sw_img + a * sw_plane + q) +
                     (dsw ? D.gamma_s * dsw[b * dsw_img + a * G.N + q] : 0.0);
        return t + D.wa[a] * (dr * (double)nsg[a * G.BN + (int64_t)b * G.N + q]);
    };
    double v = dx ? ld(dx, gp) : 0.0;
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        if (c[ax] + 1 < G.L[ax]) v += edge(a, p);
        if (c[ax] > 0) v -= edge(a, p - G.S[ax]);
    }
    rhs[gp] = v;
}

// ---------------------------------------------------------------------------
// VJP parameter gradients (per image sums over pixel blocks)
// ---------------------------------------------------------------------------

// Iso (tv_iso.cpp:1021-1066): q[a] = sum (v[p+e_a] - v[p]) u_a[p] (axes with
// L > 1), q[k] = the gamma term, q[k+1+a] = sum [H G v]_a (D_a z); the
// alpha gradient per pixel to gsw_img.
template <class T>
__global__ __launch_bounds__(kThr) void vjp_params_iso_kernel(Geo G, const DImg *di, const double *gz,
                                                              const uint8_t *flat, const int *label, const double *xr,
                                                              const T *u, int64_t u_img, int64_t u_plane, const T *sw,
                                                              int64_t sw_img, double *gsw_img, int nblk,
                                                              double *part) {
    __shared__ double sh[kThr];
    const int b = blockIdx.y;
    const DImg &D = di[b];
    double q[kMaxQ];
    for (int i = 0; i < kMaxQ; ++i) q[i] = 0.0;
    const int64_t p0 = (int64_t)blockIdx.x * kPixBlk;
    for (int64_t p = p0 + threadIdx.x; p < p0 + kPixBlk && p < G.N; p += kThr) {
        const int64_t gp = (int64_t)b * G.N + p;
        int64_t c[3];
        coords(G, p, c);
        const double v0 = xr[label[gp]];
        double gv[3], g[3];
        for (int a = 0; a < G.k; ++a) {
            const int ax = 3 - G.k + a;
            const bool nx = c[ax] + 1 < G.L[ax];
            const double v1 = nx ? xr[label[gp + G.S[ax]]] : 0.0;
            if (nx && D.use_dual) q[a] += (v1 - v0) * ldexp(ld(u, b * u_img + a * u_plane + p), -D.e);
            gv[a] = (D.wa[a] != 0.0 && nx) ? D.wa[a] * (v1 - v0) : 0.0;
            g[a] = gz[a * G.BN + gp];
        }
        if (flat[gp] & 1) {
            if (gsw_img) gsw_img[gp] = 0.0;
            continue;
        }
        double nr;
        const double hs = curv(G, D, g, sw, sw_img, b, p, &nr);
        double t = 0.0;
        for (int a = 0; a < G.k; ++a) t += gv[a] * g[a] / nr;
        if (sw) {
            q[G.k] += ld(sw, b * sw_img + p) * t;
            if (gsw_img) gsw_img[gp] = -(D.sc * (D.gamma_s * t));
        } else {
            q[G.k] += t;
        }
        hess_apply(G.k, g, hs, gv);
        for (int a = 0; a < G.k; ++a)
            if (D.wa[a] != 0.0) q[G.k + 1 + a] += gv[a] * (g[a] / D.wa[a]);
    }
    const int nq = 2 * G.k + 1;
    for (int i = 0; i < nq; ++i) {
        const double s = block_sum(q[i], sh);
        if (threadIdx.x == 0) part[((int64_t)b * nblk + blockIdx.x) * kMaxQ + i] = s;
    }
}

// Aniso (_tv_torch.py:563-585): q[a] = sum D_a(h) u_a, q[k] = sum sw w_a D_a(h) n;
// the edge-weight gradient per edge to gsw_img (k planes per image).
template <class T>
__global__ __launch_bounds__(kThr) void vjp_params_aniso_kernel(Geo G, const DImg *di, const int8_t *nsg,
                                                                const int *label, const double *xr, const T *u,
                                                                int64_t u_img, int64_t u_plane, const T *sw,
                                                                int64_t sw_img, int64_t sw_plane, double *gsw_img,
                                                                int nblk, double *part) {
    __shared__ double sh[kThr];
    const int b = blockIdx.y;
    const DImg &D = di[b];
    double q[kMaxQ];
    for (int i = 0; i < kMaxQ; ++i) q[i] = 0.0;
    const int64_t p0 = (int64_t)blockIdx.x * kPixBlk;
    for (int64_t p = p0 + threadIdx.x; p < p0 + kPixBlk && p < G.N; p += kThr) {
        const int64_t gp = (int64_t)b * G.N + p;
        int64_t c[3];
        coords(G, p, c);
        const double v0 = xr[label[gp]];
        for (int a = 0; a < G.k; ++a) {
            const int ax = 3 - G.k + a;
            const bool nx = c[ax] + 1 < G.L[ax];
            const double dh = nx ? xr[label[gp + G.S[ax]]] - v0 : 0.0;
            q[a] += dh * ld(u, b * u_img + a * u_plane + p);
            const double t = D.wa[a] * dh * (double)nsg[a * G.BN + gp];
            const double swv = sw ? ld(sw, b * sw_img + a * sw_plane + p) : 1.0;
            q[G.k] += swv * t;
            if (gsw_img) gsw_img[(int64_t)b * G.k * G.N + a * G.N + p] = -(D.gamma_s * t);
        }
    }
    const int nq = G.k + 1;
    for (int i = 0; i < nq; ++i) {
        const double s = block_sum(q[i], sh);
        if (threadIdx.x == 0) part[((int64_t)b * nblk + blockIdx.x) * kMaxQ + i] = s;
    }
}

// Per image: the gamma gradient; the axis weights' gradient summed over the
// images in order (one thread).
__global__ void vjp_combine_kernel(int B, int k, int iso, const DImg *di, const double *part, int nblk,
                                   double *ggamma, double *gw) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double tw[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < B; ++b) {
        const DImg &D = di[b];
        double q[kMaxQ];
        for (int i = 0; i < kMaxQ; ++i) q[i] = 0.0;
        for (int j = 0; j < nblk; ++j)
            for (int i = 0; i < kMaxQ; ++i) q[i] += part[((int64_t)b * nblk + j) * kMaxQ + i];
        if (ggamma) ggamma[b] = -q[k];
        for (int a = 0; a < k; ++a) {
            double acc = 0.0;
            if (iso) {
                acc -= D.sc * q[a];
                if (D.wa[a] != 0.0) acc -= D.sc * q[k + 1 + a];
            } else {
                acc = -q[a];
            }
            tw[a] += acc;
        }
    }
    if (gw)
        for (int a = 0; a < k; ++a) gw[a] = tw[a];
}

// A shared map's gradient: the per-image gradients summed in image order.
__global__ __launch_bounds__(kThr) void sum_images_kernel(int64_t n, int B, const double *src, double *dst) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= n) return;
    double t = 0.0;
    for (int b = 0; b < B; ++b) t += src[(int64_t)b * n + i];
    dst[i] = t;
}


int grid_of(int64_t n) { return (int)((n + kThr - 1) / kThr); }

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct Deriv {
    DerivSpec sp;
    Geo G{};
    int nc = 0;
    int nrb = 0;   // region blocks
    int nch = 0;   // region-sum blocks
    int npb = 0;   // pixel blocks per image
    int max_rb = 0;  // the region blocks reserved (every image's regions fit)
    std::vector<DImg> hdi, hdi2;      // reserved: B
    std::vector<int> hrb;             // reserved: B + 1
    std::vector<RBlock> hblocks;      // reserved: max_rb
    std::vector<void *> bufs;
    DImg *di = nullptr;
    RBlock *rb = nullptr;
    double *gz = nullptr, *hbuf = nullptr, *pix = nullptr;
    uint8_t *flat = nullptr;
    int8_t *nsg = nullptr;
    int *parent = nullptr, *isroot = nullptr, *rank = nullptr, *label = nullptr, *iota = nullptr;
    int *skey = nullptr, *members = nullptr, *rstart = nullptr, *rbase = nullptr, *cnt = nullptr;
    int *pos = nullptr;     // iso: each pixel's position in the member order
    double *vm = nullptr;   // iso: the apply's pixel values in the member order
    double *apart = nullptr;  // iso 2D: the fused apply's p . q partials (B x strips)
    double *pqv = nullptr;    // iso 2D: p . q per image
    int apt = 0;              // iso 2D: strips per image (0: the pq pass)
    double *inv = nullptr, *dinv = nullptr, *cb = nullptr, *xr = nullptr, *rr = nullptr, *pv = nullptr,
           *qv = nullptr;
    double *head = nullptr, *tail = nullptr, *part = nullptr, *ppart = nullptr, *gsw_img = nullptr,
           *dsw_full = nullptr;
    CgState *cg[2] = {nullptr, nullptr};
    int *hcount = nullptr;  // pinned
    CgState *hcg = nullptr; // pinned
    void *cub = nullptr, *scan_tmp = nullptr;
    size_t cub_bytes = 0, scan_bytes = 0;
    int8_t *km = nullptr;       // aniso: the kink codes of a branch override
    int64_t *nreg = nullptr;    // the region counts of a query (B)
    int cur = 0;
    bool has_result = false;

    ~Deriv() {
        for (void *p : bufs) (void)hipFree(p);
        if (hcount) (void)hipHostFree(hcount);
        if (hcg) (void)hipHostFree(hcg);
    }
    template <class P>
    hipError_t alloc(P **p, size_t bytes) {
        void *q = nullptr;
        hipError_t e = fpr::dev_alloc(&q, std::max<size_t>(bytes, 16));
        if (e) return e;
        bufs.push_back(q);
        *p = (P *)q;
        return hipSuccess;
    }
};

void deriv_free(Deriv *d) { delete d; }

namespace {

int64_t spatial_len(const Deriv &d) { return d.sp.iso ? d.G.N : d.G.k * d.G.N; }

// The branch overrides on the flat edges and signs of an aniso structure
// (defined with the kink kernels below).
template <class T>
int apply_branch(Deriv &d, hipStream_t s, const T *z, const T *u, const T *sw);

// Every buffer of a structure of d.sp's shape (B images of L, the dtype,
// iso or aniso, spatial weights shared or per image), allocated once: build()
// then writes each solution's structure into them. The sort's and scan's
// scratch are sized for any region count, the region blocks for the most an
// image's regions can take (ceil(regions / kRB) per image, summed: at most
// B N / kRB + B).
int reserve(Deriv &d) {
    const DerivSpec &sp = d.sp;
    const Geo G = d.G;
    const int64_t BN = G.BN;
    const int B = (int)sp.B;
    hipError_t e = hipSuccess;
#define FPR_A(ptr, bytes) \
    if ((e = d.alloc(&(ptr), (bytes)))) return FP_ENOMEM
    FPR_A(d.di, sizeof(DImg) * (size_t)B);
    FPR_A(d.flat, (size_t)BN);
    FPR_A(d.parent, 4 * (size_t)BN);
    FPR_A(d.isroot, 4 * (size_t)BN);
    FPR_A(d.rank, 4 * (size_t)BN);
    FPR_A(d.label, 4 * (size_t)BN);
    FPR_A(d.iota, 4 * (size_t)BN);
    FPR_A(d.skey, 4 * (size_t)BN);
    FPR_A(d.members, 4 * (size_t)BN);
    FPR_A(d.rstart, 4 * ((size_t)BN + 1));
    FPR_A(d.rbase, 4 * ((size_t)B + 1));
    FPR_A(d.cnt, 16);
    FPR_A(d.inv, 8 * (size_t)BN);
    FPR_A(d.cb, 8 * (size_t)BN);
    FPR_A(d.xr, 8 * (size_t)BN);
    FPR_A(d.pix, 8 * (size_t)BN);
    d.nch = (int)((BN + kChunk - 1) / kChunk);
    FPR_A(d.head, 8 * (size_t)d.nch);
    FPR_A(d.tail, 8 * (size_t)d.nch);
    d.npb = (int)((G.N + kPixBlk - 1) / kPixBlk);
    FPR_A(d.ppart, 8 * (size_t)kMaxQ * (size_t)B * (size_t)d.npb);
    FPR_A(d.nreg, sizeof(int64_t) * (size_t)B);
    if (sp.iso) {
        FPR_A(d.gz, 8 * (size_t)G.k * (size_t)BN);
        if (G.k == 3) FPR_A(d.hbuf, 8 * (size_t)G.k * (size_t)BN);  // 2D forms h inside apply2d_kernel
        FPR_A(d.vm, 8 * (size_t)BN);
        if (G.k == 2) {
            const int64_t ntx = (G.L[2] + kApplyCols - 1) / kApplyCols, nty = (G.L[1] + kApplyRows - 1) / kApplyRows;
            d.apt = (int)(ntx * nty);
            FPR_A(d.apart, 8 * (size_t)B * (size_t)d.apt);
            FPR_A(d.pqv, 8 * (size_t)B);
        }
        FPR_A(d.pos, 4 * (size_t)BN);
        FPR_A(d.dinv, 8 * (size_t)BN);
        FPR_A(d.rr, 8 * (size_t)BN);
        FPR_A(d.pv, 8 * (size_t)BN);
        FPR_A(d.qv, 8 * (size_t)BN);
        FPR_A(d.cg[0], sizeof(CgState) * (size_t)B);
        FPR_A(d.cg[1], sizeof(CgState) * (size_t)B);
    } else {
        FPR_A(d.nsg, (size_t)G.k * (size_t)BN);
        FPR_A(d.km, (size_t)BN);
    }
    if (sp.sw) {
        FPR_A(d.gsw_img, 8 * (size_t)B * (size_t)spatial_len(d));
        if (sp.sw_img == 0) FPR_A(d.dsw_full, 8 * (size_t)B * (size_t)spatial_len(d));
    }
    const int64_t max_rb = BN / kRB + B;
    if (max_rb >= (int64_t)1 << 31) return FP_ESHAPE;
    d.max_rb = (int)max_rb;
    FPR_A(d.rb, sizeof(RBlock) * (size_t)d.max_rb);
    FPR_A(d.part, 8 * 3 * (size_t)d.max_rb);
    size_t b1 = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, b1, d.isroot, d.rank, (int)BN)) return FP_EDEVICE;
    d.scan_bytes = b1;
    FPR_A(d.scan_tmp, b1);
    size_t most = 0;
    for (int bits = 1; bits <= 31; ++bits) {
        size_t b2 = 0;
        if (hipcub::DeviceRadixSort::SortPairs(nullptr, b2, d.label, d.skey, d.iota, d.members, (int)BN, 0, bits))
            return FP_EDEVICE;
        most = std::max(most, b2);
    }
    d.cub_bytes = most;
    FPR_A(d.cub, d.cub_bytes);
#undef FPR_A
    if (fpr::pinned_alloc((void **)&d.hcount, 16) || fpr::pinned_alloc((void **)&d.hcg, sizeof(CgState) * (size_t)B))
        return FP_ENOMEM;
    d.hdi.reserve((size_t)B);
    d.hdi2.reserve((size_t)B);
    d.hrb.reserve((size_t)B + 1);
    d.hblocks.reserve((size_t)d.max_rb);
    return FP_OK;
}

template <class T>
int build(Deriv &d, hipStream_t s) {
    const DerivSpec &sp = d.sp;
    const Geo G = d.G;
    const int64_t BN = G.BN;
    const int B = (int)sp.B;
    const T *x = (const T *)sp.x, *z = (const T *)sp.z, *u = (const T *)sp.u, *sw = (const T *)sp.sw;
    d.has_result = false;
    // Per-image constants on the host (tv_iso.cpp:342-402 for the structure's
    // view of the solve).
    d.hdi.assign((size_t)B, DImg{});
    for (int b = 0; b < B; ++b) {
        DImg &D = d.hdi[(size_t)b];
        const double gamma = sp.gamma[b];
        D.e = sp.iso && sp.e ? sp.e[b] : 0;
        D.sc = ldexp(1.0, D.e);
        D.gamma_s = ldexp(gamma, -D.e);
        D.gap_s = sp.iso && sp.gap ? ldexp(sp.gap[b], -2 * D.e) : 0.0;
        double wmax = 0.0;
        for (int a = 0; a < G.k; ++a)
            if (G.L[3 - G.k + a] > 1) wmax = std::max(wmax, sp.w[a]);
        int na = 0;
        for (int a = 0; a < 3; ++a) {
            D.wsave[a] = 0.0;
            D.wa[a] = 0.0;
        }
        for (int a = 0; a < G.k; ++a) {
            const int64_t La = G.L[3 - G.k + a];
            if (sp.iso) {
                if (gamma > 0.0 && sp.w[a] > 0.0 && La > 1 && sp.w[a] >= 1e-100 * wmax) {
                    D.wsave[a] = sp.w[a];
                    ++na;
                }
                D.wa[a] = (D.gamma_s > 0.0 && sp.w[a] > 0.0 && La > 1) ? sp.w[a] : 0.0;
            } else {
                D.wa[a] = sp.w[a];
            }
        }
        D.from_dual = na >= 2;
        D.use_dual = na >= 1;
    }
    if (hipMemcpyAsync(d.di, d.hdi.data(), sizeof(DImg) * (size_t)B, hipMemcpyHostToDevice, s)) return FP_EDEVICE;
    const double eps = sizeof(T) == 4 ? 1.1920928955078125e-07 : 2.220446049250313e-16;
    const double floor_eps = sizeof(T) == 4 ? eps : 2.2e-16;
    const double iso_eta = sizeof(T) == 4 ? std::max(kDualEta, 64.0 * eps) : kDualEta;
    hipLaunchKernelGGL(prep_kernel<T>, dim3(B), dim3(kThr), 0, s, G, (int)sp.iso, eps, floor_eps, iso_eta, x, d.di);
    const int gpx = grid_of(BN);
    if (sp.iso)
        hipLaunchKernelGGL(flat_iso_kernel<T>, dim3(gpx), dim3(kThr), 0, s, G, x, z, u, sp.u_img, sp.u_plane, sw,
                           sp.sw_img, d.di, d.gz, d.flat, d.parent);
    else
        hipLaunchKernelGGL(flat_aniso_kernel<T>, dim3(gpx), dim3(kThr), 0, s, G, x, z, u, sp.u_img, sp.u_plane, sw,
                           sp.sw_img, sp.sw_plane, d.di, d.nsg, d.flat, d.parent);
    if (!sp.iso && sp.side)
        if (int r = apply_branch<T>(d, s, z, u, sw)) return r;
    hipLaunchKernelGGL(union_kernel, dim3(gpx), dim3(kThr), 0, s, G, (int)sp.iso, d.di, d.flat, d.parent);
    hipLaunchKernelGGL(compress_kernel, dim3(gpx), dim3(kThr), 0, s, BN, d.parent, d.isroot);
    hipLaunchKernelGGL(finish_roots_kernel, dim3(gpx), dim3(kThr), 0, s, BN, d.parent);
    // the scan and the sort in their reserved scratch
    size_t b1 = d.scan_bytes, b2 = 0;
    if (hipcub::DeviceScan::ExclusiveSum(d.scan_tmp, b1, d.isroot, d.rank, (int)BN, s)) return FP_EDEVICE;
    hipLaunchKernelGGL(label_kernel, dim3(gpx), dim3(kThr), 0, s, G, d.parent, d.rank, d.isroot, d.label, d.iota,
                       d.rbase);
    std::vector<int> &rb = d.hrb;
    rb.assign((size_t)B + 1, 0);
    if (hipMemcpyAsync(rb.data(), d.rbase, 4 * ((size_t)B + 1), hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
    if (hipStreamSynchronize(s)) return FP_EDEVICE;
    d.nc = rb[(size_t)B];
    int bits = 1;
    while (bits < 31 && (1LL << bits) <= (long long)d.nc) ++bits;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, b2, d.label, d.skey, d.iota, d.members, (int)BN, 0, bits, s))
        return FP_EDEVICE;
    if (b2 > d.cub_bytes) return FP_EINTERNAL;
    b2 = d.cub_bytes;
    if (hipcub::DeviceRadixSort::SortPairs(d.cub, b2, d.label, d.skey, d.iota, d.members, (int)BN, 0, bits, s))
        return FP_EDEVICE;
    hipLaunchKernelGGL(rstart_kernel, dim3(gpx), dim3(kThr), 0, s, BN, d.skey, d.rstart);
    if (sp.iso) hipLaunchKernelGGL(pos_kernel, dim3(gpx), dim3(kThr), 0, s, BN, d.members, d.pos);
    hipLaunchKernelGGL(inv_kernel, dim3(grid_of(d.nc)), dim3(kThr), 0, s, d.nc, d.rstart, d.inv);
    // region blocks per image
    std::vector<RBlock> &blocks = d.hblocks;
    blocks.clear();
    for (int b = 0; b < B; ++b) {
        DImg &D = d.hdi[(size_t)b];
        D.rbase = rb[(size_t)b];
        D.rcount = rb[(size_t)b + 1] - rb[(size_t)b];
        D.bbase = (int)blocks.size();
        for (int c0 = D.rbase; c0 < D.rbase + D.rcount; c0 += kRB)
            blocks.push_back(RBlock{b, c0, std::min(c0 + kRB, D.rbase + D.rcount), 0});
        D.bcount = (int)blocks.size() - D.bbase;
    }
    d.nrb = (int)blocks.size();
    if (d.nrb > d.max_rb) return FP_EINTERNAL;
    if (hipMemcpyAsync(d.rb, blocks.data(), sizeof(RBlock) * blocks.size(), hipMemcpyHostToDevice, s))
        return FP_EDEVICE;
    // the region fields of the device copy (after prep's fields are final)
    std::vector<DImg> &dev = d.hdi2;
    dev.assign((size_t)B, DImg{});
    if (hipMemcpyAsync(dev.data(), d.di, sizeof(DImg) * (size_t)B, hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
    if (hipStreamSynchronize(s)) return FP_EDEVICE;
    for (int b = 0; b < B; ++b) {
        DImg &D = dev[(size_t)b];
        D.rbase = d.hdi[(size_t)b].rbase;
        D.rcount = d.hdi[(size_t)b].rcount;
        D.bbase = d.hdi[(size_t)b].bbase;
        D.bcount = d.hdi[(size_t)b].bcount;
    }
    d.hdi.swap(dev);
    if (hipMemcpyAsync(d.di, d.hdi.data(), sizeof(DImg) * (size_t)B, hipMemcpyHostToDevice, s)) return FP_EDEVICE;
    if (sp.iso) {
        // the Jacobi diagonal, once per structure
        hipLaunchKernelGGL(jacobi_own_kernel<T>, dim3(gpx), dim3(kThr), 0, s, G, d.di, d.gz, d.flat, d.label, sw,
                           sp.sw_img, d.pix);
        hipLaunchKernelGGL(rsum_kernel<FPixD>, dim3(d.nch), dim3(kThr), 0, s, FPixD{d.pix}, d.skey, d.members,
                           d.rstart, BN, d.cb, d.head, d.tail);
        hipLaunchKernelGGL(dinv_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.rstart, d.cb, d.head, d.tail, d.inv,
                           d.dinv);
    }
    if (hipGetLastError()) return FP_EDEVICE;
    return FP_OK;
}

// Solves S c = cb (region space) for every image (tv_iso.cpp:899-943).
int run_cg(Deriv &d, hipStream_t s, double tol, int max_iter) {
    const Geo G = d.G;
    const int B = (int)d.sp.B;
    hipLaunchKernelGGL(cg_init_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.rstart, d.cb, d.head, d.tail, d.inv,
                       d.dinv, d.xr, d.rr, d.pv, d.part);
    hipLaunchKernelGGL(cg_start_kernel, dim3((B + 63) / 64), dim3(64), 0, s, B, d.di, d.part, tol, max_iter,
                       d.cg[0]);
    d.cur = 0;
    const int gpx = grid_of(G.BN);
    for (int done = 0; done < max_iter;) {
        for (int j = 0; j < kCheck && done < max_iter; ++j, ++done) {
            CgState *c0 = d.cg[d.cur], *c1 = d.cg[d.cur ^ 1];
            if (G.k == 2) {
                const int ntx = (int)((G.L[2] + kApplyCols - 1) / kApplyCols);
                const int nty = (int)((G.L[1] + kApplyRows - 1) / kApplyRows);
                const int ablk = (int)(((int64_t)B * ntx * nty + 3) / 4);
                if (d.sp.tsize == 4)
                    hipLaunchKernelGGL(apply2d_kernel<float>, dim3(ablk), dim3(256), 0, s, G, d.di, c0, d.gz, d.flat,
                                       d.label, d.pos, d.pv, (const float *)d.sp.sw, d.sp.sw_img, ntx, nty, B, d.vm,
                                       d.apart);
                else
                    hipLaunchKernelGGL(apply2d_kernel<double>, dim3(ablk), dim3(256), 0, s, G, d.di, c0, d.gz,
                                       d.flat, d.label, d.pos, d.pv, (const double *)d.sp.sw, d.sp.sw_img, ntx, nty,
                                       B, d.vm, d.apart);
            } else {
                if (d.sp.tsize == 4)
                    hipLaunchKernelGGL(hgv_kernel<float>, dim3(gpx), dim3(kThr), 0, s, G, d.di, c0, d.gz, d.flat,
                                       d.label, d.pv, (const float *)d.sp.sw, d.sp.sw_img, d.hbuf);
                else
                    hipLaunchKernelGGL(hgv_kernel<double>, dim3(gpx), dim3(kThr), 0, s, G, d.di, c0, d.gz, d.flat,
                                       d.label, d.pv, (const double *)d.sp.sw, d.sp.sw_img, d.hbuf);
                hipLaunchKernelGGL(div_kernel, dim3(gpx), dim3(kThr), 0, s, G, d.di, c0, d.label, d.pos, d.pv,
                                   d.hbuf, d.vm);
            }
            hipLaunchKernelGGL((rsum_kernel<FPos, true>), dim3(d.nch), dim3(kThr), 0, s, FPos{d.vm}, d.skey,
                               d.members, d.rstart, G.BN, d.qv, d.head, d.tail);
            // 2D: p . q came with the apply (no pq pass; the update finalises the region sums)
            if (d.apt == 0)
                hipLaunchKernelGGL(cg_pq_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, c0, d.rstart, d.qv, d.head,
                                   d.tail, d.pv, d.part);
            else
                hipLaunchKernelGGL(pqsum_kernel, dim3(B), dim3(64), 0, s, c0, d.apart, d.apt, d.pqv);
            hipLaunchKernelGGL(cg_update_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.di, c0, d.pv, d.qv, d.dinv,
                               d.inv, d.xr, d.rr, d.part, d.pqv, d.apt, d.rstart, d.head, d.tail);
            hipLaunchKernelGGL(cg_dir_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.di, c0, c1, d.rr, d.dinv, d.pv,
                               d.part, d.pqv, d.apt);
            d.cur ^= 1;
        }
        hipLaunchKernelGGL(count_active_kernel, dim3(1), dim3(256), 0, s, B, d.cg[d.cur], d.cnt);
        if (hipMemcpyAsync(d.hcount, d.cnt, 4, hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
        if (hipStreamSynchronize(s)) return FP_EDEVICE;
        if (*d.hcount == 0) break;
    }
    if (hipMemcpyAsync(d.hcg, d.cg[d.cur], sizeof(CgState) * (size_t)B, hipMemcpyDeviceToHost, s)) return FP_EDEVICE;
    if (hipStreamSynchronize(s)) return FP_EDEVICE;
    d.has_result = true;
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

template <class T>
int vjp_t(Deriv &d, hipStream_t s, const void *gin, void *gxo, double *ggamma, double *gw, double *gsw, double tol,
          int max_iter) {
    const DerivSpec &sp = d.sp;
    const Geo G = d.G;
    const int B = (int)sp.B;
    const T *u = (const T *)sp.u, *sw = (const T *)sp.sw;
    hipLaunchKernelGGL(rsum_kernel<FPix<T>>, dim3(d.nch), dim3(kThr), 0, s, FPix<T>{(const T *)gin}, d.skey,
                       d.members, d.rstart, G.BN, d.cb, d.head, d.tail);
    if (sp.iso) {
        if (int r = run_cg(d, s, tol, max_iter)) return r;
    } else {
        hipLaunchKernelGGL(mean_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.rstart, d.cb, d.head, d.tail, d.inv,
                           d.xr);
        d.has_result = false;
    }
    hipLaunchKernelGGL(spread_kernel<T>, dim3(grid_of(G.BN)), dim3(kThr), 0, s, G.BN, d.label, d.xr, (T *)gxo);
    double *gsi = (gsw && sp.sw) ? d.gsw_img : nullptr;
    if (sp.iso)
        hipLaunchKernelGGL(vjp_params_iso_kernel<T>, dim3(d.npb, B), dim3(kThr), 0, s, G, d.di, d.gz, d.flat,
                           d.label, d.xr, u, sp.u_img, sp.u_plane, sw, sp.sw_img, gsi, d.npb, d.ppart);
    else
        hipLaunchKernelGGL(vjp_params_aniso_kernel<T>, dim3(d.npb, B), dim3(kThr), 0, s, G, d.di, d.nsg, d.label,
                           d.xr, u, sp.u_img, sp.u_plane, sw, sp.sw_img, sp.sw_plane, gsi, d.npb, d.ppart);
    hipLaunchKernelGGL(vjp_combine_kernel, dim3(1), dim3(64), 0, s, B, G.k, (int)sp.iso, d.di, d.ppart, d.npb,
                       ggamma, gw);
    if (gsi) {
        const int64_t n = spatial_len(d);
        if (sp.sw_img == 0)
            hipLaunchKernelGGL(sum_images_kernel, dim3(grid_of(n)), dim3(kThr), 0, s, n, B, gsi, gsw);
        else if (hipMemcpyAsync(gsw, gsi, 8 * (size_t)B * (size_t)n, hipMemcpyDeviceToDevice, s))
            return FP_EDEVICE;
    } else if (gsw && !sp.sw) {
        return FP_EINVAL;
    }
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

__global__ __launch_bounds__(kThr) void broadcast_kernel(int64_t n, int B, const double *src, double *dst) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i >= n * B) return;
    dst[i] = src[i % n];
}

template <class T>
int jvp_t(Deriv &d, hipStream_t s, const void *dx, const double *dgamma, const double *dw, const double *dsw,
          void *dz, double tol, int max_iter) {
    const DerivSpec &sp = d.sp;
    const Geo G = d.G;
    const int B = (int)sp.B;
    const T *u = (const T *)sp.u, *sw = (const T *)sp.sw;
    DwArg dwa{{0.0, 0.0, 0.0}, dw ? 1 : 0};
    if (dw)
        for (int a = 0; a < G.k; ++a) dwa.v[a] = dw[a];
    const int64_t n = spatial_len(d);
    const double *dsw_i = dsw;
    int64_t dsw_img = n;
    if (dsw && sp.sw && sp.sw_img == 0) {  // a shared map's tangent, the same for every image
        hipLaunchKernelGGL(broadcast_kernel, dim3(grid_of(n * B)), dim3(kThr), 0, s, n, B, dsw, d.dsw_full);
        dsw_i = d.dsw_full;
    }
    if (dsw && !sp.sw) return FP_EINVAL;
    if (sp.iso)
        hipLaunchKernelGGL(jvp_rhs_iso_kernel<T>, dim3(grid_of(G.BN)), dim3(kThr), 0, s, G, d.di, d.gz, d.flat,
                           (const T *)dx, u, sp.u_img, sp.u_plane, dwa, dgamma, sw, sp.sw_img, dsw_i, dsw_img, d.pix);
    else
        hipLaunchKernelGGL(jvp_rhs_aniso_kernel<T>, dim3(grid_of(G.BN)), dim3(kThr), 0, s, G, d.di, d.nsg,
                           (const T *)dx, u, sp.u_img, sp.u_plane, dwa, dgamma, sw, sp.sw_img, sp.sw_plane, dsw_i,
                           dsw_img, d.pix);
    hipLaunchKernelGGL(rsum_kernel<FPixD>, dim3(d.nch), dim3(kThr), 0, s, FPixD{d.pix}, d.skey, d.members, d.rstart,
                       G.BN, d.cb, d.head, d.tail);
    if (sp.iso) {
        if (int r = run_cg(d, s, tol, max_iter)) return r;
    } else {
        hipLaunchKernelGGL(mean_kernel, dim3(d.nrb), dim3(kThr), 0, s, d.rb, d.rstart, d.cb, d.head, d.tail, d.inv,
                           d.xr);
        d.has_result = false;
    }
    hipLaunchKernelGGL(spread_kernel<T>, dim3(grid_of(G.BN)), dim3(kThr), 0, s, G.BN, d.label, d.xr, (T *)dz);
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

}  // namespace

int deriv_reserve(const DerivSpec &spec, Deriv **out) {
    *out = nullptr;
    if (spec.B <= 0 || (spec.k != 2 && spec.k != 3)) return FP_EINVAL;
    Deriv *d = new (std::nothrow) Deriv();
    if (!d) return FP_ENOMEM;
    d->sp = spec;
    Geo &G = d->G;
    G.k = spec.k;
    for (int a = 0; a < 3; ++a) G.L[a] = spec.L[a];
    G.S[2] = 1;
    G.S[1] = G.L[2];
    G.S[0] = G.L[1] * G.L[2];
    G.N = G.L[0] * G.L[1] * G.L[2];
    G.BN = G.N * spec.B;
    if (G.BN >= (int64_t)1 << 31 || G.N <= 0) {
        delete d;
        return FP_ESHAPE;
    }
    if (int r = reserve(*d)) {
        delete d;
        return r;
    }
    *out = d;
    return FP_OK;
}

int deriv_build(Deriv *d, const DerivSpec &spec, hipStream_t s) {
    if (!d) return FP_EINVAL;
    if (!spec.x || !spec.z || !spec.gamma || !spec.u || (spec.iso && !spec.gap)) return FP_EINVAL;
    const DerivSpec &r = d->sp;
    // the solution must have the reserved shape
    if (spec.B != r.B || spec.k != r.k || spec.iso != r.iso || spec.tsize != r.tsize || !spec.sw != !r.sw ||
        (spec.sw && (spec.sw_img == 0) != (r.sw_img == 0)))
        return FP_EINVAL;
    for (int a = 0; a < 3; ++a)
        if (spec.L[a] != r.L[a]) return FP_EINVAL;
    d->sp = spec;
    return spec.tsize == 4 ? build<float>(*d, s) : build<double>(*d, s);
}

int deriv_create(const DerivSpec &spec, hipStream_t s, Deriv **out) {
    *out = nullptr;
    if (!spec.x || !spec.z || !spec.gamma || !spec.u || (spec.iso && !spec.gap)) return FP_EINVAL;
    Deriv *d = nullptr;
    if (int r = deriv_reserve(spec, &d)) return r;
    if (int r = deriv_build(d, spec, s)) {
        delete d;
        return r;
    }
    *out = d;
    return FP_OK;
}

int deriv_vjp(Deriv *d, hipStream_t s, const void *g, void *gx, double *ggamma, double *gw, double *gsw, double tol,
              int max_iter) {
    if (!d || !g || !gx) return FP_EINVAL;
    return d->sp.tsize == 4 ? vjp_t<float>(*d, s, g, gx, ggamma, gw, gsw, tol, max_iter)
                            : vjp_t<double>(*d, s, g, gx, ggamma, gw, gsw, tol, max_iter);
}

int deriv_jvp(Deriv *d, hipStream_t s, const void *dx, const double *dgamma, const double *dw, const double *dsw,
              void *dz, double tol, int max_iter) {
    if (!d || !dz) return FP_EINVAL;
    return d->sp.tsize == 4 ? jvp_t<float>(*d, s, dx, dgamma, dw, dsw, dz, tol, max_iter)
                            : jvp_t<double>(*d, s, dx, dgamma, dw, dsw, dz, tol, max_iter);
}

void deriv_image_result(const Deriv *d, double *residual, int *iters) {
    const int B = (int)d->sp.B;
    for (int b = 0; b < B; ++b) {
        if (residual) residual[b] = d->has_result ? d->hcg[b].res : 0.0;
        if (iters) iters[b] = d->has_result ? d->hcg[b].it : 0;
    }
}

void deriv_result(const Deriv *d, double *residual, int *iters, int64_t *regions) {
    double r = 0.0;
    int it = 0;
    if (d->has_result)
        for (int b = 0; b < (int)d->sp.B; ++b) {
            r = std::max(r, d->hcg[b].res);
            it = std::max(it, d->hcg[b].it);
        }
    if (residual) *residual = r;
    if (iters) *iters = it;
    if (regions) *regions = d->nc;
}

// ---------------------------------------------------------------------------
// Structure queries (fp_workspace_regions, fp_workspace_kinks)
// ---------------------------------------------------------------------------

namespace {

// Labels within each image: the union-find roots are the smallest index of
// their region and are ranked in index order, so label - rbase[b] numbers an
// image's regions in the order of their first entries
// (driver.hpp:1133-1148 renumbers the saved labels the same way).
__global__ __launch_bounds__(kThr) void region_labels_kernel(Geo G, const DImg *di, const int *label, const int *rbase,
                                                             void *lab, int lab64, int8_t *sign) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const bool nan = di[b].nan != 0;
    const int64_t v = nan ? -1 : (int64_t)(label[gp] - rbase[b]);
    if (lab) {
        if (lab64) ((int64_t *)lab)[gp] = v;
        else ((int *)lab)[gp] = (int)v;
    }
    if (sign) sign[gp] = nan ? 0 : 1;
}

__global__ void region_counts_kernel(int B, const DImg *di, const int *rbase, int64_t *out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < B) out[b] = di[b].nan ? -1 : (int64_t)(rbase[b + 1] - rbase[b]);
}

// Iso kinks (tv_iso.cpp:633-664) at the input's scale: the step as given is
// gamma_s 2^e; q is the dual as returned (|q_p| <= gamma alpha_p).
template <class T>
__global__ __launch_bounds__(kThr) void kinks_iso_kernel(Geo G, const DImg *di, const T *z, const T *q, int64_t q_img,
                                                         int64_t q_plane, const T *sw, int64_t sw_img, double tol,
                                                         int8_t *mask) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    const DImg &D = di[b];
    const double gamma = ldexp(D.gamma_s, D.e);
    double r = gamma;
    if (sw) {
        const double a = ld(sw, b * sw_img + p);
        r = a == 0.0 ? 0.0 : gamma * a;
    }
    int64_t c[3];
    coords(G, p, c);
    double gg = 0.0, uu = 0.0;
    const double z0 = ld(z, gp);
    for (int a = 0; a < G.k; ++a) {
        const int ax = 3 - G.k + a;
        if (D.wa[a] > 0.0 && c[ax] + 1 < G.L[ax]) {
            const double t = D.wa[a] * (ld(z, gp + G.S[ax]) - z0);
            gg += t * t;
        }
        const double ua = ld(q, b * q_img + a * q_plane + p);
        uu += ua * ua;
    }
    mask[gp] = (int8_t)(r > 0.0 && sqrt(gg) <= tol && sqrt(uu) >= r - tol ? 1 : 0);
}

// Aniso kinks of axis a (tv_aniso.cpp:1303-1335): one thread per fiber of
// the axis walks it, summing q_a into the edge's 1D dual. One launch per
// axis, so each entry's bit is set by one thread.
template <class T>
__global__ __launch_bounds__(kThr) void kinks_aniso_kernel(Geo G, int a, const DImg *di, const T *z, const T *q,
                                                           int64_t q_img, int64_t q_plane, const T *sw, int64_t sw_img,
                                                           int64_t sw_plane, double tol, int8_t *mask) {
    const int ax = 3 - G.k + a;
    const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;  // the other two axes, in order
    const int64_t La = G.L[ax], nf = G.N / La, B = G.BN / G.N;
    const int64_t t = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (t >= nf * B) return;
    const int b = (int)(t / nf);
    const int64_t f = t - (int64_t)b * nf;
    const int64_t st = (f / G.L[o2]) * G.S[o1] + (f % G.L[o2]) * G.S[o2];
    const DImg &D = di[b];
    const double lam = ldexp(D.gamma_s, D.e) * D.wa[a];
    if (!(lam > 0.0) || La == 1) return;
    const int64_t s = G.S[ax], zb = (int64_t)b * G.N;
    double acc = 0.0;
    for (int64_t l = 0; l + 1 < La; ++l) {
        const int64_t i = st + l * s;
        acc += ld(q, b * q_img + a * q_plane + i);
        const double r = sw ? lam * ld(sw, b * sw_img + a * sw_plane + i) : lam;
        if (!(r > 0.0) || !(fabs(ld(z, zb + i + s) - ld(z, zb + i)) <= tol) || !(fabs(acc) >= r - tol)) continue;
        mask[zb + i] = (int8_t)(mask[zb + i] | (1 << a));
    }
}

__global__ __launch_bounds__(kThr) void zero_bytes_kernel(int64_t n, int8_t *m) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (i < n) m[i] = 0;
}

__global__ __launch_bounds__(kThr) void count_nonzero_kernel(int64_t n, const int8_t *m, unsigned long long *count) {
    const int64_t i = (int64_t)blockIdx.x * kThr + threadIdx.x;
    const bool nz = i < n && m[i] != 0;
    const unsigned long long bal = __ballot(nz);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(count, (unsigned long long)__popcll(bal));
}

template <class T>
int kinks_t(Deriv &d, hipStream_t s, const void *qv, int64_t q_img, int64_t q_plane, double tol, int8_t *mask,
            int64_t *count) {
    const Geo G = d.G;
    const T *z = (const T *)d.sp.z, *q = (const T *)qv, *sw = (const T *)d.sp.sw;
    const int gpx = grid_of(G.BN);
    if (d.sp.iso) {
        hipLaunchKernelGGL(kinks_iso_kernel<T>, dim3(gpx), dim3(kThr), 0, s, G, d.di, z, q, q_img, q_plane, sw,
                           d.sp.sw_img, tol, mask);
    } else {
        hipLaunchKernelGGL(zero_bytes_kernel, dim3(gpx), dim3(kThr), 0, s, G.BN, mask);
        for (int a = 0; a < G.k; ++a) {
            const int64_t nf = G.BN / G.L[3 - G.k + a];
            hipLaunchKernelGGL(kinks_aniso_kernel<T>, dim3(grid_of(nf)), dim3(kThr), 0, s, G, a, d.di, z, q, q_img,
                               q_plane, sw, d.sp.sw_img, d.sp.sw_plane, tol, mask);
        }
    }
    unsigned long long *cnt = (unsigned long long *)d.cnt;
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(1), dim3(kThr), 0, s, (int64_t)sizeof(unsigned long long),
                       (int8_t *)cnt);
    hipLaunchKernelGGL(count_nonzero_kernel, dim3(gpx), dim3(kThr), 0, s, G.BN, mask, cnt);
    unsigned long long hc = 0;
    if (hipMemcpyAsync(&hc, cnt, sizeof hc, hipMemcpyDeviceToHost, s) || hipStreamSynchronize(s)) return FP_EDEVICE;
    if (count) *count = (int64_t)hc;
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

// At the kink edges of a set_branch call (tv_aniso.cpp:1008-1011): side -1
// joins the edge's two sides, +1 keeps them apart with the sign of its dual.
template <class T>
__global__ __launch_bounds__(kThr) void branch_aniso_kernel(Geo G, const int8_t *side, const int8_t *kink,
                                                            const T *u, int64_t u_img, int64_t u_plane, int8_t *nsg,
                                                            uint8_t *flat) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= G.BN) return;
    const int sd = side[gp], km = kink[gp];
    if (sd == 0 || km == 0) return;
    const int b = img_of(G, gp);
    const int64_t p = gp - (int64_t)b * G.N;
    uint8_t f = flat[gp];
    for (int a = 0; a < G.k; ++a) {
        if (!((km >> a) & 1)) continue;
        if (sd < 0) {
            f = (uint8_t)(f | (1u << a));
            nsg[a * G.BN + gp] = 0;
        } else {
            f = (uint8_t)(f & ~(1u << a));
            const double ua = ld(u, b * u_img + a * u_plane + p);
            nsg[a * G.BN + gp] = (int8_t)(ua > 0.0 ? 1 : (ua < 0.0 ? -1 : 0));
        }
    }
    flat[gp] = f;
}

template <class T>
int apply_branch(Deriv &d, hipStream_t s, const T *z, const T *u, const T *sw) {
    const Geo G = d.G;
    const DerivSpec &sp = d.sp;
    int8_t *km = d.km;
    const int gpx = grid_of(G.BN);
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(gpx), dim3(kThr), 0, s, G.BN, km);
    for (int a = 0; a < G.k; ++a) {
        const int64_t nf = G.BN / G.L[3 - G.k + a];
        hipLaunchKernelGGL(kinks_aniso_kernel<T>, dim3(grid_of(nf)), dim3(kThr), 0, s, G, a, d.di, z,
                           (const T *)sp.q, sp.q_img, sp.q_plane, sw, sp.sw_img, sp.sw_plane, sp.side_tol, km);
    }
    hipLaunchKernelGGL(branch_aniso_kernel<T>, dim3(gpx), dim3(kThr), 0, s, G, sp.side, km, u, sp.u_img, sp.u_plane,
                       d.nsg, d.flat);
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

}  // namespace

int deriv_regions(Deriv *d, hipStream_t s, void *labels, bool lab64, int8_t *sign, int64_t *nregions) {
    if (!d) return FP_EINVAL;
    const Geo G = d->G;
    if (labels || sign)
        hipLaunchKernelGGL(region_labels_kernel, dim3(grid_of(G.BN)), dim3(kThr), 0, s, G, d->di, d->label, d->rbase,
                           labels, lab64 ? 1 : 0, sign);
    if (nregions) {
        const int B = (int)d->sp.B;
        int64_t *dev = d->nreg;
        hipLaunchKernelGGL(region_counts_kernel, dim3((B + 255) / 256), dim3(256), 0, s, B, d->di, d->rbase, dev);
        if (hipMemcpyAsync(nregions, dev, sizeof(int64_t) * (size_t)B, hipMemcpyDeviceToHost, s) ||
            hipStreamSynchronize(s))
            return FP_EDEVICE;
    }
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

int deriv_kinks(Deriv *d, hipStream_t s, const void *q, int64_t q_img, int64_t q_plane, double tol, int8_t *mask,
                int64_t *count) {
    if (!d || !q || !mask) return FP_EINVAL;
    return d->sp.tsize == 4 ? kinks_t<float>(*d, s, q, q_img, q_plane, tol, mask, count)
                            : kinks_t<double>(*d, s, q, q_img, q_plane, tol, mask, count);
}

}  // namespace fpr
