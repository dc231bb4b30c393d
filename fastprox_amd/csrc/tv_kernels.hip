// HIP kernels of the image-TV engine for gfx950 (CDNA4, wave64).
//
// The solve of a batch is a sequence of launch pairs:
//   step      every tile of every live image runs up to TT dual-FISTA
//             iterations (tv_iso.cpp:491-537 / _tv_torch.py:259-281) from the
//             image's current state buffer into the other one, and writes the
//             restart inner products and duality gaps its levels are asked for
//             as FP64 partial sums (one slot per tile and level);
//   epilogue  one wavefront per image sums those partials over the image's
//             tiles in a fixed order, replays the schedule of the reference on
//             them (restart test, interval rule, gap check), accepts the launch
//             or rolls the image back to the first event inside it, and plans
//             the image's next launch, appending its tiles to the next list.
// A launch is speculative: it assumes no restart and no stop inside it. An
// event at a level before the last makes the image re-run just that prefix
// from its untouched input buffer, so every image follows the reference's
// iteration sequence exactly; events are rare (a stop once per image,
// restarts a few times per solve).
//
// The 2D step kernel keeps a tile in registers: a wavefront covers 64
// columns (one per lane) and RY rows (registers), with a halo of TT on each
// side, so TT iterations need one HBM round trip of the state; neighbours
// along the row come from the adjacent lane (cross-lane shifts), neighbours
// along the column from the lane's own registers. The generic one-iteration
// kernel (step1) serves 3D images and reads its neighbours through the cache.
//
// Per pixel the arithmetic follows the reference's expressions in order
// (tv_iso.cpp:251-288, 499-524), so in float64 with contraction off every
// accepted iterate has the reference's bits; the reductions differ only in
// their summation order.
//
// SPDX-License-Identifier: MIT
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "tv_common.hpp"
#include "tv_kernels.hpp"

namespace fpr {

// ---------------------------------------------------------------------------
// Small device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ double fista_t(double t) { return 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t)); }

__device__ __forceinline__ bool is_test(int sched, int n, int max_iter) {
    return sched == kSchedC ? (n - 1) % 4 == 0 : (n % 4 == 0 || n == max_iter);
}
__device__ __forceinline__ bool is_gap(int n, int max_iter) { return n % 4 == 0 || n == max_iter; }

template <class T>
__device__ __forceinline__ T lane_prev(T v) {  // the value of lane - 1 (lane 0: its own)
    return __shfl_up(v, 1, 64);
}
template <class T>
__device__ __forceinline__ T lane_next(T v) {  // the value of lane + 1 (lane 63: its own)
    return __shfl_down(v, 1, 64);
}

// The square root in a gap term. float64 iterates: the correctly rounded
// sqrt, as the reference (its stops at the same iterations). float32
// iterates: the hardware's reciprocal square root (relative error about
// 5e-8) and one Newton step on the root (about 4e-15 for normal v, far below
// the float32 iterate's own rounding), without the library sqrt's special
// cases: v is a finite sum of squares here, and v = 0 gives 0.
template <class T>
__device__ __forceinline__ double gap_sqrt(double v) {
    if constexpr (sizeof(T) == 8) {
        return sqrt(v);
    } else {
        const double y = __builtin_amdgcn_rsq(v);
        const double s = v * y;
        const double r = __builtin_fma(-s, s, v);
        return v > 0.0 ? __builtin_fma(0.5 * y, r, s) : 0.0;
    }
}

__device__ __forceinline__ double wave_sum(double v) {  // fixed butterfly: the same tree on every call
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// The images' buffers (input, output, dual, weights) are device memory: access
// through global-address-space pointers compiles to global rather than flat
// instructions (a flat access also counts against the LDS counter, so every
// LDS wait would wait for the row loads in flight as well).
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<const T> gin(const void *p) {
    return (gptr<const T>)p;
}
template <class T>
__device__ __forceinline__ gptr<T> gout(void *p) {
    return (gptr<T>)p;
}

// Element o of a plane at a uniform base, through a 32-bit byte offset
// (o * sizeof(T) < 2^32, which kStreamMaxPlaneBytes guarantees for the streaming
// 2D kernel): the access takes the scalar-base form, with no 64-bit address
// arithmetic per lane. The byte offset goes through an empty asm that keeps it
// a 32-bit register value: a caller's `in ? o : 0` would otherwise become a
// select of two 64-bit offsets, which the scalar-base form cannot take.
__device__ __forceinline__ uint32_t opaque32(uint32_t v) {
    asm("" : "+v"(v));
    return v;
}
template <class T>
__device__ __forceinline__ T ld32(gptr<const T> p, uint32_t o) {
    return *(gptr<const T>)((gptr<const char>)p + opaque32(o * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ void st32(gptr<T> p, uint32_t o, T v) {
    *(gptr<T>)((gptr<char>)p + opaque32(o * (uint32_t)sizeof(T))) = v;
}

// A pointer every lane holds the same value of, moved to scalar registers:
// accesses at it plus a 32-bit per-lane byte offset then take the
// scalar-base form (one 32-bit offset register per lane, no 64-bit address
// arithmetic), which the compiler does not find by itself when the base is
// built from fields loaded into vector registers.
template <class P>
__device__ __forceinline__ P uniform(P p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (P)(((uint64_t)hi << 32) | (uint64_t)lo);
}

// x 2^-e: exact unless the result leaves the type's range, so float64 and
// float32 give the correctly rounded product directly (one ldexp). With e = 0
// the ldexp returns v itself (denormals are kept in both precisions), so no
// select on e is needed.
template <class T>
__device__ __forceinline__ T scale_in(T v, int e) {
    if constexpr (sizeof(T) == 4) return ldexpf(v, -e);
    else return ldexp(v, -e);
}

// The radius of pixel p with pixel weights: tv_iso.cpp:180-185.
__device__ __forceinline__ double iso_radius(double gamma, double al, double bound) {
    if (al == 0.0) return 0.0;
    const double r = gamma * al;
    return r <= bound ? r : bound;
}

// ---------------------------------------------------------------------------
// Plan and walk (the schedule, per image, one thread)
// ---------------------------------------------------------------------------

__device__ void plan_next(ImgState &S, int TT) {
    if (S.mode != kIterate) return;
    const int rem = S.max_iter - S.it;
    int cnt = S.sched == kSchedC ? ((1 - S.it) % TT + TT) % TT : ((-S.it) % TT + TT) % TT;
    if (cnt == 0) cnt = TT;
    if (cnt > rem) cnt = rem;
    S.cnt = cnt;
    double t = S.t;
    uint32_t tm = 0, gm = S.pending ? 1u : 0u;
    for (int j = 1; j <= cnt; ++j) {
        const double tn = fista_t(t);
        S.beta[j] = (t - 1.0) / tn;
        t = tn;
        const int n = S.it + j;
        if (is_test(S.sched, n, S.max_iter)) tm |= 1u << j;
        if (j < cnt && is_gap(n, S.max_iter)) gm |= 1u << j;
    }
    S.test_mask = tm;
    S.gap_mask = gm;
}

__device__ void finish(ImgState &S, int64_t N, double u_round, double eta) {
    S.mode = kDone;
    const double gp = S.gp;
    if (gp != gp) {
        S.zerr = gp;
        return;
    }
    // driver.hpp:1186-1212 on the solved scale.
    const double e_ = sqrt(2.0 * (gp < 0.0 ? 0.0 : gp));
    const double a = ldexp(eta, -S.e) * sqrt((double)N);
    const double s = S.sumsq;
    if (s >= 1e-280 && s <= 1e280) {
        const double n = sqrt(s);
        S.zerr = e_ / n + (u_round * (1.0 + e_ / n) + a / n);
    } else {
        const double m = S.maxabs;  // max |x_s|; sumsq_m = sum (x_s / m)^2
        if (!(m > 0.0)) {
            S.zerr = 0.0;
            return;
        }
        const double rt = sqrt(S.sumsq_m);
        const double r = e_ / m / rt;
        S.zerr = r + (u_round * (1.0 + r) + a / m / rt);
    }
}

// The walk over the levels 0..cnt of the launch just run, with gaps[j] and
// dots[j] summed over the image's tiles; then accept or roll back, and plan.
__device__ void walk(ImgState &S, const double *gaps, const double *dots, int TT, int64_t N, double u_round,
                     double eta) {
    if (S.mode == kOutput || S.mode == kOutputNaN) {
        S.it += S.cnt;
        if (S.mode == kOutput && S.cnt > 0) {  // the output launch stored its iterate
            S.cur ^= 1;
            S.fresh = 0;
        }
        S.converged = S.mode == kOutputNaN ? 1 : (S.gp <= S.target);
        if (S.mode == kOutputNaN) S.it = 0;
        finish(S, N, u_round, eta);
        return;
    }
    if (S.mode != kIterate) return;
    const int cnt = S.cnt;
    double t = S.t, gp = S.gp;
    int since = S.since;
    int ev = 0, evj = -1;  // 1 stop, 2 restart
    if (S.gap_mask & 1u) {
        gp = gaps[0];
        if (gp <= S.target || S.it == S.max_iter) {
            ev = 1;
            evj = 0;
        }
    }
    for (int j = 1; j <= cnt && !ev; ++j) {
        t = fista_t(t);
        const int n = S.it + j;
        if (S.sched == kSchedC) ++since;
        if ((S.test_mask >> j) & 1u) {
            const double s = dots[j];
            const bool rs = S.sched == kSchedC ? (s > 0.0 && since >= S.last) : (s > 0.0);
            if (rs) {
                ev = 2;
                evj = j;
            }
        }
        if ((S.gap_mask >> j) & 1u) {
            gp = gaps[j];
            if (gp <= S.target || n == S.max_iter) {
                ev = 1;
                evj = j;
            }
        }
    }
    S.gp = gp;
    if (ev == 1) {  // stop at level evj: an output launch re-runs that prefix
        S.mode = kOutput;
        S.cnt = evj;
        S.test_mask = 0;
        S.gap_mask = 0;
        S.pending = 0;
        return;
    }
    if (ev == 2 && evj < cnt) {  // restart inside: re-run the prefix, then restart
        S.cnt = evj;
        S.test_mask = 0;
        S.gap_mask = 0;
        S.pending = 0;
        S.after = kAfterRestart;
        return;
    }
    // Accept every level.
    S.it += cnt;
    if (cnt > 0) {
        S.cur ^= 1;
        S.fresh = 0;
    }
    S.t = t;
    S.since = since;
    S.vsrc_u = 0;
    S.pending = (cnt > 0 && is_gap(S.it, S.max_iter)) ? 1 : 0;
    if (ev == 2 || S.after == kAfterRestart) {
        S.t = 1.0;
        S.vsrc_u = 1;
        if (S.sched == kSchedC) {
            S.last = S.since;
            S.since = 0;
        }
    }
    S.after = kAfterNone;
    plan_next(S, TT);
}

// ---------------------------------------------------------------------------
// Epilogue: one wavefront per image
// ---------------------------------------------------------------------------

// The epilogue of one image, run by one wavefront (lane 0..63): the per-level
// sums of its tiles' partials (the same fixed tree wherever it runs), the
// walk, and the image's tiles for the next launch. gaps and dots: kMaxTT + 1
// doubles of LDS for this wave, read and written by lane 0 only, so no
// barrier is needed.
__device__ void epilogue_img(const ImgDesc *desc, ImgState *state, const double *part, Tile *list_out,
                             int *count_out, int TT, double u_round, double eta, int img, int lane, double *gaps,
                             double *dots) {
    ImgState &S = state[img];
    const int mode = S.mode;
    if (mode == kDone || mode == kTV1D) return;
    const ImgDesc &D = desc[img];
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int cnt = S.cnt;
    const uint32_t gm = S.gap_mask, tm = S.test_mask;
    if (mode == kIterate) {
        // Every level at once (a tile's kSlots partials are contiguous), so
        // the loads and the levels' butterflies overlap; each level's sum is
        // still its tiles in list order per lane, then the fixed butterfly.
        // Levels the masks do not ask for are summed and dropped.
        double s[kSlots];
#pragma unroll
        for (int j = 0; j < kSlots; ++j) s[j] = 0.0;
        for (int i = lane; i < D.ntiles; i += 64) {
            const double2 *p = (const double2 *)(part + (int64_t)(D.tile_base + i) * kSlots);
#pragma unroll
            for (int j = 0; j <= kMaxTT; ++j) {
                const double2 v = p[j];
                s[2 * j] += v.x;
                s[2 * j + 1] += v.y;
            }
        }
#pragma unroll
        for (int j = 0; j < kSlots; ++j) s[j] = wave_sum(s[j]);
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j <= kMaxTT; ++j)
                if (j <= cnt) {
                    gaps[j] = ((gm >> j) & 1u) ? s[2 * j] : 0.0;
                    dots[j] = ((tm >> j) & 1u) ? s[2 * j + 1] : 0.0;
                }
        }
    }
    int base = -1;
    if (lane == 0) {
        if ((mode == kIterate || mode == kOutput) && cnt > 0) ++S.passes;
        walk(S, gaps, dots, TT, N, u_round, eta);
        const int m = S.mode;
        base = (m == kIterate || m == kOutput || m == kOutputNaN) ? atomicAdd(count_out, D.ntiles) : -1;
    }
    const int b = __shfl(base, 0, 64);
    if (b < 0) return;
    const int nxy = D.ntx * D.nty;
    for (int i = lane; i < D.ntiles; i += 64) {
        Tile t;
        t.img = img;
        t.tz = i / nxy;
        const int r = i - t.tz * nxy;
        t.ty = (int16_t)(r / D.ntx);
        t.tx = (int16_t)(r - t.ty * D.ntx);
        t.pad_ = i;
        list_out[b + i] = t;
    }
}

__global__ __launch_bounds__(64) void epilogue_kernel(const ImgDesc *desc, ImgState *state, const double *part,
                                                      Tile *list_out, int *count_out, int TT, double u_round,
                                                      double eta, int nimg) {
    const int img = blockIdx.x;
    if (img >= nimg) return;
    __shared__ double gaps[kMaxTT + 1], dots[kMaxTT + 1];
    epilogue_img(desc, state, part, list_out, count_out, TT, u_round, eta, img, threadIdx.x, gaps, dots);
}

// ---------------------------------------------------------------------------
// Prologue: one workgroup of 256 per image
// ---------------------------------------------------------------------------

__device__ double block_sum(double v, double *sh) {
    const int tid = threadIdx.x;
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    const double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

__device__ double block_max(double v, double *sh) {
    const int tid = threadIdx.x;
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    __syncthreads();
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    const double r = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    __syncthreads();
    return r;
}

__device__ int rescale_exponent(double m, double gamma, double step) {  // image_util.hpp:56-65
    if (!(m > 0.0)) return 0;
    int e = ilogb(m);
    if (e >= -64 && e <= 64) return 0;
    const double top = gamma < step ? step : gamma;
    if (e < 0 && top > 0.0) {
        const int lo = ilogb(top) + 1 - 960;
        const int t = e > lo ? e : lo;
        e = t < 0 ? t : 0;
    }
    return e;
}

// The prologue's reductions run over many workgroups per image: pass p
// (0: NaN, Inf and max |x|; 1: sum and sum of squares of x_s = 2^-e x;
// 2: sum |x_s - mean| and sum (x_s / max)^2) writes one partial per block,
// and a combine kernel (one thread per image) adds them in block order.
constexpr int kProBlock = 256;
constexpr int kProChunk = kProBlock * 16;  // entries per block

template <class T>
__global__ __launch_bounds__(kProBlock) void prologue_pass(const ImgDesc *desc, const ImgState *state,
                                                           double *part, int nblk, int pass) {
    const int img = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x;
    const ImgDesc &D = desc[img];
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int64_t lo = (int64_t)blk * kProChunk;
    if (lo >= N) return;
    const int64_t hi = lo + kProChunk < N ? lo + kProChunk : N;
    const auto x = gin<T>(D.d);
    __shared__ double sh[4];
    double *out = part + ((int64_t)img * nblk + blk) * 4;
    if (pass == 0) {
        double m = 0.0;
        int nan = 0, inf = 0;
        for (int64_t i = lo + tid; i < hi; i += kProBlock) {
            const double v = (double)x[i];
            nan |= v != v;
            inf |= isinf(v);
            m = fmax(m, fabs(v));
        }
        __shared__ int flags;
        if (tid == 0) flags = 0;
        __syncthreads();
        if (nan) atomicOr(&flags, 1);
        if (inf) atomicOr(&flags, 2);
        m = block_max(m, sh);
        __syncthreads();
        if (tid == 0) {
            out[0] = m;
            out[1] = (double)flags;
        }
        return;
    }
    const ImgState &S = state[img];
    if (S.status || S.mode == kOutputNaN) return;
    const int e = S.e;
    double a = 0.0, b = 0.0;
    if (pass == 1) {
        for (int64_t i = lo + tid; i < hi; i += kProBlock) {
            const double v = ldexp((double)x[i], -e);
            a += v;
            b += v * v;
        }
    } else {
        const double mean = S.t, ms = S.maxabs;  // the mean is parked in t until the combine of pass 2
        for (int64_t i = lo + tid; i < hi; i += kProBlock) {
            const double v = ldexp((double)x[i], -e);
            a += fabs(v - mean);
            const double r = ms > 0.0 ? v / ms : 0.0;
            b += r * r;
        }
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (tid == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// Pass-0 combine: flags, max, the active axes and the rescaling exponent
// (tv_iso.cpp:341-377).
__global__ void prologue_combine0(Problem P, const ImgDesc *desc, ImgState *state, const double *part, int nblk,
                                  const double *gammas, int nimg) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= nimg) return;
    const ImgDesc &D = desc[img];
    ImgState &S = state[img];
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int nb = (int)((N + kProChunk - 1) / kProChunk);
    double m = 0.0;
    int fl = 0;
    for (int j = 0; j < nb; ++j) {
        const double *q = part + ((int64_t)img * nblk + j) * 4;
        m = fmax(m, q[0]);
        fl |= (int)q[1];
    }
    const int k = P.k;
    const double gamma0 = gammas[img];
    double L[3];
    for (int a = 0; a < k; ++a) L[a] = (double)D.L[3 - k + a];
    int e = 0;
    if (P.iso && !fl) {
        double wmaxL = 0.0, wa[3] = {0.0, 0.0, 0.0};
        int na = 0;
        for (int a = 0; a < k; ++a)
            if (L[a] > 1) wmaxL = fmax(wmaxL, P.w[a]);
        for (int a = 0; a < k; ++a)
            if (gamma0 > 0.0 && P.w[a] > 0.0 && L[a] > 1 && P.w[a] >= 1e-100 * wmaxL) {
                wa[a] = P.w[a];
                ++na;
            }
        double ss = 0.0;
        for (int a = 0; a < k; ++a) ss = fmax(ss, P.w[a]);
        e = na ? rescale_exponent(m, gamma0, gamma0 * ss * S.sumsq_m) : 0;  // sumsq_m: the weight max, for now
        for (int a = 0; a < k && e; ++a)
            if (wa[a] > 0.0 && !(ldexp(gamma0 * wa[a], -e) > 0.0)) e = 0;
    }
    S.e = e;
    S.status = (fl & 2) ? 1 : 0;
    S.mode = (fl & 2) ? kDone : ((fl & 1) ? kOutputNaN : kIterate);
    S.maxabs = ldexp(m, -e);
}

// The pixel weights' max per image (the step scale of the rescaling).
template <class T>
__global__ __launch_bounds__(256) void weight_max_kernel(Problem P, const ImgDesc *desc, ImgState *state, int nimg) {
    const int img = blockIdx.x;
    if (img >= nimg) return;
    const ImgDesc &D = desc[img];
    ImgState &S = state[img];
    __shared__ double sh[4];
    double am = 0.0;
    if (P.sw && P.sw_per_image && P.iso) {
        const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
        const auto al = gin<T>(D.sw);
        for (int64_t i = threadIdx.x; i < N; i += 256) am = fmax(am, (double)al[i]);
        am = block_max(am, sh);
    } else {
        am = P.alpha_max;
    }
    if (threadIdx.x == 0) S.sumsq_m = am;
}

// Pass-1 combine: the mean (parked in t for pass 2) and |x_s|^2.
__global__ void prologue_combine1(const ImgDesc *desc, ImgState *state, const double *part, int nblk, int nimg) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= nimg) return;
    const ImgDesc &D = desc[img];
    ImgState &S = state[img];
    if (S.status || S.mode == kOutputNaN) return;
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int nb = (int)((N + kProChunk - 1) / kProChunk);
    double a = 0.0, b = 0.0;
    for (int j = 0; j < nb; ++j) {
        const double *q = part + ((int64_t)img * nblk + j) * 4;
        a += q[0];
        b += q[1];
    }
    S.t = a / (double)N;
    S.sumsq = b;
}

// Pass-2 combine: set the image up (tv_iso.cpp:333-415, 424-483;
// _tv_torch.py:229-257) and plan its first launch.
__global__ void prologue_combine2(Problem P, const ImgDesc *desc, ImgState *state, const double *part, int nblk,
                                  const double *gammas, int nimg, const double *tols, const int *maxits) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= nimg) return;
    const ImgDesc &D = desc[img];
    ImgState &S = state[img];
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int nb = (int)((N + kProChunk - 1) / kProChunk);
    double s3 = 0.0, s4 = 0.0;
    if (!S.status && S.mode != kOutputNaN)
        for (int j = 0; j < nb; ++j) {
            const double *q = part + ((int64_t)img * nblk + j) * 4;
            s3 += q[0];
            s4 += q[1];
        }
    const int k = P.k;
    const int e = S.e;
    const double gamma0 = gammas[img];
    double wa[3] = {0.0, 0.0, 0.0}, L[3];
    for (int a = 0; a < k; ++a) L[a] = (double)D.L[3 - k + a];
    int na = 0, act = -1;
    if (P.iso) {
        double wmaxL = 0.0;
        for (int a = 0; a < k; ++a)
            if (L[a] > 1) wmaxL = fmax(wmaxL, P.w[a]);
        for (int a = 0; a < k; ++a)
            if (gamma0 > 0.0 && P.w[a] > 0.0 && L[a] > 1 && P.w[a] >= 1e-100 * wmaxL) {
                wa[a] = P.w[a];
                act = a;
                ++na;
            }
    }
    const double s2 = S.sumsq;
    S.it = 0;
    S.since = 0;
    S.last = 0;
    S.cur = 0;
    S.vsrc_u = 0;
    S.fresh = 1;
    S.cnt = 0;
    S.test_mask = 0;
    S.gap_mask = 0;
    S.pending = 0;
    S.after = kAfterNone;
    S.sched = P.sched;
    S.max_iter = maxits ? maxits[img] : P.max_iter;
    S.converged = 1;
    S.t = 1.0;
    S.gp = INFINITY;
    S.sumsq_m = s4;
    S.zerr = 0.0;
    S.bound = INFINITY;
    S.wmax = 1.0;
    S.gamma_raw = ldexp(gamma0, -e);
    const double tol = tols ? tols[img] : P.tol;
    S.target = tol * 0.5 * s2;
    S.active_axis = -1;
    S.passes = 0;
    S.tv1d_lam = 0.0;
    for (int a = 0; a < 3; ++a) S.wa[a] = 0.0;
    if (S.status) {  // an infinite entry: the call fails (FP_EINVAL)
        S.mode = kDone;
        return;
    }
    if (S.mode == kOutputNaN) {  // not solved: NaN outputs, gap NaN, 0 iterations
        S.gp = NAN;
        return;
    }
    const double gamma = S.gamma_raw;
    if (P.iso) {
        if (na == 0) {  // the identity
            S.mode = kOutput;
            S.gp = 0.0;
            return;
        }
        if (na == 1 && !P.sw) {  // exactly 1D TV along the active axis
            S.mode = kTV1D;
            S.active_axis = act;
            S.tv1d_lam = gamma * P.w[act];
            S.gp = 0.0;
            return;
        }
        double wmax = 0.0, whmin = 1.0;
        for (int a = 0; a < k; ++a) wmax = wmax < wa[a] ? wa[a] : wmax;
        for (int a = 0; a < k; ++a)
            if (wa[a] != 0.0) {
                wa[a] = wa[a] / wmax;
                whmin = wa[a] < whmin ? wa[a] : whmin;
            }
        double g = gamma * wmax;
        const double bound = 2.0 * s3 / whmin + DBL_MIN;
        if (!(g <= bound) && !P.sw) g = bound;
        double L2 = 0.0;
        for (int a = 0; a < k; ++a) L2 += 4.0 * wa[a] * wa[a];
        S.tau = 1.0 / L2;
        S.gamma = g;
        S.bound = bound;
        S.wmax = wmax;
        for (int a = 0; a < k; ++a) S.wa[a] = wa[a];
        S.pending = P.warm ? 1 : 0;
    } else {
        double L2 = 0.0;
        for (int a = 0; a < k; ++a) {
            S.wa[a] = L[a] > 1 ? P.w[a] : 0.0;
            L2 += 4.0 * P.w[a] * P.w[a];
        }
        S.target = tol * s2 / 2.0;
        if (L2 == 0.0 || !(gamma > 0.0)) {  // _tv_torch.py:247-250: the zero dual, z = x
            S.mode = kOutput;
            S.gp = 0.0;
            S.gamma = 0.0;
            S.tau = 0.0;
            return;
        }
        S.tau = 1.0 / L2;
        S.gamma = gamma;
    }
    S.mode = kIterate;
    S.fresh = P.warm ? 0 : 1;
    plan_next(S, P.TT);
}

// ---------------------------------------------------------------------------
// Warm starts: the dual buffer to the initial state (buffer 0, u and v)
// ---------------------------------------------------------------------------

// Iso: u = dual * wmax * 2^-e, zero on inactive axes, past the last entry and
// where not finite, then projected onto the pixel's ball (tv_iso.cpp:456-477).
template <class T>
__global__ void warm_iso_kernel(Problem P, T *arena, const ImgDesc *desc,
                                const ImgState *state, int nimg) {
    const int img = blockIdx.y;
    if (img >= nimg) return;
    const ImgState &S = state[img];
    if (S.mode != kIterate) return;
    const ImgDesc &D = desc[img];
    const int k = P.k;
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < N; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c2 = p % D.L[2], c1 = (p / D.L[2]) % D.L[1], c0 = p / ((int64_t)D.L[2] * D.L[1]);
        const int64_t cc[3] = {c0, c1, c2};
        double u[3];
        double nn = 0.0;
        for (int a = 0; a < k; ++a) {
            const int ax = 3 - k + a;
            double v = ldexp((double)((const T *)D.q)[a * D.q_cstride + p] * S.wmax, -S.e);
            const bool has = cc[ax] + 1 < D.L[ax];
            if (S.wa[a] == 0.0 || !has || !isfinite(v)) v = 0.0;
            u[a] = v;
            nn += v * v;
        }
        double rp = S.gamma;
        if (P.sw) rp = iso_radius(S.gamma, (double)gin<T>(D.sw)[p], S.bound);
        if (sqrt(nn) > rp) {
            const double sc = rp / sqrt(nn);
            for (int a = 0; a < k; ++a) u[a] *= sc;
        }
        T *U = arena + D.st_off;
        for (int a = 0; a < k; ++a) {
            U[a * N + p] = (T)u[a];
            U[(k + a) * N + p] = (T)u[a];
        }
    }
}

// Aniso: u_a = -cumsum_a(q_a) / w_a along every fibre, clamped to the box,
// zero at the last entry (_tv_torch.py:596-606, 242-243). One thread per fibre.
template <class T>
__global__ void warm_aniso_kernel(Problem P, T *arena, const ImgDesc *desc,
                                  const ImgState *state, int nimg) {
    const int img = blockIdx.y;
    if (img >= nimg) return;
    const ImgState &S = state[img];
    if (S.mode != kIterate) return;
    const ImgDesc &D = desc[img];
    const int k = P.k;
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int64_t Ls[3] = {D.L[0], D.L[1], D.L[2]};
    const int64_t St[3] = {(int64_t)D.L[1] * D.L[2], D.L[2], 1};
    T *U = arena + D.st_off;
    for (int a = 0; a < k; ++a) {
        const int ax = 3 - k + a;
        const int64_t L = Ls[ax], s = St[ax], nf = N / L;
        for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
            // fibre f: its start has coordinate 0 along ax
            const int64_t outer = f / s, inner = f % s;
            const int64_t st = outer * L * s + inner;
            double cs = 0.0;
            for (int64_t i = 0; i < L; ++i) {
                const int64_t o = st + i * s;
                cs += (double)((const T *)D.q)[a * D.q_cstride + o];
                double v = P.w[a] > 0.0 ? -cs / P.w[a] : 0.0;
                double r = S.gamma;
                if (P.sw) r = S.gamma * (double)gin<T>(D.sw)[a * D.sw_cstride + o];
                v = r > 0.0 ? fmin(fmax(v, -r), r) : 0.0;
                if (i + 1 == L || !isfinite(v)) v = 0.0;
                U[a * N + o] = (T)v;
                U[(k + a) * N + o] = (T)v;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Exact 1D TV (the single-active-axis case): Condat's direct algorithm per
// fibre, one thread per fibre; the dual is the running sum of x - z
// (tv_iso.cpp:381-398).
// ---------------------------------------------------------------------------

template <class T>
__device__ void condat(const T *y, int64_t ys, T *x, int64_t xs, int64_t n, double lam, int e) {
    auto Y = [&](int64_t i) { return ldexp((double)y[i * ys], -e); };
    auto X = [&](int64_t i, double v) { x[i * xs] = (T)ldexp(v, e); };
    if (n == 1 || !(lam > 0.0)) {
        for (int64_t i = 0; i < n; ++i) X(i, Y(i));
        return;
    }
    int64_t k = 0, k0 = 0, kmin = 0, kplus = 0;
    double vmin = Y(0) - lam, vmax = Y(0) + lam, umin = lam, umax = -lam;
    const double twol = 2.0 * lam;
    for (;;) {
        if (k == n - 1) {
            if (umin < 0.0) {
                for (int64_t i = k0; i <= kmin; ++i) X(i, vmin);
                k = k0 = kmin = kmin + 1;
                vmin = Y(k);
                umin = lam;
                umax = Y(k) + lam - vmax;
                continue;
            } else if (umax > 0.0) {
                for (int64_t i = k0; i <= kplus; ++i) X(i, vmax);
                k = k0 = kplus = kplus + 1;
                vmax = Y(k);
                umax = -lam;
                umin = Y(k) - lam - vmin;
                continue;
            } else {
                const double v = vmin + umin / (double)(k - k0 + 1);
                for (int64_t i = k0; i <= k; ++i) X(i, v);
                return;
            }
        }
        const double yn = Y(k + 1);
        if (yn + umin < vmin - lam) {
            for (int64_t i = k0; i <= kmin; ++i) X(i, vmin);
            k = k0 = kmin = kplus = kmin + 1;
            vmin = Y(k);
            vmax = Y(k) + twol;
            umin = lam;
            umax = -lam;
        } else if (yn + umax > vmax + lam) {
            for (int64_t i = k0; i <= kplus; ++i) X(i, vmax);
            k = k0 = kmin = kplus = kplus + 1;
            vmin = Y(k) - twol;
            vmax = Y(k);
            umin = lam;
            umax = -lam;
        } else {
            k = k + 1;
            umin = umin + Y(k) - vmin;
            umax = umax + Y(k) - vmax;
            if (umin >= lam) {
                vmin = vmin + (umin - lam) / (double)(k - k0 + 1);
                umin = lam;
                kmin = k;
            }
            if (umax <= -lam) {
                vmax = vmax + (umax + lam) / (double)(k - k0 + 1);
                umax = -lam;
                kplus = k;
            }
        }
    }
}

template <class T>
__global__ void tv1d_kernel(Problem P, const ImgDesc *desc, ImgState *state, int nimg) {
    const int img = blockIdx.y;
    if (img >= nimg) return;
    ImgState &S = state[img];
    if (S.mode != kTV1D) return;
    const ImgDesc &D = desc[img];
    const int k = P.k;
    const int a = S.active_axis, ax = 3 - k + a;
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const int64_t Ls[3] = {D.L[0], D.L[1], D.L[2]};
    const int64_t St[3] = {(int64_t)D.L[1] * D.L[2], D.L[2], 1};
    const int64_t L = Ls[ax], s = St[ax], nf = N / L;
    const int e = S.e;
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nf; f += (int64_t)gridDim.x * blockDim.x) {
        const int64_t outer = f / s, inner = f % s;
        const int64_t st = outer * L * s + inner;
        const T *d = (const T *)D.d;
        T *z = (T *)D.z;
        T *q = (T *)D.q;
        const T *y = d + st;
        T *x = z + st;
        condat(y, s, x, s, L, S.tv1d_lam, e);
        double cs = 0.0;
        for (int64_t i = 0; i < L; ++i) {
            const int64_t o = st + i * s;
            cs += ldexp((double)d[o], -e) - ldexp((double)z[o], -e);
            for (int b = 0; b < k; ++b)
                q[b * D.q_cstride + o] = b == a ? (T)(i + 1 < L ? ldexp(-cs / P.w[a], e) : 0.0) : (T)0;
        }
    }
}

__global__ void tv1d_finish_kernel(const ImgDesc *desc, ImgState *state, int nimg, double u_round, double eta) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= nimg) return;
    ImgState &S = state[img];
    if (S.mode != kTV1D) return;
    const ImgDesc &D = desc[img];
    S.gp = 0.0;
    S.it = 0;
    S.converged = 1;
    finish(S, (int64_t)D.L[0] * D.L[1] * D.L[2], u_round, eta);
}

// ---------------------------------------------------------------------------
// The 2D step kernel: a wavefront per tile, TT iterations in registers.
// ---------------------------------------------------------------------------

template <class T, bool ISO, bool SW, int TT, int RY>
__device__ __forceinline__ void tile2d(const StepArgs<T> &A, const Tile tl, const int lane) {
    constexpr int TXI = 64 - 2 * TT;
    constexpr int TYI = RY - 2 * TT;
    static_assert(RY <= 32 && TYI > 0, "tile rows");
    const int img = __builtin_amdgcn_readfirstlane(tl.img);
    const int tx = __builtin_amdgcn_readfirstlane(tl.tx);
    const int ty = __builtin_amdgcn_readfirstlane(tl.ty);
    const int tloc = __builtin_amdgcn_readfirstlane(tl.pad_);
    const ImgDesc &D = A.desc[img];
    ImgState &S = A.state[img];
    const int mode = S.mode;
    const int H = D.L[1], W = D.L[2];
    const int64_t N = (int64_t)H * W;
    const int x = tx * TXI - TT + lane;
    const int y0 = ty * TYI - TT;
    const bool xin = x >= 0 && x < W;
    const bool lane_in = lane >= TT && lane < 64 - TT;
    const double w0d = S.wa[0], w1d = S.wa[1];
    const T w0 = (T)w0d, w1 = (T)w1d;
    uint32_t rowin = 0, hn0 = 0;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const int y = y0 + r;
        if (y >= 0 && y < H) rowin |= 1u << r;
        if (y >= 0 && y + 1 < H) hn0 |= 1u << r;
    }
    if (w0d == 0.0 || !xin) hn0 = 0;
    const bool hn1 = xin && x + 1 < W && w1d != 0.0;
    constexpr uint32_t kInt = ((RY - TT >= 32) ? 0xffffffffu : ((1u << (RY - TT)) - 1u)) & ~((1u << TT) - 1u);
    const uint32_t intr = lane_in && xin ? (rowin & kInt) : 0u;  // interior rows of this lane
    double *part = A.part + (int64_t)(D.tile_base + tloc) * kSlots;

    if (mode == kOutputNaN) {
        const T nanv = (T)NAN;
#pragma unroll
        for (int r = 0; r < RY; ++r)
            if ((intr >> r) & 1u) {
                const int64_t o = (int64_t)(y0 + r) * W + x;
                gout<T>(D.z)[o] = nanv;
                gout<T>(D.q)[o] = nanv;
                gout<T>(D.q)[D.q_cstride + o] = nanv;
            }
        return;
    }

    const int e = S.e;
    const int cnt = S.cnt;
    const uint32_t tm0 = __builtin_amdgcn_readfirstlane(S.test_mask), gm0 = __builtin_amdgcn_readfirstlane(S.gap_mask);
    const uint32_t tm = tm0, gm = gm0;
    const bool fresh = S.fresh;
    const T tau = (T)S.tau;
    const double gamd = S.gamma, boundd = S.bound;
    const T gam = (T)gamd;
    const auto dp = gin<T>(D.d);
    const T *U = A.arena + D.st_off + (int64_t)S.cur * 4 * N;
    const T *V = S.vsrc_u ? U : U + 2 * N;

    T d[RY], u0[RY], u1[RY], v0[RY], v1[RY];
    T ra[SW ? RY : 1], rb[(SW && !ISO) ? RY : 1];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
        const bool in = xin && ((rowin >> r) & 1u);
        const int64_t o = in ? (int64_t)(y0 + r) * W + x : 0;
        d[r] = in ? scale_in(dp[o], e) : (T)0;
        if (fresh) {
            u0[r] = u1[r] = v0[r] = v1[r] = (T)0;
        } else {
            u0[r] = in ? U[o] : (T)0;
            u1[r] = in ? U[N + o] : (T)0;
            v0[r] = in ? V[o] : (T)0;
            v1[r] = in ? V[N + o] : (T)0;
        }
        if constexpr (SW) {
            const auto sw = gin<T>(D.sw);
            if constexpr (ISO) {
                ra[r] = in ? (T)iso_radius(gamd, (double)sw[o], boundd) : (T)0;
            } else {
                ra[r] = in ? (T)(gamd * (double)sw[o]) : (T)0;
                rb[r] = in ? (T)(gamd * (double)sw[D.sw_cstride + o]) : (T)0;
            }
        }
    }

    // The gap of the state (u) on the interior, FP64 terms (tv_iso.cpp:290-306).
    auto gap_level = [&](int j) {
        double acc = 0.0;
        auto zu = [&](int r) -> double {
            const double up = r > 0 ? (double)u0[r - 1] : 0.0;
            const double ul = lane_prev((double)u1[r]);
            return ((double)d[r] + w0d * ((double)u0[r] - up)) + w1d * ((double)u1[r] - ul);
        };
        double zc = zu(TT - 1);
#pragma unroll
        for (int r = TT - 1; r < RY - TT; ++r) {
            const double zn = zu(r + 1);
            const double zr = lane_next(zc);
            if ((intr >> r) & 1u) {
                const double g0 = ((hn0 >> r) & 1u) ? w0d * (zn - zc) : 0.0;
                const double g1 = hn1 ? w1d * (zr - zc) : 0.0;
                if constexpr (ISO) {
                    const double nn = g0 * g0 + g1 * g1;
                    const double dot = g0 * (double)u0[r] + g1 * (double)u1[r];
                    const double rp = SW ? (double)ra[r] : gamd;
                    acc += rp * gap_sqrt<T>(nn) - dot;
                } else {
                    const double r0 = SW ? (double)ra[r] : gamd, r1 = SW ? (double)rb[r] : gamd;
                    acc += r0 * fabs(g0) - g0 * (double)u0[r];
                    acc += r1 * fabs(g1) - g1 * (double)u1[r];
                }
            }
            zc = zn;
        }
        acc = wave_sum(acc);
        if (lane == 0) part[2 * j] = acc;
    };

    if (gm & 1u) gap_level(0);

#pragma unroll
    for (int j = 1; j <= TT; ++j) {
        if (j > cnt) break;
        const T beta = (T)S.beta[j];
        const bool test = (tm >> j) & 1u;
        double dot = 0.0;
        // z(v) on row r + 1 before row r is updated (z[r + 1] reads the old v0[r]).
        auto zv = [&](int r) -> T {
            const T vu = r > 0 ? v0[r - 1] : (T)0;
            const T vl = lane_prev(v1[r]);
            return (d[r] + w0 * (v0[r] - vu)) + w1 * (v1[r] - vl);
        };
        T zc = zv(0);
#pragma unroll
        for (int r = 0; r < RY; ++r) {
            const T zn = r + 1 < RY ? zv(r + 1) : (T)0;
            const T zr = lane_next(zc);
            const T g0 = ((hn0 >> r) & 1u) ? w0 * (zn - zc) : (T)0;
            const T g1 = (hn1 && ((rowin >> r) & 1u)) ? w1 * (zr - zc) : (T)0;
            const T c0 = w0d != 0.0 ? v0[r] + tau * g0 : (T)0;
            const T c1 = w1d != 0.0 ? v1[r] + tau * g1 : (T)0;
            T n0, n1;
            if constexpr (ISO) {
                const T nn = c0 * c0 + c1 * c1;
                const T nr = sqrt(nn);
                const T rp = SW ? ra[r] : gam;
                const T sc = nr > rp ? rp / nr : (T)1;
                n0 = c0 * sc;
                n1 = c1 * sc;
            } else {
                const T r0 = SW ? ra[r] : gam, r1 = SW ? rb[r] : gam;
                n0 = c0 > r0 ? r0 : (c0 < -r0 ? -r0 : c0);
                n1 = c1 > r1 ? r1 : (c1 < -r1 ? -r1 : c1);
            }
            if (test && ((intr >> r) & 1u)) {
                dot += ((double)v0[r] - (double)n0) * ((double)n0 - (double)u0[r]);
                dot += ((double)v1[r] - (double)n1) * ((double)n1 - (double)u1[r]);
            }
            v0[r] = n0 + beta * (n0 - u0[r]);
            v1[r] = n1 + beta * (n1 - u1[r]);
            u0[r] = n0;
            u1[r] = n1;
            zc = zn;
        }
        if (test) {
            dot = wave_sum(dot);
            if (lane == 0) part[2 * j + 1] = dot;
        }
        if (j < TT && ((gm >> j) & 1u)) gap_level(j);
    }

    if (mode == kOutput) {
        // z(u) and the dual on the interior (tv_iso.cpp:539, 403-413; _tv_torch.py:588-593).
        const double wmax = S.wmax;
#pragma unroll
        for (int r = TT; r < RY - TT; ++r) {
            const double up = (double)u0[r - 1];
            const double ul = lane_prev((double)u1[r]);
            const double zz = ((double)d[r] + w0d * ((double)u0[r] - up)) + w1d * ((double)u1[r] - ul);
            if ((intr >> r) & 1u) {
                const int64_t o = (int64_t)(y0 + r) * W + x;
                gout<T>(D.z)[o] = (T)ldexp(zz, e);
                if constexpr (ISO) {
                    gout<T>(D.q)[o] = (T)ldexp((double)u0[r] / wmax, e);
                    gout<T>(D.q)[D.q_cstride + o] = (T)ldexp((double)u1[r] / wmax, e);
                } else {
                    gout<T>(D.q)[o] = (T)(w0d * (up - (double)u0[r]));
                    gout<T>(D.q)[D.q_cstride + o] = (T)(w1d * (ul - (double)u1[r]));
                }
            }
        }
    }
    if (cnt == 0) return;  // an output launch of cnt > 0 levels also stores the iterate (the kept dual)
    T *Uo = A.arena + D.st_off + (int64_t)(S.cur ^ 1) * 4 * N;
#pragma unroll
    for (int r = TT; r < RY - TT; ++r)
        if ((intr >> r) & 1u) {
            const int64_t o = (int64_t)(y0 + r) * W + x;
            Uo[o] = u0[r];
            Uo[N + o] = u1[r];
            Uo[2 * N + o] = v0[r];
            Uo[3 * N + o] = v1[r];
        }
}

// Tiles of one launch in chunks of 4 (one per wavefront); each XCD (workgroup
// ids congruent mod 8 share one) takes a contiguous range of the list, so an
// image's neighbouring tiles share an L2.
template <class T, bool ISO, bool SW, int TT, int RY>
__global__ __launch_bounds__(256) void step2d_kernel(StepArgs<T> A) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.count_out = 0;
    const int ntiles = *A.count_in;
    const int nchunks = (ntiles + 3) >> 2;
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, j8 = blockIdx.x >> 3;
    const int cpx = (nchunks + 7) >> 3;
    const int cend = min((xcd + 1) * cpx, nchunks);
    for (int c = xcd * cpx + j8; c < cend; c += per) {
        const int ti = c * 4 + wv;
        if (ti < ntiles) tile2d<T, ISO, SW, TT, RY>(A, A.list_in[ti], lane);
    }
}


// ---------------------------------------------------------------------------
// The streaming 2D step kernel. A wavefront owns a strip of 64 columns (one
// per lane; TT halo columns on each side) and walks down a segment of rows.
// The TT iteration levels are pipelined one row apart: at each step the wave
// loads one new row (level 0) and every level m computes the row m above it
// from level m-1's last two rows. Each level keeps only its latest row
// (v, u, z and, when the gap is due, z(u)) in registers, so TT iterations
// cost one HBM round trip of the state without a vertical halo; the rows of
// the segment's first and last TT steps are the pipeline's fill and drain.
// ---------------------------------------------------------------------------

template <class T>
__device__ __forceinline__ T from_left(T v) {  // lane i receives lane i - 1's value (lane 0: 0)
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, true));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x138, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(unsigned)lo);
    }
}
template <class T>
__device__ __forceinline__ T from_right(T v) {  // lane i receives lane i + 1's value (lane 63: 0)
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
    } else {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), 0x130, 0xf, 0xf, true);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x130, 0xf, 0xf, true);
        return __builtin_bit_cast(T, ((long long)hi << 32) | (long long)(unsigned)lo);
    }
}

// The projection factor min(1, r / |c|) of the isotropic ball. float64
// follows the reference's correctly rounded sqrt and division
// (tv_iso.cpp:509-510); float32 uses the hardware's reciprocal square root
// (about 1 ulp, one quarter-rate instruction where a square root and a
// reciprocal take two), its iterate being float32 anyway; |c| > r is tested
// as |c|^2 > r^2.
template <class T>
__device__ __forceinline__ T ball_scale(T nn, T rp) {
    if constexpr (sizeof(T) == 4) {
        return nn > rp * rp ? rp * __builtin_amdgcn_rsqf(nn) : 1.0f;
    } else {
        const T nr = sqrt(nn);
        return nr > rp ? rp / nr : (T)1;
    }
}

// a * b + c: one fused rounding in float32; in float64 the reference's two
// roundings (its builds do not contract, CMakeLists.txt:102-114), for its bits.
template <class T>
__device__ __forceinline__ T madd(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return a * b + c;
}

// Per-wave partial sums in LDS: [gap | dot][level][lane], FP64.
template <int TT>
struct StreamAcc {
    double v[2][TT + 1][64];
};

// The float32 gap term's square root: the hardware's (about 1 ulp), the
// term itself being a float32 quantity summed in FP64.
__device__ __forceinline__ float sqrt32(float v) { return __builtin_amdgcn_sqrtf(v); }

// One tile of the streaming 2D kernel. INTERIOR: the segment's rows and the
// pipeline's TT rows (and one more) on either side lie inside the image, so no
// level tests a row bound. The column bounds and inactive axes are folded
// into per-lane weights for float32 (0 where the difference leaves the
// image); float64 keeps the reference's select on the lane's masks (doubles of
// such weights would cost the registers of a resident wave). float32 computes in
// its own order (tau folded into the weights, the gap and restart terms per
// pixel in float32) and sums the partials in FP64 as float64 does: its
// iterate is float32 either way.
template <class T, bool ISO, bool SW, int TT, bool INTERIOR>
__device__ __forceinline__ void stream_tile(const StepArgs<T> &A, const Tile tl, const int lane, const int SEG,
                                            StreamAcc<TT> &acc) {
    constexpr bool F32 = sizeof(T) == 4;
    using G = std::conditional_t<F32, float, double>;  // gap and restart terms per pixel
    constexpr int TXI = 64 - 2 * TT;
    const int img = __builtin_amdgcn_readfirstlane(tl.img);
    const int tx = __builtin_amdgcn_readfirstlane(tl.tx);
    const int ty = __builtin_amdgcn_readfirstlane(tl.ty);
    const int tloc = __builtin_amdgcn_readfirstlane(tl.pad_);
    const ImgDesc &D = A.desc[img];
    ImgState &S = A.state[img];
    const int mode = S.mode;
    const int H = D.L[1], W = D.L[2];
    const int64_t N = (int64_t)H * W;
    const int x = tx * TXI - TT + lane;
    const bool xin = x >= 0 && x < W;
    const bool lane_in = lane >= TT && lane < 64 - TT && xin;
    const bool xnext = xin && x + 1 < W;
    const int y0 = ty * SEG;
    const int yend = min(y0 + SEG, H);  // rows [y0, yend) are this tile's
    double *part = A.part + (int64_t)(D.tile_base + tloc) * kSlots;

    if (mode == kOutputNaN) {
        if (lane_in)
            for (int r = y0; r < yend; ++r) {
                const int64_t o = (int64_t)r * W + x;
                gout<T>(D.z)[o] = (T)NAN;
                gout<T>(D.q)[o] = (T)NAN;
                gout<T>(D.q)[D.q_cstride + o] = (T)NAN;
            }
        return;
    }

    const int e = S.e;
    const int cnt = S.cnt;
    const uint32_t tm0 = __builtin_amdgcn_readfirstlane(S.test_mask), gm0 = __builtin_amdgcn_readfirstlane(S.gap_mask);
    const uint32_t tm = tm0, gm = gm0;
    const bool fresh = S.fresh;
    const T tau = (T)S.tau;
    const double gamd = S.gamma, boundd = S.bound;
    const T gam = (T)gamd;
    const double w0d = S.wa[0], w1d = S.wa[1];
    const T w0 = (T)w0d, w1 = (T)w1d;
    const bool act0 = w0d != 0.0, act1 = w1d != 0.0;
    const bool out_mode = mode == kOutput;
    // Per-lane weights of the two differences, 0 where the difference leaves
    // the image or the axis is inactive: of g (gw*) and, float32, of tau g (tw*).
    const T gw0 = (act0 && xin) ? w0 : (T)0, gw1 = (act1 && xnext) ? w1 : (T)0;
    const T tw0 = tau * gw0, tw1 = tau * gw1;  // float32 only
    const G gg0 = (G)gw0, gg1 = (G)gw1;        // float32 only
    // Plane bases (uniform, in scalar registers); every access below is base
    // + a 32-bit offset.
    const auto dp = uniform(gin<T>(D.d));
    const gptr<const T> U0 = uniform((gptr<const T>)(A.arena + D.st_off + (int64_t)S.cur * 4 * N));
    const gptr<const T> U1 = uniform(U0 + N);
    const gptr<const T> V0 = uniform(S.vsrc_u ? U0 : U0 + 2 * N), V1 = uniform(V0 + N);
    const gptr<T> Uo = uniform((gptr<T>)(A.arena + D.st_off + (int64_t)(S.cur ^ 1) * 4 * N));
    const gptr<T> Uo1 = uniform(Uo + N), Uo2 = uniform(Uo + 2 * N), Uo3 = uniform(Uo + 3 * N);
    const auto sw = SW ? uniform(gin<T>(D.sw)) : gptr<const T>(nullptr);
    const auto swb = SW && !ISO ? uniform(sw + D.sw_cstride) : sw;
    const auto zo = uniform(gout<T>(D.z)), qo0 = uniform(gout<T>(D.q)), qo1 = uniform(qo0 + D.q_cstride);
    T beta[TT + 1];
#pragma unroll
    for (int m = 1; m <= TT; ++m) beta[m] = (T)S.beta[m];
#pragma unroll
    for (int m = 0; m <= TT; ++m) {
        if ((gm >> m) & 1u) acc.v[0][m][lane] = 0.0;
        if ((tm >> m) & 1u) acc.v[1][m][lane] = 0.0;
    }

    // Per level m < TT: its latest row's v, u, z(v) and z(u).
    T ov0[TT], ov1[TT], ou0[TT], ou1[TT], oz[TT];
    G ozu[TT];
    T dw[TT + 2];                                          // dw[m] = x_s at row ys - m
    T ra[SW ? TT + 2 : 1], rb[(SW && !ISO) ? TT + 2 : 1];  // radii at row ys - m
#pragma unroll
    for (int m = 0; m < TT; ++m) {
        ov0[m] = ov1[m] = ou0[m] = ou1[m] = oz[m] = (T)0;
        ozu[m] = (G)0;
    }
#pragma unroll
    for (int m = 0; m < TT + 2; ++m) {
        dw[m] = (T)0;
        if constexpr (SW) ra[m] = (T)0;
        if constexpr (SW && !ISO) rb[m] = (T)0;
    }
    T uf0 = (T)0;  // the final level's previous row, u0

    // The gap term of one pixel (tv_iso.cpp:290-306; _tv_torch.py:179-186).
    auto gap_term = [&](G g0, G g1, G u0, G u1, G r0, G r1) -> G {
        if constexpr (ISO) {
            if constexpr (F32) return r0 * sqrt32(g0 * g0 + g1 * g1) - (g0 * u0 + g1 * u1);
            else return r0 * gap_sqrt<T>(g0 * g0 + g1 * g1) - (g0 * u0 + g1 * u1);
        } else {
            return (r0 * fabs(g0) - g0 * u0) + (r1 * fabs(g1) - g1 * u1);
        }
    };
    // z of row r from its v (or u) and the row above's: float32 and float64
    // alike, d + w0 (a0 - a0_above) + w1 (a1 - a1_left).
    auto zrow = [&](G d, G a0, G a0up, G a1) -> G {
        if constexpr (F32) return madd(w1, a1 - from_left(a1), madd(w0, a0 - a0up, d));
        else return (d + w0d * (a0 - a0up)) + w1d * (a1 - from_left(a1));
    };
    // Row-bound tests of the general tiles (constant true for interior ones).
    auto row_g0 = [&](int r) { return INTERIOR || (r >= 0 && r + 1 < H); };
    auto row_g1 = [&](int r) { return INTERIOR || (r >= 0 && r < H); };

    const int ystart = y0 - TT, ylast = yend - 1 + TT;
    // A row's input, loaded ahead of its step; two slots when rows are
    // loaded two steps ahead.
    struct RowIn {
        T d, u0, u1, v0, v1, a, b;
    };
    RowIn q[2] = {};
    auto load_row = [&](int ys, RowIn &Q) {
        const bool in = xin && (INTERIOR || (ys >= 0 && ys < H));
        const uint32_t o = in ? (uint32_t)(ys * W + x) : 0u;
        Q.d = in ? ld32<T>(dp, o) : (T)0;
        Q.u0 = Q.u1 = Q.v0 = Q.v1 = (T)0;
        if (!fresh && in) {
            Q.u0 = ld32<T>(U0, o);
            Q.u1 = ld32<T>(U1, o);
            Q.v0 = ld32<T>(V0, o);
            Q.v1 = ld32<T>(V1, o);
        }
        if constexpr (SW) {
            Q.a = in ? ld32<T>(sw, o) : (T)0;
            if constexpr (!ISO) Q.b = in ? ld32<T>(swb, o) : (T)0;
        }
    };
    // The interior float32 tiles of a full launch (an even number of rows:
    // SEG + 2 TT) load each row two steps ahead, through two fixed slots
    // that alternate; the others one step ahead through slot 0.
    const bool deep = F32 && INTERIOR && cnt == TT && (SEG & 1) == 0;
    load_row(ystart, q[0]);
    if (deep) load_row(ystart + 1, q[1]);
    // One step of the row pipeline: row ys enters at level 0 and each level
    // m produces row ys - m. STEADY: every level is past its skipped rows
    // (ys >= ystart + 2 TT) and the launch runs all TT levels, so no level
    // tests either (uniform tests, but each costs scalar instructions and a
    // branch per level and row). SL: the slot row ys was loaded into, or -1
    // (slot 0, the next row loaded one step ahead).
    auto row_step = [&](const int ys, auto steady, auto slot) {
        constexpr bool STEADY = decltype(steady)::value;
        constexpr int SL = decltype(slot)::value;
        RowIn &Q = q[SL < 0 ? 0 : SL];
        // The level masks, made opaque at each row (an empty asm on their
        // scalar registers): each level's test is then a scalar bit test and
        // branch in the row, where the compiler would otherwise hoist the
        // 2 TT tests out of the row loop as per-lane masks (two scalar
        // registers each, rebuilt through vector instructions at every use).
        uint32_t gm = gm0, tm = tm0;
        asm volatile("" : "+s"(gm), "+s"(tm));
#pragma unroll
        for (int m = TT + 1; m >= 1; --m) {
            dw[m] = dw[m - 1];
            if constexpr (SW) ra[m] = ra[m - 1];
            if constexpr (SW && !ISO) rb[m] = rb[m - 1];
        }
        // Level 0: the new row.
        const T nv0 = Q.v0, nv1 = Q.v1, nu0 = Q.u0, nu1 = Q.u1;
        dw[0] = scale_in(Q.d, e);
        if constexpr (SW) {
            if constexpr (ISO) {
                ra[0] = (T)iso_radius(gamd, (double)Q.a, boundd);
            } else {
                ra[0] = (T)(gamd * (double)Q.a);
                rb[0] = (T)(gamd * (double)Q.b);
            }
        }
        if constexpr (SL < 0) {
            if (ys < ylast) load_row(ys + 1, Q);
        } else {
            if (ys + 2 <= ylast) load_row(ys + 2, Q);
        }
        T zn = madd(w1, nv1 - from_left(nv1), madd(w0, nv0 - ov0[0], dw[0]));
        G zun = (G)0;
        // The gap of level 0 (the loaded state) on row ys - 1.
        if (gm & 1u) {
            zun = zrow((G)dw[0], (G)nu0, (G)ou0[0], (G)nu1);
            const int r = ys - 1;
            const G zc = ozu[0], zr = from_right(zc);
            const G g0 = F32 ? (row_g0(r) ? gg0 * (zun - zc) : (G)0)
                             : ((act0 && xin && row_g0(r)) ? (G)w0d * (zun - zc) : (G)0);
            const G g1 = F32 ? (row_g1(r) ? gg1 * (zr - zc) : (G)0)
                             : ((act1 && xnext && row_g1(r)) ? (G)w1d * (zr - zc) : (G)0);
            const G gp = gap_term(g0, g1, (G)ou0[0], (G)ou1[0], SW ? (G)ra[1] : (G)gamd,
                                  SW && !ISO ? (G)rb[1] : (G)gamd);
            if (lane_in && r >= y0 && r < yend) acc.v[0][0][lane] += (double)gp;
        }
        // The new row's values rise through the levels.
        T pv0 = nv0, pv1 = nv1, pu0 = nu0, pu1 = nu1;  // level m - 1's new row
        T pz = zn;
        G pzu = zun;
#pragma unroll
        for (int m = 1; m <= TT; ++m) {
            const int r = ys - m;  // the row level m produces
            // Level m's rows above ystart + m feed nothing: the final level's
            // rows need level m on rows from y0 - (TT - m) on, reached at
            // ys = ystart + 2m; until then the level is skipped (its values
            // are never read by a level, a partial sum or a store).
            const bool need = STEADY || ys >= ystart + 2 * m;
            const T zc = oz[m - 1];
            T v0n, v1n, u0n, u1n;
            if ((STEADY || m <= cnt) && need) {
                const T zr = from_right(zc);  // inside the branch: the shift folds into the subtraction (DPP)
                // An inactive axis has a 0 weight and v = 0 (cold start, or
                // zeroed by the warm start), so its c is 0 as the reference's
                // select makes it.
                T c0, c1;
                if constexpr (F32) {
                    c0 = madd(row_g0(r) ? tw0 : (T)0, pz - zc, ov0[m - 1]);
                    c1 = madd(row_g1(r) ? tw1 : (T)0, zr - zc, ov1[m - 1]);
                } else {
                    // float64: the lane masks (doubles of folded weights would
                    // cost the registers of a resident wave)
                    const T g0 = (act0 && xin && row_g0(r)) ? w0 * (pz - zc) : (T)0;
                    const T g1 = (act1 && xnext && row_g1(r)) ? w1 * (zr - zc) : (T)0;
                    c0 = madd(tau, g0, ov0[m - 1]);
                    c1 = madd(tau, g1, ov1[m - 1]);
                }
                T n0, n1;
                if constexpr (ISO) {
                    const T sc = ball_scale(madd(c1, c1, c0 * c0), SW ? ra[m] : gam);
                    n0 = c0 * sc;
                    n1 = c1 * sc;
                } else {
                    const T r0 = SW ? ra[m] : gam, r1 = SW ? rb[m] : gam;
                    n0 = c0 > r0 ? r0 : (c0 < -r0 ? -r0 : c0);
                    n1 = c1 > r1 ? r1 : (c1 < -r1 ? -r1 : c1);
                }
                if ((tm >> m) & 1u) {
                    const G dd = ((G)ov0[m - 1] - (G)n0) * ((G)n0 - (G)ou0[m - 1]) +
                                 ((G)ov1[m - 1] - (G)n1) * ((G)n1 - (G)ou1[m - 1]);
                    if (lane_in && r >= y0 && r < yend) acc.v[1][m][lane] += (double)dd;
                }
                v0n = madd(beta[m], n0 - ou0[m - 1], n0);
                v1n = madd(beta[m], n1 - ou1[m - 1], n1);
                u0n = n0;
                u1n = n1;
            } else {
                v0n = ov0[m - 1];
                v1n = ov1[m - 1];
                u0n = ou0[m - 1];
                u1n = ou1[m - 1];
            }
            // Level m - 1's latest row is now the one that just rose from it.
            ov0[m - 1] = pv0;
            ov1[m - 1] = pv1;
            ou0[m - 1] = pu0;
            ou1[m - 1] = pu1;
            oz[m - 1] = pz;
            ozu[m - 1] = pzu;
            if (m < TT) {
                if ((STEADY || m < cnt) && need) pz = madd(w1, v1n - from_left(v1n), madd(w0, v0n - ov0[m], dw[m]));
                if (((gm >> m) & 1u) && need) {
                    // The gap of level m on row r - 1.
                    pzu = zrow((G)dw[m], (G)u0n, (G)ou0[m], (G)u1n);
                    const int rg = r - 1;
                    const G zc2 = ozu[m], zr2 = from_right(zc2);
                    const G g0 = F32 ? (row_g0(rg) ? gg0 * (pzu - zc2) : (G)0)
                                     : ((act0 && xin && row_g0(rg)) ? (G)w0d * (pzu - zc2) : (G)0);
                    const G g1 = F32 ? (row_g1(rg) ? gg1 * (zr2 - zc2) : (G)0)
                                     : ((act1 && xnext && row_g1(rg)) ? (G)w1d * (zr2 - zc2) : (G)0);
                    const G gp = gap_term(g0, g1, (G)ou0[m], (G)ou1[m], SW ? (G)ra[m + 1] : (G)gamd,
                                          SW && !ISO ? (G)rb[m + 1] : (G)gamd);
                    if (lane_in && rg >= y0 && rg < yend) acc.v[0][m][lane] += (double)gp;
                }
                pv0 = v0n;
                pv1 = v1n;
                pu0 = u0n;
                pu1 = u1n;
            } else {
                // The final level: row r leaves the wave.
                const bool st = lane_in && r >= y0 && r < yend;
                const uint32_t so = st ? (uint32_t)(r * W + x) : 0u;
                if (out_mode) {
                    const double ul = from_left((double)u1n);
                    const double zz =
                        ((double)dw[TT] + w0d * ((double)u0n - (double)uf0)) + w1d * ((double)u1n - ul);
                    if (st) {
                        st32<T>(zo, so, (T)ldexp(zz, e));
                        if constexpr (ISO) {
                            st32<T>(qo0, so, (T)ldexp((double)u0n / S.wmax, e));
                            st32<T>(qo1, so, (T)ldexp((double)u1n / S.wmax, e));
                        } else {
                            st32<T>(qo0, so, (T)(w0d * ((double)uf0 - (double)u0n)));
                            st32<T>(qo1, so, (T)(w1d * (ul - (double)u1n)));
                        }
                    }
                }
                if (st && cnt > 0) {  // the iterate (also of an output launch: the kept dual)
                    st32<T>(Uo, so, u0n);
                    st32<T>(Uo1, so, u1n);
                    st32<T>(Uo2, so, v0n);
                    st32<T>(Uo3, so, v1n);
                }
                uf0 = u0n;
            }
        }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    int ys = ystart;
    if (deep) {
        for (; ys < ystart + 2 * TT; ys += 2) {
            row_step(ys, std::false_type{}, S0{});
            row_step(ys + 1, std::false_type{}, S1{});
        }
        for (; ys <= ylast; ys += 2) {
            row_step(ys, std::true_type{}, S0{});
            row_step(ys + 1, std::true_type{}, S1{});
        }
    } else {
        for (int y = ystart; y <= ylast; ++y) row_step(y, std::false_type{}, std::integral_constant<int, -1>{});
    }
    // The tile's partial sums of the flagged levels.
#pragma unroll
    for (int m = 0; m <= TT; ++m) {
        if ((gm >> m) & 1u) {
            const double v = wave_sum(acc.v[0][m][lane]);
            if (lane == 0) part[2 * m] = v;
        }
        if ((tm >> m) & 1u) {
            const double v = wave_sum(acc.v[1][m][lane]);
            if (lane == 0) part[2 * m + 1] = v;
        }
    }
}

// Waves per SIMD the register allocation is held to (float32: four where
// the variant fits in 128 VGPRs without spilling, three otherwise; float64:
// the compiler's choice). More resident waves hide the per-row load latency.
template <class T, bool ISO, bool SW, int TT>
constexpr int stream2d_waves() {
    if (sizeof(T) != 4) return (!SW && TT <= 6) ? 3 : 1;  // float64: three where 168 registers suffice
    if (!SW) return TT <= 6 ? 4 : 3;
    return TT <= (ISO ? 2 : 4) ? 4 : 3;
}

// One launch's tiles (the list A.list_in of *A.count_in tiles): groups of
// four consecutive tiles per workgroup, one per wave, each XCD taking a
// contiguous range of the groups.
template <class T, bool ISO, bool SW, int TT>
__device__ __forceinline__ void stream2d_body(const StepArgs<T> &A, int SEG, StreamAcc<TT> *accs) {
    // the wave's index is the same on all its lanes: in a scalar register, so
    // that the tile loop compiles as uniform control flow
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.count_out = 0;
    const int ntiles = __hip_atomic_load(A.count_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int nchunks = (ntiles + 3) >> 2;
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, j8 = blockIdx.x >> 3;
    const int cpx = (nchunks + 7) >> 3;
    const int cend = min((xcd + 1) * cpx, nchunks);
    for (int c = xcd * cpx + j8; c < cend; c += per) {
        const int ti = c * 4 + wv;
        if (ti >= ntiles) continue;
        const Tile tl = A.list_in[ti];
        // rows y0 - TT - 1 .. yend + TT inside the image: the interior form
        const int H = A.desc[__builtin_amdgcn_readfirstlane(tl.img)].L[1];
        const int y0 = __builtin_amdgcn_readfirstlane(tl.ty) * SEG;
        if (y0 - TT - 1 >= 0 && y0 + SEG + TT < H) stream_tile<T, ISO, SW, TT, true>(A, tl, lane, SEG, accs[wv]);
        else stream_tile<T, ISO, SW, TT, false>(A, tl, lane, SEG, accs[wv]);
    }
}

template <class T, bool ISO, bool SW, int TT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(stream2d_waves<T, ISO, SW, TT>())))
void stream2d_kernel(StepArgs<T> A, int SEG) {
    __shared__ StreamAcc<TT> accs[4];
    stream2d_body<T, ISO, SW, TT>(A, SEG, accs);
}

// ---------------------------------------------------------------------------
// The generic step kernel: one iteration per launch, one thread per pixel,
// neighbours read through the cache; 2D and 3D. Tiles of 64 x 4 pixels of
// one slice (a workgroup each).
// ---------------------------------------------------------------------------

template <class T>
struct Pix {
    int64_t L[3], S[3];
    int k;
    // z = d_s + sum_a w_a (f_a[p] - f_a[p - e_a]) at p (in the image), f = u or v planes.
    __device__ __forceinline__ T zval(const T *dp, int e, const T *F, int64_t N, const T *w, int64_t p,
                                      const int64_t *c) const {
        T z = scale_in(dp[p], e);
        for (int a = 0; a < k; ++a) {
            const int ax = 3 - k + a;
            const T fp = c[ax] > 0 ? F[a * N + p - S[ax]] : (T)0;
            z = z + w[a] * (F[a * N + p] - fp);
        }
        return z;
    }
};

template <class T, bool ISO, bool SW>
__global__ __launch_bounds__(256) void step1_kernel(StepArgs<T> A) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.count_out = 0;
    const int ntiles = *A.count_in;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    __shared__ double red[4][2 * 2];
    for (int ti = blockIdx.x; ti < ntiles; ti += gridDim.x) {
        const Tile tl = A.list_in[ti];
        const int img = tl.img;
        const ImgDesc &D = A.desc[img];
        ImgState &S = A.state[img];
        const int mode = S.mode;
        const int k = A.k;
        Pix<T> G;
        G.k = k;
        for (int a = 0; a < 3; ++a) G.L[a] = D.L[a];
        G.S[2] = 1;
        G.S[1] = D.L[2];
        G.S[0] = (int64_t)D.L[1] * D.L[2];
        const int64_t N = G.S[0] * D.L[0];
        const int64_t c[3] = {tl.tz, (int64_t)tl.ty * 4 + wv, (int64_t)tl.tx * 64 + lane};
        const bool in = c[0] < D.L[0] && c[1] < D.L[1] && c[2] < D.L[2];
        const int64_t p = in ? c[0] * G.S[0] + c[1] * G.S[1] + c[2] : 0;
        double *part = A.part + (int64_t)(D.tile_base + tl.pad_) * kSlots;
        if (mode == kOutputNaN) {
            if (in) {
                gout<T>(D.z)[p] = (T)NAN;
                for (int a = 0; a < k; ++a) gout<T>(D.q)[a * D.q_cstride + p] = (T)NAN;
            }
            continue;
        }
        const int e = S.e, cnt = S.cnt;
        const uint32_t tm = S.test_mask, gm = S.gap_mask;
        T w[3];
        double wd[3];
        bool has[3];
        for (int a = 0; a < 3; ++a) {
            wd[a] = a < k ? S.wa[a] : 0.0;
            w[a] = (T)wd[a];
        }
        for (int a = 0; a < k; ++a) has[a] = in && c[3 - k + a] + 1 < D.L[3 - k + a] && wd[a] != 0.0;
        const T *dp = (const T *)D.d;  // G.zval reads through a generic pointer
        const bool fresh = S.fresh;
        const T *U = A.arena + D.st_off + (int64_t)S.cur * 2 * k * N;
        const T *V = S.vsrc_u ? U : U + k * N;
        const auto sw = SW ? gin<T>(D.sw) : gptr<const T>(nullptr);
        const double gamd = S.gamma;
        double rad[3] = {gamd, gamd, gamd};
        if (SW && in) {
            if (ISO) rad[0] = iso_radius(gamd, (double)sw[p], S.bound);
            else
                for (int a = 0; a < k; ++a) rad[a] = gamd * (double)sw[a * D.sw_cstride + p];
        }
        // The gap of the u planes at p, FP64 (tv_iso.cpp:290-306).
        auto gap_at = [&]() -> double {
            if (!in || fresh) {
                if (!in) return 0.0;
            }
            double zc = 0.0, g[3] = {0.0, 0.0, 0.0};
            auto zu = [&](int64_t q, const int64_t *cq) -> double {
                double z = ldexp((double)dp[q], -e);
                for (int a = 0; a < k; ++a) {
                    const int ax = 3 - k + a;
                    const double fp = (fresh || cq[ax] == 0) ? 0.0 : (double)U[a * N + q - G.S[ax]];
                    const double fc = fresh ? 0.0 : (double)U[a * N + q];
                    z = z + wd[a] * (fc - fp);
                }
                return z;
            };
            zc = zu(p, c);
            for (int a = 0; a < k; ++a) {
                if (!has[a]) continue;
                const int ax = 3 - k + a;
                int64_t cn[3] = {c[0], c[1], c[2]};
                cn[ax] += 1;
                g[a] = wd[a] * (zu(p + G.S[ax], cn) - zc);
            }
            double acc = 0.0;
            if (ISO) {
                double nn = 0.0, dot = 0.0;
                for (int a = 0; a < k; ++a) {
                    nn += g[a] * g[a];
                    dot += g[a] * (fresh ? 0.0 : (double)U[a * N + p]);
                }
                acc = rad[0] * gap_sqrt<T>(nn) - dot;
            } else {
                for (int a = 0; a < k; ++a) acc += rad[a] * fabs(g[a]) - g[a] * (fresh ? 0.0 : (double)U[a * N + p]);
            }
            return acc;
        };
        double gsum = 0.0, dsum = 0.0;
        const bool g0 = gm & 1u;
        if (g0) gsum = gap_at();
        T un[3] = {0, 0, 0}, vn[3] = {0, 0, 0};
        const bool step = cnt >= 1 && mode == kIterate;
        if (step && in) {
            // z(v) at p and its forward neighbours, then the update (tv_iso.cpp:499-524).
            T vv[3], uu[3];
            for (int a = 0; a < k; ++a) {
                vv[a] = fresh ? (T)0 : V[a * N + p];
                uu[a] = fresh ? (T)0 : U[a * N + p];
            }
            auto zv = [&](int64_t q, const int64_t *cq) -> T {
                if (fresh) return scale_in(dp[q], e);
                return G.zval(dp, e, V, N, w, q, cq);
            };
            const T zc = zv(p, c);
            T g[3] = {0, 0, 0};
            for (int a = 0; a < k; ++a) {
                if (!has[a]) continue;
                const int ax = 3 - k + a;
                int64_t cn[3] = {c[0], c[1], c[2]};
                cn[ax] += 1;
                g[a] = w[a] * (zv(p + G.S[ax], cn) - zc);
            }
            const T tau = (T)S.tau, beta = (T)S.beta[1];
            T cv[3];
            T nn = 0;
            for (int a = 0; a < k; ++a) {
                cv[a] = wd[a] != 0.0 ? vv[a] + tau * g[a] : (T)0;
                nn += cv[a] * cv[a];
            }
            if (ISO) {
                const T sc = ball_scale(nn, (T)rad[0]);
                for (int a = 0; a < k; ++a) un[a] = cv[a] * sc;
            } else {
                for (int a = 0; a < k; ++a) {
                    const T r = (T)rad[a];
                    un[a] = cv[a] > r ? r : (cv[a] < -r ? -r : cv[a]);
                }
            }
            for (int a = 0; a < k; ++a) {
                if ((tm >> 1) & 1u) dsum += ((double)vv[a] - (double)un[a]) * ((double)un[a] - (double)uu[a]);
                vn[a] = un[a] + beta * (un[a] - uu[a]);
            }
        }
        // Reductions: a fixed tree over the workgroup.
        const bool dt = step && ((tm >> 1) & 1u);
        if (g0 || dt) {
            const double a0 = wave_sum(gsum), a1 = wave_sum(dsum);
            if (lane == 0) {
                red[wv][0] = a0;
                red[wv][1] = a1;
            }
            __syncthreads();
            if (tid == 0) {
                if (g0) part[0] = (red[0][0] + red[1][0]) + (red[2][0] + red[3][0]);
                if (dt) part[3] = (red[0][1] + red[1][1]) + (red[2][1] + red[3][1]);
            }
            __syncthreads();
        }
        if (!in) continue;
        if (mode == kOutput) {
            // cnt is 0 here: z(u) and the dual at p.
            double zz = ldexp((double)dp[p], -e);
            for (int a = 0; a < k; ++a) {
                const int ax = 3 - k + a;
                const double fp = (fresh || c[ax] == 0) ? 0.0 : (double)U[a * N + p - G.S[ax]];
                const double fc = fresh ? 0.0 : (double)U[a * N + p];
                zz = zz + wd[a] * (fc - fp);
                double qv;
                if (ISO) qv = ldexp(fc / S.wmax, e);
                else qv = wd[a] * (fp - fc);
                gout<T>(D.q)[a * D.q_cstride + p] = (T)qv;
            }
            gout<T>(D.z)[p] = (T)ldexp(zz, e);
            continue;
        }
        if (!step) continue;
        T *Uo = A.arena + D.st_off + (int64_t)(S.cur ^ 1) * 2 * k * N;
        for (int a = 0; a < k; ++a) {
            Uo[a * N + p] = un[a];
            Uo[(k + a) * N + p] = vn[a];
        }
    }
}

// ---------------------------------------------------------------------------
// The 3D streaming step kernel: one iteration per launch, one HBM pass.
// A workgroup of NW waves owns a tile of 62 x (NW - 1) pixels of every slice
// of a slab of SEGZ slices: wave w is row y0 + w (the last wave a halo row
// that only supplies z at y + 1), lane l column x0 - 1 + l (lanes 0 and 63
// halo columns). The slab is walked slice by slice with a lag of one: at
// slice s every lane forms z(v) at its voxel (v_2 at x - 1 by a DPP shift,
// v_1 at y - 1 from the row above, v_0 at s - 1 from the previous step),
// publishes it in a triple-buffered LDS row, and updates slice s - 1, whose
// forward neighbours are then all known: s from the registers, y + 1 from
// the LDS row of the wave below, x + 1 by a DPP shift. The per-voxel
// arithmetic is step1_kernel's, expression for expression.
// ---------------------------------------------------------------------------

template <class T, bool ISO, bool SW, int NW>
__global__ __launch_bounds__(64 * NW) void stream3d_kernel(StepArgs<T> A, int SEGZ) {
    constexpr int RYU = NW - 1;
    __shared__ T zrow[3][NW][64];
    __shared__ double zurow[3][NW][64];
    __shared__ double red[NW][2];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (blockIdx.x == 0 && tid == 0) *A.count_out = 0;
    const int ntiles = *A.count_in;
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, j8 = blockIdx.x >> 3;
    const int cpx = (ntiles + 7) >> 3;
    const int cend = min((xcd + 1) * cpx, ntiles);
    for (int ti = xcd * cpx + j8; ti < cend; ti += per) {
        const Tile tl = A.list_in[ti];
        const int img = tl.img;
        const ImgDesc &D = A.desc[img];
        ImgState &S = A.state[img];
        const int mode = S.mode;
        const int L0 = D.L[0], H = D.L[1], W = D.L[2];
        const int64_t SY = W, SZ = (int64_t)H * W, N = SZ * L0;
        const int x = tl.tx * 62 - 1 + lane;
        const int y = tl.ty * RYU + wv;
        const int z0 = tl.tz * SEGZ, z1 = min(z0 + SEGZ, L0);
        const bool xin = x >= 0 && x < W, yin = y < H, in = xin && yin;
        const bool upd = in && lane >= 1 && lane <= 62 && wv < RYU;
        double *part = A.part + (int64_t)(D.tile_base + tl.pad_) * kSlots;
        if (mode == kOutputNaN) {
            if (upd)
                for (int t = z0; t < z1; ++t) {
                    const int64_t o = t * SZ + (int64_t)y * SY + x;
                    gout<T>(D.z)[o] = (T)NAN;
                    for (int a = 0; a < 3; ++a) gout<T>(D.q)[a * D.q_cstride + o] = (T)NAN;
                }
            continue;
        }
        const int e = S.e, cnt = S.cnt;
        const uint32_t tm = S.test_mask, gm = S.gap_mask;
        const bool g0on = gm & 1u;
        const bool step = cnt >= 1 && mode == kIterate;
        const bool dton = step && ((tm >> 1) & 1u);
        const bool out_mode = mode == kOutput;
        const bool fresh = S.fresh;
        T w[3];
        double wd[3];
        for (int a = 0; a < 3; ++a) {
            wd[a] = S.wa[a];
            w[a] = (T)wd[a];
        }
        const T tau = (T)S.tau, beta = (T)S.beta[1];
        const double gamd = S.gamma;
        const auto dp = gin<T>(D.d);
        const T *U = A.arena + D.st_off + (int64_t)S.cur * 6 * N;
        const T *V = S.vsrc_u ? U : U + 3 * N;
        T *Uo = A.arena + D.st_off + (int64_t)(S.cur ^ 1) * 6 * N;
        const auto sw = SW ? gin<T>(D.sw) : gptr<const T>(nullptr);
        const int64_t oxy = in ? (int64_t)y * SY + x : 0;
        const bool yup = in && y > 0;
        // The previous slice's state at this voxel, and z at it.
        T pv[3] = {0, 0, 0}, pu[3] = {0, 0, 0}, zvp = 0;
        double zup = 0.0;
        double gacc = 0.0, dacc = 0.0;
        auto load = [&](int s, T *v, T *u, T &v1m, T &u1m, T &dd) {
            const bool ok = in && s >= 0 && s < L0;
            const int64_t o = ok ? s * SZ + oxy : 0;
            dd = ok ? dp[o] : (T)0;
            for (int a = 0; a < 3; ++a) {
                v[a] = (ok && !fresh) ? V[a * N + o] : (T)0;
                u[a] = (ok && !fresh) ? U[a * N + o] : (T)0;
            }
            v1m = (ok && !fresh && yup) ? V[N + o - SY] : (T)0;
            u1m = (ok && !fresh && yup) ? U[N + o - SY] : (T)0;
        };
        {  // the slice before the slab: only its axis-0 planes are needed
            T v[3], u[3], v1m, u1m, dd;
            load(z0 - 1, v, u, v1m, u1m, dd);
            pv[0] = v[0];
            pu[0] = u[0];
        }
        // Slice s + 1's input is loaded one step ahead, so that its latency
        // overlaps slice s's work (the barriers wait on LDS only).
        T nv[3], nu[3], nv1m, nu1m, nd;
        load(z0, nv, nu, nv1m, nu1m, nd);
        for (int s = z0; s <= z1; ++s) {
            const bool sin = s < L0;
            T cv[3], cu[3];
            for (int a = 0; a < 3; ++a) {
                cv[a] = nv[a];
                cu[a] = nu[a];
            }
            const T v1m = nv1m, u1m = nu1m, dd = nd;
            if (s < z1) load(s + 1, nv, nu, nv1m, nu1m, nd);
            // z(v) and, on gap launches, z(u) at slice s (step1_kernel's zval).
            T zv = 0;
            double zu = 0.0;
            const T v2l = from_left(cv[2]);
            const double u2l = (g0on || out_mode) ? from_left((double)cu[2]) : 0.0;
            if (in && sin) {
                zv = scale_in(dd, e);
                zv = zv + w[0] * (cv[0] - (s > 0 ? pv[0] : (T)0));
                zv = zv + w[1] * (cv[1] - v1m);
                zv = zv + w[2] * (cv[2] - (x > 0 ? v2l : (T)0));
                if (g0on || out_mode) {
                    zu = ldexp((double)dd, -e);
                    zu = zu + wd[0] * ((double)cu[0] - (s > 0 ? (double)pu[0] : 0.0));
                    zu = zu + wd[1] * ((double)cu[1] - (double)u1m);
                    zu = zu + wd[2] * ((double)cu[2] - (x > 0 ? u2l : 0.0));
                }
            }
            const int b = s % 3, bp = (s + 2) % 3;
            zrow[b][wv][lane] = zv;
            if (g0on) zurow[b][wv][lane] = zu;
            __syncthreads();
            // Slice t = s - 1 of the slab: its update, gap and output.
            const int t = s - 1;
            const T zr = from_right(zvp);
            const double zur = (g0on) ? from_right(zup) : 0.0;
            if (t >= z0 && upd) {
                const int64_t o = t * SZ + oxy;
                const bool has0 = wd[0] != 0.0 && t + 1 < L0;
                const bool has1 = wd[1] != 0.0 && y + 1 < H;
                const bool has2 = wd[2] != 0.0 && x + 1 < W;
                double rad[3] = {gamd, gamd, gamd};
                if (SW) {
                    if (ISO) rad[0] = iso_radius(gamd, (double)sw[o], S.bound);
                    else
                        for (int a = 0; a < 3; ++a) rad[a] = gamd * (double)sw[a * D.sw_cstride + o];
                }
                if (g0on) {
                    double g[3] = {0.0, 0.0, 0.0};
                    if (has0) g[0] = wd[0] * (zu - zup);
                    if (has1) g[1] = wd[1] * (zurow[bp][wv + 1][lane] - zup);
                    if (has2) g[2] = wd[2] * (zur - zup);
                    double acc = 0.0;
                    if (ISO) {
                        double nn = 0.0, dot = 0.0;
                        for (int a = 0; a < 3; ++a) {
                            nn += g[a] * g[a];
                            dot += g[a] * (fresh ? 0.0 : (double)pu[a]);
                        }
                        acc = rad[0] * gap_sqrt<T>(nn) - dot;
                    } else {
                        for (int a = 0; a < 3; ++a)
                            acc += rad[a] * fabs(g[a]) - g[a] * (fresh ? 0.0 : (double)pu[a]);
                    }
                    gacc += acc;
                }
                if (out_mode) {
                    // cnt is 0 here: z(u) and the dual at the voxel.
                    const double ul1 = y > 0 ? (double)U[N + o - SY] : 0.0;
                    const double ul2 = x > 0 ? (double)U[2 * N + o - 1] : 0.0;
                    const double ul0 = t > 0 ? (double)U[o - SZ] : 0.0;
                    const double fpv[3] = {fresh ? 0.0 : ul0, fresh ? 0.0 : ul1, fresh ? 0.0 : ul2};
                    for (int a = 0; a < 3; ++a) {
                        const double fc = fresh ? 0.0 : (double)pu[a];
                        double qv;
                        if (ISO) qv = ldexp(fc / S.wmax, e);
                        else qv = wd[a] * (fpv[a] - fc);
                        gout<T>(D.q)[a * D.q_cstride + o] = (T)qv;
                    }
                    gout<T>(D.z)[o] = (T)ldexp(zup, e);
                } else if (step) {
                    T g[3] = {0, 0, 0};
                    if (has0) g[0] = w[0] * (zv - zvp);
                    if (has1) g[1] = w[1] * (zrow[bp][wv + 1][lane] - zvp);
                    if (has2) g[2] = w[2] * (zr - zvp);
                    T c[3];
                    T nn = 0;
                    for (int a = 0; a < 3; ++a) {
                        c[a] = wd[a] != 0.0 ? pv[a] + tau * g[a] : (T)0;
                        nn += c[a] * c[a];
                    }
                    T un[3], vn[3];
                    if (ISO) {
                        const T sc = ball_scale(nn, (T)rad[0]);
                        for (int a = 0; a < 3; ++a) un[a] = c[a] * sc;
                    } else {
                        for (int a = 0; a < 3; ++a) {
                            const T r = (T)rad[a];
                            un[a] = c[a] > r ? r : (c[a] < -r ? -r : c[a]);
                        }
                    }
                    for (int a = 0; a < 3; ++a) {
                        if (dton) dacc += ((double)pv[a] - (double)un[a]) * ((double)un[a] - (double)pu[a]);
                        vn[a] = un[a] + beta * (un[a] - pu[a]);
                    }
                    for (int a = 0; a < 3; ++a) {
                        Uo[a * N + o] = un[a];
                        Uo[(3 + a) * N + o] = vn[a];
                    }
                }
            }
            for (int a = 0; a < 3; ++a) {
                pv[a] = cv[a];
                pu[a] = cu[a];
            }
            zvp = zv;
            zup = zu;
        }
        // The tile's partial sums: a fixed tree over the workgroup.
        if (g0on || dton) {
            const double a0 = wave_sum(gacc), a1 = wave_sum(dacc);
            if (lane == 0) {
                red[wv][0] = a0;
                red[wv][1] = a1;
            }
            __syncthreads();
            if (tid == 0) {
                double s0 = 0.0, s1 = 0.0;
                for (int i = 0; i < NW; ++i) {
                    s0 += red[i][0];
                    s1 += red[i][1];
                }
                if (g0on) part[0] = s0;
                if (dton) part[3] = s1;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The 3D temporally tiled step kernel (float32, no pixel weights): TT
// iterations per launch, one HBM pass. A workgroup of kS3TWaves waves owns
// a tile of (64 - 2 TT) x (kS3TWaves - 2 TT) voxels of every slice of a slab
// of SEGZ slices, with TT halo columns and rows on each side: wave w is row
// y0 - TT + w, lane l column x0 - TT + l. The slab is walked slice by slice;
// the TT levels are pipelined two slices apart, so that within one step no
// level waits for another wave: at step s level m updates slice s - 2m (its
// gradient needs z of level m - 1 at that slice and the next, both already
// computed, and at the next row, which the wave below published at step
// s - 1) and forms z of its previous slice s - 2m - 1 (whose v_1 at the row
// above the wave above published at step s - 1). Each step publishes, per
// level, v_1 of the slice it updated and z of the slice it formed (and, for
// a gap level, u_1 and z(u)) in a double-buffered LDS exchange, and ends
// with the one barrier of the step. The halo shrinks by one column, row and
// slice per level, so the tile's own voxels are exact after TT levels.
// Per voxel the arithmetic is step1_kernel's (float32, fused where the
// expression allows).
// ---------------------------------------------------------------------------

template <int TT>
struct S3TLds {
    float v1[2][kS3TWaves][TT][64];
    float z[2][kS3TWaves][TT][64];
    float u1[2][kS3TWaves][TT + 1][64];
    float zu[2][kS3TWaves][TT][64];
    double red[kS3TWaves][2 * (TT + 1)];
};

template <bool ISO, int TT>
__global__ __launch_bounds__(64 * kS3TWaves) void stream3t_kernel(StepArgs<float> A, int SEGZ) {
    using T = float;
    constexpr int NW = kS3TWaves;
    constexpr int TXI = 64 - 2 * TT, TYI = NW - 2 * TT;
    constexpr int ND = 2 * TT + 2;  // the d delay line: d at slices s .. s - 2 TT - 1
    __shared__ S3TLds<TT> X;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (blockIdx.x == 0 && tid == 0) *A.count_out = 0;
    const int ntiles = __hip_atomic_load(A.count_in, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const int xcd = blockIdx.x & 7, per = gridDim.x >> 3, j8 = blockIdx.x >> 3;
    const int cpx = (ntiles + 7) >> 3;
    const int cend = min((xcd + 1) * cpx, ntiles);
    for (int ti = xcd * cpx + j8; ti < cend; ti += per) {
        const Tile tl = A.list_in[ti];
        const int img = __builtin_amdgcn_readfirstlane(tl.img);
        const ImgDesc &D = A.desc[img];
        ImgState &S = A.state[img];
        const int mode = S.mode;
        const int L0 = D.L[0], H = D.L[1], W = D.L[2];
        const int64_t SZ = (int64_t)H * W, N = SZ * L0;
        const int x = __builtin_amdgcn_readfirstlane(tl.tx) * TXI - TT + lane;
        const int y = __builtin_amdgcn_readfirstlane(tl.ty) * TYI - TT + wv;
        const int z0 = __builtin_amdgcn_readfirstlane(tl.tz) * SEGZ, z1 = min(z0 + SEGZ, L0);
        const bool xin = x >= 0 && x < W, yin = y >= 0 && y < H, inxy = xin && yin;
        const bool own = inxy && lane >= TT && lane < 64 - TT && wv >= TT && wv < NW - TT;
        const uint32_t oxy = inxy ? (uint32_t)(y * W + x) : 0u;
        double *part = A.part + (int64_t)(D.tile_base + __builtin_amdgcn_readfirstlane(tl.pad_)) * kSlots;
        // Plane bases (uniform, in scalar registers); every access is a base
        // plus a 32-bit offset (N sizeof(T) < 2^32: kernel_available_s3).
        const auto zo = uniform(gout<T>(D.z));
        const gptr<T> qo[3] = {uniform(gout<T>(D.q)), uniform(gout<T>(D.q) + D.q_cstride),
                               uniform(gout<T>(D.q) + 2 * D.q_cstride)};
        if (mode == kOutputNaN) {
            if (own)
                for (int t = z0; t < z1; ++t) {
                    const uint32_t o = (uint32_t)(t * SZ) + oxy;
                    st32<T>(zo, o, (T)NAN);
                    for (int a = 0; a < 3; ++a) st32<T>(qo[a], o, (T)NAN);
                }
            continue;
        }
        const int e = S.e, cnt = S.cnt;
        const uint32_t tm0 = __builtin_amdgcn_readfirstlane(S.test_mask),
                       gm0 = __builtin_amdgcn_readfirstlane(S.gap_mask);
        const uint32_t tm = tm0, gm = gm0;
        const bool out_mode = mode == kOutput;
        const bool fresh = S.fresh;
        const double wd0 = S.wa[0], wd1 = S.wa[1], wd2 = S.wa[2];
        const T w0 = (T)wd0, w1 = (T)wd1, w2 = (T)wd2;
        const T tau = (T)S.tau;
        const double gamd = S.gamma;
        const T gam = (T)gamd;
        T beta[TT + 1];
#pragma unroll
        for (int m = 1; m <= TT; ++m) beta[m] = (T)S.beta[m];
        // Per-lane weights of the in-plane forward differences (0 where the
        // neighbour leaves the image or the axis is inactive); of tau g for
        // the update (t*), of g for the gap (g*).
        const T gw1 = (wd1 != 0.0 && inxy && y + 1 < H) ? w1 : (T)0;
        const T gw2 = (wd2 != 0.0 && inxy && x + 1 < W) ? w2 : (T)0;
        const T tw1 = tau * gw1, tw2 = tau * gw2;
        const bool act0 = wd0 != 0.0;
        const auto dp = uniform(gin<T>(D.d));
        const gptr<const T> U = uniform((gptr<const T>)(A.arena + D.st_off + (int64_t)S.cur * 6 * N));
        const gptr<const T> Ub[3] = {U, uniform(U + N), uniform(U + 2 * N)};
        const gptr<const T> V = S.vsrc_u ? U : uniform(U + 3 * N);
        const gptr<const T> Vb[3] = {V, uniform(V + N), uniform(V + 2 * N)};
        const gptr<T> Uo = uniform((gptr<T>)(A.arena + D.st_off + (int64_t)(S.cur ^ 1) * 6 * N));
        const gptr<T> Ob[6] = {Uo, uniform(Uo + N), uniform(Uo + 2 * N), uniform(Uo + 3 * N), uniform(Uo + 4 * N),
                               uniform(Uo + 5 * N)};
        // the axis-0 weight of a slice: 0 outside the image or at its last slice
        auto gw0_at = [&](int t) -> T { return (act0 && inxy && t >= 0 && t + 1 < L0) ? w0 : (T)0; };
        auto in_slice = [&](int t) { return inxy && t >= 0 && t < L0; };

        // Level m's ring: A = slice t_m - 2, B = slice t_m - 1 (t_m = s - 2 m).
        T av[TT + 1][3], au[TT + 1][3], bv[TT + 1][3], bu[TT + 1][3];
        T za[TT], zb[TT];    // z of level m at slices t_m - 2 and t_m - 1
        T zua[TT], zub[TT];  // z(u) of a gap level, the same slices
        T dq[ND];            // scaled d at slices s - k
        double gacc[TT], dacc[TT + 1];
#pragma unroll
        for (int m = 0; m <= TT; ++m) {
#pragma unroll
            for (int a = 0; a < 3; ++a) av[m][a] = au[m][a] = bv[m][a] = bu[m][a] = (T)0;
            dacc[m] = 0.0;
            if (m < TT) {
                za[m] = zb[m] = zua[m] = zub[m] = (T)0;
                gacc[m] = 0.0;
            }
        }
#pragma unroll
        for (int k = 0; k < ND; ++k) dq[k] = (T)0;

        // Level 0's input, loaded one step ahead.
        T nv[3], nu[3], nd;
        const int s_last_load = z1 - 1 + TT;
        auto load_slice = [&](int s) {
            const bool ok = in_slice(s) && s <= s_last_load;
            const uint32_t o = ok ? (uint32_t)(s * SZ) + oxy : 0u;
            nd = ok ? ld32<T>(dp, o) : (T)0;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                nv[a] = (ok && !fresh) ? ld32<T>(Vb[a], o) : (T)0;
                nu[a] = (ok && !fresh) ? ld32<T>(Ub[a], o) : (T)0;
            }
        };
        const int sbeg = z0 - TT;
        const int send = z1 - 1 + 2 * TT + (out_mode ? 1 : 0);
        load_slice(sbeg);
        // The slice steps, two per trip at TT 2: the rings' and the delay
        // line's values are renamed across the pair instead of copied at every
        // slice (at TT 3 the pair would spill).
        auto slice_step = [&](const int s) {
                const int p = s & 1, pp = p ^ 1;
                // the level masks, opaque at each slice: scalar bit tests here
                // rather than per-lane masks hoisted out of the loop
                uint32_t gm = gm0, tm = tm0;
                asm volatile("" : "+s"(gm), "+s"(tm));
                // Level 0: slice s.
                T cv[TT + 1][3], cu[TT + 1][3];
    #pragma unroll
                for (int a = 0; a < 3; ++a) {
                    cv[0][a] = nv[a];
                    cu[0][a] = nu[a];
                }
    #pragma unroll
                for (int k = ND - 1; k >= 1; --k) dq[k] = dq[k - 1];
                dq[0] = scale_in(nd, e);
                if (s < send) load_slice(s + 1);
    #pragma unroll
                for (int m = 0; m <= TT; ++m) {
                    const int t = s - 2 * m;  // the slice level m updates
                    if (m >= 1) {
                        // UPDATE(m) at slice t from level m - 1 (its ring's A is slice t).
                        const bool upd = t >= z0 - (TT - m) && t <= z1 - 1 + (TT - m);
                        if (upd && m <= cnt) {
                            const T zc = za[m - 1];
                            const T zyn = wv + 1 < NW ? X.z[pp][wv + 1][m - 1][lane] : (T)0;
                            const T zxn = from_right(zc);
                            // outside the image every c is 0, so v stays 0 there
                            const bool tin = in_slice(t);
                            const T g0 = gw0_at(t) * (zb[m - 1] - zc);
                            const T c0 = act0 && tin ? madd(tau, g0, av[m - 1][0]) : (T)0;
                            const T c1 = tin ? madd(tw1, zyn - zc, av[m - 1][1]) : (T)0;
                            const T c2 = tin ? madd(tw2, zxn - zc, av[m - 1][2]) : (T)0;
                            T n0, n1, n2;
                            if constexpr (ISO) {
                                const T sc = ball_scale(madd(c2, c2, madd(c1, c1, c0 * c0)), gam);
                                n0 = c0 * sc;
                                n1 = c1 * sc;
                                n2 = c2 * sc;
                            } else {
                                n0 = c0 > gam ? gam : (c0 < -gam ? -gam : c0);
                                n1 = c1 > gam ? gam : (c1 < -gam ? -gam : c1);
                                n2 = c2 > gam ? gam : (c2 < -gam ? -gam : c2);
                            }
                            if ((tm >> m) & 1u) {
                                const T dd = (av[m - 1][0] - n0) * (n0 - au[m - 1][0]) +
                                             (av[m - 1][1] - n1) * (n1 - au[m - 1][1]) +
                                             (av[m - 1][2] - n2) * (n2 - au[m - 1][2]);
                                if (own && t >= z0 && t < z1) dacc[m] += (double)dd;
                            }
                            cv[m][0] = madd(beta[m], n0 - au[m - 1][0], n0);
                            cv[m][1] = madd(beta[m], n1 - au[m - 1][1], n1);
                            cv[m][2] = madd(beta[m], n2 - au[m - 1][2], n2);
                            cu[m][0] = n0;
                            cu[m][1] = n1;
                            cu[m][2] = n2;
                        } else {
    #pragma unroll
                            for (int a = 0; a < 3; ++a) {
                                cv[m][a] = av[m - 1][a];
                                cu[m][a] = au[m - 1][a];
                            }
                        }
                    }
                    if (m < TT) {
                        // ZCOMP(m): z (and on a gap level z(u)) of slice t - 1 from
                        // the ring's B (t - 1) and A (t - 2) and the row above.
                        const int tz = t - 1;
                        // needed by UPDATE(m + 1) on its slices and the next, and by the gap
                        const bool zon = tz >= z0 - TT + m + 1 && tz <= z1 - 1 + (TT - m);
                        if (zon) {
                            const T v1u = wv > 0 ? X.v1[pp][wv - 1][m][lane] : (T)0;
                            zb[m] = madd(w2, bv[m][2] - from_left(bv[m][2]),
                                         madd(w1, bv[m][1] - v1u, madd(w0, bv[m][0] - av[m][0], dq[2 * m + 1])));
                            if ((gm >> m) & 1u) {
                                const T u1u = wv > 0 ? X.u1[pp][wv - 1][m][lane] : (T)0;
                                zub[m] = madd(w2, bu[m][2] - from_left(bu[m][2]),
                                              madd(w1, bu[m][1] - u1u, madd(w0, bu[m][0] - au[m][0], dq[2 * m + 1])));
                            }
                            // The gap of level m at slice t - 2 (its ring's A), a slab slice.
                            const int tg = t - 2;
                            if (((gm >> m) & 1u) && tg >= z0 && tg < z1) {
                                const T zc = zua[m];
                                const T zyn = wv + 1 < NW ? X.zu[pp][wv + 1][m][lane] : (T)0;
                                const T zxn = from_right(zc);  // every lane, before the per-lane selects
                                const bool tin = in_slice(tg);
                                const T g0 = tin ? gw0_at(tg) * (zub[m] - zc) : (T)0;
                                const T g1 = tin ? gw1 * (zyn - zc) : (T)0;
                                const T g2 = tin ? gw2 * (zxn - zc) : (T)0;
                                T gp;
                                if constexpr (ISO) {
                                    gp = gam * sqrt32(g0 * g0 + g1 * g1 + g2 * g2) -
                                         (g0 * au[m][0] + g1 * au[m][1] + g2 * au[m][2]);
                                } else {
                                    gp = (gam * fabsf(g0) - g0 * au[m][0]) + (gam * fabsf(g1) - g1 * au[m][1]) +
                                         (gam * fabsf(g2) - g2 * au[m][2]);
                                }
                                if (own) gacc[m] += (double)gp;
                            }
                        }
                    }
                }
                // The final level: the iterate of slice s - 2 TT; on an output
                // launch, z and the dual of slice s - 2 TT - 1 (from the ring).
                {
                    const int t = s - 2 * TT;
                    if (cnt > 0 && own && t >= z0 && t < z1) {
                        const uint32_t o = (uint32_t)(t * SZ) + oxy;
    #pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            st32<T>(Ob[a], o, cu[TT][a]);
                            st32<T>(Ob[3 + a], o, cv[TT][a]);
                        }
                    }
                    if (out_mode) {
                        const int to = t - 1;
                        const T u1u = wv > 0 ? X.u1[pp][wv - 1][TT][lane] : (T)0;
                        const double ul2 = from_left((double)bu[TT][2]);
                        if (own && to >= z0 && to < z1) {
                            const uint32_t o = (uint32_t)(to * SZ) + oxy;
                            const double zz = (((double)dq[2 * TT + 1] + wd0 * ((double)bu[TT][0] - (double)au[TT][0])) +
                                               wd1 * ((double)bu[TT][1] - (double)u1u)) +
                                              wd2 * ((double)bu[TT][2] - ul2);
                            st32<T>(zo, o, (T)ldexp(zz, e));
                            if constexpr (ISO) {
    #pragma unroll
                                for (int a = 0; a < 3; ++a) st32<T>(qo[a], o, (T)ldexp((double)bu[TT][a] / S.wmax, e));
                            } else {
                                st32<T>(qo[0], o, (T)(wd0 * ((double)au[TT][0] - (double)bu[TT][0])));
                                st32<T>(qo[1], o, (T)(wd1 * ((double)u1u - (double)bu[TT][1])));
                                st32<T>(qo[2], o, (T)(wd2 * (ul2 - (double)bu[TT][2])));
                            }
                        }
                    }
                }
                // Publish this step's values, rotate the rings, one barrier.
    #pragma unroll
                for (int m = 0; m < TT; ++m) {
                    X.v1[p][wv][m][lane] = cv[m][1];
                    X.z[p][wv][m][lane] = zb[m];
                    if ((gm >> m) & 1u) {
                        X.u1[p][wv][m][lane] = cu[m][1];
                        X.zu[p][wv][m][lane] = zub[m];
                    }
                }
                if (out_mode) X.u1[p][wv][TT][lane] = cu[TT][1];
    #pragma unroll
                for (int m = 0; m <= TT; ++m) {
    #pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        av[m][a] = bv[m][a];
                        au[m][a] = bu[m][a];
                        bv[m][a] = cv[m][a];
                        bu[m][a] = cu[m][a];
                    }
                    if (m < TT) {
                        za[m] = zb[m];
                        zua[m] = zub[m];
                    }
                }
                __syncthreads();
        };
        int s = sbeg;
        if constexpr (TT == 2)
            for (; s + 1 <= send; s += 2) {
                slice_step(s);
                slice_step(s + 1);
            }
        for (; s <= send; ++s) slice_step(s);
        // The tile's partial sums: per wave, then over the waves in order.
#pragma unroll
        for (int m = 0; m <= TT; ++m) {
            if (m < TT && ((gm >> m) & 1u)) {
                const double v = wave_sum(gacc[m]);
                if (lane == 0) X.red[wv][2 * m] = v;
            }
            if (m >= 1 && ((tm >> m) & 1u)) {
                const double v = wave_sum(dacc[m]);
                if (lane == 0) X.red[wv][2 * m + 1] = v;
            }
        }
        __syncthreads();
        if (tid < 2 * (TT + 1)) {
            const int m = tid >> 1;
            const bool on = (tid & 1) ? (m >= 1 && ((tm >> m) & 1u)) : (m < TT && ((gm >> m) & 1u));
            if (on) {
                double sum = 0.0;
                for (int i = 0; i < NW; ++i) sum += X.red[i][tid];
                part[tid] = sum;
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Bookkeeping kernels
// ---------------------------------------------------------------------------

__global__ void reset_counts_kernel(int *counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0;
}

__global__ __launch_bounds__(64) void initial_list_kernel(const ImgDesc *desc, const ImgState *state, Tile *list,
                                                          int *count, int nimg, int img0) {
    const int img = blockIdx.x;
    if (img >= nimg) return;
    const ImgState &S = state[img];
    const int m = S.mode;
    if (!(m == kIterate || m == kOutput || m == kOutputNaN)) return;
    const ImgDesc &D = desc[img];
    __shared__ int base;
    if (threadIdx.x == 0) base = atomicAdd(count, D.ntiles);
    __syncthreads();
    const int nxy = D.ntx * D.nty;
    for (int i = threadIdx.x; i < D.ntiles; i += 64) {
        Tile t;
        t.img = img0 + img;
        t.tz = i / nxy;
        const int r = i - t.tz * nxy;
        t.ty = (int16_t)(r / D.ntx);
        t.tx = (int16_t)(r - t.ty * D.ntx);
        t.pad_ = i;
        list[base + i] = t;
    }
}

template <class T>
__global__ void strided_copy_kernel(const T *src, T *dst, int64_t n, int nd, StridedShape sh, int src_strided) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i, off = 0;
        for (int d = nd - 1; d >= 0; --d) {
            const int64_t c = r % sh.shape[d];
            r /= sh.shape[d];
            off += c * sh.stride[d];
        }
        if (src_strided) dst[i] = src[off];
        else dst[off] = src[i];
    }
}

// The inputs of a group of admitted requests, one image per blockIdx.y, into
// their pool slots (contiguous), each from its own strides.
template <class T>
__global__ void gather_group_kernel(const GatherEntry *tab, int G) {
    const int j = blockIdx.y;
    if (j >= G) return;
    const GatherEntry &g = tab[j];
    const auto src = gin<T>(g.src);
    const auto dst = gout<T>(g.dst);
    const int64_t n = g.n;
    const int64_t step = (int64_t)gridDim.x * blockDim.x;
    if (g.contiguous) {
        for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += step) dst[i] = src[i];
        return;
    }
    const int64_t L1 = g.shape[1], L2 = g.shape[2];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += step) {
        const int64_t c2 = i % L2, r = i / L2;
        const int64_t c1 = r % L1, c0 = r / L1;
        dst[i] = src[c0 * g.stride[0] + c1 * g.stride[1] + c2 * g.stride[2]];
    }
}

// A group's set-up, staged in sdesc/sstate/sgam, moves to its slots map[j]
// (one block per image), and the tiles of every image the prologue left to
// iterate or output are appended to the list, as initial_list_kernel does.
__global__ __launch_bounds__(64) void admit_commit_kernel(const ImgDesc *sdesc, const ImgState *sstate,
                                                          const double *sgam, const int *map, ImgDesc *desc,
                                                          ImgState *state, double *gam, Tile *list, int *count,
                                                          int G) {
    const int j = blockIdx.x;
    if (j >= G) return;
    const int slot = map[j];
    static_assert(sizeof(ImgDesc) % 4 == 0 && sizeof(ImgState) % 4 == 0, "word copies");
    {
        const uint32_t *a = (const uint32_t *)(sdesc + j);
        uint32_t *b = (uint32_t *)(desc + slot);
        for (int w = threadIdx.x; w < (int)(sizeof(ImgDesc) / 4); w += 64) b[w] = a[w];
        const uint32_t *c = (const uint32_t *)(sstate + j);
        uint32_t *d = (uint32_t *)(state + slot);
        for (int w = threadIdx.x; w < (int)(sizeof(ImgState) / 4); w += 64) d[w] = c[w];
        if (threadIdx.x == 0) gam[slot] = sgam[j];
    }
    const ImgState &S = sstate[j];
    const int m = S.mode;
    if (!(m == kIterate || m == kOutput || m == kOutputNaN)) return;
    const ImgDesc &D = sdesc[j];
    __shared__ int base;
    if (threadIdx.x == 0) base = atomicAdd(count, D.ntiles);
    __syncthreads();
    const int nxy = D.ntx * D.nty;
    for (int i = threadIdx.x; i < D.ntiles; i += 64) {
        Tile t;
        t.img = slot;
        t.tz = i / nxy;
        const int r = i - t.tz * nxy;
        t.ty = (int16_t)(r / D.ntx);
        t.tx = (int16_t)(r - t.ty * D.ntx);
        t.pad_ = i;
        list[base + i] = t;
    }
}

// src (contiguous, C order) converted into dst with its strides.
template <class S, class D>
__global__ void convert_scatter_kernel(const S *src, D *dst, int64_t n, int nd, StridedShape sh) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t r = i, off = 0;
        for (int d = nd - 1; d >= 0; --d) {
            const int64_t c = r % sh.shape[d];
            r /= sh.shape[d];
            off += c * sh.stride[d];
        }
        dst[off] = (D)src[i];
    }
}

// Zeros n bytes.
__global__ void zero_bytes_kernel(unsigned char *p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

template <class S, class D>
__global__ void convert_kernel(const S *src, D *dst, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (D)src[i];
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

static int grid_for(int64_t n, int block, int cap = 65535) {
    int64_t g = (n + block - 1) / block;
    if (g < 1) g = 1;
    return (int)(g > cap ? cap : g);
}

int prologue_blocks(int64_t maxN) { return (int)((maxN + kProChunk - 1) / kProChunk); }

template <class T>
void launch_prologue(const Problem &P, const ImgDesc *desc, ImgState *state, const double *gammas, int nimg,
                     double *part, int64_t maxN, hipStream_t s, const double *tols, const int *maxits) {
    const int nblk = prologue_blocks(maxN);
    const dim3 grid(nblk, nimg);
    const int g1 = grid_for(nimg, 64);
    hipLaunchKernelGGL(weight_max_kernel<T>, dim3(nimg), dim3(256), 0, s, P, desc, state, nimg);
    hipLaunchKernelGGL(prologue_pass<T>, grid, dim3(kProBlock), 0, s, desc, state, part, nblk, 0);
    hipLaunchKernelGGL(prologue_combine0, dim3(g1), dim3(64), 0, s, P, desc, state, part, nblk, gammas, nimg);
    hipLaunchKernelGGL(prologue_pass<T>, grid, dim3(kProBlock), 0, s, desc, state, part, nblk, 1);
    hipLaunchKernelGGL(prologue_combine1, dim3(g1), dim3(64), 0, s, desc, state, part, nblk, nimg);
    hipLaunchKernelGGL(prologue_pass<T>, grid, dim3(kProBlock), 0, s, desc, state, part, nblk, 2);
    hipLaunchKernelGGL(prologue_combine2, dim3(g1), dim3(64), 0, s, P, desc, state, part, nblk, gammas, nimg, tols,
                       maxits);
}

void launch_reset_counts(int *counts, hipStream_t s) {
    hipLaunchKernelGGL(reset_counts_kernel, dim3(1), dim3(64), 0, s, counts);
}

void launch_initial_list(const ImgDesc *desc, const ImgState *state, Tile *list, int *count, int nimg,
                         hipStream_t s, int img0) {
    hipLaunchKernelGGL(initial_list_kernel, dim3(nimg), dim3(64), 0, s, desc, state, list, count, nimg, img0);
}

template <class T>
void launch_warm(const Problem &P, T *arena, const ImgDesc *desc, const ImgState *state, int nimg, int64_t maxN,
                 hipStream_t s) {
    if (P.iso)
        hipLaunchKernelGGL(warm_iso_kernel<T>, dim3(grid_for(maxN, 256, 4096), nimg), dim3(256), 0, s, P, arena,
                           desc, state, nimg);
    else
        hipLaunchKernelGGL(warm_aniso_kernel<T>, dim3(grid_for(maxN, 256, 1024), nimg), dim3(256), 0, s, P, arena,
                           desc, state, nimg);
}

template <class T>
void launch_tv1d(const Problem &P, const ImgDesc *desc, ImgState *state, int nimg, int64_t maxN, double u_round,
                 double eta, hipStream_t s) {
    hipLaunchKernelGGL(tv1d_kernel<T>, dim3(grid_for(maxN, 64, 4096), nimg), dim3(64), 0, s, P, desc, state,
                       nimg);
    hipLaunchKernelGGL(tv1d_finish_kernel, dim3(grid_for(nimg, 64)), dim3(64), 0, s, desc, state, nimg, u_round,
                       eta);
}

// A launch that stamps t0 when the kernel starts and t1 when it completes
// (either may be null) through its own dispatch, with no separate event
// record in the stream; without events, a plain launch (the form graph
// capture records).
template <class F, class... Args>
static void launch_timed(F kernel, dim3 grid, dim3 block, hipStream_t s, hipEvent_t t0, hipEvent_t t1, Args... args) {
    if (t0 || t1) hipExtLaunchKernelGGL(kernel, grid, block, 0, s, t0, t1, 0, args...);
    else hipLaunchKernelGGL(kernel, grid, block, 0, s, args...);
}

// t1 (optional): an event stamped when the epilogue completes.
void launch_epilogue(const ImgDesc *desc, ImgState *state, const double *part, Tile *list_out, int *count_out,
                     int TT, double u_round, double eta, int nimg, hipStream_t s, hipEvent_t t1) {
    launch_timed(epilogue_kernel, dim3(nimg), dim3(64), s, nullptr, t1, desc, state, part, list_out, count_out, TT,
                 u_round, eta, nimg);
}

// t0, t1 (optional): events stamped when the step kernel starts and when it
// completes, by its own dispatch.
template <class T, bool ISO, bool SW>
static void step_dispatch(const KernelChoice &kc, const StepArgs<T> &A, int grid, hipStream_t s, hipEvent_t t0,
                          hipEvent_t t1) {
    if (kc.s3) {
        if (kc.TT >= 2) {
            if constexpr (sizeof(T) == 4 && !SW) {
                if (kc.TT == 2)
                    launch_timed(stream3t_kernel<ISO, 2>, dim3(grid), dim3(64 * kS3TWaves), s, t0, t1, A, kc.RY);
                else if (kc.TT == 3)
                    launch_timed(stream3t_kernel<ISO, 3>, dim3(grid), dim3(64 * kS3TWaves), s, t0, t1, A, kc.RY);
            }
            return;  // kernel_available_s3 admits no other combination
        }
        launch_timed(stream3d_kernel<T, ISO, SW, kS3Waves>, dim3(grid), dim3(64 * kS3Waves), s, t0, t1, A, kc.RY);
        return;
    }
    if (kc.generic) {
        launch_timed(step1_kernel<T, ISO, SW>, dim3(grid), dim3(256), s, t0, t1, A);
        return;
    }
    if (kc.stream) {
#define FPR_SCASE(TTv)                                                                                          \
    if (kc.TT == TTv) {                                                                                         \
        launch_timed(stream2d_kernel<T, ISO, SW, TTv>, dim3(grid), dim3(256), s, t0, t1, A, kc.RY);            \
        return;                                                                                                 \
    }
        FPR_SCASE(2)
        FPR_SCASE(4)
        FPR_SCASE(6)
        FPR_SCASE(8)
#undef FPR_SCASE
        return;
    }
#define FPR_CASE(TTv, RYv)                                                                              \
    if (kc.TT == TTv && kc.RY == RYv) {                                                                 \
        launch_timed(step2d_kernel<T, ISO, SW, TTv, RYv>, dim3(grid), dim3(256), s, t0, t1, A);          \
        return;                                                                                         \
    }
    if constexpr (sizeof(T) == 4) {
        FPR_CASE(2, 32)
        FPR_CASE(4, 32)
    } else {
        FPR_CASE(2, 16)
    }
#undef FPR_CASE
}

template <class T>
void launch_step(const KernelChoice &kc, bool iso, bool sw, const StepArgs<T> &A, int grid, hipStream_t s,
                 hipEvent_t t0, hipEvent_t t1) {
    if (iso) {
        if (sw) step_dispatch<T, true, true>(kc, A, grid, s, t0, t1);
        else step_dispatch<T, true, false>(kc, A, grid, s, t0, t1);
    } else {
        if (sw) step_dispatch<T, false, true>(kc, A, grid, s, t0, t1);
        else step_dispatch<T, false, false>(kc, A, grid, s, t0, t1);
    }
}

bool kernel_available_s3(int RY, int TT, int dtype_bytes, bool sw, int64_t maxN) {
    if (RY < 1 || RY > 4096) return false;
    return TT == 1 || ((TT == 2 || TT == 3) && dtype_bytes == 4 && !sw && stream_fits(dtype_bytes, maxN));
}

bool kernel_available(int dtype_bytes, bool stream, int TT, int RY) {
    if (stream) return (TT == 2 || TT == 4 || TT == 6 || TT == 8) && RY >= 1 && RY <= 4096;
    if (dtype_bytes == 4) return (TT == 2 && RY == 32) || (TT == 4 && RY == 32);
    return TT == 2 && RY == 16;
}

// The dual iterate u (k planes) of every image out of the state arena, to
// dst + i * img_stride (zero for an image whose state was never written).
template <class T>
__global__ __launch_bounds__(256) void copy_u_kernel(const T *arena, const ImgDesc *desc, const ImgState *state,
                                                     int k, T *dst, int64_t img_stride) {
    const int img = blockIdx.y;
    const ImgDesc &D = desc[img];
    const ImgState &S = state[img];
    const int64_t N = (int64_t)D.L[0] * D.L[1] * D.L[2];
    const T *U = arena + D.st_off + (int64_t)S.cur * 2 * k * N;
    T *out = dst + img * img_stride;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)k * N;
         i += (int64_t)gridDim.x * blockDim.x)
        out[i] = S.fresh ? (T)0 : U[i];
}

template <class T>
void launch_copy_u(const T *arena, const ImgDesc *desc, const ImgState *state, int nimg, int k, int64_t maxN, T *dst,
                   int64_t img_stride, hipStream_t s) {
    if (nimg <= 0) return;
    hipLaunchKernelGGL(copy_u_kernel<T>, dim3(grid_for(k * maxN, 256, 1024), nimg), dim3(256), 0, s, arena, desc,
                       state, k, dst, img_stride);
}

template <class T>
void launch_gather(const T *src, T *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s) {
    hipLaunchKernelGGL(strided_copy_kernel<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n, nd, sh, 1);
}
template <class T>
void launch_scatter(const T *src, T *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s) {
    hipLaunchKernelGGL(strided_copy_kernel<T>, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n, nd, sh, 0);
}
template <class T>
void launch_gather_group(const GatherEntry *tab, int G, int64_t maxN, hipStream_t s) {
    hipLaunchKernelGGL(gather_group_kernel<T>, dim3(grid_for(maxN, 256, 1024), G), dim3(256), 0, s, tab, G);
}
void launch_admit_commit(const ImgDesc *sdesc, const ImgState *sstate, const double *sgam, const int *map,
                         ImgDesc *desc, ImgState *state, double *gam, Tile *list, int *count, int G, hipStream_t s) {
    hipLaunchKernelGGL(admit_commit_kernel, dim3(G), dim3(64), 0, s, sdesc, sstate, sgam, map, desc, state, gam,
                       list, count, G);
}
template <class S, class D>
void launch_convert(const S *src, D *dst, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL((convert_kernel<S, D>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n);
}
template <class S, class D>
void launch_convert_scatter(const S *src, D *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s) {
    hipLaunchKernelGGL((convert_scatter_kernel<S, D>), dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, src, dst, n, nd,
                       sh);
}
void launch_zero_bytes(void *p, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(zero_bytes_kernel, dim3(grid_for((int64_t)n, 256, 8192)), dim3(256), 0, s, (unsigned char *)p,
                       n);
}

#define FPR_INST(T)                                                                                              \
    template void launch_prologue<T>(const Problem &, const ImgDesc *, ImgState *, const double *, int, double *, \
                                     int64_t, hipStream_t, const double *, const int *);                                 \
    template void launch_gather_group<T>(const GatherEntry *, int, int64_t, hipStream_t);                             \
    template void launch_warm<T>(const Problem &, T *, const ImgDesc *, const ImgState *, int, int64_t, hipStream_t);                                                         \
    template void launch_tv1d<T>(const Problem &, const ImgDesc *, ImgState *, int, int64_t, double, double,      \
                                 hipStream_t);                                                  \
    template void launch_step<T>(const KernelChoice &, bool, bool, const StepArgs<T> &, int, hipStream_t,        \
                                 hipEvent_t, hipEvent_t);                                                        \
    template void launch_gather<T>(const T *, T *, int64_t, int, const StridedShape &, hipStream_t);            \
    template void launch_scatter<T>(const T *, T *, int64_t, int, const StridedShape &, hipStream_t);            \
    template void launch_copy_u<T>(const T *, const ImgDesc *, const ImgState *, int, int, int64_t, T *, int64_t,  \
                                   hipStream_t);
FPR_INST(float)
FPR_INST(double)
#undef FPR_INST
// Integer outputs of the structure queries (labels, signs, states, masks).
template void launch_scatter<int8_t>(const int8_t *, int8_t *, int64_t, int, const StridedShape &, hipStream_t);
template void launch_gather<int8_t>(const int8_t *, int8_t *, int64_t, int, const StridedShape &, hipStream_t);
template void launch_scatter<int32_t>(const int32_t *, int32_t *, int64_t, int, const StridedShape &, hipStream_t);
template void launch_scatter<int64_t>(const int64_t *, int64_t *, int64_t, int, const StridedShape &, hipStream_t);
template void launch_convert<float, double>(const float *, double *, int64_t, hipStream_t);
template void launch_convert<double, float>(const double *, float *, int64_t, hipStream_t);
template void launch_convert<float, float>(const float *, float *, int64_t, hipStream_t);
template void launch_convert<double, double>(const double *, double *, int64_t, hipStream_t);
template void launch_convert_scatter<float, double>(const float *, double *, int64_t, int, const StridedShape &,
                                                    hipStream_t);
template void launch_convert_scatter<double, float>(const double *, float *, int64_t, int, const StridedShape &,
                                                    hipStream_t);

}  // namespace fpr
