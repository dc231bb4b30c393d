// Anisotropic TV on the device by the reference CPU library's method
// (src/ops/tv_aniso.cpp:259-355 prox_img, :376-530 the schedules and adbpg,
// :534-616 fdpg3, :630-675 feasible_dual, :686-727 mark_edges, :886-953
// try_polish; src/ops/tv1d.cpp for the 1D steps).
//
// Method. The dual of the prox splits into one block per axis, and each block
// step is an exact 1D TV prox on every fibre of that axis:
//   * two active axes a1 (the one with the smaller stride) and a2: ADBPG,
//       u  = prox_{g1}(d + w),  y+ = w - u + prox_{g2}(u - w),
//       w  = y+ + (t_k - 1)/t_{k+1} (y+ - y),
//     the point certified being the second block step z = prox_{g2}(u - w);
//   * three active axes: FDPG with u = d + sum_i w_i,
//       y_i+ = w_i - u/3 + prox_{3 g_i}(u - 3 w_i)/3,
//     the primal the block step z = prox_{g_2}(d + y_0 + y_1);
//   * one active axis: one exact pass.
// Restarts every 4th iteration on the sign of <w - y+, y+ - y>, gap checks
// on the reference's schedules (CheckSchedule in 2 axes, every 4th iteration
// in 3), and the polish: edges whose 1D dual p_e (the running sum of q_a
// along the fibre, negated) lies strictly inside lam (1 - 1e-9) join their
// pixels into regions, on which the candidate is the closed-form mean; it is
// taken when its gap (against the same dual) meets the target and is at
// most 1/100 of the iterate's.
//
// The 1D prox of a fibre is Condat's direct algorithm in the reference's
// generalised form (the first three samples of a segment by the paper's
// bounds, then by the running sum; reciprocals 1/m), with the reference's
// range tests, radius clamp, power-of-two scaling of huge rows and constant
// mean, and the taut string when Condat has rescanned more than 3n + 1024
// steps. One thread solves one fibre; a fibre past the budget is queued for a
// second kernel that runs the taut string in a scratch slot. Every
// arithmetic step of the 1D solves, the block updates, the dual, the edge
// marks and the region means is the reference's, in its order (the library
// is built with contraction off), so z follows the reference's iterates; the
// reductions (gaps, restart dots, region sums) are fixed-order FP64 trees,
// which differ from the reference's sequential sums by rounding only.
//
// Both dtypes compute in FP64, as the reference does (tv_aniso.cpp:274-281:
// its working arrays are double); outputs are rounded to the dtype at the end.
//
// Layout: a uniform batch of B images of N pixels, image-major; planes of an
// image at + p N. Arrays (double): d (the input at the solved scale 2^-e x),
// u, z, zp (the polish candidate), y and w (1 plane, 3 for three axes), tmp
// (3 planes, three axes: the block steps before their update), q (k planes,
// the C-form dual of the reference's c.dual). One thread per image runs the
// control on the device (phases below); the host launches each phase's
// kernels for every image (a block returns at once for an image in another
// phase) and reads the phase counts once per iteration.
//
// SPDX-License-Identifier: MIT
#include "adbpg.hpp"

#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/fastprox_rocm.h"
#include "alloc.hpp"
#include "regions.hpp"

namespace fpr {
namespace {

constexpr int kCheck = 4;              // tv_aniso.cpp:78
constexpr double kPolishDrop = 0.1;    // :80
constexpr double kPolishEta = 1e-9;    // :83
constexpr double kPolishRatio = 0.01;  // :88
constexpr double kMaxScaledStep = 0x1p960;  // op.hpp:48
constexpr int kThr = 256;
constexpr int kFib = 64;          // fibres per block of a pass
constexpr int kElem = kThr * 8;   // pixels per block of the per-image sums
constexpr int kMaxSlots = 1024;   // taut-string scratch slots
constexpr int kRing = 128;        // recent samples of each fibre in LDS (Condat's rescans read them)
constexpr int kChunkS = 16;       // samples per load batch
constexpr int kSweep = 8;         // entries per load batch of the updates
constexpr int kQueue = 16;        // finished segments of a fibre waiting to be written

// Per-image phases.
enum : int {
    kPhDone = 0,
    kPhIter,    // one iteration: the block steps and the control
    kPhGap,     // the gap of z (three axes: after the primal pass)
    kPhPolish,  // a polish attempt
    kPhFinal,   // ADBPG: the dual's own point d - q1 - q2 and its gap
    kPhOut,     // outputs
    kPhWarm,    // a warm start's first point
    kPhInit,    // ADBPG: the first block step z(y) = prox_{g1}(d + y)
    kPhExact,   // one active axis: its exact pass
    kPhCount
};
// Where a gap check returns to.
enum : int { kCtxWarm = 0, kCtxInit = 1, kCtxLoop = 2, kCtxAfter = 3, kCtxAfterGap = 4 };

struct AdImg {
    double lam[3];  // gamma w_a at the solved scale, 0 on inactive axes
    double target, gp, gz, pz, accept, t, beta, ck_prev, sched_next;
    double s2, s2m, maxabs, zerr;  // |d|^2, |d / max|d||^2, max |d|; the certified error
    int e, na, a1, a2;
    int it, max_iter, since, last, ck_next, ck_prev_it;
    int phase, ctx, test, wv_is_y, polished, warm, status, nan;
    int need_dual, take_zp, primal, gap_zp, exact_pending, pad_;
};

struct Geo {
    int k, B, N;
    int L[3], S[3];  // padded extents and strides (axis 2 contiguous)
};

struct Seg {
    int dx, pad_;
    double dy, slope;
};

struct Bufs {
    double *d, *u, *z, *zp, *y, *wv, *tmp, *q;
    int P;  // planes of y, w and tmp
    uint8_t *flag;
    int *par, *isroot, *rank, *label, *iota, *skey, *members, *rstart;
    double *rsum, *rhead, *rtail;
    double *part;  // per image pp slots
    int pp;
    int *queue;  // fallback entries (image, axis, fibre) packed
    int *ctr;    // [0]: queued, [1]: blocks done, [2]: total queued
    Seg *segs;
    int nslots, segstride;
    AdImg *st;
    int *phase_count;
    void **img;  // per image: x, z, q (3 B device pointers)
};

__device__ __forceinline__ double min_nn(double a, double v) { return v < a ? v : a; }  // op.hpp:1198
__device__ __forceinline__ double max_nn(double a, double v) { return v > a ? v : a; }  // op.hpp:1211
__device__ __forceinline__ double fista_t(double t) { return 0.5 * (1.0 + sqrt(1.0 + 4.0 * t * t)); }

__device__ __forceinline__ int fibres_of(const Geo &G, int ax) { return G.N / G.L[ax]; }
__device__ __forceinline__ int fibre_base(const Geo &G, int ax, int f) {
    const int s = G.S[ax];
    const int outer = (int)((unsigned)f / (unsigned)s);
    return outer * G.L[ax] * s + (f - outer * s);
}
__device__ __forceinline__ void coords(const Geo &G, int o, int c[3]) {
    const unsigned q = (unsigned)o, L2 = (unsigned)G.L[2], L1 = (unsigned)G.L[1];
    const unsigned t = q / L2, v = t / L1;
    c[2] = (int)(q - t * L2);
    c[1] = (int)(t - v * L1);
    c[0] = (int)v;
}

// ---------------------------------------------------------------------------
// Fixed-order block sums
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

// ---------------------------------------------------------------------------
// The 1D prox of one fibre (tv1d.cpp:505-537 prox / prox_finite with unit
// weight, :72-217 condat, :285-378 taut_string)
// ---------------------------------------------------------------------------

enum : int { kPrepId = 0, kPrepMean = 1, kPrepSolve = 2, kPrepNaN = 3 };

struct Prep {
    int kind, sk;
    double lam, lo, hi;
};

// The fibre's range, the scaling of huge rows (tv1d.cpp:238-244), and which
// path solves it.
template <class IO>
__device__ Prep tv1d_prep(const IO &io, int n, double gamma) {
    Prep P;
    P.sk = 0;
    P.lam = 0.0;
    double lo = io.in(0), hi = lo;
    bool bad = false;
    // in chunks of independent loads (the fibre's entries are far apart in
    // memory; one at a time, each would wait for its own)
    for (int i0 = 0; i0 < n; i0 += kChunkS) {
        double v[kChunkS];
#pragma unroll
        for (int j = 0; j < kChunkS; ++j) v[j] = io.in(i0 + j < n ? i0 + j : n - 1);  // unconditional: one wait
#pragma unroll
        for (int j = 0; j < kChunkS; ++j) {
            bad |= !isfinite(v[j]);
            lo = min_nn(lo, v[j]);
            hi = max_nn(hi, v[j]);
        }
    }
    if (bad) {
        P.kind = kPrepNaN;
        return P;
    }
    const double m = -lo < hi ? hi : -lo;
    if (!(m < 0x1p960)) {
        int e;
        frexp(m, &e);
        P.sk = e - 960;
        lo = ldexp(lo, -P.sk);
        hi = ldexp(hi, -P.sk);
        gamma = ldexp(gamma, -P.sk);
    }
    P.lo = lo;
    P.hi = hi;
    const double range = hi - lo;
    if (gamma == 0.0 || range == 0.0) {
        P.kind = kPrepId;
        return P;
    }
    if (gamma >= (double)(n / 2) * range) {
        P.kind = kPrepMean;
        return P;
    }
    const double cap = 2.0 * (double)(n / 2 > 1 ? n / 2 : 1) * range;  // inactive_radius (tv1d.cpp:225)
    const double lam = cap < gamma ? cap : gamma;
    P.kind = lam == 0.0 ? kPrepId : kPrepSolve;
    P.lam = lam;
    return P;
}

template <class IO>
__device__ __forceinline__ double scaled_in(const IO &io, int i, int sk) {
    const double v = io.in(i);
    return sk ? ldexp(v, -sk) : v;
}

// The identity and the compensated constant mean (tv1d.cpp:249-261).
template <class IO>
__device__ void tv1d_simple(const IO &io, int n, const Prep &P) {
    if (P.kind == kPrepNaN) {
        for (int i = 0; i < n; ++i) io.put(i, __builtin_nan(""));
        return;
    }
    if (P.kind == kPrepId) {
        for (int i = 0; i < n; ++i) io.put(i, scaled_in(io, i, P.sk));
        return;
    }
    double hi = 0.0, lo = 0.0;
    for (int i = 0; i < n; ++i) {
        const double v = scaled_in(io, i, P.sk);
        const double s = hi + v;
        const double bb = s - hi;
        lo += (hi - (s - bb)) + (v - bb);
        hi = s;
    }
    const double mean = (hi + lo) / (double)n;
    for (int i = 0; i < n; ++i) io.put(i, mean);
}

// Back from the scaling of a huge row (tv1d.cpp:415-423).
template <class IO>
__device__ void tv1d_unscale(const IO &io, int n, const Prep &P) {
    if (!P.sk || P.kind == kPrepNaN) return;
    for (int i = 0; i < n; ++i) {
        const double z = io.got(i);
        const double v = ldexp(z, P.sk);
        const double c = z < P.lo ? P.lo : z;
        io.put(i, isinf(v) ? ldexp(P.hi < c ? P.hi : c, P.sk) : v);
    }
}

// Condat's algorithm with a constant radius lam (tv1d.cpp:72-217, the
// generalised form: the paper's bounds for the first three samples of a
// segment, then the running sum), step for step, with its input staged
// through LDS and its output written in lockstep. Every lane of the wave
// solves its own fibre. The wave loads the next kChunkS samples of all its
// fibres at once (issued before the lanes scan the samples already staged,
// so the loads overlap the scan) into a ring of the last kRing samples of
// each fibre; a rescan further back than the ring reads global memory. A
// finished segment goes into a short per-lane queue, and each lockstep
// iteration writes one entry of every lane's queue, so the segments' stores
// are issued together rather than one lane's at a time. act: this lane's
// fibre is solved here (every lane of the wave runs the loop). Returns 2
// when solved, 4 past the budget (the output is then incomplete), 3 for an
// inactive lane.
struct RingLds {  // 76 KiB: two blocks per CU
    double ring[kRing][kFib];
    double qv[kQueue][kFib];
    int qe[kQueue][kFib];
};

template <class IO>
__device__ int condat_ring(const IO &io, int n, double lam, int sk, int budget, bool act, RingLds &L) {
    const int lane = threadIdx.x;
    int k = 0, k0 = 0, km = 0, kp = 0, rescanned = 0;
    double vmin = 0.0, vmax = 0.0, umin = lam, umax = -lam, ref = 0.0, s = 0.0;
    int state = act ? 0 : 3;  // 0 not started, 1 scanning, 2 solved, 3 inactive, 4 past the budget
    int F = 0;                // samples staged so far (uniform over the wave)
    int head = 0, tail = 0, w = 0;  // the segment queue and the next entry to write
    double pf[kChunkS];
#pragma unroll
    for (int j = 0; j < kChunkS; ++j) pf[j] = scaled_in(io, j < n ? j : n - 1, sk);  // every lane: in bounds
    auto rcp = [&](int m) -> double { return 1.0 / (double)m; };  // the reference's table of 1/m (tv1d.cpp:49-57)
    auto Y = [&](int i) -> double {
        return i >= F - kRing ? L.ring[i & (kRing - 1)][lane] : scaled_in(io, i, sk);
    };
    auto push = [&](int end, double v) {
        L.qe[tail & (kQueue - 1)][lane] = end;
        L.qv[tail & (kQueue - 1)][lane] = v;
        ++tail;
    };
    auto drain = [&]() {
        if (head < tail) {
            const int h = head & (kQueue - 1);
            io.put(w, L.qv[h][lane]);
            if (w == L.qe[h][lane]) ++head;
            ++w;
        }
    };
    auto jump_down = [&]() {
        push(km, vmin);
        rescanned += k - km;
        k = k0 = kp = km = km + 1;
        const double yk = Y(k < n ? k : n - 1);  // k < n: no jump starts past the last sample
        vmin = yk + (lam - lam);
        vmax = yk + (lam + lam);
        umin = lam;
        umax = -lam;
    };
    auto jump_up = [&]() {
        push(kp, vmax);
        rescanned += k - kp;
        k = k0 = km = kp = kp + 1;
        const double yk = Y(k < n ? k : n - 1);
        vmin = yk - (lam + lam);
        vmax = yk - (lam - lam);
        umin = lam;
        umax = -lam;
    };
    // One step of the scan (the body of the reference's loops). Both forms'
    // tests and updates are formed and the lane's one selected, so the wave
    // runs two paths per step (advance or jump) rather than one per form and
    // outcome; every value selected is the reference's expression.
    auto step = [&]() {
        if (k < n - 1) {
            const double yn = Y(k + 1);
            const double ln = lam;
            const int m = k - k0 + 2;
            const bool shrt = m <= 3;
            const double r = rcp(m);
            // the running-sum form's candidates
            const double s1 = s + (yn - ref);
            const double ca = ref + (s1 - ln) * r, cb = ref + (s1 + ln) * r;
            const bool down = shrt ? (yn + umin < vmin - ln) : (cb < vmin);
            const bool up = !down && (shrt ? (yn + umax > vmax + ln) : (ca > vmax));
            if (down || up) {
                const int e = down ? km : kp;
                push(e, down ? vmin : vmax);
                rescanned += k - e;
                k = k0 = km = kp = e + 1;
                const double yk = Y(k < n ? k : n - 1);  // k < n: no jump starts past the last sample
                vmin = down ? yk + (lam - lam) : yk - (lam + lam);
                vmax = down ? yk + (lam + lam) : yk - (lam - lam);
                umin = lam;
                umax = -lam;
                if (rescanned > budget) state = 4;
                return;
            }
            ++k;
            // the first samples' form
            const double un = umin + (yn - vmin), ux = umax + (yn - vmax);
            const bool cm = un >= ln, cp = ux <= -ln;
            const double dm = max_nn(un - ln, 0.0), dp = min_nn(ux + ln, 0.0);
            const double vn_s = vmin + dm * r, vx_s = vmax + dp * r;
            const double un_s = min_nn(un, ln), ux_s = max_nn(ux, -ln);
            // the running-sum form's
            const double vn_l = max_nn(vmin, ca), vx_l = min_nn(vmax, cb);
            km = shrt ? (cm ? k : km) : (ca >= vmin ? k : km);
            kp = shrt ? (cp ? k : kp) : (cb <= vmax ? k : kp);
            vmin = shrt ? vn_s : vn_l;
            vmax = shrt ? vx_s : vx_l;
            umin = shrt ? un_s : umin;
            umax = shrt ? ux_s : umax;
            s = shrt ? s : s1;
            if (m == 3) {  // to the running sum
                ref = Y(k0);
                s = umin + 3.0 * (vmin - ref);
            }
        } else {
            const int m = k - k0 + 1;
            bool down, up;
            double v;
            if (m <= 3) {
                down = umin < 0.0;
                up = umax > 0.0;
                v = vmin + umin / (double)m;
            } else {
                v = ref + s * rcp(m);
                down = v < vmin;
                up = v > vmax;
            }
            if (down) {
                jump_down();
            } else if (up) {
                jump_up();
            } else {
                push(n - 1, v);
                state = 2;
                return;
            }
            if (rescanned > budget) state = 4;
        }
    };
    for (;;) {
#pragma unroll
        for (int j = 0; j < kChunkS; ++j)
            if (F + j < n) L.ring[(F + j) & (kRing - 1)][lane] = pf[j];
        F += kChunkS;
#pragma unroll
        for (int j = 0; j < kChunkS; ++j) pf[j] = scaled_in(io, F + j < n ? F + j : n - 1, sk);
        if (state == 0) {
            const double y0 = Y(0);
            vmin = y0 - lam;
            vmax = y0 + lam;
            state = 1;
        }
        for (;;) {
            // a step needs its next sample staged and room for one segment
            const bool can = state == 1 && (k >= n - 1 || k + 1 < F) && tail - head < kQueue;
            if (state == 4) head = tail;  // the taut string rewrites the fibre
            if (!__any(can || head < tail)) break;
            if (can) step();
            drain();
        }
        if (!__any(state <= 1)) break;
    }
    return state;
}

// The taut string with a constant radius lam; up and lo hold n + 1
// segments each.
template <class IO>
__device__ void taut_string(const IO &io, int m, double lam, int sk, Seg *up, Seg *lo) {
    if (m == 1) {
        io.put(0, scaled_in(io, 0, sk));
        return;
    }
    int uhd = 0, utl = 0, lhd = 0, ltl = 0;
    auto radius = [&](int k) { return (k <= 0 || k >= m) ? 0.0 : lam; };
    int pos = 0, dx = 0;
    double du = 0.0, dl = 0.0;
    auto emit = [&](const Seg &sg) {
        const double v = sg.slope;
        for (int i = 0; i < sg.dx; ++i) io.put(pos + i, v);
        pos += sg.dx;
        dx -= sg.dx;
    };
    for (int k = 1; k <= m; ++k) {
        const double yk = scaled_in(io, k - 1, sk);
        const double dr = radius(k) - radius(k - 1);
        ++dx;
        du += yk + dr;
        dl += yk - dr;
        Seg s{1, 0, yk + dr, yk + dr};
        if (utl > uhd && s.slope < up[utl - 1].slope) {
            do {
                --utl;
                s.dx += up[utl].dx;
                s.dy += up[utl].dy;
            } while (utl > uhd && s.dy < (double)s.dx * up[utl - 1].slope);
            s.slope = s.dy / (double)s.dx;
        }
        up[utl++] = s;
        s = Seg{1, 0, yk - dr, yk - dr};
        if (ltl > lhd && s.slope > lo[ltl - 1].slope) {
            do {
                --ltl;
                s.dx += lo[ltl].dx;
                s.dy += lo[ltl].dy;
            } while (ltl > lhd && s.dy > (double)s.dx * lo[ltl - 1].slope);
            s.slope = s.dy / (double)s.dx;
        }
        lo[ltl++] = s;
        while (dx > 0 && up[uhd].slope < lo[lhd].slope) {
            if (up[uhd].dx < lo[lhd].dx) {
                const Seg knot = up[uhd++];
                emit(knot);
                du -= knot.dy;
                dl -= knot.dy;
                lo[0] = Seg{dx, 0, dl, dl / (double)dx};
                lhd = 0;
                ltl = 1;
            } else {
                const Seg knot = lo[lhd++];
                emit(knot);
                du -= knot.dy;
                dl -= knot.dy;
                up[0] = Seg{dx, 0, du, du / (double)dx};
                uhd = 0;
                utl = 1;
            }
        }
    }
    if (dx == 0) return;
    if (utl - uhd > 1) {
        for (int j = uhd; j < utl; ++j) emit(up[j]);
    } else {
        for (int j = lhd; j < ltl; ++j) emit(lo[j]);
    }
}

// ---------------------------------------------------------------------------
// Fibre passes
// ---------------------------------------------------------------------------

// What a pass reads and does with the block step zo of each entry.
enum : int {
    kPassA = 0,     // ADBPG: u = prox_{g1}(d + w)
    kPassB = 1,     // ADBPG: z = prox_{g2}(u - w), then y, w
    kPassInit = 2,  // ADBPG: z = prox_{g1}(d + y)
    kPassExact = 3, // one active axis: z = prox(d), q_a = d - z
    kPass3 = 4,     // FDPG: y_a, w_a from prox_{3 g_a}(u - 3 w_a)
    kPassPrimal3 = 5  // FDPG: z = prox_{g_2}(d + y_0 + y_1)
};

struct PassArgs {
    Geo G;
    Bufs b;
    int phase;     // images in this phase
    int axis[3];   // image axes (0..k-1) of the launch's z blocks; -1: from the image (ADBPG a1/a2)
    int which;     // ADBPG: 1 = a1, 2 = a2
    int qslot;     // the queue counter (2 per launch pair)
};

template <int MODE>
struct FibreIO {
    const double *d, *u, *w, *w2, *y;  // per mode
    double *buf;
    int o0, s;
    double p;
    __device__ double in(int i) const {
        const int o = o0 + i * s;
        if (MODE == kPassA) return d[o] + w[o];
        if (MODE == kPassB) return u[o] - w[o];
        if (MODE == kPassInit) return d[o] + y[o];
        if (MODE == kPassExact) return d[o];
        if (MODE == kPass3) return u[o] - p * w[o];
        return d[o] + w[o] + w2[o];  // kPassPrimal3: d + y_0 + y_1
    }
    __device__ void put(int i, double v) const { buf[o0 + i * s] = v; }
    __device__ double got(int i) const { return buf[o0 + i * s]; }
};

// The image-axis of a launch's z block for image S.
__device__ __forceinline__ int pass_axis(const PassArgs &A, const AdImg &S) {
    const int a = A.axis[blockIdx.z];
    if (a >= 0) return a;
    return A.which == 1 ? S.a1 : S.a2;
}

template <int MODE>
__device__ FibreIO<MODE> make_io(const PassArgs &A, const AdImg &S, int img, int a, int o0, int s) {
    const Geo &G = A.G;
    const Bufs &B = A.b;
    const int64_t N = G.N, P = B.P;
    const int64_t base = (int64_t)img * N;
    FibreIO<MODE> io;
    io.o0 = o0;
    io.s = s;
    io.p = 3.0;
    io.d = B.d + base;
    io.u = B.u + base;
    io.y = B.y + (int64_t)img * P * N;
    const double *wv = (S.wv_is_y ? B.y : B.wv) + (int64_t)img * P * N;
    io.w = wv;
    io.w2 = nullptr;
    if (MODE == kPassA) io.buf = B.u + base;
    if (MODE == kPassB || MODE == kPassInit || MODE == kPassExact || MODE == kPassPrimal3) io.buf = B.z + base;
    if (MODE == kPass3) {
        io.w = wv + (int64_t)a * N;
        io.buf = B.tmp + ((int64_t)img * P + a) * N;
    }
    if (MODE == kPassPrimal3) {
        io.w = io.y;
        io.w2 = io.y + N;
    }
    return io;
}

// The entry updates of a solved fibre (tv_aniso.cpp:469-488, 578-593) and
// its restart inner product.
template <int MODE>
__device__ double sweep(const PassArgs &A, const AdImg &S, const FibreIO<MODE> &io, int img, int a, int n) {
    const Geo &G = A.G;
    const Bufs &B = A.b;
    const int64_t N = G.N, P = B.P;
    double acc = 0.0;
    if (MODE == kPassB) {
        double *y = B.y + (int64_t)img * P * N;
        double *wv = B.wv + (int64_t)img * P * N;
        const double *wr = S.wv_is_y ? y : wv;
        const double beta = S.beta;
        for (int i0 = 0; i0 < n; i0 += kSweep) {
            double zo[kSweep], wo[kSweep], yo[kSweep], uo[kSweep];
#pragma unroll
            for (int j = 0; j < kSweep; ++j) {
                const int o = io.o0 + (i0 + j < n ? i0 + j : n - 1) * io.s;
                zo[j] = io.buf[o];
                wo[j] = wr[o];
                yo[j] = y[o];
                uo[j] = io.u[o];
            }
#pragma unroll
            for (int j = 0; j < kSweep; ++j) {
                if (i0 + j >= n) break;
                const int o = io.o0 + (i0 + j) * io.s;
                const double v = uo[j] - wo[j];
                const double yn = zo[j] - v;
                const double r = (wo[j] - yn) * (yn - yo[j]);
                wv[o] = yn + beta * (yn - yo[j]);
                y[o] = yn;
                acc += r;
            }
        }
    } else if (MODE == kPass3) {
        double *y = B.y + ((int64_t)img * P + a) * N;
        double *wv = B.wv + ((int64_t)img * P + a) * N;
        const double *wr = S.wv_is_y ? y : wv;
        const double beta = S.beta;
        for (int i0 = 0; i0 < n; i0 += kSweep) {
            double zo[kSweep], wo[kSweep], yo[kSweep], uo[kSweep];
#pragma unroll
            for (int j = 0; j < kSweep; ++j) {
                const int o = io.o0 + (i0 + j < n ? i0 + j : n - 1) * io.s;
                zo[j] = io.buf[o];
                wo[j] = wr[o];
                yo[j] = y[o];
                uo[j] = io.u[o];
            }
#pragma unroll
            for (int j = 0; j < kSweep; ++j) {
                if (i0 + j >= n) break;
                const int o = io.o0 + (i0 + j) * io.s;
                const double v = uo[j] - 3.0 * wo[j];
                const double yn = (zo[j] - v) / 3.0;
                const double r = (wo[j] - yn) * (yn - yo[j]);
                wv[o] = yn + beta * (yn - yo[j]);
                y[o] = yn;
                acc += r;
            }
        }
    } else if (MODE == kPassExact) {
        double *qa = B.q + ((int64_t)img * G.k + a) * N;
        for (int i = 0; i < n; ++i) {
            const int o = io.o0 + i * io.s;
            qa[o] = io.d[o] - io.buf[o];
        }
    }
    return acc;
}

// The slot of fibre f along image axis a in the image's partials.
__device__ __forceinline__ int part_slot(const Geo &G, int a, int f) {
    int off = 0;
    for (int b = 0; b < a; ++b) off += G.N / G.L[3 - G.k + b];
    return off + f;
}

template <int MODE>
__global__ __launch_bounds__(kFib) void pass_kernel(PassArgs A) {
    __shared__ RingLds ring;
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != A.phase) return;
    if (MODE == kPassPrimal3 && !S.primal) return;
    if (MODE == kPass3 && S.na != 3) return;
    if ((MODE == kPassA || MODE == kPassB || MODE == kPassInit) && S.na != 2) return;
    const Geo &G = A.G;
    const int a = pass_axis(A, S);
    const int ax = 3 - G.k + a;
    const int n = G.L[ax], s = G.S[ax];
    const int f = blockIdx.x * kFib + threadIdx.x;
    const int nf = fibres_of(G, ax);
    if (blockIdx.x * kFib >= nf) return;  // (uniform over the block)
    // every lane runs the staged scan's loop; lanes past the last fibre idle
    const bool valid = f < nf;
    const FibreIO<MODE> io = make_io<MODE>(A, S, img, a, valid ? fibre_base(G, ax, f) : 0, s);
    const double gamma = MODE == kPass3 ? 3.0 * S.lam[a] : S.lam[a];
    Prep P;
    P.kind = kPrepId;
    P.sk = 0;
    P.lam = 0.0;
    if (valid) P = tv1d_prep(io, n, gamma);
    const bool act = valid && P.kind == kPrepSolve;
    if (valid && !act) tv1d_simple(io, n, P);
    const int st = condat_ring(io, n, P.lam, P.sk, 3 * n + 1024, act, ring);
    if (!valid) return;
    if (st == 4) {
        const int at = atomicAdd(&A.b.ctr[A.qslot], 1);
        atomicAdd(&A.b.ctr[2], 1);
        A.b.queue[at * 3 + 0] = img;
        A.b.queue[at * 3 + 1] = a;
        A.b.queue[at * 3 + 2] = f;
        return;
    }
    tv1d_unscale(io, n, P);
    const double acc = sweep<MODE>(A, S, io, img, a, n);
    if (MODE == kPassB || MODE == kPass3) A.b.part[(int64_t)img * A.b.pp + part_slot(G, a, f)] = acc;
}

// The fibres past Condat's budget: the taut string, one slot per thread
// (tv1d.cpp:467-470). The last block to finish resets the queue.
template <int MODE>
__global__ __launch_bounds__(kThr) void fallback_kernel(PassArgs A) {
    const int count = __atomic_load_n(&A.b.ctr[A.qslot], __ATOMIC_RELAXED);
    const int t = blockIdx.x * kThr + threadIdx.x;
    if (t < A.b.nslots) {
        Seg *up = A.b.segs + (int64_t)t * 2 * A.b.segstride;
        Seg *lo = up + A.b.segstride;
        for (int j = t; j < count; j += A.b.nslots) {
            const int img = A.b.queue[j * 3 + 0], a = A.b.queue[j * 3 + 1], f = A.b.queue[j * 3 + 2];
            const AdImg &S = A.b.st[img];
            const Geo &G = A.G;
            const int ax = 3 - G.k + a;
            const int n = G.L[ax], s = G.S[ax];
            const FibreIO<MODE> io = make_io<MODE>(A, S, img, a, fibre_base(G, ax, f), s);
            const double gamma = MODE == kPass3 ? 3.0 * S.lam[a] : S.lam[a];
            const Prep P = tv1d_prep(io, n, gamma);
            taut_string(io, n, P.lam, P.sk, up, lo);
            tv1d_unscale(io, n, P);
            const double acc = sweep<MODE>(A, S, io, img, a, n);
            if (MODE == kPassB || MODE == kPass3) A.b.part[(int64_t)img * A.b.pp + part_slot(G, a, f)] = acc;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&A.b.ctr[A.qslot + 1], 1) == (int)gridDim.x - 1) {
            A.b.ctr[A.qslot] = 0;
            A.b.ctr[A.qslot + 1] = 0;
            __threadfence();
        }
    }
}

// ---------------------------------------------------------------------------
// Setup
// ---------------------------------------------------------------------------

struct SetupArgs {
    Geo G;
    Bufs b;
    int64_t q_plane;
    const double *gam;
    double w[3];
    double tol;
    int max_iter, warm, nblk;
};

// Per block of an image: NaN and Inf flags and max |x|.
template <class T>
__global__ __launch_bounds__(kThr) void scan_kernel(SetupArgs A) {
    __shared__ double sh[kThr];
    const int img = blockIdx.y;
    const int64_t N = A.G.N;
    const T *x = (const T *)A.b.img[3 * img];
    const int lo = blockIdx.x * kElem;
    double m = 0.0;
    int nan = 0, inf = 0;
    for (int i = lo + threadIdx.x; i < lo + kElem && i < N; i += kThr) {
        const double v = (double)x[i];
        nan |= v != v;
        inf |= isinf(v);
        m = fmax(m, fabs(v));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int d = kThr / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + d]);
        __syncthreads();
    }
    const double mx = sh[0];
    __syncthreads();
    const double fl = block_sum((double)(nan | (inf << 1)), sh);
    if (threadIdx.x == 0) {
        double *p = A.b.part + (int64_t)img * A.b.pp + 2 * blockIdx.x;
        p[0] = mx;
        p[1] = fl;
    }
}

__device__ int rescale_exponent(double m, double gamma, double step) {  // image_util.hpp:56-65
    if (!(m > 0.0)) return 0;
    int e = ilogb(m);
    if (e >= -64 && e <= 64) return 0;
    const double top = gamma < step ? step : gamma;
    if (e < 0 && top > 0.0) {
        const int lo = ilogb(top) + 1 - ilogb(kMaxScaledStep);
        const int t = e > lo ? e : lo;
        e = t < 0 ? t : 0;
    }
    return e;
}

// Per image (tv_aniso.cpp:282-315): flags, the active axes and their
// radii, the scale exponent.
__global__ void setup_kernel(SetupArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    const double *p = A.b.part + (int64_t)img * A.b.pp;
    double m = 0.0;
    int fl = 0;
    for (int j = 0; j < A.nblk; ++j) {
        m = fmax(m, p[2 * j]);
        fl |= (int)p[2 * j + 1];
    }
    const Geo &G = A.G;
    const double gamma = A.gam[img];
    S.status = (fl & 2) ? 1 : 0;
    S.nan = (!S.status && (fl & 1)) ? 1 : 0;
    int na = 0, act[3] = {0, 0, 0};
    double lam[3] = {0.0, 0.0, 0.0};
    double wmax = 0.0;
    for (int a = 0; a < 3; ++a) wmax = fmax(wmax, a < G.k ? A.w[a] : 0.0);
    for (int a = 0; a < G.k; ++a) {
        lam[a] = gamma * A.w[a];
        if (lam[a] > 0.0 && G.L[3 - G.k + a] > 1) act[na++] = a;
        else lam[a] = 0.0;
    }
    int e = (na && !S.status && !S.nan) ? rescale_exponent(m, gamma, gamma * wmax) : 0;
    for (int a = 0; a < G.k && e; ++a)
        if (lam[a] > 0.0 && !(ldexp(lam[a], -e) > 0.0)) e = 0;
    for (int a = 0; a < 3; ++a) S.lam[a] = e ? ldexp(lam[a], -e) : lam[a];
    S.e = e;
    S.maxabs = e ? ldexp(m, -e) : m;
    S.na = na;
    // two active axes: a1 the one with the smaller stride (the later axis)
    S.a1 = na >= 2 ? act[1] : act[0];
    S.a2 = act[0];
    S.warm = A.warm;
    S.max_iter = A.max_iter;
}

// d = 2^-e x, |d|^2 per block, the dual's planes (a warm start's at the
// solved scale, with inactive planes zero; else zero), y and w zero.
template <class T>
__global__ __launch_bounds__(kThr) void load_kernel(SetupArgs A) {
    __shared__ double sh[kThr];
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    const Geo &G = A.G;
    const int64_t N = G.N, P = A.b.P;
    const T *x = (const T *)A.b.img[3 * img];
    const T *qin = (const T *)A.b.img[3 * img + 2];
    double *d = A.b.d + (int64_t)img * N;
    const int lo = blockIdx.x * kElem;
    const bool live = !S.status && !S.nan;
    const double ms = S.maxabs;
    double s2 = 0.0, s2m = 0.0;
    for (int i = lo + threadIdx.x; i < lo + kElem && i < N; i += kThr) {
        double v = (double)x[i];
        if (live && S.e) v = ldexp(v, -S.e);
        d[i] = v;
        s2 += v * v;
        const double r = ms > 0.0 ? v / ms : 0.0;
        s2m += r * r;
        for (int a = 0; a < G.k; ++a) {
            double qv = 0.0;
            const bool keep = S.warm && live && S.na >= 2 && S.lam[a] > 0.0;
            if (keep) {
                qv = (double)qin[a * A.q_plane + i];
                if (S.e) qv = ldexp(qv, -S.e);
            }
            A.b.q[((int64_t)img * G.k + a) * N + i] = qv;
        }
        for (int pl = 0; pl < P; ++pl) {
            A.b.y[((int64_t)img * P + pl) * N + i] = 0.0;
            A.b.wv[((int64_t)img * P + pl) * N + i] = 0.0;
        }
    }
    s2 = block_sum(s2, sh);
    s2m = block_sum(s2m, sh);
    if (threadIdx.x == 0) {
        A.b.part[(int64_t)img * A.b.pp + 2 * blockIdx.x] = s2;
        A.b.part[(int64_t)img * A.b.pp + 2 * blockIdx.x + 1] = s2m;
    }
}

// The stopping target and the first phase (tv_aniso.cpp:316-337, 441-458,
// 548-551).
__global__ void init_kernel(SetupArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    const double *p = A.b.part + (int64_t)img * A.b.pp;
    double s2 = 0.0, s2m = 0.0;
    for (int j = 0; j < A.nblk; ++j) {
        s2 += p[2 * j];
        s2m += p[2 * j + 1];
    }
    S.s2 = s2;
    S.s2m = s2m;
    S.zerr = 0.0;
    S.target = A.tol * 0.5 * s2;
    S.gp = INFINITY;
    S.gz = INFINITY;
    S.pz = 0.0;
    S.t = 1.0;
    const double tn = fista_t(1.0);
    S.beta = (1.0 - 1.0) / tn;
    S.t = tn;
    S.it = 0;
    S.since = 0;
    S.last = 0;
    S.ck_next = 1;
    S.ck_prev_it = -1;
    S.ck_prev = 0.0;
    S.sched_next = INFINITY;
    S.test = 1;  // it = 0
    S.wv_is_y = 0;
    S.polished = 0;
    S.need_dual = 0;
    S.take_zp = 0;
    S.primal = 0;
    S.gap_zp = 0;
    S.exact_pending = 0;
    S.ctx = kCtxLoop;
    if (S.status || S.nan || S.na == 0) {
        S.phase = kPhOut;
        return;
    }
    if (S.na == 1) {
        S.phase = kPhExact;
        return;
    }
    if (S.warm) {
        S.phase = kPhWarm;
        return;
    }
    if (S.na == 2 && (S.max_iter == 0 || !(S.target < INFINITY))) {
        S.phase = kPhInit;
        return;
    }
    S.phase = kPhIter;
}

// ---------------------------------------------------------------------------
// Warm starts (tv_aniso.cpp:630-675 feasible_dual, 441-451, 548-551)
// ---------------------------------------------------------------------------

struct ElemArgs {
    Geo G;
    Bufs b;
    int phase;
};

// Makes q_a feasible along each fibre of a warm start's active axis a.
__global__ __launch_bounds__(kFib) void feasible_kernel(ElemArgs A, int a) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhWarm || !(S.lam[a] > 0.0)) return;
    const Geo &G = A.G;
    const int ax = 3 - G.k + a;
    const int L = G.L[ax], s = G.S[ax];
    const int f = blockIdx.x * kFib + threadIdx.x;
    if (f >= fibres_of(G, ax)) return;
    const int b0 = fibre_base(G, ax, f);
    double *qa = A.b.q + ((int64_t)img * G.k + a) * G.N;
    const double lam = S.lam[a];
    double acc = 0.0, mass = 0.0, peak = -lam;
    for (int l = 0; l < L; ++l) {
        const int o = b0 + l * s;
        acc += qa[o];
        mass += fabs(qa[o]);
        const double e = l + 1 < L ? fabs(acc) - lam : -lam;
        peak = e > peak || e != e ? e : peak;
    }
    const double tau = 1e-10 * lam + 16.0 * 2.220446049250313e-16 * mass;
    if (isfinite(mass) && peak <= tau && fabs(acc) <= tau) return;
    double run = 0.0, prev = 0.0;
    for (int l = 0; l < L; ++l) {
        double &qi = qa[b0 + l * s];
        run += qi;
        double pe = 0.0;
        if (l + 1 < L && isfinite(run)) pe = fmin(fmax(-run, -lam), lam);
        qi = prev - pe;
        prev = pe;
    }
}

// A warm start's point: ADBPG y = -q2, z = d - q1 - q2; FDPG y_i = -q_i.
__global__ __launch_bounds__(kThr) void warm_point_kernel(ElemArgs A) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhWarm) return;
    const Geo &G = A.G;
    const int64_t N = G.N, P = A.b.P;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const double *q = A.b.q + (int64_t)img * G.k * N;
    double *y = A.b.y + (int64_t)img * P * N;
    double *wv = A.b.wv + (int64_t)img * P * N;
    if (S.na == 2) {
        const double q1 = q[(int64_t)S.a1 * N + i], q2 = q[(int64_t)S.a2 * N + i];
        y[i] = -q2;
        wv[i] = -q2;
        A.b.z[(int64_t)img * N + i] = A.b.d[(int64_t)img * N + i] - q1 - q2;
    } else {
        for (int a = 0; a < 3; ++a) {
            y[(int64_t)a * N + i] = -q[(int64_t)a * N + i];
            wv[(int64_t)a * N + i] = -q[(int64_t)a * N + i];
        }
    }
}

// ---------------------------------------------------------------------------
// Iteration pieces
// ---------------------------------------------------------------------------

// FDPG: u = d + w_0 + w_1 + w_2.
__global__ __launch_bounds__(kThr) void fdpg_u_kernel(ElemArgs A) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhIter || S.na != 3) return;
    const int64_t N = A.G.N, P = A.b.P;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const double *w = (S.wv_is_y ? A.b.y : A.b.wv) + (int64_t)img * P * N;
    A.b.u[(int64_t)img * N + i] = A.b.d[(int64_t)img * N + i] + w[i] + w[N + i] + w[2 * N + i];
}

// CheckSchedule::done (tv_aniso.cpp:397-413).
__device__ void checks_done(AdImg &S, double gp) {
    int step = 1;
    if (S.ck_prev_it >= 0) {
        step = kCheck;
        if (gp < S.ck_prev && gp > S.target) {
            const double n = log(S.target / gp) / log(gp / S.ck_prev) * (double)(S.it - S.ck_prev_it);
            if (n < kCheck) step = n >= 1.0 ? (int)n : 1;
        }
    }
    S.ck_prev_it = S.it;
    S.ck_prev = gp;
    S.ck_next = S.it + step;
}

// FDPG after its loop, with a finite gap (tv_aniso.cpp:607-611): the dual,
// and a last polish unless the gap meets the target.
__device__ void after_loop3(AdImg &S) {
    S.ctx = kCtxAfter;
    if (!S.polished) {
        S.need_dual = 1;
        if (!(S.gp <= S.target)) {
            S.accept = S.gp;
            S.gz = S.gp;
            S.phase = kPhPolish;
            return;
        }
    }
    S.phase = kPhOut;
}

// Where an image goes after a gap check (and its polish) returns gp.
__device__ void after_check(AdImg &S) {
    if (S.na == 2) {
        if (S.ctx == kCtxWarm) {
            if ((!(S.gp <= S.target) && (S.warm || S.max_iter == 0)) || !(S.target < INFINITY)) {
                S.phase = kPhInit;
                return;
            }
        } else if (S.ctx == kCtxLoop) {
            checks_done(S, S.gp);
        } else if (S.ctx == kCtxAfter) {  // FINAL's gap of d - q1 - q2 is handled by final_ctrl
            S.phase = kPhOut;
            return;
        }
        if (!(S.gp <= S.target) && S.it < S.max_iter) {
            S.phase = kPhIter;
        } else if (!S.polished && S.it > 0) {
            S.phase = kPhFinal;
            S.gap_zp = 1;
        } else {
            S.phase = kPhOut;
        }
        return;
    }
    // three axes
    if (S.ctx == kCtxAfterGap) {
        after_loop3(S);
        return;
    }
    if (S.ctx == kCtxAfter) {
        S.phase = kPhOut;
        return;
    }
    if (!(S.gp <= S.target) && S.it < S.max_iter) {
        S.phase = kPhIter;
        return;
    }
    // after the loop (tv_aniso.cpp:606-611)
    if (!isfinite(S.gp)) {
        S.phase = kPhGap;
        S.primal = 1;
        S.ctx = kCtxAfterGap;
        return;
    }
    after_loop3(S);
}

// After an iteration's block steps: the restart test (tv_aniso.cpp:382-391),
// the iteration count, whether a gap check is due, and the next
// iteration's momentum.
__global__ void iter_ctrl_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhIter) return;
    const Geo &G = A.G;
    double up = 0.0;
    if (S.test) {
        const double *p = A.b.part + (int64_t)img * A.b.pp;
        if (S.na == 2) {
            const int nf = G.N / G.L[3 - G.k + S.a2];
            const int off = part_slot(G, S.a2, 0);
            for (int f = 0; f < nf; ++f) up += p[off + f];
        } else {
            for (int a = 0; a < 3; ++a) {
                const int nf = G.N / G.L[a];
                const int off = part_slot(G, a, 0);
                double sa = 0.0;
                for (int f = 0; f < nf; ++f) sa += p[off + f];
                up += sa;
            }
        }
    }
    ++S.since;
    const bool rs = S.test && up > 0.0 && S.since >= S.last;
    if (rs) {
        S.last = S.since;
        S.since = 0;
        S.t = 1.0;
    }
    S.wv_is_y = rs ? 1 : 0;
    ++S.it;
    const double tn = fista_t(S.t);
    S.beta = (S.t - 1.0) / tn;
    S.t = tn;
    S.test = S.it % kCheck == 0;
    const bool check = S.na == 2 ? (S.it == S.ck_next || S.it == S.max_iter)
                                 : (S.it % kCheck == 0 || S.it == S.max_iter);
    if (check) {
        S.phase = kPhGap;
        S.ctx = kCtxLoop;
        S.primal = S.na == 3;
    }
}

// The gap of z (or of zp) per block: sum_a lam_a TV_a - <z, d - z>
// (tv_aniso.cpp:191-219).
__global__ __launch_bounds__(kThr) void gap_kernel(ElemArgs A) {
    __shared__ double sh[kThr];
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (!(S.phase == kPhGap || (S.phase == kPhFinal && S.gap_zp))) return;
    const Geo &G = A.G;
    const int64_t N = G.N;
    const double *z = (S.phase == kPhFinal ? A.b.zp : A.b.z) + (int64_t)img * N;
    const double *d = A.b.d + (int64_t)img * N;
    const int lo = blockIdx.x * kElem;
    double acc = 0.0;
    for (int i = lo + threadIdx.x; i < lo + kElem && i < N; i += kThr) {
        int c[3];
        coords(G, i, c);
        const double zi = z[i];
        double s = 0.0;
        for (int a = 0; a < G.k; ++a) {
            const int ax = 3 - G.k + a;
            if (S.lam[a] > 0.0 && c[ax] + 1 < G.L[ax]) s += S.lam[a] * fabs(z[i + G.S[ax]] - zi);
        }
        acc += s - zi * (d[i] - zi);
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) A.b.part[(int64_t)img * A.b.pp + blockIdx.x] = acc;
}

// A gap check (tv_aniso.cpp:459-468 polish, 567-578 check): whether a polish
// is due.
__global__ void gap_ctrl_kernel(ElemArgs A, int nblk) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhGap) return;
    const double *p = A.b.part + (int64_t)img * A.b.pp;
    double gz = 0.0;
    for (int j = 0; j < nblk; ++j) gz += p[j];
    S.primal = 0;
    if (S.na == 2) {
        const bool last = S.it == S.max_iter;
        const bool due = !(gz <= S.target) && (last || ((S.warm || S.it >= kCheck) && gz <= S.sched_next));
        if (due) {
            S.accept = last ? gz : S.target;
            S.gz = gz;
            S.need_dual = S.it > 0;
            S.phase = kPhPolish;
            return;
        }
        S.gp = gz;
        after_check(S);
        return;
    }
    if (S.ctx == kCtxAfterGap) {  // the primal of a gap that was never finite
        S.gp = gz;
        after_check(S);
        return;
    }
    if (!(gz <= S.target) && gz <= S.sched_next) {
        S.accept = S.target;
        S.gz = gz;
        S.need_dual = 1;
        S.phase = kPhPolish;
        return;
    }
    S.gp = gz;
    after_check(S);
}

// The dual of the iterate (tv_aniso.cpp:452-459, 556-563): ADBPG q1 = d + y
// - z, q2 = -y and the point d - q1 - q2 (into z before a polish, into zp
// at the end); FDPG q_0 = -y_0, q_1 = -y_1, q_2 = d + y_0 + y_1 - z.
__global__ __launch_bounds__(kThr) void dual_kernel(ElemArgs A) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    const bool final2 = S.phase == kPhFinal && S.gap_zp;
    if (!((S.phase == kPhPolish || S.phase == kPhOut) && S.need_dual) && !final2) return;
    const Geo &G = A.G;
    const int64_t N = G.N, P = A.b.P;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const double *d = A.b.d + (int64_t)img * N;
    const double *y = A.b.y + (int64_t)img * P * N;
    double *q = A.b.q + (int64_t)img * G.k * N;
    if (S.na == 2) {
        const double zi = A.b.z[(int64_t)img * N + i];
        const double q1 = d[i] + y[i] - zi;
        const double q2 = -y[i];
        q[(int64_t)S.a1 * N + i] = q1;
        q[(int64_t)S.a2 * N + i] = q2;
        double *dst = (final2 ? A.b.zp : A.b.z) + (int64_t)img * N;
        dst[i] = d[i] - q1 - q2;
    } else {
        const double zi = A.b.z[(int64_t)img * N + i];
        q[i] = -y[i];
        q[N + i] = -y[N + i];
        q[2 * N + i] = d[i] + y[i] + y[N + i] - zi;
    }
}

// The end of an ADBPG solve (tv_aniso.cpp:521-529): d - q1 - q2 in place of
// z unless rounding alone takes its gap over the target.
__global__ void final_ctrl_kernel(ElemArgs A, int nblk) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhFinal) return;
    const double *p = A.b.part + (int64_t)img * A.b.pp;
    double gw = 0.0;
    for (int j = 0; j < nblk; ++j) gw += p[j];
    if (gw <= S.target || !(S.gp <= S.target)) {
        S.take_zp = 1;
        S.gp = gw;
    }
    S.gap_zp = 0;
    S.phase = kPhOut;
}

// z = zp for the images whose candidate was taken.
__global__ __launch_bounds__(kThr) void take_kernel(ElemArgs A) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (!S.take_zp) return;
    const int64_t N = A.G.N;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    A.b.z[(int64_t)img * N + i] = A.b.zp[(int64_t)img * N + i];
}

__global__ void untake_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    A.b.st[img].take_zp = 0;
}

// ---------------------------------------------------------------------------
// The polish (tv_aniso.cpp:686-727 mark_edges, :740-880 components, :886-953
// try_polish)
// ---------------------------------------------------------------------------

// zp = d and no flags, for the images polishing.
__global__ __launch_bounds__(kThr) void polish_init_kernel(ElemArgs A) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    const int64_t BN = (int64_t)A.G.B * A.G.N;
    if (gp >= BN) return;
    const int img = (int)((uint64_t)gp / (uint64_t)A.G.N);
    A.b.par[gp] = (int)gp;
    if (A.b.st[img].phase != kPhPolish) return;
    A.b.zp[gp] = A.b.d[gp];
    A.b.flag[gp] = 0;
}

// Along each fibre of axis a: the 1D dual p_e = -(running sum of q_a), a
// flag where |p_e| < lam (1 - eta), else lam sign(p_e) into zp at i and out
// of it at j.
__global__ __launch_bounds__(kFib) void mark_kernel(ElemArgs A, int a) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhPolish || !(S.lam[a] > 0.0)) return;
    const Geo &G = A.G;
    const int ax = 3 - G.k + a;
    const int L = G.L[ax], s = G.S[ax];
    const int f = blockIdx.x * kFib + threadIdx.x;
    if (f >= fibres_of(G, ax)) return;
    const int b0 = fibre_base(G, ax, f);
    const int64_t base = (int64_t)img * G.N;
    const double *qa = A.b.q + ((int64_t)img * G.k + a) * G.N;
    uint8_t *flag = A.b.flag + base;
    double *v = A.b.zp + base;
    const double lb = S.lam[a];
    const uint8_t bit = (uint8_t)(1u << a);
    double acc = 0.0;
    for (int l = 0; l + 1 < L; ++l) {
        const int o = b0 + l * s;
        acc += qa[o];
        const double pe = -acc;
        const bool inside = fabs(pe) < lb * (1.0 - kPolishEta);
        const double ls = inside ? 0.0 : copysign(lb, pe);
        if (inside) flag[o] |= bit;
        v[o] += ls;
        v[o + s] -= ls;
    }
}

// Joins the flagged edges of the images polishing.
__global__ __launch_bounds__(kThr) void polish_union_kernel(ElemArgs A) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    const Geo &G = A.G;
    if (gp >= (int64_t)G.B * G.N) return;
    const int img = (int)((uint64_t)gp / (uint64_t)G.N);
    if (A.b.st[img].phase != kPhPolish) return;
    const uint8_t f = A.b.flag[gp];
    if (!f) return;
    for (int a = 0; a < G.k; ++a)
        if ((f >> a) & 1) uf_unite(A.b.par, (int)gp, (int)(gp + G.S[3 - G.k + a]));
}

__global__ __launch_bounds__(kThr) void polish_compress_kernel(int64_t BN, int *par, int *isroot) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= BN) return;
    const int r = uf_find(par, (int)gp);
    isroot[gp] = r == (int)gp ? 1 : 0;
}

__global__ __launch_bounds__(kThr) void polish_label_kernel(int64_t BN, const int *par, const int *rank, int *label,
                                                            int *iota) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (gp >= BN) return;
    label[gp] = rank[uf_find(const_cast<int *>(par), (int)gp)];
    iota[gp] = (int)gp;
}

__global__ __launch_bounds__(kThr) void polish_rstart_kernel(int64_t BN, const int *skey, int *rstart) {
    const int64_t m = (int64_t)blockIdx.x * kThr + threadIdx.x;
    if (m >= BN) return;
    if (m == 0 || skey[m] != skey[m - 1]) rstart[skey[m]] = (int)m;
    if (m == BN - 1) rstart[skey[m] + 1] = (int)BN;
}

struct FVal {  // the polish's adjusted values
    const double *v;
    __device__ double operator()(int p, int) const { return v[p]; }
};

// The closed-form candidate: zp = the region's mean of d plus the boundary
// terms (tv_aniso.cpp:929-931).
__global__ __launch_bounds__(kThr) void polish_mean_kernel(ElemArgs A, double *zpn) {
    const int64_t gp = (int64_t)blockIdx.x * kThr + threadIdx.x;
    const Geo &G = A.G;
    if (gp >= (int64_t)G.B * G.N) return;
    const int img = (int)((uint64_t)gp / (uint64_t)G.N);
    if (A.b.st[img].phase != kPhPolish) return;
    const int C = A.b.label[gp];
    const double s = region_value(C, A.b.rsum, A.b.rstart, A.b.rhead, A.b.rtail);
    zpn[gp] = s / (double)(A.b.rstart[C + 1] - A.b.rstart[C]);
}

// The candidate's gap against the dual: P(zp) - D(q) (tv_aniso.cpp:933-946).
__global__ __launch_bounds__(kThr) void pz_kernel(ElemArgs A, const double *zpn) {
    __shared__ double sh[kThr];
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhPolish) return;
    const Geo &G = A.G;
    const int64_t N = G.N;
    const double *z = zpn + (int64_t)img * N;
    const double *d = A.b.d + (int64_t)img * N;
    const double *q = A.b.q + (int64_t)img * G.k * N;
    const int lo = blockIdx.x * kElem;
    double acc = 0.0;
    for (int i = lo + threadIdx.x; i < lo + kElem && i < N; i += kThr) {
        int c[3];
        coords(G, i, c);
        const double zi = z[i];
        const double e = zi - d[i];
        double sj = q[i];
        for (int a = 1; a < G.k; ++a) sj += q[(int64_t)a * N + i];
        double t = 0.5 * e * e - (d[i] * sj - 0.5 * sj * sj);
        for (int a = 0; a < G.k; ++a) {
            const int ax = 3 - G.k + a;
            if (S.lam[a] > 0.0 && c[ax] + 1 < G.L[ax]) t += S.lam[a] * fabs(z[i + G.S[ax]] - zi);
        }
        acc += t;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) A.b.part[(int64_t)img * A.b.pp + blockIdx.x] = acc;
}

// Takes the candidate or schedules the next attempt (tv_aniso.cpp:359-372
// PolishSchedule, :947-952).
__global__ void polish_ctrl_kernel(ElemArgs A, int nblk) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhPolish) return;
    const double *p = A.b.part + (int64_t)img * A.b.pp;
    double pz = 0.0;
    for (int j = 0; j < nblk; ++j) pz += p[j];
    S.pz = pz;
    S.need_dual = 0;
    double gz = S.gz;
    if (pz <= S.accept && pz <= kPolishRatio * gz) {
        S.polished = 1;
        S.take_zp = 1;
        gz = pz;
    } else {
        const bool eager = S.na == 3;
        double next = gz * kPolishDrop;
        if (pz > 0.0 && pz < (eager ? 1.0 : kPolishRatio) * gz) {
            const double c = S.target * gz / pz;
            next = next < c ? c : next;
        }
        S.sched_next = next;
    }
    S.gp = gz;
    after_check(S);
}

// ---------------------------------------------------------------------------
// Outputs
// ---------------------------------------------------------------------------

struct OutArgs {
    Geo G;
    Bufs b;
    void *usave;
    int64_t q_plane, u_img;
    double w[3];
    int any_phase;  // the per-edge dual of every image (after the solve), else of those in kPhOut
};

// z and the dual at the input's scale (tv_aniso.cpp:288-299, 339-354).
template <class T>
__global__ __launch_bounds__(kThr) void out_kernel(OutArgs A) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhOut) return;
    const Geo &G = A.G;
    const int64_t N = G.N;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const int64_t gi = (int64_t)img * N + i;
    T *z = (T *)A.b.img[3 * img + 1];
    T *q = (T *)A.b.img[3 * img + 2];
    if (S.status) return;
    if (S.nan) {
        z[i] = (T)__builtin_nan("");
        for (int a = 0; a < G.k; ++a) q[a * A.q_plane + i] = (T)__builtin_nan("");
        return;
    }
    if (S.na == 0) {
        z[i] = ((const T *)A.b.img[3 * img])[i];
        for (int a = 0; a < G.k; ++a) q[a * A.q_plane + i] = (T)0;
        return;
    }
    const double zi = A.b.z[gi];
    z[i] = (T)(S.e ? ldexp(zi, S.e) : zi);
    for (int a = 0; a < G.k; ++a) {
        const double qv = A.b.q[((int64_t)img * G.k + a) * N + i];
        q[a * A.q_plane + i] = (T)(S.e ? ldexp(qv, S.e) : qv);
    }
}

// The per-edge dual u_a = p_a / w_a at the input's scale (0 at the end of a
// fibre and on inactive axes), for the derivatives.
template <class T>
__global__ __launch_bounds__(kFib) void usave_kernel(OutArgs A, int a) {
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if ((!A.any_phase && S.phase != kPhOut) || S.status) return;
    const Geo &G = A.G;
    const int ax = 3 - G.k + a;
    const int L = G.L[ax], s = G.S[ax];
    const int f = blockIdx.x * kFib + threadIdx.x;
    if (f >= fibres_of(G, ax)) return;
    const int b0 = fibre_base(G, ax, f);
    T *u = (T *)A.usave + img * A.u_img + (int64_t)a * G.N;
    const double *qa = A.b.q + ((int64_t)img * G.k + a) * G.N;
    const bool on = !S.nan && S.na > 0 && S.lam[a] > 0.0 && A.w[a] > 0.0;
    double acc = 0.0;
    for (int l = 0; l < L; ++l) {
        const int o = b0 + l * s;
        double v = 0.0;
        if (on && l + 1 < L) {
            acc += qa[o];
            v = -acc;
            if (S.e) v = ldexp(v, S.e);
            v = v / A.w[a];
        }
        u[o] = (T)(S.nan ? __builtin_nan("") : v);
    }
}

// Phase moves after the launches of a phase.
__global__ void exact_done_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhExact) return;
    S.gp = 0.0;
    S.phase = kPhOut;
}

__global__ void warm_done_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhWarm) return;
    S.phase = kPhGap;
    S.ctx = kCtxWarm;
    S.primal = S.na == 3;
}

// After the first block step z(y) (tv_aniso.cpp:505-508): its dual q1 = d +
// y - z, q2 = -y, and z = d - q1 - q2.
__global__ __launch_bounds__(kThr) void init_dual_kernel(ElemArgs A) {
    const int64_t N = A.G.N;
    const int img = blockIdx.y;
    const AdImg &S = A.b.st[img];
    if (S.phase != kPhInit) return;
    const int i = blockIdx.x * kThr + threadIdx.x;
    if (i >= N) return;
    const double *d = A.b.d + (int64_t)img * N;
    const double *y = A.b.y + (int64_t)img * N;
    double *z = A.b.z + (int64_t)img * N;
    double *q = A.b.q + (int64_t)img * A.G.k * N;
    const double q1 = d[i] + y[i] - z[i], q2 = -y[i];
    q[(int64_t)S.a1 * N + i] = q1;
    q[(int64_t)S.a2 * N + i] = q2;
    z[i] = d[i] - q1 - q2;
}

__global__ void init_done_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase != kPhInit) return;
    S.phase = kPhGap;
    S.ctx = kCtxInit;
}

// The certified error |z - z*| / |x| <= sqrt(2 gap) / |x| plus the output's
// rounding (driver.hpp:1186-1212), on the solved scale.
__global__ void result_kernel(ElemArgs A, double u_round, double eta) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.nan) {
        S.zerr = __builtin_nan("");
        return;
    }
    const double gp = S.na < 2 ? 0.0 : (S.gp < 0.0 ? 0.0 : S.gp);
    const double e_ = sqrt(2.0 * gp);
    const double a = ldexp(eta, -S.e) * sqrt((double)A.G.N);
    if (S.s2 >= 1e-280 && S.s2 <= 1e280) {
        const double n = sqrt(S.s2);
        S.zerr = e_ / n + (u_round * (1.0 + e_ / n) + a / n);
    } else if (!(S.maxabs > 0.0)) {
        S.zerr = 0.0;
    } else {
        const double rt = sqrt(S.s2m);
        const double r = e_ / S.maxabs / rt;
        S.zerr = r + (u_round * (1.0 + r) + a / S.maxabs / rt);
    }
}

__global__ void done_kernel(ElemArgs A) {
    const int img = blockIdx.x * blockDim.x + threadIdx.x;
    if (img >= A.G.B) return;
    AdImg &S = A.b.st[img];
    if (S.phase == kPhOut) S.phase = kPhDone;
}

// The images in each phase.
__global__ void count_kernel(ElemArgs A) {
    __shared__ int c[kPhCount];
    if (threadIdx.x < kPhCount) c[threadIdx.x] = 0;
    __syncthreads();
    for (int i = threadIdx.x; i < A.G.B; i += blockDim.x) atomicAdd(&c[A.b.st[i].phase], 1);
    __syncthreads();
    if (threadIdx.x < kPhCount) A.b.phase_count[threadIdx.x] = c[threadIdx.x];
}

__global__ void zero_ints_kernel(int *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

int blocks(int64_t n, int per) { return (int)((n + per - 1) / per); }

}  // namespace

struct AdSolver {
    AdSpec sp;
    Geo G{};
    Bufs b{};
    void *scan_tmp = nullptr, *sort_tmp = nullptr;
    size_t scan_bytes = 0, sort_bytes = 0;
    int sort_bits = 1;
    int nch = 0;  // region-sum blocks
    double *zpn = nullptr;
    double *gam = nullptr;
    double *hgam = nullptr;   // pinned
    AdImg *hst = nullptr;     // pinned
    int *hcount = nullptr;    // pinned
    void **himg = nullptr;    // pinned: per image x, z, q
    std::vector<AdResult> res;
    int launches = 0, syncs = 0, fallbacks = 0;
    std::vector<void *> owned;
    ~AdSolver() {
        for (void *p : owned) (void)hipFree(p);
        if (hgam) (void)hipHostFree(hgam);
        if (hst) (void)hipHostFree(hst);
        if (hcount) (void)hipHostFree(hcount);
        if (himg) (void)hipHostFree(himg);
    }
    template <class P>
    int dev(P **p, size_t bytes) {
        void *q = nullptr;
        if (dev_alloc_bytes(&q, bytes < 16 ? 16 : bytes)) return FP_ENOMEM;
        owned.push_back(q);
        *p = (P *)q;
        return FP_OK;
    }
};

int ad_reserve(const AdSpec &spec, AdSolver **out) {
    *out = nullptr;
    if (spec.k != 2 && spec.k != 3) return FP_EINVAL;
    const int64_t N = spec.L[0] * spec.L[1] * spec.L[2];
    const int64_t BN = N * spec.B;
    if (spec.B < 1 || N < 1 || BN >= ((int64_t)1 << 31) - 1) return FP_ESHAPE;
    std::unique_ptr<AdSolver> s(new AdSolver());
    s->sp = spec;
    Geo &G = s->G;
    G.k = spec.k;
    G.B = (int)spec.B;
    G.N = (int)N;
    for (int a = 0; a < 3; ++a) G.L[a] = (int)spec.L[a];
    G.S[2] = 1;
    G.S[1] = G.L[2];
    G.S[0] = G.L[1] * G.L[2];
    Bufs &b = s->b;
    b.P = spec.k == 3 ? 3 : 1;
    const size_t nd = (size_t)BN * sizeof(double);
    int r = 0;
    if ((r = s->dev(&b.d, nd)) || (r = s->dev(&b.u, nd)) || (r = s->dev(&b.z, nd)) || (r = s->dev(&b.zp, nd)) ||
        (r = s->dev(&b.y, nd * b.P)) || (r = s->dev(&b.wv, nd * b.P)) || (r = s->dev(&b.q, nd * spec.k)) ||
        (r = s->dev(&s->zpn, nd)))
        return r;
    if (spec.k == 3 && (r = s->dev(&b.tmp, nd * 3))) return r;
    if ((r = s->dev(&b.flag, (size_t)BN)) || (r = s->dev(&b.par, 4 * (size_t)BN)) ||
        (r = s->dev(&b.isroot, 4 * (size_t)BN)) || (r = s->dev(&b.rank, 4 * (size_t)BN)) ||
        (r = s->dev(&b.label, 4 * (size_t)BN)) || (r = s->dev(&b.iota, 4 * (size_t)BN)) ||
        (r = s->dev(&b.skey, 4 * (size_t)BN)) || (r = s->dev(&b.members, 4 * (size_t)BN)) ||
        (r = s->dev(&b.rstart, 4 * ((size_t)BN + 1))) || (r = s->dev(&b.rsum, nd)))
        return r;
    s->nch = (int)((BN + kRsChunk - 1) / kRsChunk);
    if ((r = s->dev(&b.rhead, (size_t)s->nch * 8)) || (r = s->dev(&b.rtail, (size_t)s->nch * 8))) return r;
    // partial slots per image: the per-block sums, the restart sums of every
    // fibre (three axes: of all three)
    int64_t pp = 2 * blocks(N, kElem);
    int64_t fib = 0;
    for (int a = 0; a < spec.k; ++a) fib += N / spec.L[3 - spec.k + a];
    pp = std::max(pp, fib);
    b.pp = (int)pp;
    if ((r = s->dev(&b.part, (size_t)pp * spec.B * 8))) return r;
    int64_t maxfib = 0, Lmax = 1;
    for (int a = 0; a < spec.k; ++a) {
        maxfib = std::max(maxfib, N / spec.L[3 - spec.k + a]);
        Lmax = std::max(Lmax, spec.L[3 - spec.k + a]);
    }
    const int64_t q_entries = maxfib * spec.B * (spec.k == 3 ? 3 : 1);
    if ((r = s->dev(&b.queue, (size_t)q_entries * 3 * 4)) || (r = s->dev(&b.ctr, 8 * 4))) return r;
    b.nslots = (int)std::min<int64_t>(kMaxSlots, q_entries);
    b.segstride = (int)(Lmax + 1);
    if ((r = s->dev(&b.segs, (size_t)b.nslots * 2 * b.segstride * sizeof(Seg)))) return r;
    if ((r = s->dev(&b.st, sizeof(AdImg) * spec.B)) || (r = s->dev(&b.phase_count, 4 * kPhCount)) ||
        (r = s->dev(&s->gam, 8 * spec.B)) || (r = s->dev(&b.img, 3 * sizeof(void *) * spec.B)))
        return r;
    if (pinned_alloc(&s->hgam, 8 * spec.B) || pinned_alloc(&s->hst, sizeof(AdImg) * spec.B) ||
        pinned_alloc(&s->hcount, 4 * kPhCount) || pinned_alloc(&s->himg, 3 * sizeof(void *) * spec.B))
        return FP_ENOMEM;
    // hipcub temporaries at the largest size
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, s->scan_bytes, b.isroot, b.rank, (int)BN)) return FP_EDEVICE;
    while (((int64_t)1 << s->sort_bits) <= BN) ++s->sort_bits;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, s->sort_bytes, b.label, b.skey, b.iota, b.members, (int)BN, 0,
                                           s->sort_bits))
        return FP_EDEVICE;
    if ((r = s->dev(&s->scan_tmp, s->scan_bytes)) || (r = s->dev(&s->sort_tmp, s->sort_bytes))) return r;
    s->res.assign((size_t)spec.B, AdResult{});
    *out = s.release();
    return FP_OK;
}

void ad_free(AdSolver *s) { delete s; }

const AdResult *ad_results(const AdSolver *s) { return s->res.data(); }

void ad_stats(const AdSolver *s, int *launches, int *host_syncs, int *fallbacks) {
    if (launches) *launches = s->launches;
    if (host_syncs) *host_syncs = s->syncs;
    if (fallbacks) *fallbacks = s->fallbacks;
}

namespace {

template <int MODE>
int launch_pass(AdSolver &s, hipStream_t st, int phase, int a0, int a1, int a2, int nz, int which, int qslot) {
    const Geo &G = s.G;
    PassArgs A;
    A.G = G;
    A.b = s.b;
    A.phase = phase;
    A.axis[0] = a0;
    A.axis[1] = a1;
    A.axis[2] = a2;
    A.which = which;
    A.qslot = qslot;
    // the grid covers the most fibres of any axis
    int maxf = 0;
    for (int a = 0; a < G.k; ++a) maxf = std::max(maxf, G.N / G.L[3 - G.k + a]);
    hipLaunchKernelGGL(pass_kernel<MODE>, dim3(blocks(maxf, kFib), G.B, nz), dim3(kFib), 0, st, A);
    hipLaunchKernelGGL(fallback_kernel<MODE>, dim3(blocks(s.b.nslots, kThr)), dim3(kThr), 0, st, A);
    s.launches += 2;
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

template <class T>
int solve_t(AdSolver &s, hipStream_t st, const AdCall &c) {
    const Geo &G = s.G;
    const int64_t N = G.N, BN = (int64_t)G.B * N;
    const int nblk = blocks(N, kElem);
    const int nelem = blocks(N, kThr);
    const int img_blocks = blocks(G.B, 64);
    for (int i = 0; i < G.B; ++i) {
        s.hgam[i] = c.gammas[i];
        s.himg[3 * i] = const_cast<void *>(c.xs[i]);
        s.himg[3 * i + 1] = c.zs[i];
        s.himg[3 * i + 2] = c.qs[i];
    }
    if (hipMemcpyAsync(s.gam, s.hgam, 8 * (size_t)G.B, hipMemcpyHostToDevice, st)) return FP_EDEVICE;
    if (hipMemcpyAsync(s.b.img, s.himg, 3 * sizeof(void *) * (size_t)G.B, hipMemcpyHostToDevice, st))
        return FP_EDEVICE;
    hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, st, s.b.ctr, 8);
    SetupArgs SA;
    SA.G = G;
    SA.b = s.b;
    SA.q_plane = c.q_plane;
    SA.gam = s.gam;
    for (int a = 0; a < 3; ++a) SA.w[a] = a < G.k ? s.sp.w[a] : 0.0;
    SA.tol = c.tol;
    SA.max_iter = c.max_iter;
    SA.warm = c.warm ? 1 : 0;
    SA.nblk = nblk;
    hipLaunchKernelGGL(scan_kernel<T>, dim3(nblk, G.B), dim3(kThr), 0, st, SA);
    hipLaunchKernelGGL(setup_kernel, dim3(img_blocks), dim3(64), 0, st, SA);
    hipLaunchKernelGGL(load_kernel<T>, dim3(nblk, G.B), dim3(kThr), 0, st, SA);
    hipLaunchKernelGGL(init_kernel, dim3(img_blocks), dim3(64), 0, st, SA);
    s.launches += 5;
    ElemArgs E;
    E.G = G;
    E.b = s.b;
    E.phase = 0;
    auto count = [&]() -> int {
        hipLaunchKernelGGL(count_kernel, dim3(1), dim3(256), 0, st, E);
        ++s.launches;
        if (hipMemcpyAsync(s.hcount, s.b.phase_count, 4 * kPhCount, hipMemcpyDeviceToHost, st)) return FP_EDEVICE;
        if (hipStreamSynchronize(st)) return FP_EDEVICE;
        ++s.syncs;
        return FP_OK;
    };
    auto fib_grid = [&](int a) { return dim3(blocks(G.N / G.L[3 - G.k + a], kFib), G.B); };
    int r = count();
    if (r) return r;
    // the one-axis exact passes, and warm starts' first points
    if (s.hcount[kPhExact]) {
        // the axis is the image's only active one: a1 = a2 = act[0]
        if ((r = launch_pass<kPassExact>(s, st, kPhExact, -1, -1, -1, 1, 2, 4))) return r;
        hipLaunchKernelGGL(exact_done_kernel, dim3(img_blocks), dim3(64), 0, st, E);
        ++s.launches;
    }
    if (s.hcount[kPhWarm]) {
        for (int a = 0; a < G.k; ++a) hipLaunchKernelGGL(feasible_kernel, fib_grid(a), dim3(kFib), 0, st, E, a);
        hipLaunchKernelGGL(warm_point_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
        hipLaunchKernelGGL(warm_done_kernel, dim3(img_blocks), dim3(64), 0, st, E);
        s.launches += G.k + 2;
    }
    const int hard_cap = 4 * (c.max_iter + 8);
    for (int round = 0;; ++round) {
        if (round > hard_cap) return FP_EDEVICE;  // the control never ends (cannot happen)
        if ((r = count())) return r;
        const int *h = s.hcount;
        const int live = h[kPhIter] + h[kPhGap] + h[kPhPolish] + h[kPhFinal] + h[kPhInit];
        if (!live) break;
        if (h[kPhInit]) {
            if ((r = launch_pass<kPassInit>(s, st, kPhInit, -1, -1, -1, 1, 1, 4))) return r;
            hipLaunchKernelGGL(init_dual_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(init_done_kernel, dim3(img_blocks), dim3(64), 0, st, E);
            s.launches += 2;
        }
        if (h[kPhIter]) {
            if (G.k == 3) {
                hipLaunchKernelGGL(fdpg_u_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
                ++s.launches;
                if ((r = launch_pass<kPass3>(s, st, kPhIter, 0, 1, 2, 3, 0, 4))) return r;
            }
            if ((r = launch_pass<kPassA>(s, st, kPhIter, -1, -1, -1, 1, 1, 4))) return r;
            if ((r = launch_pass<kPassB>(s, st, kPhIter, -1, -1, -1, 1, 2, 4))) return r;
            hipLaunchKernelGGL(iter_ctrl_kernel, dim3(img_blocks), dim3(64), 0, st, E);
            ++s.launches;
        }
        // gap checks (three axes: after the primal pass)
        if (G.k == 3 && (r = launch_pass<kPassPrimal3>(s, st, kPhGap, 2, -1, -1, 1, 0, 4))) return r;
        hipLaunchKernelGGL(gap_kernel, dim3(nblk, G.B), dim3(kThr), 0, st, E);
        hipLaunchKernelGGL(gap_ctrl_kernel, dim3(img_blocks), dim3(64), 0, st, E, nblk);
        s.launches += 2;
        if ((r = count())) return r;
        if (h[kPhPolish]) {
            hipLaunchKernelGGL(dual_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(polish_init_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, E);
            for (int a = 0; a < G.k; ++a) hipLaunchKernelGGL(mark_kernel, fib_grid(a), dim3(kFib), 0, st, E, a);
            hipLaunchKernelGGL(polish_union_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(polish_compress_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, BN, s.b.par,
                               s.b.isroot);
            size_t b1 = s.scan_bytes;
            if (hipcub::DeviceScan::ExclusiveSum(s.scan_tmp, b1, s.b.isroot, s.b.rank, (int)BN, st)) return FP_EDEVICE;
            hipLaunchKernelGGL(polish_label_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, BN, s.b.par, s.b.rank,
                               s.b.label, s.b.iota);
            size_t b2 = s.sort_bytes;
            if (hipcub::DeviceRadixSort::SortPairs(s.sort_tmp, b2, s.b.label, s.b.skey, s.b.iota, s.b.members, (int)BN,
                                                   0, s.sort_bits, st))
                return FP_EDEVICE;
            hipLaunchKernelGGL(polish_rstart_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, BN, s.b.skey,
                               s.b.rstart);
            hipLaunchKernelGGL(rsum_kernel<FVal>, dim3(s.nch), dim3(kRsThr), 0, st, FVal{s.b.zp}, s.b.skey,
                               s.b.members, s.b.rstart, BN, s.b.rsum, s.b.rhead, s.b.rtail);
            hipLaunchKernelGGL(polish_mean_kernel, dim3(blocks(BN, kThr)), dim3(kThr), 0, st, E, s.zpn);
            hipLaunchKernelGGL(pz_kernel, dim3(nblk, G.B), dim3(kThr), 0, st, E, (const double *)s.zpn);
            hipLaunchKernelGGL(polish_ctrl_kernel, dim3(img_blocks), dim3(64), 0, st, E, nblk);
            s.launches += 12 + G.k;
            // a taken candidate: z = the candidate
            Bufs bz = s.b;
            bz.zp = s.zpn;
            ElemArgs Ez = E;
            Ez.b = bz;
            hipLaunchKernelGGL(take_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, Ez);
            hipLaunchKernelGGL(untake_kernel, dim3(img_blocks), dim3(64), 0, st, E);
            s.launches += 2;
        }
        {
            // ADBPG's end: the dual's point d - q1 - q2 into zp and its gap
            hipLaunchKernelGGL(dual_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(gap_kernel, dim3(nblk, G.B), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(final_ctrl_kernel, dim3(img_blocks), dim3(64), 0, st, E, nblk);
            hipLaunchKernelGGL(take_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
            hipLaunchKernelGGL(untake_kernel, dim3(img_blocks), dim3(64), 0, st, E);
            s.launches += 5;
        }
        if (hipGetLastError()) return FP_EDEVICE;
    }
    // outputs: the FDPG dual first where it is due
    hipLaunchKernelGGL(dual_kernel, dim3(nelem, G.B), dim3(kThr), 0, st, E);
    OutArgs O;
    O.G = G;
    O.b = s.b;
    O.usave = c.usave;
    O.any_phase = 0;
    O.q_plane = c.q_plane;
    O.u_img = c.u_img;
    for (int a = 0; a < 3; ++a) O.w[a] = a < G.k ? s.sp.w[a] : 0.0;
    hipLaunchKernelGGL(out_kernel<T>, dim3(nelem, G.B), dim3(kThr), 0, st, O);
    s.launches += 2;
    if (c.usave) {
        for (int a = 0; a < G.k; ++a) hipLaunchKernelGGL(usave_kernel<T>, fib_grid(a), dim3(kFib), 0, st, O, a);
        s.launches += G.k;
    }
    const double u_round = 0.5 * (sizeof(T) == 4 ? (double)FLT_EPSILON : DBL_EPSILON);
    const double eta = 0.5 * (sizeof(T) == 4 ? (double)FLT_TRUE_MIN : DBL_TRUE_MIN);
    hipLaunchKernelGGL(result_kernel, dim3(img_blocks), dim3(64), 0, st, E, u_round, eta);
    if (hipMemcpyAsync(s.hst, s.b.st, sizeof(AdImg) * (size_t)G.B, hipMemcpyDeviceToHost, st)) return FP_EDEVICE;
    hipLaunchKernelGGL(done_kernel, dim3(img_blocks), dim3(64), 0, st, E);
    int hctr[8] = {0};
    if (hipMemcpyAsync(hctr, s.b.ctr, sizeof(hctr), hipMemcpyDeviceToHost, st)) return FP_EDEVICE;
    if (hipStreamSynchronize(st) || hipGetLastError()) return FP_EDEVICE;
    ++s.syncs;
    s.fallbacks = hctr[2];
    for (int i = 0; i < G.B; ++i) {
        const AdImg &S = s.hst[i];
        AdResult &R = s.res[(size_t)i];
        double g = S.gp;
        if (S.nan) g = NAN;
        else if (S.na < 2) g = 0.0;
        else g = g < 0.0 ? 0.0 : g;
        if (S.e && g == g) g = std::ldexp(g, 2 * S.e);
        R.gap = g;
        R.iters = S.na >= 2 ? S.it : 0;
        R.converged = (S.nan || S.na < 2) ? 1 : (S.gp <= S.target ? 1 : 0);
        R.status = S.status;
        R.e = S.e;
        R.zerr = S.zerr;
        R.polished = S.polished;
    }
    return FP_OK;
}

}  // namespace

template <class T>
int copy_edge_dual_t(AdSolver &s, hipStream_t st, void *usave, int64_t u_img) {
    const Geo &G = s.G;
    OutArgs O;
    O.G = G;
    O.b = s.b;
    O.usave = usave;
    O.q_plane = 0;
    O.u_img = u_img;
    for (int a = 0; a < 3; ++a) O.w[a] = a < G.k ? s.sp.w[a] : 0.0;
    O.any_phase = 1;
    for (int a = 0; a < G.k; ++a)
        hipLaunchKernelGGL(usave_kernel<T>, dim3(blocks(G.N / G.L[3 - G.k + a], kFib), G.B), dim3(kFib), 0, st, O,
                           a);
    return hipGetLastError() ? FP_EDEVICE : FP_OK;
}

int ad_copy_edge_dual(AdSolver *s, hipStream_t st, void *usave, int64_t u_img) {
    return s->sp.tsize == 4 ? copy_edge_dual_t<float>(*s, st, usave, u_img)
                            : copy_edge_dual_t<double>(*s, st, usave, u_img);
}

int ad_solve(AdSolver *s, hipStream_t st, const AdCall &c) {
    s->launches = s->syncs = s->fallbacks = 0;
    return s->sp.tsize == 4 ? solve_t<float>(*s, st, c) : solve_t<double>(*s, st, c);
}

}  // namespace fpr
