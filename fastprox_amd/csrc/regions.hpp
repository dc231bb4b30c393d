// Connected regions of pixel graphs and fixed-order sums over them, shared by
// the derivative structure (deriv.hip) and the anisotropic polish
// (adbpg.hip): union-find with atomic hooks of the larger root under the
// smaller, so that a component's smallest index ends as its root whatever
// the order of the hooks, and per-region sums over members sorted by
// (region, index) in blocks of kRsChunk positions, every sum in a fixed order
// (no floating-point atomics), so that results repeat bit for bit.
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fpr {
namespace {

constexpr int kRsThr = 256;
constexpr int kRsPer = 16;
constexpr int kRsChunk = kRsThr * kRsPer;  // member positions per region-sum block

__device__ int uf_find(int *par, int x) {
    while (true) {
        const int p = __atomic_load_n(&par[x], __ATOMIC_RELAXED);
        if (p == x) return x;
        const int q = __atomic_load_n(&par[p], __ATOMIC_RELAXED);
        if (q == p) return p;
        __atomic_store_n(&par[x], q, __ATOMIC_RELAXED);  // path halving: q is an ancestor of x
        x = q;
    }
}

// Hooks the larger root under the smaller until a and b share a root.
__device__ void uf_unite(int *par, int a, int b) {
    while (true) {
        a = uf_find(par, a);
        b = uf_find(par, b);
        if (a == b) return;
        if (a > b) {
            const int t = a;
            a = b;
            b = t;
        }
        if (atomicCAS(&par[b], b, a) == b) return;
    }
}

// out[C] = the sum of f(p, C) over C's members for every region inside one
// block of kRsChunk member positions; head[j] / tail[j]: the parts of the
// regions that cross into block j from before / out of it after.
// The block's keys and member values are first loaded into LDS with one
// member per thread per pass (coalesced), then each thread walks kRsPer
// consecutive positions from LDS in order; rows of kRsPer + 1 entries keep
// the walk's LDS accesses on distinct banks. The sums are the same, term for
// term and in the same order, as a walk over global memory.
// BYPOS: f takes the member's position (values already in member order).
template <class F, bool BYPOS = false>
__global__ __launch_bounds__(kRsThr) void rsum_kernel(F f, const int *skey, const int *members, const int *rstart,
                                                    int64_t M, double *out, double *head, double *tail) {
    constexpr int kRow = kRsPer + 1;
    __shared__ double lv[kRsThr * kRow];
    __shared__ int lk[kRsThr * kRow];
    __shared__ double sv[kRsThr];
    __shared__ int sf[kRsThr];
    const int64_t base = (int64_t)blockIdx.x * kRsChunk;
    const int t = threadIdx.x;
    const int64_t cend = base + kRsChunk < M ? base + kRsChunk : M;
    const int n = (int)(cend - base);
    for (int i = t; i < n; i += kRsThr) {
        const int C = skey[base + i];
        const int j = (i / kRsPer) * kRow + i % kRsPer;
        lk[j] = C;
        lv[j] = BYPOS ? f((int)(base + i), C) : f(members[base + i], C);
    }
    __syncthreads();
    const int64_t m0 = base + (int64_t)t * kRsPer;
    const int64_t m1 = m0 + kRsPer < M ? m0 + kRsPer : M;
    double open = 0.0;
    int reset = 1;
    int has_head = 0, head_key = -1, last_key = -1, last_after = 0;
    double head_sum = 0.0;
    if (m0 < m1) {
        const int *rk = lk + t * kRow;
        const double *rv = lv + t * kRow;
        const int kprev = m0 > 0 ? (t > 0 ? lk[(t - 1) * kRow + kRsPer - 1] : skey[m0 - 1]) : -1;
        const int knext = m1 < M ? (m1 < cend ? lk[(t + 1) * kRow] : skey[m1]) : -1;
        const int cnt = (int)(m1 - m0);
        int cur = rk[0];
        double s = 0.0;
        bool first = true;
        for (int i = 0; i < cnt; ++i) {
            const int C = rk[i];
            if (C != cur) {
                if (first && cur == kprev) {
                    has_head = 1;
                    head_sum = s;
                    head_key = cur;
                } else {
                    out[cur] = s;
                }
                first = false;
                cur = C;
                s = 0.0;
            }
            s += rv[i];
        }
        const bool before = first && cur == kprev;
        const bool after = cur == knext;
        if (!before && !after) {
            out[cur] = s;
        } else if (before && !after) {
            has_head = 1;
            head_sum = s;
            head_key = cur;
        } else {
            open = s;
            reset = before ? 0 : 1;
        }
        last_key = cur;
        last_after = after ? 1 : 0;
    }
    // Inclusive segmented scan of the open runs over the threads, in a fixed
    // tree: sv[t] = the sum of the run open at thread t's end, from its start
    // (or the block's).
    sv[t] = open;
    sf[t] = reset;
    __syncthreads();
    for (int d = 1; d < kRsThr; d <<= 1) {
        double v = sv[t];
        int fl = sf[t];
        if (t >= d && !fl) {
            v = sv[t - d] + v;
            fl = sf[t - d];
        }
        __syncthreads();
        sv[t] = v;
        sf[t] = fl;
        __syncthreads();
    }
    if (has_head) {
        const double tot = (t > 0 ? sv[t - 1] : 0.0) + head_sum;
        if (rstart[head_key] >= base) out[head_key] = tot;
        else head[blockIdx.x] = tot;
    }
    const int tl = (int)((cend - base - 1) / kRsPer);
    if (t == tl && last_after) {
        if (rstart[last_key] >= base) tail[blockIdx.x] = sv[t];
        else head[blockIdx.x] = sv[t];  // the region covers the whole block
    }
}

// A region's sum: its block's, or its pieces added in block order.
__device__ __forceinline__ double region_value(int C, const double *out, const int *rstart, const double *head,
                                               const double *tail) {
    const int64_t j0 = rstart[C] / kRsChunk, j1 = (rstart[C + 1] - 1) / kRsChunk;
    if (j0 == j1) return out[C];
    double t = tail[j0];
    for (int64_t j = j0 + 1; j <= j1; ++j) t += head[j];
    return t;
}

}  // namespace
}  // namespace fpr
