// Host launchers of the image-TV kernels (tv_kernels.hip).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include "tv_common.hpp"

namespace fpr {

constexpr int kMaxStrided = 32;

struct StridedShape {
    int64_t shape[kMaxStrided];
    int64_t stride[kMaxStrided];
};

// The streaming 2D kernel addresses each plane of an image through 32-bit
// byte offsets: it takes images of at most kStreamMaxPlaneBytes per plane
// (2^30 pixels in float32, 2^29 in float64); larger ones take the generic
// kernel.
constexpr int64_t kStreamMaxPlaneBytes = int64_t(1) << 32;
inline bool stream_fits(int tsize, int64_t maxN) { return maxN * (int64_t)tsize <= kStreamMaxPlaneBytes; }

// One image of a grouped gather: n entries of extents shape (3 axes) and
// element strides stride from src into the contiguous dst.
struct GatherEntry {
    const void *src;
    void *dst;
    int64_t n;
    int64_t shape[3];
    int64_t stride[3];
    int32_t contiguous;
    int32_t pad_;
};

// The step kernel of a solve: the register-tile 2D kernel with TT iterations
// per launch on tiles of (64 - 2 TT) x (RY - 2 TT) pixels, or the generic
// one-iteration kernel (2D and 3D).
// stream: the streaming kernel (TT iterations, segments of RY rows); else
// the block kernel (tiles of RY rows in registers, TT iterations).
// s3: the 3D streaming kernels. TT 1: one iteration per launch, tiles of
// 62 x (kS3Waves - 1) voxels by RY slices; TT >= 2 (float32, no pixel
// weights): TT iterations per launch, tiles of (64 - 2 TT) x (kS3TWaves -
// 2 TT) voxels by RY slices.
constexpr int kS3Waves = 8;
constexpr int kS3TWaves = 16;
struct KernelChoice {
    bool generic = false;
    bool stream = true;
    bool s3 = false;
    int TT = 4;
    int RY = 64;
    bool s3t() const { return s3 && TT >= 2; }
    int tile_x() const { return s3 ? (TT >= 2 ? 64 - 2 * TT : 62) : (generic ? 64 : 64 - 2 * TT); }
    int tile_y() const {
        return s3 ? (TT >= 2 ? kS3TWaves - 2 * TT : kS3Waves - 1) : (generic ? 4 : (stream ? RY : RY - 2 * TT));
    }
    int64_t tiles_z(int64_t L0) const { return s3 ? (L0 + RY - 1) / RY : (generic ? L0 : 1); }
    int iters() const { return generic ? 1 : (s3 ? (TT >= 2 ? TT : 1) : TT); }
};

bool kernel_available(int dtype_bytes, bool stream, int TT, int RY);
// The 3D streaming kernel with TT iterations per launch (TT 1: any dtype and
// weights; TT 2 or 3: float32 without pixel weights, volumes of at most
// kStreamMaxPlaneBytes per plane) and slabs of RY slices.
bool kernel_available_s3(int RY, int TT = 1, int dtype_bytes = 8, bool sw = false, int64_t maxN = 0);

// The prologue's per-block partials: 4 doubles per block, prologue_blocks(maxN) blocks per image.
int prologue_blocks(int64_t maxN);
// tols / maxits: per-image tolerance and iteration cap (null: P's).
template <class T>
void launch_prologue(const Problem &P, const ImgDesc *desc, ImgState *state, const double *gammas, int nimg,
                     double *part, int64_t maxN, hipStream_t s, const double *tols = nullptr,
                     const int *maxits = nullptr);
void launch_reset_counts(int *counts, hipStream_t s);
// desc and state may point at image img0 of the arrays the step kernels
// read: the tiles carry img0 + i.
void launch_initial_list(const ImgDesc *desc, const ImgState *state, Tile *list, int *count, int nimg,
                         hipStream_t s, int img0 = 0);
template <class T>
void launch_warm(const Problem &P, T *arena, const ImgDesc *desc, const ImgState *state, int nimg, int64_t maxN,
                 hipStream_t s);
template <class T>
void launch_tv1d(const Problem &P, const ImgDesc *desc, ImgState *state, int nimg, int64_t maxN, double u_round,
                 double eta, hipStream_t s);
// t0 / t1 (optional): events the launch stamps when the kernel starts /
// completes (hipExtLaunchKernel), for timing without separate records.
template <class T>
void launch_step(const KernelChoice &kc, bool iso, bool sw, const StepArgs<T> &A, int grid, hipStream_t s,
                 hipEvent_t t0 = nullptr, hipEvent_t t1 = nullptr);
void launch_epilogue(const ImgDesc *desc, ImgState *state, const double *part, Tile *list_out, int *count_out,
                     int TT, double u_round, double eta, int nimg, hipStream_t s, hipEvent_t t1 = nullptr);
template <class T>
void launch_gather(const T *src, T *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s);
template <class T>
void launch_scatter(const T *src, T *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s);
template <class T>
void launch_copy_u(const T *arena, const ImgDesc *desc, const ImgState *state, int nimg, int k, int64_t maxN, T *dst,
                   int64_t img_stride, hipStream_t s);
template <class S, class D>
void launch_convert(const S *src, D *dst, int64_t n, hipStream_t s);
template <class S, class D>
void launch_convert_scatter(const S *src, D *dst, int64_t n, int nd, const StridedShape &sh, hipStream_t s);
void launch_zero_bytes(void *p, size_t n, hipStream_t s);
template <class T>
void launch_gather_group(const GatherEntry *tab, int G, int64_t maxN, hipStream_t s);
void launch_admit_commit(const ImgDesc *sdesc, const ImgState *sstate, const double *sgam, const int *map,
                         ImgDesc *desc, ImgState *state, double *gam, Tile *list, int *count, int G, hipStream_t s);

}  // namespace fpr
