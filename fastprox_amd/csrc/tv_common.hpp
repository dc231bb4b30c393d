// Shared device/host layout of the image-TV engine: per-image descriptors and
// solver state, the tile list, and the arguments of the kernels.
//
// One solve is a batch of images (uniform, from fp_prox, or of uneven sizes).
// Every image keeps its dual iterate in HBM as two ping-pong buffers
// (state s in {0, 1}: k planes of u then k planes of v, each N entries of T),
// so that tiles with halos read the old state while others write the new one.
// A step launch runs up to TT dual-FISTA iterations on every tile of every
// live image; an epilogue launch then reduces each image's per-tile partial
// sums (restart inner products and duality gaps, FP64) in a fixed order,
// replays the reference's schedule on them and plans the image's next launch
// (see tv_kernels.hip).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <stdint.h>

namespace fpr {

constexpr int kMaxTT = 8;                 // most iterations per step launch
constexpr int kSlots = 2 * (kMaxTT + 1);  // per tile: gap and dot at each level 0..TT

// Image modes.
enum : int {
    kDone = 0,      // finished: outputs written
    kIterate = 1,   // running dual FISTA
    kOutput = 2,    // next step launch runs cnt iterations and writes z and the dual
    kOutputNaN = 3, // an image holding a NaN: next launch writes NaN outputs
    kTV1D = 4,      // one active axis: solved by the exact 1D kernel
};

// Schedules: the C kernel's (tv_iso.cpp:489-537: restart test on 0-based
// iterations it % 4 == 0 with the interval rule, gap after it % 4 == 0) and
// the device path's (_tv_torch.py:259-281: test and gap on 1-based
// iterations n % 4 == 0, no interval rule).
enum : int { kSchedC = 0, kSchedDevice = 1 };

// After-actions of a launch that re-runs a prefix up to a known event.
enum : int { kAfterNone = 0, kAfterRestart = 1 };

struct ImgDesc {
    const void *d;      // the image's input (contiguous, C order)
    void *z;            // its output z
    void *q;            // component 0 of its dual
    const void *sw;     // its spatial weights (alpha: N; beta: k planes), null for none
    int64_t st_off;     // its state (2 buffers x 2k planes x N) in the state arena
    int64_t q_cstride;  // distance between dual components
    int64_t sw_cstride; // distance between edge-weight planes
    int32_t L[3];       // extents, padded in front with 1 (L[2] the contiguous axis)
    int32_t tile_base;  // index of the image's first tile in the partials
    int32_t ntx, nty, ntz;  // tiles along x, y and slices
    int32_t ntiles;
    int32_t pad_;
};

struct ImgState {
    // fixed for the solve (set by the prologue)
    double gamma;      // the projection radius (iso: gamma w_max, clamped; aniso: gamma)
    double gamma_raw;  // the step as given, rescaled by 2^-e
    double bound;      // iso: the cap of every pixel radius (pixel weights); else +inf
    double tau;        // 1 / (4 sum wa^2)
    double wa[3];      // the weights inside G = W D, per axis (0 for inactive axes)
    double wmax;       // iso: the weight normalisation (dual out = u / wmax); aniso 1
    double target;     // stop when gap <= target
    double sumsq;      // |x_s|^2, x_s = 2^-e x (for the certified error)
    double sumsq_m;    // sum (x_s / maxabs)^2 (the certified error far from unit scale)
    double maxabs;     // max |x_s|
    // iteration state
    double t;          // FISTA t at the current state
    double gp;         // the latest gap evaluated
    double beta[kMaxTT + 1];  // beta[j]: momentum of level j of the next launch
    int32_t e;         // power-of-two rescaling exponent
    int32_t it, since, last;
    int32_t cur;       // current state buffer
    int32_t vsrc_u;    // the next launch reads v from the u planes (a restart)
    int32_t fresh;     // the state is all zero (cold start, nothing written yet)
    int32_t cnt;       // iterations of the next launch
    uint32_t test_mask;  // bit j: level j of the next launch is a restart test
    uint32_t gap_mask;   // bit j: evaluate the gap of level j in the next launch
    int32_t pending;   // the gap of the current state is due (landed on a launch's last level)
    int32_t after;     // action once the next launch is accepted (kAfter*)
    int32_t mode;
    int32_t sched;
    int32_t max_iter;
    int32_t converged;
    int32_t status;    // 0 ok; 1 an infinite entry (FP_EINVAL)
    int32_t active_axis;  // kTV1D: the active axis
    int32_t passes;    // step launches that iterated the image (each one HBM pass over its state)
    double tv1d_lam;   // kTV1D: gamma_s w[a]
    double zerr;       // certified relative error (filled when done)
};

struct Tile {
    int32_t img;
    int16_t tx, ty;
    int32_t tz;
    int32_t pad_;
};

// Problem constants shared by every image of a solve.
struct Problem {
    int k;            // image axes (2 or 3)
    int iso;          // isotropic (ball) or anisotropic (box)
    int sw;           // spatial weights present (iso: alpha, aniso: beta)
    int sw_per_image; // spatial weights differ per image (else one shared map)
    double w[3];      // axis weights as given (last k entries used)
    double tol;
    int max_iter;
    int sched;
    int warm;         // start from the dual buffer
    int TT;           // iterations per step launch of the chosen kernel
    int wmax_norm;    // iso: normalise the weights by their max (C kernel)
    double alpha_max; // largest pixel weight (step scale), 1 without
};

template <class T>
struct StepArgs {
    T *arena;            // states
    const ImgDesc *desc;
    ImgState *state;
    double *part;        // kSlots per tile
    const Tile *list_in; // tiles of this launch
    const int *count_in;
    int *count_out;      // reset here for the epilogue that follows
    int k;
};

}  // namespace fpr
