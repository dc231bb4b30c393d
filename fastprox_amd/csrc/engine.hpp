// The solve engine: device buffers for a batch of images and the host loop
// that drives the prologue, the step/epilogue launch pairs and the readback.
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "tv_common.hpp"
#include "tv_kernels.hpp"

namespace fpr {

// One image of a batch: its extents and the layout of its dual and weights.
struct ImageSpec {
    int64_t L[3];        // extents, padded in front with 1
    int64_t q_cstride;   // distance between the components of its dual
    int64_t sw_cstride;  // distance between its edge-weight planes
};

// Where one image's data lives for a solve (device pointers of the dtype).
struct ImagePtrs {
    const void *d;   // input, contiguous
    void *z;         // output, contiguous
    void *q;         // dual (component 0), read first on a warm start
    const void *sw;  // spatial weights or null
};

struct SolveStats {
    int launches = 0;      // step launches
    int host_syncs = 0;
    double step_ms = 0.0;  // summed step-kernel time (HIP events), when timing is on
    double epi_ms = 0.0;   // summed epilogue time
    int timed = 0;         // step launches timed
    // Time at least one step kernel ran, over both streams of a split
    // batch (the union of the timed step intervals); equal to step_ms
    // when one stream runs them.
    double busy_ms = 0.0;
};

class Engine {
public:
    Engine() = default;
    ~Engine();
    Engine(const Engine &) = delete;
    Engine &operator=(const Engine &) = delete;

    // Allocates every device buffer for the batch: the states, the per-image
    // descriptors and results, the tile lists and the partial sums. kc picks
    // the step kernel. Returns a hipError_t.
    int setup(int tsize, const Problem &P, const KernelChoice &kc, const std::vector<ImageSpec> &imgs);

    // Solves every image, ptrs[i] its data and gammas[i] its step (host
    // arrays of nimg()). Blocks until done; results in state(). On a stream
    // that is being captured into a hipGraph, it instead enqueues the whole
    // static schedule (every step/epilogue pair that max_iter allows, each
    // returning at once for the images already finished: the same iterates),
    // with no host synchronisation and no host read: the results are in
    // state() once the caller has synchronised and called collect().
    int solve(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol, int max_iter);
    // After a static (captured) solve has run: gathers the parts' states.
    void collect();
    bool last_static() const { return static_; }

    // The tail kernel: once the live tiles of a solve drop below frac of the
    // batch's tiles, the solve continues with kernel kc on its own tiling
    // (every kernel computes the same iterates, so results do not change;
    // a latency-oriented tiling shortens the launches of the last images).
    // frac 0 turns it off; kc must run at least as many iterations per launch
    // as the main kernel. Call after setup. Returns a hipError_t.
    int set_tail(const KernelChoice &kc, double frac);
    const KernelChoice &tail_kernel() const { return tail_kc_; }
    double tail_fraction() const { return tail_frac_; }

    // Copies every image's dual iterate u (k planes of N entries, the solver's
    // own scaling) to dst + i * img_stride elements, after a solve.
    int copy_dual_iterate(hipStream_t s, void *dst, int64_t img_stride) const;

    const ImgState *state() const { return hstate_; }
    int nimg() const { return own() + (half_ ? half_->nimg() : 0); }
    int total_tiles() const { return total_tiles_; }
    const KernelChoice &kernel() const { return kc_; }
    const Problem &problem() const { return P_; }
    SolveStats stats;
    int steps_per_sync = 8;
    bool timing = false;   // time every step and epilogue launch with HIP events
    // Replay each chunk of step/epilogue launches from a captured hipGraph
    // (captured on first use, kept until the engine's configuration changes);
    // off, or with timing on, every kernel is launched on its own. Results
    // are the same either way.
    bool graphs = true;

    // The graph of one chunk: the launches depend only on the engine's
    // buffers and on (tail, slot, p, grid, steps), so the key names them.
    struct ChunkGraph {
        int tail, slot, p, grid, steps, timed;
        hipGraphExec_t exec;
    };
    hipGraphExec_t find_graph(int tail, int slot, int p, int grid, int steps, int timed) const;
    void add_graph(const ChunkGraph &g) {
        if (graphs_.size() >= 64) drop_graphs();
        graphs_.push_back(g);
    }
    hipStream_t capture_stream();  // a stream of its own that chunks are captured on (null if none)
    // The solve loop's events, kept across solves: with timing, 3 per launch
    // of two chunks of steps launches; and each chunk's completion. Returns a
    // hipError_t.
    int ensure_events(int steps, bool timed);
    std::vector<hipEvent_t> tev;
    hipEvent_t done_ev[2] = {nullptr, nullptr};
    // With timing: the event every step interval is measured from (recorded
    // on the caller's stream before either half starts), and the intervals
    // [start, end) in ms of this engine's timed step launches.
    hipEvent_t ref_ = nullptr;
    std::vector<std::pair<float, float>> spans_;

    // A batch of at least kSplitTiles tiles is solved as parts (two by
    // default) on as many streams, each part after the first driven by a
    // worker thread: while one part waits for its epilogue, a launch or its
    // host sync, another's step kernel fills the GPU. Every image's result is
    // the same either way. FASTPROX_ROCM_SPLIT=n asks for n parts (at most
    // kMaxParts; 0 or 1 turns the split off), FASTPROX_ROCM_SPLIT_TILES
    // another threshold. Two parts at most: a three-part batch (a split half
    // split again) was slower (256 x 512^2 -12%, the uneven batch -18%) and
    // its process crashed on the host in the uneven batch's captured-graph
    // run, for a cause not found; that nested path is not reachable.
    static constexpr int kSplitTiles = 2048;
    static constexpr int kMaxParts = 2;
    bool allow_split = true;

private:
    bool static_ = false;  // this solve enqueues the static schedule (its stream is capturing)
    int own() const { return (int)specs_.size(); }  // the images of this engine itself
    int solve_own(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol, int max_iter);
    int solve_parts(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol, int max_iter);
    int split_parts_ = 0;               // parts this engine's images are split into (0: the environment's)
    std::unique_ptr<Engine> half_;      // images [own(), nimg()) when split
    hipStream_t half_stream_ = nullptr;
    hipEvent_t half_ev_[2] = {nullptr, nullptr};
    hipEvent_t ref_ev_ = nullptr;  // owned reference event of timed solves
    void release();
    void drop_graphs();
    std::vector<ChunkGraph> graphs_;
    hipStream_t cap_stream_ = nullptr;
    int tile_layout(const KernelChoice &kc, std::vector<ImgDesc> &descs) const;
    int tsize_ = 4;
    KernelChoice tail_kc_{};
    double tail_frac_ = 0.0;
    int total_tiles_tail_ = 0;
    int tiles_cap_ = 0;           // capacity of the partials and the lists
    std::vector<ImgDesc> hdesc_tail_;
    ImgDesc *desc_tail_ = nullptr;
    ImgDesc *hdesc_tail_pinned_ = nullptr;
    Problem P_{};
    KernelChoice kc_{};
    std::vector<ImageSpec> specs_;
    std::vector<ImgDesc> hdesc_;
    int total_tiles_ = 0;
    int64_t maxN_ = 0;
    bool tv1d_possible_ = false;
    void *arena_ = nullptr;
    ImgDesc *desc_ = nullptr;
    ImgState *state_ = nullptr;
    double *part_ = nullptr;
    Tile *list_[2] = {nullptr, nullptr};
    int *counts_ = nullptr;
    double *gam_ = nullptr;
    double *propart_ = nullptr;
    ImgDesc *hdesc_pinned_ = nullptr;  // pinned
    ImgState *hstate_ = nullptr;  // pinned
    int *hcount_ = nullptr;       // pinned
    double *hgam_ = nullptr;      // pinned
};

// The solve loop of an engine over buffers the caller provides (laid out as
// Engine::setup lays out its own: the states, descriptors, partials, the two
// tile lists, the counts, the steps and the prologue's partials; hcount two
// host ints), with no tail kernel: the prologue, then chunks of step and
// epilogue launches until no tile is live. E lends its events; set its
// graphs and timing off to launch every kernel directly. Returns a
// hipError_t, or -1 when the schedule does not finish.
template <class T>
int run_solve_in(Engine &E, const Problem &P, const KernelChoice &kc, hipStream_t s, void *arena, ImgDesc *desc,
                 ImgState *state, double *part, Tile **list, int *counts, double *gam, double *propart, int B,
                 int total_tiles, int64_t maxN, bool tv1d, int *hcount, int steps_per_sync, int max_iter);

}  // namespace fpr
