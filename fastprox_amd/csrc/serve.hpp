// A serving loop for image-TV requests on one GPU: continuous batching over
// the solve engine's kernels (serve.cpp).
//
// SPDX-License-Identifier: MIT
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "tv_common.hpp"
#include "tv_kernels.hpp"

namespace fpr {

// A first-fit allocator of [0, capacity) with coalescing frees, in units.
class RangeAlloc {
public:
    void reset(int64_t capacity);
    int64_t alloc(int64_t n);  // offset, or -1
    bool fits(int64_t n) const;  // whether alloc(n) would succeed now
    void free(int64_t off, int64_t n);
    int64_t free_total() const { return free_total_; }

private:
    std::map<int64_t, int64_t> free_;  // offset -> size
    int64_t free_total_ = 0;
};

struct ServeRequest {
    int64_t id = -1;
    int k = 2;
    int64_t L[3] = {1, 1, 1};    // extents, padded in front with 1
    const void *x = nullptr;     // device, the server's dtype
    bool has_xs = false;         // xs holds x's strides (else x is in C order)
    int64_t xs[3] = {0, 0, 0};   // element strides of x (padded like L)
    hipStream_t caller = nullptr;  // x is produced and z consumed on this stream
    hipEvent_t ready = nullptr;    // recorded on caller at the submit
    void *z = nullptr;           // device output (contiguous) or null
    double gamma = 0.0, tol = 1e-6;
    int max_iter = 100000;
    int keep_dual = 0;
    int64_t warm_from = -1;
    double t_submit = 0.0;       // seconds on the server clock
};

struct ServeResult {
    int64_t id;
    int status;     // 0, or an fp_status
    int converged;
    int iters;
    double gap;     // at the input's scale
    double zerr;
    double t_submit, t_admit, t_done;
};

// One class of requests (one k and one step kernel) with its own stream,
// slots, memory pool and tile pool.
class SlotPool {
public:
    ~SlotPool();
    int setup(const Problem &P, const KernelChoice &kc, int tsize, int slots, int64_t pool_elems,
              int64_t tile_cap, int64_t max_pixels);
    // What an admission found short: a slot, tile range or memory.
    enum : int { kFits = 0, kShortSlot = 1, kShortTiles = 2, kShortMem = 3 };
    // Admits r into a free slot if a slot, memory and tiles are free (and
    // enqueues its set-up); false otherwise, with *why what was short.
    bool try_admit(const ServeRequest &r, double now, int *err, int *why);
    // Grouped admission: stage() takes r's slot, memory and tiles like
    // try_admit, and leaves its set-up to flush_group(), which sets up every
    // staged request at once: one copy of their descriptors, one gather of
    // their inputs, one prologue, and one kernel that moves the results into
    // their slots and lists their tiles (instead of about ten small
    // operations per request). Requests with a warm start, exact 1D TV, or
    // more than kGroupBlocks / kGroupMax prologue blocks take try_admit.
    static constexpr int kGroupMax = 64;
    static constexpr int kGroupBlocks = 64 * 512;
    bool groupable(const ServeRequest &r) const;
    bool exact_1d(const ServeRequest &r) const;
    bool stage(const ServeRequest &r, double now, int *err, int *why);
    int flush_group();
    // Enqueues a chunk of step/epilogue launches and a state snapshot when
    // any slot is live; returns whether one was enqueued.
    int enqueue_chunk(bool *enqueued);
    // Waits for the oldest chunk in flight and finishes the requests it shows
    // done (results appended to out).
    int retire(std::vector<ServeResult> &out, double (*clock)(const void *), const void *ctx);
    // Whether the oldest chunk in flight has finished (without waiting).
    bool oldest_done() const;
    int in_flight() const { return (int)flight_.size(); }
    int live() const { return live_; }
    int64_t live_tiles() const { return live_tiles_; }
    bool has_kept(int64_t id) const { return kept_.count(id) != 0; }
    bool fits_ever(const ServeRequest &r) const;
    int64_t footprint(const ServeRequest &r) const;
    bool release_dual(int64_t id);
    // Frees the oldest kept dual other than keep's; false if there is none.
    bool evict_one_dual(int64_t keep = -1);
    const KernelChoice &kernel() const { return kc_; }
    int tsize() const { return tsize_; }
    // Launch pairs per chunk: a finished request is seen when its chunk's
    // snapshot is read, so shorter chunks cut latency (4 against 8: served
    // p50 0.67 against 0.78 ms, the same p99; 2 lengthens the p99).
    int steps_per_sync = 4;
    hipStream_t stream = nullptr;
    int64_t launches = 0;

private:
    struct Slot {
        bool used = false;
        bool done = false;
        ServeRequest req;
        double t_admit = 0.0;
        int64_t off_main = -1, n_main = 0;  // x, z, state
        int64_t off_dual = -1, n_dual = 0;  // q
        int64_t tile_off = -1, ntiles = 0;
        int64_t N = 0;
        int64_t admit_seq = 0;  // chunks enqueued before the request's set-up
    };
    struct Kept {
        int64_t off, n, L[3];
        uint64_t stamp;
    };
    int finish_slot(int i, const ImgState &S, std::vector<ServeResult> &out, double now);
    // Orders the callers' streams of the finished slots after the chunk
    // whose output launch wrote their z.
    int order_out(const std::vector<int> &done, hipEvent_t chunk);
    int reserve(const ServeRequest &r, int *slot, ImgDesc *D, int64_t *offs);  // kFits or what is short
    void occupy(int i, const ServeRequest &r, const ImgDesc &D, const int64_t *offs, double now);
    void release_reserved(const ImgDesc &D, const int64_t *offs);
    // The staged group (pinned, two buffers: one may still be in a copy) and
    // its device copy; the prologue's states and partials of the group.
    struct Stage {
        ImgDesc desc[kGroupMax];
        double gam[kGroupMax];
        double tol[kGroupMax];
        int maxit[kGroupMax];
        int map[kGroupMax];
        GatherEntry gat[kGroupMax];
    };
    Stage *hstage_[2] = {nullptr, nullptr};
    hipEvent_t stage_ev_[2] = {nullptr, nullptr};
    int stage_buf_ = 0;
    int staged_ = 0;
    int64_t staged_maxN_ = 0;
    Stage *dstage_ = nullptr;
    ImgState *sstate_ = nullptr;
    double *gpart_ = nullptr;
    Problem P_{};
    KernelChoice kc_{};
    int tsize_ = 4;
    int nslots_ = 0;
    std::vector<Slot> slots_;
    RangeAlloc mem_, tiles_;
    int64_t pool_elems_ = 0, tile_cap_ = 0, max_pixels_ = 0;
    void *pool_ = nullptr;
    ImgDesc *desc_ = nullptr;
    ImgState *state_ = nullptr;
    double *part_ = nullptr, *gam_ = nullptr, *propart_ = nullptr;
    Tile *list_[2] = {nullptr, nullptr};
    int *counts_ = nullptr;
    ImgDesc *hdesc_ = nullptr;     // pinned, per slot
    double *hgam_ = nullptr;       // pinned, per slot
    ImgState *hsnap_[2] = {nullptr, nullptr};  // pinned snapshots of the chunks in flight
    struct Flight {
        hipEvent_t ev;
        int buf;
        int top;      // slots covered by the snapshot
        int64_t seq;  // the chunk's number (from 1)
    };
    int64_t seq_ = 0;
    std::deque<Flight> flight_;
    std::vector<hipEvent_t> events_;
    int p_ = 0;       // the list the next step reads
    int live_ = 0;
    int top_ = 0;     // 1 + the highest slot ever used
    int64_t live_tiles_ = 0;
    std::map<int64_t, Kept> kept_;
    uint64_t stamp_ = 0;
    std::vector<int> retire_buf_;  // the slots a retirement finishes
};

class Server {
public:
    ~Server();
    // ops[0]: 2D, ops[1]: 3D problems (k = 0 for a class not served).
    int setup(const Problem *probs, const KernelChoice *kcs, int tsize, int slots, size_t pool_bytes);
    int submit(ServeRequest r, int64_t *id);
    // One round: admissions, a chunk of launches per busy class, and the
    // retirement of the oldest chunk in flight. live/queued after the round.
    int step(int *live, int *queued);
    int64_t poll(ServeResult *out, int64_t max);
    double now() const;
    bool release_dual(int64_t id);
    int64_t launches() const;
    // The step kernel of class c (0: 2D, 1: 3D); null when c is not served.
    const KernelChoice *kernel(int c) const { return pools_[c] ? &pools_[c]->kernel() : nullptr; }

private:
    // pools_[0]: 2D, [1]: 3D, [2]: a second 2D lane. A server of 2D requests
    // alone, with a pool of at least kLanePoolBytes and kLaneSlots slots,
    // runs them as two lanes, each with half the slots and the memory, on its
    // own stream: while one lane waits for its epilogue or its chunk's
    // snapshot, the other's step kernel runs (as the batch engine's two
    // halves do). Measured on 2048 queued 512^2 requests: 4,368 -> 4,838
    // images/s. A server of both classes keeps one 2D lane: on the mixed
    // 2D/3D trace (Poisson 200 requests/s) a second 2D lane raised the 3D
    // requests' p99 from 59 to 77 ms, whether or not it took work while 3D
    // requests were live.
    static constexpr size_t kLanePoolBytes = size_t(1) << 30;
    static constexpr int kLaneSlots = 32;
    static constexpr int64_t kLaneSpill = 4096;
    std::unique_ptr<SlotPool> pools_[3];
    int lanes(int c) const { return c == 0 && pools_[2] ? 2 : 1; }
    SlotPool *lane(int c, int i) const { return i == 0 ? pools_[c].get() : pools_[2].get(); }
    std::deque<ServeRequest> queue_[2];
    std::vector<ServeResult> done_;
    // Events recorded on the callers' streams at submission, waited on by
    // the admission; recycled once the wait is enqueued.
    std::vector<hipEvent_t> ev_free_, ev_all_;
    int64_t next_id_ = 0;
    std::chrono::steady_clock::time_point t0_ = std::chrono::steady_clock::now();
};

}  // namespace fpr
