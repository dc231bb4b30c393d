// The solve engine (engine.hpp).
//
// SPDX-License-Identifier: MIT
#include "alloc.hpp"
#include "engine.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

namespace fpr {

Engine::~Engine() {
    release();
    for (auto &x : half_ev_)
        if (x) (void)hipEventDestroy(x);
    if (half_stream_) (void)hipStreamDestroy(half_stream_);
    if (ref_ev_) (void)hipEventDestroy(ref_ev_);
    for (auto &x : tev) (void)hipEventDestroy(x);
    for (auto &x : done_ev)
        if (x) (void)hipEventDestroy(x);
    if (cap_stream_) (void)hipStreamDestroy(cap_stream_);
}

int Engine::ensure_events(int steps, bool timed) {
    hipError_t e;
    for (auto &x : done_ev)
        if (!x && (e = hipEventCreateWithFlags(&x, hipEventDisableTiming))) return e;
    if (timed && tev.size() != (size_t)(6 * steps)) {
        drop_graphs();  // graphs record into the events they were captured with
        for (auto &x : tev) (void)hipEventDestroy(x);
        tev.assign((size_t)(6 * steps), nullptr);
        for (auto &x : tev)  // timestamps only: no system-scope fence at each record
            if ((e = hipEventCreateWithFlags(&x, hipEventDisableSystemFence))) return e;
    }
    return hipSuccess;
}

void Engine::drop_graphs() {
    for (auto &g : graphs_) (void)hipGraphExecDestroy(g.exec);
    graphs_.clear();
}

hipGraphExec_t Engine::find_graph(int tail, int slot, int p, int grid, int steps, int timed) const {
    for (const auto &g : graphs_)
        if (g.tail == tail && g.slot == slot && g.p == p && g.grid == grid && g.steps == steps && g.timed == timed)
            return g.exec;
    return nullptr;
}

hipStream_t Engine::capture_stream() {
    if (!cap_stream_ && hipStreamCreateWithFlags(&cap_stream_, hipStreamNonBlocking)) cap_stream_ = nullptr;
    return cap_stream_;
}

void Engine::release() {
    drop_graphs();
    half_.reset();
    auto f = [](void *p) {
        if (p) (void)hipFree(p);
    };
    f(arena_);
    f(desc_);
    f(state_);
    f(part_);
    f(list_[0]);
    f(list_[1]);
    f(counts_);
    f(gam_);
    f(propart_);
    f(desc_tail_);
    if (hdesc_tail_pinned_) (void)hipHostFree(hdesc_tail_pinned_);
    if (hstate_) (void)hipHostFree(hstate_);
    if (hdesc_pinned_) (void)hipHostFree(hdesc_pinned_);
    if (hcount_) (void)hipHostFree(hcount_);
    if (hgam_) (void)hipHostFree(hgam_);
    arena_ = nullptr;
    desc_ = nullptr;
    state_ = nullptr;
    part_ = nullptr;
    list_[0] = list_[1] = nullptr;
    counts_ = nullptr;
    gam_ = nullptr;
    propart_ = nullptr;
    desc_tail_ = nullptr;
    hdesc_tail_pinned_ = nullptr;
    tail_frac_ = 0.0;
    total_tiles_tail_ = 0;
    tiles_cap_ = 0;
    hstate_ = nullptr;
    hdesc_pinned_ = nullptr;
    hcount_ = nullptr;
    hgam_ = nullptr;
}

int Engine::setup(int tsize, const Problem &P, const KernelChoice &kc, const std::vector<ImageSpec> &all) {
    release();
    tsize_ = tsize;
    P_ = P;
    kc_ = kc;
    P_.TT = kc.iters();
    // The split: the first images up to 1/parts of the batch's tiles stay
    // here, the others go to a second engine (which splits them again into
    // parts - 1 when parts > 2).
    static const int split_env = [] {
        const char *e = std::getenv("FASTPROX_ROCM_SPLIT");
        const int v = e ? std::atoi(e) : 2;
        return std::min(std::max(v, 1), kMaxParts);
    }();
    const int parts = split_parts_ > 0 ? split_parts_ : split_env;
    size_t keep = all.size();
    if (allow_split && parts >= 2 && all.size() >= 2) {
        std::vector<int64_t> t(all.size());
        int64_t tot = 0;
        for (size_t b = 0; b < all.size(); ++b) {
            const ImageSpec &s = all[b];
            t[b] = ((s.L[2] + kc.tile_x() - 1) / kc.tile_x()) * ((s.L[1] + kc.tile_y() - 1) / kc.tile_y()) *
                   kc.tiles_z(s.L[0]);
            tot += t[b];
        }
        static const int64_t split_tiles = [] {
            const char *e = std::getenv("FASTPROX_ROCM_SPLIT_TILES");
            const long long v = e ? std::atoll(e) : 0;
            return v > 0 ? (int64_t)v : (int64_t)kSplitTiles;
        }();
        if (split_parts_ > 0 || tot >= split_tiles) {
            int64_t acc = 0;
            keep = 0;
            while (keep + 1 < all.size() && parts * (acc + t[keep]) <= tot) acc += t[keep++];
            if (keep == 0) keep = 1;
        }
    }
    if (keep < all.size()) {
        half_.reset(new Engine());
        half_->split_parts_ = parts - 1;
        half_->allow_split = parts > 2;
        const std::vector<ImageSpec> rest(all.begin() + (std::ptrdiff_t)keep, all.end());
        if (int e = half_->setup(tsize, P, kc, rest)) return e;
        hipError_t e;
        if (!half_stream_ && (e = hipStreamCreateWithFlags(&half_stream_, hipStreamNonBlocking))) return e;
        for (auto &x : half_ev_)
            if (!x && (e = hipEventCreateWithFlags(&x, hipEventDisableTiming))) return e;
    }
    specs_.assign(all.begin(), all.begin() + (std::ptrdiff_t)keep);
    const std::vector<ImageSpec> &imgs = specs_;
    const int B = (int)imgs.size();
    const int Ball = (int)all.size();
    hdesc_.assign((size_t)B, ImgDesc{});
    int64_t arena = 0;
    int tiles = 0;
    maxN_ = 0;
    tv1d_possible_ = false;
    const int k = P.k;
    for (int b = 0; b < B; ++b) {
        const ImageSpec &s = imgs[(size_t)b];
        ImgDesc &D = hdesc_[(size_t)b];
        for (int a = 0; a < 3; ++a) D.L[a] = (int32_t)s.L[a];
        const int64_t N = s.L[0] * s.L[1] * s.L[2];
        maxN_ = std::max(maxN_, N);
        D.d = nullptr;
        D.z = nullptr;
        D.q = nullptr;
        D.sw = nullptr;
        D.q_cstride = s.q_cstride;
        D.sw_cstride = s.sw_cstride;
        D.st_off = arena;
        arena += 4 * (int64_t)k * N;
        D.ntx = (int32_t)((s.L[2] + kc.tile_x() - 1) / kc.tile_x());
        D.nty = (int32_t)((s.L[1] + kc.tile_y() - 1) / kc.tile_y());
        D.ntz = (int32_t)kc.tiles_z(s.L[0]);
        D.ntiles = D.ntx * D.nty * D.ntz;
        D.tile_base = tiles;
        tiles += D.ntiles;
        if (P.iso && !P.sw) {
            int active = 0;
            for (int a = 0; a < k; ++a)
                if (s.L[3 - k + a] > 1 && P.w[a] > 0.0) ++active;
            if (active == 1) tv1d_possible_ = true;
        }
    }
    total_tiles_ = tiles;
    tiles_cap_ = tiles;
    // the streaming 2D kernel's 32-bit plane offsets
    if (((kc.stream && !kc.generic) || kc.s3t()) && !stream_fits(tsize, maxN_)) return hipErrorInvalidValue;
    hipError_t e;
    if ((e = fpr::dev_alloc(&arena_, (size_t)std::max<int64_t>(arena, 1) * (size_t)tsize))) return e;
    if ((e = fpr::dev_alloc((void **)&desc_, sizeof(ImgDesc) * (size_t)std::max(B, 1)))) return e;
    if ((e = fpr::dev_alloc((void **)&state_, sizeof(ImgState) * (size_t)std::max(B, 1)))) return e;
    if ((e = fpr::dev_alloc((void **)&part_, sizeof(double) * kSlots * (size_t)std::max(tiles, 1)))) return e;
    for (int i = 0; i < 2; ++i)
        if ((e = fpr::dev_alloc((void **)&list_[i], sizeof(Tile) * (size_t)std::max(tiles, 1)))) return e;
    if ((e = fpr::dev_alloc((void **)&counts_, 64))) return e;
    if ((e = fpr::dev_alloc((void **)&gam_, sizeof(double) * (size_t)std::max(B, 1)))) return e;
    if ((e = fpr::dev_alloc((void **)&propart_,
                       sizeof(double) * 4 * (size_t)std::max(B, 1) * (size_t)prologue_blocks(std::max<int64_t>(maxN_, 1)))))
        return e;
    if ((e = fpr::pinned_alloc((void **)&hstate_, sizeof(ImgState) * (size_t)std::max(Ball, 1)))) return e;
    if ((e = fpr::pinned_alloc((void **)&hdesc_pinned_, sizeof(ImgDesc) * (size_t)std::max(B, 1)))) return e;
    if ((e = fpr::pinned_alloc((void **)&hcount_, 64))) return e;
    if ((e = fpr::pinned_alloc((void **)&hgam_, sizeof(double) * (size_t)std::max(B, 1)))) return e;
    return hipSuccess;
}

// The tiles of every image under kc: descs[b]'s tile fields and the total.
int Engine::tile_layout(const KernelChoice &kc, std::vector<ImgDesc> &descs) const {
    int tiles = 0;
    for (size_t b = 0; b < descs.size(); ++b) {
        ImgDesc &D = descs[b];
        D.ntx = (int32_t)((D.L[2] + kc.tile_x() - 1) / kc.tile_x());
        D.nty = (int32_t)((D.L[1] + kc.tile_y() - 1) / kc.tile_y());
        D.ntz = (int32_t)kc.tiles_z(D.L[0]);
        D.ntiles = D.ntx * D.nty * D.ntz;
        D.tile_base = tiles;
        tiles += D.ntiles;
    }
    return tiles;
}

int Engine::set_tail(const KernelChoice &kc, double frac) {
    if (half_)
        if (int e = half_->set_tail(kc, frac)) return e;
    drop_graphs();
    tail_frac_ = 0.0;
    if (!(frac > 0.0) || kc.generic != kc_.generic || P_.k != 2) return hipSuccess;
    if (!kernel_available(tsize_, kc.stream, kc.TT, kc.RY)) return hipErrorInvalidValue;
    // The last launch before the switch was planned for up to kc_.TT
    // iterations; the tail kernel must be able to run them.
    if (kc.iters() < kc_.iters()) return hipErrorInvalidValue;
    hdesc_tail_ = hdesc_;
    const int tiles = tile_layout(kc, hdesc_tail_);
    hipError_t e;
    if (!desc_tail_) {
        const size_t B = std::max<size_t>(hdesc_.size(), 1);
        if ((e = fpr::dev_alloc((void **)&desc_tail_, sizeof(ImgDesc) * B))) return e;
        if ((e = fpr::pinned_alloc((void **)&hdesc_tail_pinned_, sizeof(ImgDesc) * B))) return e;
    }
    if (tiles > tiles_cap_) {  // the partials and the lists hold either tiling
        for (void *p : {(void *)part_, (void *)list_[0], (void *)list_[1]}) (void)hipFree(p);
        part_ = nullptr;
        list_[0] = list_[1] = nullptr;
        if ((e = fpr::dev_alloc((void **)&part_, sizeof(double) * kSlots * (size_t)tiles))) return e;
        for (int i = 0; i < 2; ++i)
            if ((e = fpr::dev_alloc((void **)&list_[i], sizeof(Tile) * (size_t)tiles))) return e;
        tiles_cap_ = tiles;
    }
    tail_kc_ = kc;
    tail_kc_.stream = kc.stream;
    total_tiles_tail_ = tiles;
    tail_frac_ = frac;
    return hipSuccess;
}

template <class T>
static int run_solve(Engine &E, const Problem &P, const KernelChoice &kc, hipStream_t s, void *arena,
                     ImgDesc *desc, ImgState *state, double *part, Tile **list,
                     int *counts, double *gam, double *propart, int B, int total_tiles, int64_t maxN, bool tv1d,
                     int *hcount, int steps_per_sync, int max_iter, const KernelChoice &tail_kc,
                     ImgDesc *desc_tail, int total_tiles_tail, double tail_frac) {
    const double u_round = 0.5 * (sizeof(T) == 4 ? (double)FLT_EPSILON : DBL_EPSILON);
    const double eta = 0.5 * (sizeof(T) == 4 ? (double)FLT_TRUE_MIN : DBL_TRUE_MIN);
    launch_prologue<T>(P, desc, state, gam, B, propart, maxN, s);
    launch_reset_counts(counts, s);
    launch_initial_list(desc, state, list[0], counts, B, s);
    if (P.warm) launch_warm<T>(P, (T *)arena, desc, state, B, maxN, s);
    if (tv1d) launch_tv1d<T>(P, desc, state, B, maxN, u_round, eta, s);
    StepArgs<T> A;
    A.arena = (T *)arena;
    A.desc = desc;
    A.state = state;
    A.part = part;
    A.k = P.k;
    auto grid_of = [](const KernelChoice &c, int tiles) {
        if (c.generic) return std::max(1, std::min(tiles, 8192));
        if (c.s3) return std::max(8, std::min((tiles + 7) / 8 * 8, 4096));
        return std::max(8, std::min(((tiles + 3) / 4 + 7) / 8 * 8, 2048));
    };
    int grid = grid_of(kc, total_tiles);
    KernelChoice cur_kc = kc;
    const ImgDesc *cur_desc = desc;
    int TT = kc.iters();
    bool tail = false, want_tail = false;
    // A bound on the launches any image can need: every accepted launch runs
    // at least one iteration or ends the image, and each restart or stop
    // costs at most two more.
    const int64_t cap = 8LL * ((int64_t)max_iter / std::max(1, std::min(TT, tail_kc.iters())) + 8) + 64;
    int p = 0;
    int64_t launched = 0;
    (void)0;
    hipError_t err = hipGetLastError();
    if (err) return err;
    if (E.last_static()) {
        // The whole schedule, with no host read: every pair after an image
        // has finished returns at once for it (an empty list).
        const int64_t chunks = (cap + steps_per_sync - 1) / steps_per_sync;
        for (int64_t c = 0; c < chunks; ++c)
            for (int i = 0; i < steps_per_sync; ++i) {
                A.list_in = list[p];
                A.count_in = counts + p;
                A.count_out = counts + (p ^ 1);
                launch_step<T>(cur_kc, P.iso != 0, P.sw != 0, A, grid, s, nullptr, nullptr);
                launch_epilogue(cur_desc, state, part, list[p ^ 1], counts + (p ^ 1), TT, u_round, eta, B, s,
                                nullptr);
                p ^= 1;
            }
        E.stats.launches += (int)(chunks * steps_per_sync);
        return hipGetLastError();
    }
    // Optional timing: an event before and after every step and epilogue,
    // read back after each chunk of launches.
    const bool timing = E.timing;
    if ((err = (hipError_t)E.ensure_events(steps_per_sync, timing))) return err;
    const std::vector<hipEvent_t> &ev = E.tev;
    hipEvent_t *done_ev = E.done_ev;
    // Chunks of launches are pipelined: chunk c + 1 is enqueued before the
    // host reads chunk c's count of live tiles, so the device never waits on
    // the host; the chunk enqueued after the last one only meets empty lists.
    // Graphs: each chunk's launches are captured once per key and replayed.
    static const bool graph_env = [] {
        const char *e = std::getenv("FASTPROX_ROCM_GRAPH");
        return !(e && e[0] == '0');
    }();
    const bool use_graphs = E.graphs && graph_env && !timing;  // timing events stay outside graphs
    // The chunk's launches and its count readback, on stream st.
    auto record = [&](hipStream_t st, int slot) -> hipError_t {
        for (int i = 0; i < steps_per_sync; ++i) {
            A.list_in = list[p];
            A.count_in = counts + p;
            A.count_out = counts + (p ^ 1);
            // With timing, the step kernel's start and end and the epilogue's
            // end are stamped by the launches themselves.
            const size_t eb = (size_t)(3 * (slot * steps_per_sync + i));
            launch_step<T>(cur_kc, P.iso != 0, P.sw != 0, A, grid, st, timing ? ev[eb] : nullptr,
                           timing ? ev[eb + 1] : nullptr);
            launch_epilogue(cur_desc, state, part, list[p ^ 1], counts + (p ^ 1), TT, u_round, eta, B, st,
                            timing ? ev[eb + 2] : nullptr);
            p ^= 1;
        }
        return hipMemcpyAsync(hcount + slot, counts + p, sizeof(int), hipMemcpyDeviceToHost, st);
    };
    auto enqueue = [&](int slot) -> hipError_t {
        if (want_tail && !tail) {
            // Relist the live images on the tail kernel's tiling: the list
            // the next step reads is rebuilt from the image modes.
            tail = true;
            cur_kc = tail_kc;
            cur_desc = desc_tail;
            A.desc = desc_tail;
            TT = tail_kc.iters();
            const int live = hcount[slot] > 0 ? hcount[slot] : total_tiles;  // the count just read
            // The tail's tiles, estimated and rounded up to a power of two
            // (few distinct grids, so few captured graphs).
            int64_t est = (int64_t)live * total_tiles_tail / std::max(1, total_tiles) + 64, r = 64;
            while (r < est) r <<= 1;
            grid = grid_of(cur_kc, (int)std::min<int64_t>(total_tiles_tail, r));
            launch_reset_counts(counts, s);
            launch_initial_list(desc_tail, state, list[p], counts + p, B, s);
        }
        launched += steps_per_sync;
        E.stats.launches += steps_per_sync;
        hipError_t e2 = hipSuccess;
        hipStream_t cs = use_graphs ? E.capture_stream() : nullptr;
        if (cs) {
            const int t = tail ? 1 : 0, p0 = p;
            hipGraphExec_t ex = E.find_graph(t, slot, p0, grid, steps_per_sync, timing ? 1 : 0);
            if (ex) {
                p ^= steps_per_sync & 1;
            } else {
                hipGraph_t g = nullptr;
                hipError_t e1 = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
                if (!e1) {
                    e1 = record(cs, slot);
                    const hipError_t e3 = hipStreamEndCapture(cs, &g);
                    if (!e1) e1 = e3;
                }
                if (!e1) e1 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e1) {
                    // No graph: this engine launches every kernel from now on.
                    (void)hipGetLastError();
                    E.graphs = false;
                    ex = nullptr;
                    p = p0;
                } else {
                    E.add_graph(Engine::ChunkGraph{t, slot, p0, grid, steps_per_sync, timing ? 1 : 0, ex});
                }
            }
            e2 = ex ? hipGraphLaunch(ex, s) : record(s, slot);
        } else {
            e2 = record(s, slot);
        }
        if (!e2) e2 = hipEventRecord(done_ev[slot], s);
        return e2;
    };
    auto read_timing = [&](int slot) {
        if (!timing) return;
        for (int i = 0; i < steps_per_sync; ++i) {
            const size_t eb = (size_t)(3 * (slot * steps_per_sync + i));
            float a = 0.f, b = 0.f;
            (void)hipEventElapsedTime(&a, ev[eb], ev[eb + 1]);
            (void)hipEventElapsedTime(&b, ev[eb + 1], ev[eb + 2]);
            E.stats.step_ms += a;
            E.stats.epi_ms += b;
            ++E.stats.timed;
            float t0 = 0.f, t1 = 0.f;
            if (E.ref_ && !hipEventElapsedTime(&t0, E.ref_, ev[eb]) && !hipEventElapsedTime(&t1, E.ref_, ev[eb + 1]))
                E.spans_.emplace_back(t0, t1);
        }
    };
    int slot = 0;
    if ((err = enqueue(slot))) return err;
    for (;;) {
        if ((err = enqueue(slot ^ 1))) break;
        if ((err = hipEventSynchronize(done_ev[slot]))) break;
        ++E.stats.host_syncs;
        read_timing(slot);
        if (hcount[slot] == 0) {
            // The chunk in flight finds no tiles; let it drain.
            if ((err = hipEventSynchronize(done_ev[slot ^ 1]))) break;
            read_timing(slot ^ 1);
            break;
        }
        if (launched > cap) {
            err = (hipError_t)-1;  // the schedule did not finish: an internal error
            break;
        }
        if (tail_frac > 0.0 && !tail && (double)hcount[slot] < tail_frac * (double)total_tiles) want_tail = true;
        slot ^= 1;
    }
    return err == (hipError_t)-1 ? -1 : err;
}

template <class T>
int run_solve_in(Engine &E, const Problem &P, const KernelChoice &kc, hipStream_t s, void *arena, ImgDesc *desc,
                 ImgState *state, double *part, Tile **list, int *counts, double *gam, double *propart, int B,
                 int total_tiles, int64_t maxN, bool tv1d, int *hcount, int steps_per_sync, int max_iter) {
    return run_solve<T>(E, P, kc, s, arena, desc, state, part, list, counts, gam, propart, B, total_tiles, maxN, tv1d,
                        hcount, steps_per_sync, max_iter, kc, nullptr, 0, 0.0);
}
template int run_solve_in<float>(Engine &, const Problem &, const KernelChoice &, hipStream_t, void *, ImgDesc *,
                                 ImgState *, double *, Tile **, int *, double *, double *, int, int, int64_t, bool,
                                 int *, int, int);
template int run_solve_in<double>(Engine &, const Problem &, const KernelChoice &, hipStream_t, void *, ImgDesc *,
                                  ImgState *, double *, Tile **, int *, double *, double *, int, int, int64_t, bool,
                                  int *, int, int);

namespace {

// Length of the union of the intervals.
double union_ms(std::vector<std::pair<float, float>> &v) {
    std::sort(v.begin(), v.end());
    double tot = 0.0, lo = 0.0, hi = -1.0;
    bool open = false;
    for (const auto &x : v) {
        if (open && x.first <= hi) {
            hi = std::max(hi, (double)x.second);
            continue;
        }
        if (open) tot += hi - lo;
        lo = x.first;
        hi = x.second;
        open = true;
    }
    if (open) tot += hi - lo;
    return tot;
}

}  // namespace

int Engine::solve(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol,
                  int max_iter) {
    spans_.clear();
    ref_ = nullptr;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipError_t e = hipStreamIsCapturing(s, &cs)) return e;
    static_ = cs == hipStreamCaptureStatusActive;
    for (Engine *h = half_.get(); h; h = h->half_.get()) h->static_ = static_;
    if (static_ && timing) return hipErrorInvalidValue;  // timing reads events back: not under capture
    if (timing) {
        hipError_t e = ref_ev_ ? hipSuccess : hipEventCreate(&ref_ev_);
        if (!e) e = hipEventRecord(ref_ev_, s);
        if (e) return e;
        ref_ = ref_ev_;
    }
    const int r = solve_parts(s, ptrs, gammas, warm, tol, max_iter);
    if (!r) stats.busy_ms = union_ms(spans_);
    return r;
}

void Engine::collect() {
    if (!half_) return;
    half_->collect();
    std::memcpy(hstate_ + own(), half_->state(), sizeof(ImgState) * (size_t)half_->nimg());
}

int Engine::solve_parts(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol,
                        int max_iter) {
    if (!half_) return solve_own(s, ptrs, gammas, warm, tol, max_iter);
    if (static_) {
        // Under capture both parts are enqueued from this thread: the second
        // part's stream joins the capture through an event and rejoins s at
        // the end, so the graph runs the parts side by side as the eager
        // solve does.
        const int B0 = own();
        half_->steps_per_sync = steps_per_sync;
        hipError_t e = hipEventRecord(half_ev_[0], s);
        if (!e) e = hipStreamWaitEvent(half_stream_, half_ev_[0], 0);
        if (e) return e;
        const int r0 = solve_own(s, ptrs, gammas, warm, tol, max_iter);
        const int r1 = half_->solve_parts(half_stream_, ptrs + B0, gammas + B0, warm, tol, max_iter);
        e = hipEventRecord(half_ev_[1], half_stream_);
        if (!e) e = hipStreamWaitEvent(s, half_ev_[1], 0);
        if (r0 || r1) return r0 ? r0 : r1;
        return e;
    }
    const int B0 = own();
    half_->spans_.clear();
    half_->ref_ = ref_;
    half_->timing = timing;
    half_->graphs = graphs;
    half_->steps_per_sync = steps_per_sync;
    // The second half's stream starts after the caller's earlier work on s.
    hipError_t e = hipEventRecord(half_ev_[0], s);
    if (!e) e = hipStreamWaitEvent(half_stream_, half_ev_[0], 0);
    if (e) return e;
    int dev = 0;
    if ((e = hipGetDevice(&dev))) return e;
    int r1 = 0;
    std::thread worker([&] {
        r1 = hipSetDevice(dev);
        if (!r1) r1 = half_->solve_parts(half_stream_, ptrs + B0, gammas + B0, warm, tol, max_iter);
    });
    const int r0 = solve_own(s, ptrs, gammas, warm, tol, max_iter);
    worker.join();
    // Later work on s follows the second half's (which has finished: its
    // solve synchronises its stream).
    if (!r1) {
        e = hipEventRecord(half_ev_[1], half_stream_);
        if (!e) e = hipStreamWaitEvent(s, half_ev_[1], 0);
        if (e) r1 = e;
    }
    if (r0 || r1) return r0 ? r0 : r1;
    std::memcpy(hstate_ + B0, half_->state(), sizeof(ImgState) * (size_t)half_->nimg());
    const SolveStats &h = half_->stats;
    stats.launches += h.launches;
    stats.host_syncs += h.host_syncs;
    stats.step_ms += h.step_ms;
    stats.epi_ms += h.epi_ms;
    stats.timed += h.timed;
    spans_.insert(spans_.end(), half_->spans_.begin(), half_->spans_.end());
    return hipSuccess;
}

int Engine::solve_own(hipStream_t s, const ImagePtrs *ptrs, const double *gammas, bool warm, double tol,
                      int max_iter) {
    const int B = own();
    stats = SolveStats{};
    if (B == 0) return hipSuccess;
    P_.warm = warm ? 1 : 0;
    P_.tol = tol;
    P_.max_iter = max_iter;
    std::memcpy(hgam_, gammas, sizeof(double) * (size_t)B);
    for (int b = 0; b < B; ++b) {
        ImgDesc D = hdesc_[(size_t)b];
        D.d = ptrs[b].d;
        D.z = ptrs[b].z;
        D.q = ptrs[b].q;
        D.sw = ptrs[b].sw;
        hdesc_pinned_[b] = D;
    }
    hipError_t e = hipMemcpyAsync(gam_, hgam_, sizeof(double) * (size_t)B, hipMemcpyHostToDevice, s);
    if (e) return e;
    if ((e = hipMemcpyAsync(desc_, hdesc_pinned_, sizeof(ImgDesc) * (size_t)B, hipMemcpyHostToDevice, s))) return e;
    if (tail_frac_ > 0.0) {
        for (int b = 0; b < B; ++b) {
            ImgDesc D = hdesc_tail_[(size_t)b];
            D.d = ptrs[b].d;
            D.z = ptrs[b].z;
            D.q = ptrs[b].q;
            D.sw = ptrs[b].sw;
            hdesc_tail_pinned_[b] = D;
        }
        if ((e = hipMemcpyAsync(desc_tail_, hdesc_tail_pinned_, sizeof(ImgDesc) * (size_t)B, hipMemcpyHostToDevice,
                                s)))
            return e;
    }
    int r;
    if (tsize_ == 4)
        r = run_solve<float>(*this, P_, kc_, s, arena_, desc_, state_, part_, list_, counts_, gam_, propart_, B,
                             total_tiles_, maxN_, tv1d_possible_, hcount_, steps_per_sync, max_iter, tail_kc_,
                             desc_tail_, total_tiles_tail_, tail_frac_);
    else
        r = run_solve<double>(*this, P_, kc_, s, arena_, desc_, state_, part_, list_, counts_, gam_, propart_, B,
                              total_tiles_, maxN_, tv1d_possible_, hcount_, steps_per_sync, max_iter, tail_kc_,
                              desc_tail_, total_tiles_tail_, tail_frac_);
    if (r) return r;
    if ((e = hipMemcpyAsync(hstate_, state_, sizeof(ImgState) * (size_t)B, hipMemcpyDeviceToHost, s))) return e;
    return static_ ? hipSuccess : hipStreamSynchronize(s);
}

int Engine::copy_dual_iterate(hipStream_t s, void *dst, int64_t img_stride) const {
    if (!arena_) return hipErrorInvalidValue;
    if (tsize_ == 4)
        launch_copy_u<float>((const float *)arena_, desc_, state_, own(), P_.k, maxN_, (float *)dst, img_stride, s);
    else
        launch_copy_u<double>((const double *)arena_, desc_, state_, own(), P_.k, maxN_, (double *)dst, img_stride,
                              s);
    if (hipError_t e = hipGetLastError()) return e;
    if (half_)
        return half_->copy_dual_iterate(s, (char *)dst + (size_t)own() * (size_t)img_stride * (size_t)tsize_,
                                        img_stride);
    return hipSuccess;
}

}  // namespace fpr
