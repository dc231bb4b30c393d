// A serving loop for image-TV requests: continuous batching on one GPU.
//
// Every request class (2D, 3D) owns a fixed table of slots, one device pool
// of the dtype for the requests' buffers, a pool of tiles for the per-tile
// partial sums, and a stream. The step and epilogue kernels of the solve
// engine run over the whole slot table: an empty or finished slot is in mode
// kDone and gets no tiles, so one launch serves every live request of the
// class whatever its size, step, tolerance or age. A round of the loop
//   1. admits queued requests, first come first served, while a slot, the
//      request's memory (x, z and the state: (2 + 4k) N entries; the dual:
//      k N) and its tiles fit, evicting kept duals (least recently used)
//      when memory is short: the set-up (input copy, the prologue, a warm
//      start) is enqueued on the class's stream and the request's tiles are
//      appended to the list that the next step launch reads;
//   2. enqueues a chunk of step/epilogue launch pairs and a snapshot of the
//      slot states, keeping two chunks in flight so the GPU never waits on
//      the host;
//   3. waits for a class's older chunk (with both classes served, of the
//      first class to finish one) and retires the requests its snapshot shows
//      done: their output is copied out, their memory and tiles return to
//      the pools, their dual stays resident when asked (a warm start for a
//      later request of the same shape).
// A request's result depends only on its own data: the kernels compute each
// image's tiles and fixed-order reductions independently of the others, so
// it is the same alone and in any mix.
//
// SPDX-License-Identifier: MIT
#include "alloc.hpp"
#include "serve.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <thread>

#include "../../include/fastprox_rocm.h"

namespace fpr {

// ---------------------------------------------------------------------------
// RangeAlloc
// ---------------------------------------------------------------------------

void RangeAlloc::reset(int64_t capacity) {
    free_.clear();
    if (capacity > 0) free_[0] = capacity;
    free_total_ = capacity;
}

int64_t RangeAlloc::alloc(int64_t n) {
    if (n <= 0) return -1;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second < n) continue;
        const int64_t off = it->first, rest = it->second - n;
        free_.erase(it);
        if (rest) free_[off + n] = rest;
        free_total_ -= n;
        return off;
    }
    return -1;
}

bool RangeAlloc::fits(int64_t n) const {
    if (n <= 0) return false;
    for (const auto &r : free_)
        if (r.second >= n) return true;
    return false;
}

void RangeAlloc::free(int64_t off, int64_t n) {
    if (off < 0 || n <= 0) return;
    auto it = free_.emplace(off, n).first;
    free_total_ += n;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// ---------------------------------------------------------------------------
// SlotPool
// ---------------------------------------------------------------------------

SlotPool::~SlotPool() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (void *p : {pool_, (void *)desc_, (void *)state_, (void *)part_, (void *)gam_, (void *)propart_,
                    (void *)list_[0], (void *)list_[1], (void *)counts_, (void *)dstage_, (void *)sstate_,
                    (void *)gpart_})
        if (p) (void)hipFree(p);
    for (void *p : {(void *)hdesc_, (void *)hgam_, (void *)hsnap_[0], (void *)hsnap_[1], (void *)hstage_[0],
                    (void *)hstage_[1]})
        if (p) (void)hipHostFree(p);
    for (hipEvent_t e : stage_ev_)
        if (e) (void)hipEventDestroy(e);
    for (hipEvent_t e : events_) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
}

int SlotPool::setup(const Problem &P, const KernelChoice &kc, int tsize, int slots, int64_t pool_elems,
                    int64_t tile_cap, int64_t max_pixels) {
    P_ = P;
    kc_ = kc;
    P_.TT = kc.iters();
    tsize_ = tsize;
    nslots_ = slots;
    pool_elems_ = pool_elems;
    tile_cap_ = tile_cap;
    max_pixels_ = max_pixels;
    slots_.assign((size_t)slots, Slot{});
    mem_.reset(pool_elems);
    tiles_.reset(tile_cap);
    hipError_t e;
    // non-blocking: a caller on the null stream is ordered by the events
    // alone (a blocking pool stream would make every submit's event and every
    // retirement's wait on that stream serialise with all chunks in flight)
    if ((e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking))) return e;
    if ((e = fpr::dev_alloc(&pool_, (size_t)pool_elems * (size_t)tsize))) return e;
    if ((e = fpr::dev_alloc((void **)&desc_, sizeof(ImgDesc) * (size_t)slots))) return e;
    if ((e = fpr::dev_alloc((void **)&state_, sizeof(ImgState) * (size_t)slots))) return e;
    if ((e = fpr::dev_alloc((void **)&part_, sizeof(double) * kSlots * (size_t)tile_cap))) return e;
    if ((e = fpr::dev_alloc((void **)&gam_, sizeof(double) * (size_t)slots))) return e;
    if ((e = fpr::dev_alloc((void **)&propart_, sizeof(double) * 4 * (size_t)prologue_blocks(max_pixels)))) return e;
    for (int i = 0; i < 2; ++i)
        if ((e = fpr::dev_alloc((void **)&list_[i], sizeof(Tile) * (size_t)tile_cap))) return e;
    if ((e = fpr::dev_alloc((void **)&counts_, 64))) return e;
    if ((e = fpr::dev_alloc((void **)&dstage_, sizeof(Stage)))) return e;
    if ((e = fpr::dev_alloc((void **)&sstate_, sizeof(ImgState) * kGroupMax))) return e;
    if ((e = fpr::dev_alloc((void **)&gpart_, sizeof(double) * 4 * (size_t)kGroupBlocks))) return e;
    for (int i = 0; i < 2; ++i) {
        if ((e = fpr::pinned_alloc((void **)&hstage_[i], sizeof(Stage)))) return e;
        if ((e = hipEventCreateWithFlags(&stage_ev_[i], hipEventDisableTiming))) return e;
    }
    if ((e = fpr::pinned_alloc((void **)&hdesc_, sizeof(ImgDesc) * (size_t)slots))) return e;
    if ((e = fpr::pinned_alloc((void **)&hgam_, sizeof(double) * (size_t)slots))) return e;
    for (int i = 0; i < 2; ++i)
        if ((e = fpr::pinned_alloc((void **)&hsnap_[i], sizeof(ImgState) * (size_t)slots))) return e;
    // every slot starts finished (mode kDone = 0); both lists empty
    if ((e = hipMemsetAsync(state_, 0, sizeof(ImgState) * (size_t)slots, stream))) return e;
    if ((e = hipMemsetAsync(desc_, 0, sizeof(ImgDesc) * (size_t)slots, stream))) return e;
    launch_reset_counts(counts_, stream);
    events_.resize(4);
    for (auto &ev : events_)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming))) return e;
    return hipStreamSynchronize(stream);
}

int64_t SlotPool::footprint(const ServeRequest &r) const {
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    return (2 + 5 * (int64_t)P_.k) * N;
}

bool SlotPool::fits_ever(const ServeRequest &r) const {
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    return N <= max_pixels_ && footprint(r) <= pool_elems_;
}

bool SlotPool::release_dual(int64_t id) {
    auto it = kept_.find(id);
    if (it == kept_.end()) return false;
    mem_.free(it->second.off, it->second.n);
    kept_.erase(it);
    return true;
}

bool SlotPool::evict_one_dual(int64_t keep) {
    auto old = kept_.end();
    for (auto it = kept_.begin(); it != kept_.end(); ++it)
        if (it->first != keep && (old == kept_.end() || it->second.stamp < old->second.stamp)) old = it;
    if (old == kept_.end()) return false;
    mem_.free(old->second.off, old->second.n);
    kept_.erase(old);
    return true;
}

template <class T>
static void launch_setup(const Problem &Pi, void *pool, ImgDesc *desc, ImgState *state, double *gam, double *propart,
                         int64_t N, bool warm, bool tv1d, hipStream_t s) {
    const double u_round = 0.5 * (sizeof(T) == 4 ? (double)FLT_EPSILON : DBL_EPSILON);
    const double eta = 0.5 * (sizeof(T) == 4 ? (double)FLT_TRUE_MIN : DBL_TRUE_MIN);
    launch_prologue<T>(Pi, desc, state, gam, 1, propart, N, s);
    if (warm) launch_warm<T>(Pi, (T *)pool, desc, state, 1, N, s);
    if (tv1d) launch_tv1d<T>(Pi, desc, state, 1, N, u_round, eta, s);
}

// A free slot, the request's pool memory (x, z, state; the dual) and its
// tiles; its descriptor. kFits, or the first of them (slot, tiles, memory)
// that is not free now; only a shortage of memory can be cured by evicting
// kept duals.
int SlotPool::reserve(const ServeRequest &r, int *slot, ImgDesc *Dout, int64_t *offs) {
    int i = -1;
    for (int j = 0; j < nslots_; ++j)
        if (!slots_[(size_t)j].used) {
            i = j;
            break;
        }
    if (i < 0) return kShortSlot;
    const int k = P_.k;
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    ImgDesc D{};
    for (int a = 0; a < 3; ++a) D.L[a] = (int32_t)r.L[a];
    D.ntx = (int32_t)((r.L[2] + kc_.tile_x() - 1) / kc_.tile_x());
    D.nty = (int32_t)((r.L[1] + kc_.tile_y() - 1) / kc_.tile_y());
    D.ntz = (int32_t)kc_.tiles_z(r.L[0]);
    D.ntiles = D.ntx * D.nty * D.ntz;
    const int64_t n_main = (2 + 4 * (int64_t)k) * N, n_dual = (int64_t)k * N;
    if (!tiles_.fits(D.ntiles)) return kShortTiles;
    const int64_t off_main = mem_.alloc(n_main);
    if (off_main < 0) return kShortMem;
    const int64_t off_dual = mem_.alloc(n_dual);
    if (off_dual < 0) {
        mem_.free(off_main, n_main);
        return kShortMem;
    }
    const int64_t tile_off = tiles_.alloc(D.ntiles);  // fits: checked above
    char *base = (char *)pool_;
    const size_t ts = (size_t)tsize_;
    D.d = base + (size_t)off_main * ts;
    // the output launch writes z straight into the caller's buffer; the
    // pool's z block is the target only when the caller asked for none
    D.z = r.z ? r.z : base + (size_t)(off_main + N) * ts;
    D.st_off = off_main + 2 * N;
    D.q = base + (size_t)off_dual * ts;
    D.sw = nullptr;
    D.q_cstride = N;
    D.sw_cstride = N;
    D.tile_base = (int32_t)tile_off;
    *slot = i;
    *Dout = D;
    offs[0] = off_main;
    offs[1] = n_main;
    offs[2] = off_dual;
    offs[3] = n_dual;
    offs[4] = tile_off;
    return kFits;
}

void SlotPool::occupy(int i, const ServeRequest &r, const ImgDesc &D, const int64_t *offs, double now) {
    Slot &S = slots_[(size_t)i];
    S.used = true;
    S.done = false;
    S.req = r;
    S.t_admit = now;
    S.off_main = offs[0];
    S.n_main = offs[1];
    S.off_dual = offs[2];
    S.n_dual = offs[3];
    S.tile_off = offs[4];
    S.ntiles = D.ntiles;
    S.N = r.L[0] * r.L[1] * r.L[2];
    S.admit_seq = seq_;
    ++live_;
    live_tiles_ += D.ntiles;
    top_ = std::max(top_, i + 1);
}

// x's strides when given (any values, zero for a broadcast axis), else the
// C-order ones. An axis of extent 1 is never stepped along, so it takes its
// C-order stride, which lets a contiguous input be recognised as one.
static void input_strides(const ServeRequest &r, int64_t st[3]) {
    const int64_t c[3] = {r.L[1] * r.L[2], r.L[2], 1};
    for (int a = 0; a < 3; ++a) st[a] = (r.has_xs && r.L[a] > 1) ? r.xs[a] : c[a];
}

// Whether the prologue will solve r as exact 1D TV (one active axis, the
// rule of prologue_combine2), which needs the 1D kernels after it.
bool SlotPool::exact_1d(const ServeRequest &r) const {
    if (!P_.iso || P_.sw || !(r.gamma > 0.0)) return false;
    const int k = P_.k;
    double wmaxL = 0.0;
    for (int a = 0; a < k; ++a)
        if (r.L[3 - k + a] > 1) wmaxL = std::max(wmaxL, P_.w[a]);
    int active = 0;
    for (int a = 0; a < k; ++a)
        if (P_.w[a] > 0.0 && r.L[3 - k + a] > 1 && P_.w[a] >= 1e-100 * wmaxL) ++active;
    return active == 1;
}

bool SlotPool::groupable(const ServeRequest &r) const {
    if (r.warm_from >= 0 || exact_1d(r)) return false;
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    return prologue_blocks(N) <= kGroupBlocks / kGroupMax;
}

bool SlotPool::stage(const ServeRequest &r, double now, int *err, int *why) {
    *err = 0;
    if (staged_ == kGroupMax)
        if ((*err = flush_group())) return false;
    int i = -1;
    ImgDesc D{};
    int64_t offs[5];
    if ((*why = reserve(r, &i, &D, offs)) != kFits) return false;
    // the group's gather reads x after the work on its caller's stream
    // before the submission
    if (r.ready && (*err = hipStreamWaitEvent(stream, r.ready, 0))) {
        release_reserved(D, offs);
        return false;
    }
    if (staged_ == 0) {
        // the staging buffer's previous copy has been read
        if ((*err = hipEventSynchronize(stage_ev_[stage_buf_]))) return false;
        staged_maxN_ = 0;
    }
    Stage &G = *hstage_[stage_buf_];
    const int j = staged_++;
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    G.desc[j] = D;
    G.gam[j] = r.gamma;
    G.tol[j] = r.tol;
    G.maxit[j] = r.max_iter;
    G.map[j] = i;
    GatherEntry &g = G.gat[j];
    int64_t st[3];
    input_strides(r, st);
    g.src = r.x;
    g.dst = (void *)D.d;
    g.n = N;
    for (int a = 0; a < 3; ++a) {
        g.shape[a] = r.L[a];
        g.stride[a] = st[a];
    }
    g.contiguous = st[2] == 1 && st[1] == r.L[2] && st[0] == r.L[1] * r.L[2];
    g.pad_ = 0;
    staged_maxN_ = std::max(staged_maxN_, N);
    occupy(i, r, D, offs, now);
    return true;
}

int SlotPool::flush_group() {
    if (staged_ == 0) return 0;
    const int G = staged_;
    staged_ = 0;
    hipStream_t s = stream;
    const int b = stage_buf_;
    stage_buf_ ^= 1;
    hipError_t e = hipMemcpyAsync(dstage_, hstage_[b], sizeof(Stage), hipMemcpyHostToDevice, s);
    if (e) return e;
    if ((e = hipEventRecord(stage_ev_[b], s))) return e;
    Problem Pi = P_;
    Pi.warm = 0;
    if (tsize_ == 4) {
        launch_gather_group<float>(dstage_->gat, G, staged_maxN_, s);
        launch_prologue<float>(Pi, dstage_->desc, sstate_, dstage_->gam, G, gpart_, staged_maxN_, s, dstage_->tol,
                               dstage_->maxit);
    } else {
        launch_gather_group<double>(dstage_->gat, G, staged_maxN_, s);
        launch_prologue<double>(Pi, dstage_->desc, sstate_, dstage_->gam, G, gpart_, staged_maxN_, s, dstage_->tol,
                                dstage_->maxit);
    }
    launch_admit_commit(dstage_->desc, sstate_, dstage_->gam, dstage_->map, desc_, state_, gam_, list_[p_],
                        counts_ + p_, G, s);
    return hipGetLastError();
}

bool SlotPool::try_admit(const ServeRequest &r, double now, int *err, int *why) {
    *err = 0;
    if ((*err = flush_group())) return false;  // admissions stay in order
    int i = -1;
    ImgDesc D{};
    int64_t offs[5];
    if ((*why = reserve(r, &i, &D, offs)) != kFits) return false;
    if (r.ready && (*err = hipStreamWaitEvent(stream, r.ready, 0))) {
        release_reserved(D, offs);
        return false;
    }
    const int64_t N = r.L[0] * r.L[1] * r.L[2];
    const int64_t n_dual = offs[3];
    hipStream_t s = stream;
    char *base = (char *)pool_;
    const size_t ts = (size_t)tsize_;
    // the input, gathered from its strides
    StridedShape sh{};
    int64_t st[3];
    input_strides(r, st);
    for (int a = 0; a < 3; ++a) {
        sh.shape[a] = r.L[a];
        sh.stride[a] = st[a];
    }
    if (tsize_ == 4) launch_gather<float>((const float *)r.x, (float *)D.d, N, 3, sh, s);
    else launch_gather<double>((const double *)r.x, (double *)D.d, N, 3, sh, s);
    // a warm start from a kept dual of the same shape
    bool warm = false;
    if (r.warm_from >= 0) {
        auto it = kept_.find(r.warm_from);
        if (it != kept_.end() && it->second.L[0] == r.L[0] && it->second.L[1] == r.L[1] &&
            it->second.L[2] == r.L[2]) {
            if ((*err = hipMemcpyAsync(D.q, base + (size_t)it->second.off * ts, (size_t)n_dual * ts,
                                       hipMemcpyDeviceToDevice, s))) {
                release_reserved(D, offs);
                return false;
            }
            it->second.stamp = ++stamp_;
            warm = true;
        }
    }
    hdesc_[i] = D;
    hgam_[i] = r.gamma;
    if ((*err = hipMemcpyAsync(desc_ + i, hdesc_ + i, sizeof(ImgDesc), hipMemcpyHostToDevice, s)) ||
        (*err = hipMemcpyAsync(gam_ + i, hgam_ + i, sizeof(double), hipMemcpyHostToDevice, s))) {
        release_reserved(D, offs);
        return false;
    }
    Problem Pi = P_;
    Pi.tol = r.tol;
    Pi.max_iter = r.max_iter;
    Pi.warm = warm ? 1 : 0;
    const bool tv1d = exact_1d(r);
    if (tsize_ == 4) launch_setup<float>(Pi, pool_, desc_ + i, state_ + i, gam_ + i, propart_, N, warm, tv1d, s);
    else launch_setup<double>(Pi, pool_, desc_ + i, state_ + i, gam_ + i, propart_, N, warm, tv1d, s);
    launch_initial_list(desc_ + i, state_ + i, list_[p_], counts_ + p_, 1, s, i);
    if ((*err = hipGetLastError())) {
        release_reserved(D, offs);
        return false;
    }
    occupy(i, r, D, offs, now);
    return true;
}

void SlotPool::release_reserved(const ImgDesc &D, const int64_t *offs) {
    mem_.free(offs[0], offs[1]);
    mem_.free(offs[2], offs[3]);
    tiles_.free(offs[4], D.ntiles);
}

int SlotPool::enqueue_chunk(bool *enqueued) {
    *enqueued = false;
    if (int e = flush_group()) return e;
    if (live_ == 0 || flight_.size() >= 2) return 0;
    hipStream_t s = stream;
    const int TT = kc_.iters();
    const double u_round = 0.5 * (tsize_ == 4 ? (double)FLT_EPSILON : DBL_EPSILON);
    const double eta = 0.5 * (tsize_ == 4 ? (double)FLT_TRUE_MIN : DBL_TRUE_MIN);
    const int64_t tl = std::max<int64_t>(live_tiles_, 1);
    int grid;
    if (kc_.generic) grid = (int)std::max<int64_t>(1, std::min<int64_t>(tl, 8192));
    else if (kc_.s3) grid = (int)std::max<int64_t>(8, std::min<int64_t>((tl + 7) / 8 * 8, 4096));
    else grid = (int)std::max<int64_t>(8, std::min<int64_t>(((tl + 3) / 4 + 7) / 8 * 8, 2048));
    for (int j = 0; j < steps_per_sync; ++j) {
        if (tsize_ == 4) {
            StepArgs<float> A{(float *)pool_, desc_, state_, part_, list_[p_], counts_ + p_, counts_ + (p_ ^ 1),
                              P_.k};
            launch_step<float>(kc_, P_.iso != 0, P_.sw != 0, A, grid, s);
        } else {
            StepArgs<double> A{(double *)pool_, desc_, state_, part_, list_[p_], counts_ + p_, counts_ + (p_ ^ 1),
                               P_.k};
            launch_step<double>(kc_, P_.iso != 0, P_.sw != 0, A, grid, s);
        }
        launch_epilogue(desc_, state_, part_, list_[p_ ^ 1], counts_ + (p_ ^ 1), TT, u_round, eta, top_, s);
        p_ ^= 1;
        ++launches;
    }
    const int buf = flight_.empty() ? 0 : (flight_.back().buf ^ 1);
    hipError_t e = hipMemcpyAsync(hsnap_[buf], state_, sizeof(ImgState) * (size_t)top_, hipMemcpyDeviceToHost, s);
    if (e) return e;
    ++seq_;
    hipEvent_t ev = events_[(size_t)(seq_ % (int64_t)events_.size())];
    if ((e = hipEventRecord(ev, s))) return e;
    flight_.push_back(Flight{ev, buf, top_, seq_});
    *enqueued = true;
    return 0;
}

int SlotPool::finish_slot(int i, const ImgState &St, std::vector<ServeResult> &out, double now) {
    Slot &S = slots_[(size_t)i];
    ServeResult R{};
    R.id = S.req.id;
    R.status = St.status ? FP_EINVAL : 0;
    R.converged = St.converged;
    R.iters = St.it;
    double g = St.gp;
    if (P_.iso && g == g) g = g < 0.0 ? 0.0 : g;
    if (P_.iso && St.e) g = std::ldexp(g, 2 * St.e);
    R.gap = g;
    R.zerr = St.zerr;
    R.t_submit = S.req.t_submit;
    R.t_admit = S.t_admit;
    R.t_done = now;
    mem_.free(S.off_main, S.n_main);
    tiles_.free(S.tile_off, S.ntiles);
    if (S.req.keep_dual && !R.status) {
        Kept K{S.off_dual, S.n_dual, {S.req.L[0], S.req.L[1], S.req.L[2]}, ++stamp_};
        kept_[S.req.id] = K;
    } else {
        mem_.free(S.off_dual, S.n_dual);
    }
    --live_;
    live_tiles_ -= S.ntiles;
    S = Slot{};
    out.push_back(R);
    return 0;
}

bool SlotPool::oldest_done() const {
    // an error counts as done: retire() then reports it
    return !flight_.empty() && hipEventQuery(flight_.front().ev) != hipErrorNotReady;
}

int SlotPool::order_out(const std::vector<int> &done, hipEvent_t chunk) {
    // The outputs are in the callers' buffers once the chunk has run: each
    // distinct caller stream waits for its event once.
    std::vector<hipStream_t> streams;
    for (int i : done) {
        const Slot &S = slots_[(size_t)i];
        if (S.req.z && std::find(streams.begin(), streams.end(), S.req.caller) == streams.end())
            streams.push_back(S.req.caller);
    }
    for (hipStream_t cs : streams)
        if (hipError_t e = hipStreamWaitEvent(cs, chunk, 0)) return e;
    return 0;
}

int SlotPool::retire(std::vector<ServeResult> &out, double (*clock)(const void *), const void *ctx) {
    if (flight_.empty()) return 0;
    const Flight f = flight_.front();
    flight_.pop_front();
    if (hipError_t e = hipEventSynchronize(f.ev)) return e;
    const double now = clock(ctx);
    const ImgState *snap = hsnap_[f.buf];
    std::vector<int> &done = retire_buf_;
    done.clear();
    for (int i = 0; i < f.top; ++i) {
        const Slot &S = slots_[(size_t)i];
        // a slot admitted after the chunk was enqueued is not in its snapshot
        if (!S.used || S.admit_seq >= f.seq) continue;
        if (snap[i].mode != kDone) continue;
        done.push_back(i);
    }
    if (int e = order_out(done, f.ev)) return e;
    for (int i : done)
        if (int e = finish_slot(i, snap[i], out, now)) return e;
    return 0;
}

// ---------------------------------------------------------------------------
// Server
// ---------------------------------------------------------------------------

Server::~Server() {
    for (auto &p : pools_) p.reset();  // each waits for its stream
    for (hipEvent_t e : ev_all_) (void)hipEventDestroy(e);
}

int Server::setup(const Problem *probs, const KernelChoice *kcs, int tsize, int slots, size_t pool_bytes) {
    for (int c = 0; c < 2; ++c) {
        if (probs[c].k == 0) continue;
        const bool two = c == 0 && probs[1].k == 0 && pool_bytes >= kLanePoolBytes && slots >= kLaneSlots;
        const int nl = two ? 2 : 1;
        const int64_t elems = (int64_t)(pool_bytes / (size_t)tsize) / nl;
        const int k = probs[c].k;
        const int64_t max_pixels = std::min<int64_t>(elems / (2 + 5 * k), (int64_t)1 << 28);
        const int64_t tile_cap = std::max<int64_t>(elems / 1024, 4096);
        for (int i = 0; i < nl; ++i) {
            std::unique_ptr<SlotPool> &P = i == 0 ? pools_[c] : pools_[2];
            P.reset(new SlotPool());
            if (int e = P->setup(probs[c], kcs[c], tsize, (slots + nl - 1) / nl, elems, tile_cap, max_pixels))
                return e;
        }
    }
    return 0;
}

double Server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

static double server_clock(const void *ctx) { return ((const Server *)ctx)->now(); }

int Server::submit(ServeRequest r, int64_t *id) {
    const int c = r.k == 3 ? 1 : 0;
    if ((r.k != 2 && r.k != 3) || !pools_[c]) return FP_EINVAL;
    for (int a = 0; a < 3; ++a)
        if (r.L[a] < 1) return FP_ESHAPE;
    if (!r.x) return FP_EINVAL;
    if (!(r.gamma >= 0.0) || !std::isfinite(r.gamma) || !(r.tol >= 0.0) || !std::isfinite(r.tol) || r.max_iter < 1)
        return FP_EPARAM;
    {
        // the streaming kernels address each plane with 32-bit byte offsets
        const KernelChoice &kc = pools_[c]->kernel();
        if ((kc.stream || kc.s3t()) && !stream_fits(pools_[c]->tsize(), r.L[0] * r.L[1] * r.L[2])) return FP_ESHAPE;
    }
    if (!pools_[c]->fits_ever(r)) return FP_ENOMEM;
    // the admission's gather waits for the work on the caller's stream
    // before this point (the work that produces x)
    if (ev_free_.empty()) {
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming)) return FP_EDEVICE;
        ev_all_.push_back(ev);
        ev_free_.push_back(ev);
    }
    if (hipEventRecord(ev_free_.back(), r.caller)) return FP_EDEVICE;
    r.ready = ev_free_.back();
    ev_free_.pop_back();
    r.id = next_id_++;
    if (r.t_submit < 0.0) r.t_submit = now();
    queue_[c].push_back(r);
    if (id) *id = r.id;
    return FP_OK;
}

int Server::step(int *live, int *queued) {
    for (int c = 0; c < 2; ++c) {
        if (!pools_[c]) continue;
        const int nl = lanes(c);
        // admissions, first come first served; most are set up in groups. A
        // warm start goes to the lane that keeps its dual; others to the first
        // lane while it holds fewer than kLaneSpill live tiles (a light load
        // runs on one stream, as without lanes), then to the lane with fewer;
        // the second lane also takes what the first has no room for.
        while (!queue_[c].empty()) {
            int err = 0;
            const ServeRequest &r = queue_[c].front();
            SlotPool *order[2] = {lane(c, 0), nl > 1 ? lane(c, 1) : nullptr};
            if (nl > 1) {
                const bool swap = r.warm_from >= 0 ? order[1]->has_kept(r.warm_from)
                                                   : (order[0]->live_tiles() >= kLaneSpill &&
                                                      order[1]->live_tiles() < order[0]->live_tiles());
                if (swap) std::swap(order[0], order[1]);
                if (r.warm_from >= 0 && order[0]->has_kept(r.warm_from)) order[1] = nullptr;
            }
            bool ok = false;
            int why[2] = {SlotPool::kShortSlot, SlotPool::kShortSlot};
            for (int j = 0; j < 2 && !ok; ++j) {
                SlotPool *P = order[j];
                if (!P) continue;
                ok = P->groupable(r) ? P->stage(r, now(), &err, &why[j]) : P->try_admit(r, now(), &err, &why[j]);
                if (err) return err;
            }
            if (ok) {
                ev_free_.push_back(r.ready);  // its wait is enqueued
                queue_[c].pop_front();
                continue;
            }
            // Room from kept duals, in a lane whose only shortage was memory
            // (a slot and tiles were free): first duals other than the one
            // this request warms from; that one goes too when nothing else
            // would ever free room (its lane has no request live or in
            // flight): the request then starts cold instead of waiting for
            // good. A lane short of slots or tiles waits for completions and
            // keeps its duals.
            bool evicted = false;
            for (int j = 0; j < 2 && !evicted; ++j)
                if (order[j] && why[j] == SlotPool::kShortMem) evicted = order[j]->evict_one_dual(r.warm_from);
            if (!evicted && r.warm_from >= 0)
                for (int j = 0; j < 2 && !evicted; ++j) {
                    SlotPool *P = order[j];
                    if (P && why[j] == SlotPool::kShortMem && P->live() == 0 && P->in_flight() == 0)
                        evicted = P->evict_one_dual();
                }
            if (evicted) continue;
            break;
        }
        for (int i = 0; i < nl; ++i) {
            bool enq = false;
            if (int e = lane(c, i)->enqueue_chunk(&enq)) return e;
        }
    }
    // Retirement keeps one chunk in flight per pool while the older one is
    // retired (both, once nothing is live). In a server of both classes a
    // pool whose older chunk is done is retired at once, and while none is,
    // every pool due is polled until one is, so the 2D class's long chunks
    // (TT iterations per launch) do not hold back the 3D class's short ones
    // (one per launch): on the mixed trace (Poisson 200 requests/s) p99 37.3
    // -> 34.6 ms, the 3D requests' 55 -> 52 ms. Pools of one class (the two
    // 2D lanes) have alike chunks and are retired in turn, waiting for each:
    // polling them cost 2.5% of the served rate on 2048 queued 512^2 requests.
    const bool poll = pools_[0] && pools_[1];
    for (;;) {
        bool due = false, retired = false;
        for (int c = 0; c < 3; ++c) {
            SlotPool *P = pools_[c].get();
            if (!P || !(P->in_flight() >= 2 || (P->in_flight() > 0 && P->live() == 0))) continue;
            due = true;
            if (poll && !P->oldest_done()) continue;
            if (int e = P->retire(done_, server_clock, this)) return e;
            retired = true;
        }
        if (!due || retired) break;
        std::this_thread::yield();
    }
    int l = 0, q = 0;
    for (int c = 0; c < 3; ++c)
        if (pools_[c]) l += pools_[c]->live() + pools_[c]->in_flight();
    for (int c = 0; c < 2; ++c) q += (int)queue_[c].size();
    if (live) *live = l;
    if (queued) *queued = q;
    return 0;
}

int64_t Server::poll(ServeResult *out, int64_t max) {
    const int64_t n = std::min<int64_t>(max, (int64_t)done_.size());
    for (int64_t i = 0; i < n; ++i) out[i] = done_[(size_t)i];
    done_.erase(done_.begin(), done_.begin() + n);
    return n;
}

bool Server::release_dual(int64_t id) {
    for (auto &p : pools_)
        if (p && p->release_dual(id)) return true;
    return false;
}

int64_t Server::launches() const {
    int64_t n = 0;
    for (auto &p : pools_)
        if (p) n += p->launches;
    return n;
}

}  // namespace fpr
