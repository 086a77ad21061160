// server.cpp -- serving loop for tree builds (no reference counterpart: the
// reference builds one cloud per blocking call, src/main.rs:60-170).
//
// Requests name device point buffers.  A scheduler thread owned by the server
// keeps n_streams batch slots, each with its own context, stream and batch
// workspace, and a HIP graph of the whole segmented pipeline captured once.
// Whenever a slot is free it takes the requests waiting at that moment (in
// arrival order, within the batch limits and the memory budget; sched.hpp),
// uploads their plan (per-cloud input buffers and result buffers) and replays
// the graph.  Requests arriving meanwhile go out on the next free slot, so
// batches overlap on the device.  Result memory comes from one arena sized by
// the budget: the worst case is reserved at admission and cut back to the
// actual node count when the batch ends; the client returns it with
// tc_server_release.
#include <hip/hip_runtime.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/tc.h"
#include "result_layout.hpp"
#include "sched.hpp"

namespace {

using Clock = std::chrono::steady_clock;

struct Live {
    SchedReq r;
    const float* pts;
    float min_coord, range;
    uint64_t off;  // result allocation in the arena
    double t_start;
};

struct Slot {
    tc_ctx* ctx = nullptr;
    hipStream_t stream = nullptr;
    tc_batch* batch = nullptr;
    hipEvent_t done = nullptr;
    uint32_t* h_ub = nullptr;  // pinned [max_clouds + 1]
    uint32_t* h_nb = nullptr;
    bool busy = false;
    std::vector<Live> reqs;
    // Launch staging.
    std::vector<uint32_t> counts;
    std::vector<float> mins, ranges;
    std::vector<const float*> in;
    std::vector<tc_tree_ptrs> out;
};

struct Input {
    const float* pts;
    float min_coord, range;
};

struct Held {
    uint64_t off, bytes;
};

}  // namespace

struct tc_server {
    tc_server_config cfg{};
    int device = 0;
    Clock::time_point t0;
    std::vector<Slot> slots;
    char* arena_base = nullptr;
    Arena arena;
    Scheduler sched{SchedLimits{1, 1, 0}};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    uint64_t gen = 0;  // bumped by submit / release / stop
    bool stop = false;
    std::thread thr;
    std::atomic<uint64_t> next_id{1};
    std::unordered_map<uint64_t, Input> inputs;
    std::vector<tc_server_result> finished;
    std::unordered_map<uint64_t, Held> held;
    // Statistics.
    uint64_t submitted = 0, completed = 0, rejected = 0, cancelled = 0, batches = 0, clouds_launched = 0,
             points_done = 0;
    std::vector<double> lat_ms;
    double first_submit = -1.0, last_done = 0.0;

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

tc_tree_ptrs ptrs_at(tc_server* s, uint64_t off, uint32_t n) {
    const uint64_t nodes = n >= 2 ? tc_octree_node_bound(n, s->cfg.code_bits) : 0;
    return TreeResultLayout(n, s->cfg.code_bits, nodes).at(s->arena_base + off);
}

// Called with the lock held.  Returns false if the slot could not be launched
// (its requests are then failed).
bool launch(tc_server* s, Slot& sl, std::vector<SchedReq>& batch, std::vector<uint64_t>& offs,
            std::unique_lock<std::mutex>& lk) {
    const uint32_t nc = (uint32_t)batch.size();
    sl.reqs.clear();
    sl.counts.resize(nc);
    sl.mins.resize(nc);
    sl.ranges.resize(nc);
    sl.in.resize(nc);
    sl.out.resize(nc);
    for (uint32_t j = 0; j < nc; ++j) {
        const Input in = s->inputs[batch[j].id];
        s->inputs.erase(batch[j].id);
        sl.reqs.push_back({batch[j], in.pts, in.min_coord, in.range, offs[j], 0.0});
        sl.counts[j] = batch[j].n;
        sl.mins[j] = in.min_coord;
        sl.ranges[j] = in.range;
        sl.in[j] = in.pts;
        sl.out[j] = ptrs_at(s, offs[j], batch[j].n);
    }
    sl.busy = true;
    lk.unlock();
    const double t = s->now();
    int st = tc_batch_set_io(sl.batch, nc, sl.counts.data(), sl.mins.data(), sl.ranges.data(), sl.in.data(),
                             sl.out.data(), sl.stream);
    if (st == TC_OK) st = tc_batch_replay(sl.batch, sl.stream);
    if (st == TC_OK) st = tc_batch_read_info_async(sl.batch, sl.h_ub, sl.h_nb, sl.stream);
    if (st == TC_OK && hipEventRecord(sl.done, sl.stream) != hipSuccess) st = TC_ERR_HIP;
    // A launch that failed part-way may have queued kernels that still write
    // into the requests' result ranges: wait for the stream before the caller
    // returns those ranges to the arena.
    if (st != TC_OK) (void)hipStreamSynchronize(sl.stream);
    lk.lock();
    for (Live& l : sl.reqs) l.t_start = t;
    ++s->batches;
    s->clouds_launched += nc;
    return st == TC_OK;
}

// Called with the lock held, once the slot's event has completed (ok) or its
// launch failed (!ok).
void complete(tc_server* s, Slot& sl, bool ok) {
    const double t = s->now();
    for (uint32_t j = 0; j < sl.reqs.size(); ++j) {
        const Live& l = sl.reqs[j];
        tc_server_result res{};
        res.id = l.r.id;
        res.n_points = l.r.n;
        res.t_submit = l.r.t_submit;
        res.t_start = l.t_start;
        res.t_done = t;
        uint64_t keep = 0;
        if (ok) {
            res.status = TC_OK;
            res.n_unique = sl.h_ub[j + 1] - sl.h_ub[j];
            res.n_nodes = sl.h_nb[j + 1] - sl.h_nb[j];
            keep = TreeResultLayout(l.r.n, s->cfg.code_bits, res.n_nodes).total;
            res.tree = ptrs_at(s, l.off, l.r.n);
        } else {
            res.status = TC_ERR_HIP;
        }
        // Cut the worst-case reservation back to what the result occupies.
        s->arena.shrink(l.off, l.r.reserve, keep);
        s->sched.release(l.r.reserve - std::min(l.r.reserve, Arena::up(keep ? keep : 1)));
        if (ok) {
            s->held[l.r.id] = {l.off, Arena::up(keep ? keep : 1)};
        } else {
            s->arena.release(l.off, Arena::up(1));
            s->sched.release(Arena::up(1));
        }
        s->finished.push_back(res);
        ++s->completed;
        s->points_done += l.r.n;
        s->lat_ms.push_back((t - l.r.t_submit) * 1e3);
        s->last_done = t;
    }
    sl.reqs.clear();
    sl.busy = false;
    s->cv_done.notify_all();
}

void scheduler_loop(tc_server* s) {
    (void)hipSetDevice(s->device);
    std::vector<SchedReq> batch;
    std::vector<uint64_t> offs;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        bool progressed = false, any_busy = false;
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                complete(s, sl, true);
                progressed = true;
            } else if (q != hipErrorNotReady) {
                (void)hipStreamSynchronize(sl.stream);  // nothing may still write the result ranges
                complete(s, sl, false);
                progressed = true;
            } else {
                any_busy = true;
            }
        }
        for (Slot& sl : s->slots) {
            if (sl.busy || s->stop) continue;
            offs.clear();
            s->sched.take(batch, [&](const SchedReq& r) {
                uint64_t off;
                if (!s->arena.alloc(r.reserve, &off)) return false;
                offs.push_back(off);
                return true;
            });
            if (batch.empty()) break;
            if (!launch(s, sl, batch, offs, lk)) complete(s, sl, false);
            progressed = true;
            any_busy = any_busy || sl.busy;
        }
        if (s->stop && !any_busy) break;  // requests still queued are dropped
        if (progressed) continue;
        if (any_busy) {
            // A batch is running: poll its event again after a short sleep,
            // or at once when a request or a release comes in.  The wait
            // gives the lock up for the whole sleep; an unlock-yield-relock
            // spin holds it almost all the time and can win it back before a
            // waiting submitter or poller gets it (std::mutex is not fair).
            const uint64_t g = s->gen;
            s->cv_work.wait_for(lk, std::chrono::microseconds(20), [&] { return s->gen != g; });
        } else {
            const uint64_t g = s->gen;
            s->cv_work.wait(lk, [&] { return s->gen != g; });
        }
    }
}

void destroy_slots(tc_server* s) {
    for (Slot& sl : s->slots) {
        if (sl.batch) tc_batch_destroy(sl.batch);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_ub) (void)hipHostFree(sl.h_ub);
        if (sl.h_nb) (void)hipHostFree(sl.h_nb);
        if (sl.ctx) tc_destroy(sl.ctx);
    }
    s->slots.clear();
}

double percentile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    const size_t k = (size_t)(q * (double)(v.size() - 1) + 0.5);
    std::nth_element(v.begin(), v.begin() + (ptrdiff_t)k, v.end());
    return v[k];
}

void spin_until(Clock::time_point t) {
    while (true) {
        const auto now = Clock::now();
        if (now >= t) return;
        if (t - now > std::chrono::microseconds(300))
            std::this_thread::sleep_for(t - now - std::chrono::microseconds(200));
        else
            std::this_thread::yield();
    }
}

}  // namespace

extern "C" void tc_server_config_default(tc_server_config* cfg) {
    if (!cfg) return;
    cfg->mem_budget_bytes = 64ull << 30;
    cfg->n_streams = 2;
    cfg->max_batch_clouds = 256;
    cfg->max_batch_points = 1u << 24;
    cfg->code_bits = 30;
    cfg->link_leaves = 1;
    cfg->max_queued = 1u << 16;
}

extern "C" int tc_server_create(tc_ctx* ctx, const tc_server_config* cfg_in, tc_server** out) {
    if (!ctx || !out) return TC_ERR_INVALID;
    *out = nullptr;
    tc_server_config cfg;
    tc_server_config_default(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (cfg.n_streams == 0 || cfg.n_streams > 16 || cfg.max_batch_clouds == 0 || cfg.max_batch_points == 0 ||
        (cfg.code_bits != 30 && cfg.code_bits != 63))
        return TC_ERR_INVALID;
    if (cfg.max_batch_points > TC_MAX_SORT_N) return TC_ERR_TOO_LARGE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return TC_ERR_HIP;
    const uint64_t ws = (uint64_t)cfg.n_streams *
                        tc_batch_memory_bytes(cfg.max_batch_clouds, cfg.max_batch_points, cfg.code_bits);
    if (ws >= cfg.mem_budget_bytes) return TC_ERR_OVER_BUDGET;
    const uint64_t arena_bytes = (cfg.mem_budget_bytes - ws) & ~(uint64_t)255;
    tc_server* s = new (std::nothrow) tc_server();
    if (!s) return TC_ERR_NOMEM;
    s->cfg = cfg;
    s->device = dev;
    s->t0 = Clock::now();
    s->sched = Scheduler(SchedLimits{cfg.max_batch_clouds, cfg.max_batch_points, arena_bytes, cfg.max_queued});
    int st = TC_OK;
    if (hipMalloc((void**)&s->arena_base, arena_bytes) != hipSuccess) st = TC_ERR_NOMEM;
    s->arena.init(arena_bytes);
    s->slots.resize(cfg.n_streams);
    for (Slot& sl : s->slots) {
        if (st != TC_OK) break;
        st = tc_create(&sl.ctx, dev);
        if (st == TC_OK && hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess) st = TC_ERR_HIP;
        if (st == TC_OK && hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) != hipSuccess) st = TC_ERR_HIP;
        if (st == TC_OK && (hipHostMalloc((void**)&sl.h_ub, ((size_t)cfg.max_batch_clouds + 1) * 4) != hipSuccess ||
                            hipHostMalloc((void**)&sl.h_nb, ((size_t)cfg.max_batch_clouds + 1) * 4) != hipSuccess))
            st = TC_ERR_NOMEM;
        if (st == TC_OK) st = tc_batch_create(sl.ctx, cfg.max_batch_clouds, cfg.max_batch_points, cfg.code_bits,
                                              &sl.batch);
        // Capture the pipeline once in the per-cloud buffer mode; every batch
        // then uploads its plan and replays the graph.
        const float* dummy_in = nullptr;
        tc_tree_ptrs dummy_out{};
        if (st == TC_OK) st = tc_batch_set_io(sl.batch, 0, nullptr, nullptr, nullptr, &dummy_in, &dummy_out, sl.stream);
        if (st == TC_OK) st = tc_batch_capture(sl.batch, nullptr, cfg.link_leaves, sl.stream);
        if (st == TC_OK && hipStreamSynchronize(sl.stream) != hipSuccess) st = TC_ERR_HIP;
    }
    if (st != TC_OK) {
        destroy_slots(s);
        if (s->arena_base) (void)hipFree(s->arena_base);
        delete s;
        return st;
    }
    s->thr = std::thread(scheduler_loop, s);
    *out = s;
    return TC_OK;
}

extern "C" int tc_server_drain(tc_server* s) {
    if (!s) return TC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->completed >= s->submitted; });
    return TC_OK;
}

extern "C" int tc_server_destroy(tc_server* s) {
    if (!s) return TC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        ++s->gen;
    }
    s->cv_work.notify_all();
    if (s->thr.joinable()) s->thr.join();
    destroy_slots(s);
    if (s->arena_base) (void)hipFree(s->arena_base);
    delete s;
    return TC_OK;
}

extern "C" int tc_server_submit(tc_server* s, const float* pts, uint32_t n, float min_coord, float range,
                                uint64_t* id) {
    if (!s || !id || (n && (!pts || (reinterpret_cast<uintptr_t>(pts) & 15u))) || !(range != 0.0f))
        return TC_ERR_INVALID;
    const uint64_t reserve = tc_tree_result_bytes(n, s->cfg.code_bits);
    std::lock_guard<std::mutex> lk(s->mu);
    if (n > s->cfg.max_batch_points) {
        ++s->rejected;
        return TC_ERR_CAPACITY;
    }
    if (reserve > s->sched.limits().budget) {
        ++s->rejected;
        return TC_ERR_OVER_BUDGET;
    }
    const double t = s->now();
    const uint64_t rid = s->next_id;
    if (!s->sched.push({rid, n, reserve, t})) {
        ++s->rejected;
        return TC_ERR_BUSY;  // back-pressure: max_queued requests already wait
    }
    ++s->next_id;
    s->inputs[rid] = {pts, min_coord, range};
    if (s->first_submit < 0) s->first_submit = t;
    ++s->submitted;
    ++s->gen;
    s->cv_work.notify_one();
    *id = rid;
    return TC_OK;
}

extern "C" int tc_server_cancel(tc_server* s, uint64_t id) {
    if (!s) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->sched.remove(id)) {
        s->inputs.erase(id);
        tc_server_result res{};
        res.id = id;
        res.status = TC_ERR_CANCELLED;
        res.t_done = s->now();
        s->finished.push_back(res);
        ++s->cancelled;
        ++s->completed;
        s->cv_done.notify_all();
        return TC_OK;
    }
    for (const Slot& sl : s->slots)
        for (const Live& l : sl.reqs)
            if (l.r.id == id) return TC_ERR_BUSY;  // being built
    return TC_ERR_NOT_FOUND;
}

extern "C" int tc_server_poll(tc_server* s, tc_server_result* out, uint32_t max, uint32_t* n_out, int wait_ms) {
    if (!s || !n_out || (max && !out)) return TC_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->finished.empty() && wait_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), [&] { return !s->finished.empty(); });
    const uint32_t k = (uint32_t)std::min<size_t>(max, s->finished.size());
    for (uint32_t i = 0; i < k; ++i) out[i] = s->finished[i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + k);
    *n_out = k;
    return TC_OK;
}

extern "C" int tc_server_release(tc_server* s, uint64_t id) {
    if (!s) return TC_ERR_INVALID;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->held.find(id);
        if (it == s->held.end()) return TC_ERR_NOT_FOUND;
        s->arena.release(it->second.off, it->second.bytes);
        s->sched.release(it->second.bytes);
        s->held.erase(it);
        ++s->gen;
    }
    s->cv_work.notify_one();
    return TC_OK;
}

extern "C" int tc_server_stats_get(tc_server* s, tc_server_stats* st) {
    if (!s || !st) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    *st = tc_server_stats{};
    st->submitted = s->submitted;
    st->completed = s->completed;
    st->rejected = s->rejected;
    st->batches = s->batches;
    st->points_completed = s->points_done;
    st->mean_clouds_per_batch = s->batches ? (double)s->clouds_launched / (double)s->batches : 0.0;
    st->latency_p50_ms = percentile(s->lat_ms, 0.50);
    st->latency_p99_ms = percentile(s->lat_ms, 0.99);
    double sum = 0, mx = 0;
    for (double v : s->lat_ms) {
        sum += v;
        mx = std::max(mx, v);
    }
    st->latency_mean_ms = s->lat_ms.empty() ? 0.0 : sum / (double)s->lat_ms.size();
    st->latency_max_ms = mx;
    st->first_submit_s = s->first_submit;
    st->last_done_s = s->last_done;
    st->reserved_bytes = s->sched.reserved();
    st->peak_reserved_bytes = s->sched.peak();
    st->result_budget_bytes = s->sched.limits().budget;
    st->cancelled = s->cancelled;
    st->queued = s->sched.queued();
    return TC_OK;
}

extern "C" int tc_server_stats_reset(tc_server* s) {
    if (!s) return TC_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    s->submitted -= s->completed;
    s->completed = s->rejected = s->cancelled = s->batches = s->clouds_launched = s->points_done = 0;
    s->lat_ms.clear();
    s->first_submit = -1.0;
    s->last_done = 0.0;
    return TC_OK;
}

extern "C" int tc_server_run_trace(tc_server* s, uint32_t n_req, const double* t_arrive, const float* const* pts,
                                   const uint32_t* n, float min_coord, float range, double* latency_s) {
    if (!s || (n_req && (!t_arrive || !pts || !n || !latency_s))) return TC_ERR_INVALID;
    std::mutex idm;
    std::unordered_map<uint64_t, uint32_t> index;
    std::vector<double> t_sub(n_req, 0.0);
    std::atomic<int> err{TC_OK};
    const Clock::time_point start = Clock::now();
    const double base = s->now();
    std::thread submitter([&] {
        (void)hipSetDevice(s->device);
        for (uint32_t i = 0; i < n_req; ++i) {
            spin_until(start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t_arrive[i])));
            uint64_t id = 0;
            // Hold the index lock across the submission, so the poller can
            // map the id even when the result comes back at once.
            std::lock_guard<std::mutex> lk(idm);
            const int st = tc_server_submit(s, pts[i], n[i], min_coord, range, &id);
            if (st != TC_OK) {
                err = st;
                break;
            }
            index[id] = i;
        }
    });
    std::vector<tc_server_result> res(256);
    uint32_t done = 0;
    while (done < n_req && err == TC_OK) {
        uint32_t k = 0;
        tc_server_poll(s, res.data(), (uint32_t)res.size(), &k, 2);
        for (uint32_t j = 0; j < k; ++j) {
            uint32_t i;
            {
                std::lock_guard<std::mutex> lk(idm);
                auto it = index.find(res[j].id);
                if (it == index.end()) {  // not part of this trace (the server should be otherwise idle)
                    tc_server_release(s, res[j].id);
                    continue;
                }
                i = it->second;
            }
            latency_s[i] = res[j].t_done - (base + t_arrive[i]);
            if (res[j].status != TC_OK) err = res[j].status;
            tc_server_release(s, res[j].id);
            ++done;
        }
    }
    submitter.join();
    return err;
}

extern "C" int tc_plain_run_trace(tc_ctx* ctx, uint32_t n_req, const double* t_arrive, const float* const* pts,
                                  const uint32_t* n, float min_coord, float range, int code_bits,
                                  double* latency_s) {
    if (!ctx || (n_req && (!t_arrive || !pts || !n || !latency_s))) return TC_ERR_INVALID;
    uint32_t cap = 1;
    for (uint32_t i = 0; i < n_req; ++i) cap = std::max(cap, n[i]);
    tc_tree tree{};
    int st = tc_tree_alloc(ctx, cap, code_bits, &tree);
    if (st != TC_OK) return st;
    hipStream_t stream = nullptr;
    if (hipStreamCreateWithFlags(&stream, hipStreamNonBlocking) != hipSuccess) {
        tc_tree_free(ctx, &tree);
        return TC_ERR_HIP;
    }
    const Clock::time_point start = Clock::now();
    for (uint32_t i = 0; i < n_req && st == TC_OK; ++i) {
        const Clock::time_point arrive =
            start + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(t_arrive[i]));
        spin_until(arrive);  // returns at once when the loop is behind
        st = tc_build_tree(ctx, pts[i], n[i], min_coord, range, 1, &tree, stream);
        if (st == TC_OK && hipStreamSynchronize(stream) != hipSuccess) st = TC_ERR_HIP;
        latency_s[i] = std::chrono::duration<double>(Clock::now() - arrive).count();
    }
    (void)hipStreamDestroy(stream);
    tc_tree_free(ctx, &tree);
    return st;
}
