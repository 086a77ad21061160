// sched_sim.cpp -- discrete-event simulation of the server's policy
// (sched.hpp) with a modelled device, and the size functions (octree node
// bound, result bytes) shared with the live server and the pipeline.  No GPU
// needed: this is how the policy is tested on a CPU.
#include <algorithm>
#include <cmath>
#include <limits>
#include <queue>
#include <vector>

#include "../../include/tc.h"
#include "result_layout.hpp"
#include "sched.hpp"

// Octree node bound: 1 + sum over levels of min(8^L, n_keys/2) (each level has
// at most 8^L cells and at most n/2 nodes, since every node has at least two
// keys below it).
extern "C" uint64_t tc_octree_node_bound(uint64_t n_keys, int code_bits) {
    if (n_keys < 2) return 0;
    const uint64_t half = n_keys / 2;
    uint64_t total = 1, p = 1;
    const int levels = (code_bits - 1) / 3;
    for (int L = 1; L <= levels; ++L) {
        p = p < (1ull << 60) ? p * 8 : p;
        total += std::min(p, half);
    }
    return total;
}

extern "C" uint64_t tc_tree_result_bytes(uint32_t n, int code_bits) {
    const uint64_t nodes = n >= 2 ? tc_octree_node_bound(n, code_bits) : 0;
    return TreeResultLayout(n, code_bits, nodes).total;
}

extern "C" int tc_server_simulate_ex(const tc_sim_config* cfg, uint32_t n_req, const double* t_arrive,
                                     const uint32_t* n, const double* t_cancel, double* t_start, double* t_done,
                                     uint32_t* batch_of, int32_t* status, uint64_t* peak_reserved) {
    if (!cfg || (n_req && (!t_arrive || !n || !t_start || !t_done || !batch_of))) return TC_ERR_INVALID;
    if (cfg->n_streams == 0 || cfg->max_batch_clouds == 0 || (cfg->code_bits != 30 && cfg->code_bits != 63))
        return TC_ERR_INVALID;
    for (uint32_t i = 1; i < n_req; ++i)
        if (t_arrive[i] < t_arrive[i - 1]) return TC_ERR_INVALID;  // arrivals in time order
    SchedLimits lim{cfg->max_batch_clouds, cfg->max_batch_points, cfg->result_budget_bytes};
    lim.max_queued = cfg->max_queued;
    Scheduler sched(lim);
    // The result arena, as in the server (worst case held until release),
    // with injected failures.
    Arena arena;
    arena.init(cfg->result_budget_bytes + 255);
    uint64_t allocs = 0, fails_left = cfg->alloc_fail_count;
    const bool fail_forever = cfg->alloc_fail_count == 0xFFFFFFFFu;
    std::vector<uint64_t> off(n_req, 0);
    bool injected = false;  // a placement failed by injection since the last check
    auto place = [&](const SchedReq& r) {
        if (allocs >= cfg->alloc_fail_after && (fail_forever || fails_left > 0)) {
            if (!fail_forever) --fails_left;
            injected = true;
            return false;
        }
        uint64_t o;
        if (!arena.alloc(r.reserve, &o)) return false;
        off[r.id] = o;
        ++allocs;
        return true;
    };
    struct SlotState {
        bool busy = false;
        double until = 0;
        std::vector<SchedReq> reqs;
    };
    struct Rel {
        double t;
        uint64_t id, bytes;
        bool operator>(const Rel& o) const { return t > o.t; }
    };
    std::vector<SlotState> slots(cfg->n_streams);
    std::priority_queue<Rel, std::vector<Rel>, std::greater<Rel>> releases;
    const double inf = std::numeric_limits<double>::infinity();
    // Cancellations in time order.
    std::vector<uint32_t> cancels;
    if (t_cancel) {
        for (uint32_t i = 0; i < n_req; ++i)
            if (t_cancel[i] < inf) cancels.push_back(i);
        std::sort(cancels.begin(), cancels.end(), [&](uint32_t a, uint32_t b) { return t_cancel[a] < t_cancel[b]; });
    }
    uint32_t next = 0, next_cancel = 0, batches = 0;
    std::vector<SchedReq> batch;
    for (uint32_t i = 0; i < n_req; ++i) {
        t_start[i] = t_done[i] = -1.0;
        batch_of[i] = 0xFFFFFFFFu;
        if (status) status[i] = TC_OK;
    }
    int result = TC_OK;
    double t = n_req ? t_arrive[0] : 0.0;
    while (true) {
        // Arrivals: rejected when they can never be served or the queue is full.
        while (next < n_req && t_arrive[next] <= t) {
            const uint64_t r = tc_tree_result_bytes(n[next], cfg->code_bits);
            int st = TC_OK;
            if (n[next] > cfg->max_batch_points) st = TC_ERR_CAPACITY;
            else if (r > cfg->result_budget_bytes) st = TC_ERR_OVER_BUDGET;
            else if (!sched.push({next, n[next], r, t_arrive[next]})) st = TC_ERR_BUSY;
            if (status) status[next] = st;
            ++next;
        }
        // Cancellations of requests still waiting.
        while (next_cancel < cancels.size() && t_cancel[cancels[next_cancel]] <= t) {
            const uint32_t i = cancels[next_cancel++];
            if (sched.remove(i) && status) status[i] = TC_ERR_CANCELLED;
        }
        // Completions, then releases due.
        for (SlotState& s : slots) {
            if (!s.busy || s.until > t) continue;
            for (const SchedReq& r : s.reqs) {
                t_done[r.id] = s.until;
                releases.push({s.until + cfg->hold_s, r.id, r.reserve});
            }
            s.reqs.clear();
            s.busy = false;
        }
        while (!releases.empty() && releases.top().t <= t) {
            arena.release(off[releases.top().id], releases.top().bytes);
            sched.release(releases.top().bytes);
            releases.pop();
        }
        // Launches on free streams.
        bool any_busy = false;
        for (SlotState& s : slots) {
            if (s.busy) {
                any_busy = true;
                continue;
            }
            sched.take(batch, place);
            if (batch.empty()) continue;
            uint64_t pts = 0;
            for (const SchedReq& r : batch) {
                pts += r.n;
                t_start[r.id] = t;
                batch_of[r.id] = batches;
            }
            ++batches;
            s.busy = true;
            any_busy = true;
            s.until = t + cfg->launch_s + cfg->per_point_s * (double)pts;
            s.reqs = batch;
        }
        // Next event.
        double tn = next < n_req ? t_arrive[next] : inf;
        if (next_cancel < cancels.size()) tn = std::min(tn, t_cancel[cancels[next_cancel]]);
        for (const SlotState& s : slots)
            if (s.busy) tn = std::min(tn, s.until);
        if (!releases.empty()) tn = std::min(tn, releases.top().t);
        if (tn == inf) {
            if (sched.queued() == 0) break;
            // Requests wait and no event is left that could free memory: after
            // a finite run of injected failures the placement is retried at
            // once (the live server retries on its next wake-up); a permanent
            // one leaves them.
            if (injected && !fail_forever) {
                injected = false;
                continue;
            }
            result = TC_ERR_NOMEM;
            if (status) {
                for (uint32_t i = 0; i < n_req; ++i)
                    if (status[i] == TC_OK && t_start[i] < 0) status[i] = TC_ERR_NOMEM;
            }
            break;
        }
        t = std::max(t, tn);
    }
    if (peak_reserved) *peak_reserved = sched.peak();
    return result;
}

extern "C" int tc_server_simulate(const tc_sim_config* cfg, uint32_t n_req, const double* t_arrive,
                                  const uint32_t* n, double* t_start, double* t_done, uint32_t* batch_of,
                                  uint64_t* peak_reserved) {
    return tc_server_simulate_ex(cfg, n_req, t_arrive, n, nullptr, t_start, t_done, batch_of, nullptr,
                                 peak_reserved);
}
