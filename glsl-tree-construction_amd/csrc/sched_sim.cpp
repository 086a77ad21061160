// sched_sim.cpp -- discrete-event simulation of the server's policy
// (sched.hpp) with a modelled device, and the result-size function shared with
// the live server.  No GPU needed: this is how the policy is tested on a CPU.
#include <cmath>
#include <limits>
#include <queue>

#include "../../include/tc.h"
#include "result_layout.hpp"
#include "sched.hpp"

extern "C" uint64_t tc_tree_result_bytes(uint32_t n, int code_bits) {
    const uint64_t nodes = n >= 2 ? tc_octree_node_bound(n, code_bits) : 0;
    return TreeResultLayout(n, code_bits, nodes).total;
}

extern "C" int tc_server_simulate(const tc_sim_config* cfg, uint32_t n_req, const double* t_arrive,
                                  const uint32_t* n, double* t_start, double* t_done, uint32_t* batch_of,
                                  uint64_t* peak_reserved) {
    if (!cfg || (n_req && (!t_arrive || !n || !t_start || !t_done || !batch_of))) return TC_ERR_INVALID;
    if (cfg->n_streams == 0 || cfg->max_batch_clouds == 0 || (cfg->code_bits != 30 && cfg->code_bits != 63))
        return TC_ERR_INVALID;
    for (uint32_t i = 1; i < n_req; ++i)
        if (t_arrive[i] < t_arrive[i - 1]) return TC_ERR_INVALID;  // arrivals in time order
    Scheduler sched({cfg->max_batch_clouds, cfg->max_batch_points, cfg->result_budget_bytes});
    struct SlotState {
        bool busy = false;
        double until = 0;
        std::vector<SchedReq> reqs;
    };
    std::vector<SlotState> slots(cfg->n_streams);
    using Rel = std::pair<double, uint64_t>;  // (time, bytes)
    std::priority_queue<Rel, std::vector<Rel>, std::greater<Rel>> releases;
    const double inf = std::numeric_limits<double>::infinity();
    uint32_t next = 0, batches = 0;
    std::vector<SchedReq> batch;
    for (uint32_t i = 0; i < n_req; ++i) {
        t_start[i] = t_done[i] = -1.0;
        batch_of[i] = 0xFFFFFFFFu;
    }
    double t = n_req ? t_arrive[0] : 0.0;
    while (true) {
        // Arrivals (requests that can never be served are left at -1).
        while (next < n_req && t_arrive[next] <= t) {
            const uint64_t r = tc_tree_result_bytes(n[next], cfg->code_bits);
            if (n[next] <= cfg->max_batch_points && r <= cfg->result_budget_bytes)
                sched.push({next, n[next], r, t_arrive[next]});
            ++next;
        }
        // Completions, then releases due.
        for (SlotState& s : slots) {
            if (!s.busy || s.until > t) continue;
            for (const SchedReq& r : s.reqs) {
                t_done[r.id] = s.until;
                releases.push({s.until + cfg->hold_s, r.reserve});
            }
            s.reqs.clear();
            s.busy = false;
        }
        while (!releases.empty() && releases.top().first <= t) {
            sched.release(releases.top().second);
            releases.pop();
        }
        // Launches on free streams.
        for (SlotState& s : slots) {
            if (s.busy) continue;
            sched.take(batch, [](const SchedReq&) { return true; });
            if (batch.empty()) break;
            uint64_t pts = 0;
            for (const SchedReq& r : batch) {
                pts += r.n;
                t_start[r.id] = t;
                batch_of[r.id] = batches;
            }
            ++batches;
            s.busy = true;
            s.until = t + cfg->launch_s + cfg->per_point_s * (double)pts;
            s.reqs = batch;
        }
        // Next event.
        double tn = next < n_req ? t_arrive[next] : inf;
        for (const SlotState& s : slots)
            if (s.busy) tn = std::min(tn, s.until);
        if (!releases.empty()) tn = std::min(tn, releases.top().first);
        if (tn == inf) break;
        t = std::max(t, tn);
    }
    if (peak_reserved) *peak_reserved = sched.peak();
    return TC_OK;
}
