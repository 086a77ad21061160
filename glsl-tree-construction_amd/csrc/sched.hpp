// sched.hpp -- admission and batching policy of the tree-build server, and the
// allocator of its result memory.  Plain C++ (no HIP), shared by the live
// server (server.cpp) and its discrete-event simulation (sched_sim.cpp).
#pragma once

#include <algorithm>
#include <cstdint>
#include <deque>
#include <map>
#include <vector>

struct SchedReq {
    uint64_t id;
    uint32_t n;        // points
    uint64_t reserve;  // result bytes held from admission until release (worst case until the build ends)
    double t_submit;
};

struct SchedLimits {
    uint32_t max_clouds;  // clouds per batch
    uint32_t max_points;  // points per batch
    uint64_t budget;      // bytes of results that may be reserved at once
    uint32_t max_queued = 0;  // requests waiting for launch (0: unbounded)
};

// First come, first served: a batch is the longest prefix of the waiting queue
// that fits the batch limits and the remaining budget.  A request that does
// not fit waits at the head (later ones never overtake it), so no request
// starves and results come out in arrival order within each stream.
class Scheduler {
  public:
    explicit Scheduler(SchedLimits lim) : lim_(lim) {}

    // Queues a request; false (and nothing queued) when max_queued requests
    // are already waiting: the caller reports back-pressure.
    bool push(const SchedReq& r) {
        if (lim_.max_queued && q_.size() >= lim_.max_queued) return false;
        q_.push_back(r);
        return true;
    }

    // Removes a waiting request (cancellation); false if it is not queued.
    bool remove(uint64_t id) {
        for (auto it = q_.begin(); it != q_.end(); ++it) {
            if (it->id == id) {
                q_.erase(it);
                return true;
            }
        }
        return false;
    }

    // Appends the next batch to out (cleared first).  can_place(r) may refuse
    // a request (its memory could not be placed): the batch then ends there.
    template <typename F>
    void take(std::vector<SchedReq>& out, F&& can_place) {
        out.clear();
        uint64_t pts = 0;
        while (!q_.empty() && out.size() < lim_.max_clouds) {
            const SchedReq& r = q_.front();
            if (pts + r.n > lim_.max_points) break;
            if (reserved_ + r.reserve > lim_.budget) break;
            if (!can_place(r)) break;
            reserved_ += r.reserve;
            peak_ = std::max(peak_, reserved_);
            pts += r.n;
            out.push_back(r);
            q_.pop_front();
        }
    }

    void release(uint64_t bytes) { reserved_ -= std::min(bytes, reserved_); }
    size_t queued() const { return q_.size(); }
    uint64_t reserved() const { return reserved_; }
    uint64_t peak() const { return peak_; }
    const SchedLimits& limits() const { return lim_; }

  private:
    SchedLimits lim_;
    std::deque<SchedReq> q_;
    uint64_t reserved_ = 0, peak_ = 0;
};

// First-fit allocator over [0, size) in 256-byte units, coalescing on free.
class Arena {
  public:
    static uint64_t up(uint64_t x) { return (x + 255) & ~(uint64_t)255; }

    void init(uint64_t size) {
        free_.clear();
        size_ = size & ~(uint64_t)255;
        if (size_) free_[0] = size_;
    }

    bool alloc(uint64_t bytes, uint64_t* off) {
        bytes = up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }

    void release(uint64_t off, uint64_t bytes) {
        bytes = up(bytes ? bytes : 1);
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + bytes == next->first) {
            bytes += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += bytes;
                return;
            }
        }
        free_[off] = bytes;
    }

    // Keep the first new_bytes of an allocation, return the tail.
    void shrink(uint64_t off, uint64_t old_bytes, uint64_t new_bytes) {
        old_bytes = up(old_bytes ? old_bytes : 1);
        new_bytes = up(new_bytes ? new_bytes : 1);
        if (new_bytes < old_bytes) release(off + new_bytes, old_bytes - new_bytes);
    }

    uint64_t free_bytes() const {
        uint64_t t = 0;
        for (const auto& kv : free_) t += kv.second;
        return t;
    }
    uint64_t size() const { return size_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0;
};
