// xc_sched.hpp -- host-only pieces of the serving loop's scheduler (no HIP,
// so they build and are tested on the CPU as well): the first-fit arena
// allocator and the request queue (arrival order, deadlines, cancellation).
#pragma once

#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <iterator>
#include <map>
#include <unordered_map>

#include "../../include/xcorr.h"

namespace xc {
namespace sched {

using Clock = std::chrono::steady_clock;

inline double us_between(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::micro>(b - a).count();
}

// First-fit allocator over one device allocation (offsets, 256-byte aligned);
// freed blocks merge with their free neighbours.
struct ArenaAlloc {
    char *base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    std::map<size_t, size_t> free_;  // offset -> bytes
    void init(char *b, size_t c)
    {
        base = b;
        cap = c & ~(size_t)255;
        used = peak = 0;
        free_.clear();
        if (cap) free_[0] = cap;
    }
    static size_t round(size_t n) { return (n + 255) & ~(size_t)255; }
    bool alloc(size_t n, size_t &off)
    {
        n = round(n);
        if (n == 0) n = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                off = it->first;
                const size_t rest = it->second - n;
                free_.erase(it);
                if (rest) free_[off + n] = rest;
                used += n;
                peak = std::max(peak, used);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = round(n);
        if (n == 0) n = 256;
        used -= n;
        auto it = free_.emplace(off, n).first;
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
};

struct Req {
    xc_pair pair;
    Clock::time_point t_submit, deadline;
    bool has_deadline = false;
    bool launched = false;  // handed to a batch (no longer cancellable)
    bool done = false;
    double latency_us = 0.0;
    xc_result res;
};

// Requests by ticket, and the queue of those not yet launched, in arrival order.
struct Queue {
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Req> reqs;
    uint64_t next_id = 1;

    // deadline_us <= 0: none.
    uint64_t submit(const xc_pair &p, Clock::time_point now, int64_t deadline_us)
    {
        const uint64_t id = next_id++;
        Req r;
        r.pair = p;
        r.t_submit = now;
        r.has_deadline = deadline_us > 0;
        r.deadline = now + std::chrono::microseconds(deadline_us > 0 ? deadline_us : 0);
        memset(&r.res, 0, sizeof r.res);
        reqs[id] = r;
        pending.push_back(id);
        return id;
    }
    void complete(uint64_t id, int status, Clock::time_point now)
    {
        Req &r = reqs[id];
        memset(&r.res, 0, sizeof r.res);
        r.res.status = status;
        r.latency_us = us_between(r.t_submit, now);
        r.done = true;
    }
    // Complete every queued request whose deadline has passed with XC_E_TIMEOUT; returns how many.
    int expire(Clock::time_point now)
    {
        int n = 0;
        for (auto it = pending.begin(); it != pending.end();) {
            Req &r = reqs[*it];
            if (r.has_deadline && now >= r.deadline) {
                complete(*it, XC_E_TIMEOUT, now);
                it = pending.erase(it);
                ++n;
            } else {
                ++it;
            }
        }
        return n;
    }
    // XC_OK: the request was still queued and is now completed with XC_E_CANCELLED; XC_E_NOTREADY: it is on
    // the device or done (its result stands); XC_E_ARG: unknown ticket.
    int cancel(uint64_t id, Clock::time_point now)
    {
        auto it = reqs.find(id);
        if (it == reqs.end()) return XC_E_ARG;
        if (it->second.done || it->second.launched) return XC_E_NOTREADY;
        auto q = std::find(pending.begin(), pending.end(), id);
        if (q == pending.end()) return XC_E_NOTREADY;
        pending.erase(q);
        complete(id, XC_E_CANCELLED, now);
        return XC_OK;
    }
    // Earliest deadline among queued requests (has = false when none has one).
    Clock::time_point next_deadline(bool &has) const
    {
        has = false;
        Clock::time_point t{};
        for (uint64_t id : pending) {
            const Req &r = reqs.at(id);
            if (r.has_deadline && (!has || r.deadline < t)) {
                t = r.deadline;
                has = true;
            }
        }
        return t;
    }
};

}  // namespace sched
}  // namespace xc
