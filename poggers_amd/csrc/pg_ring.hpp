// pg_ring.hpp -- FIFO ring allocator over the serving staging arena (host
// only).  Batches allocate one contiguous block at the head and are released
// from the tail in launch order, so the live blocks are always one or two
// contiguous runs of the arena; nothing is allocated on the device after
// start-up.  A request is admitted only when its batch's block fits here,
// which is how the server keeps its device memory within the HBM budget.
#pragma once

#include <cstdint>
#include <deque>
#include <utility>

namespace pg {

struct Ring {
    uint64_t cap = 0, head = 0;
    std::deque<std::pair<uint64_t, uint64_t>> live;  // [begin, end) per batch, oldest first

    // Largest block alloc() can return now.
    uint64_t largest_free() const {
        if (live.empty()) return cap;
        const uint64_t t = live.front().first;
        if (live.back().first < t) return t - head;  // wrapped: free is [head, t)
        const uint64_t a = cap - head, b = t;       // free is [head, cap) and [0, t)
        return a > b ? a : b;
    }
    bool alloc(uint64_t need, uint64_t &at) {
        if (need > cap) return false;
        if (live.empty()) {
            at = 0;
        } else {
            const uint64_t t = live.front().first;
            if (live.back().first < t) {
                if (t - head < need) return false;
                at = head;
            } else if (cap - head >= need) {
                at = head;
            } else if (t >= need) {
                at = 0;
            } else {
                return false;
            }
        }
        head = at + need;
        live.push_back({at, head});
        return true;
    }
    void release_front() {
        live.pop_front();
        if (live.empty()) head = 0;
    }
    // Undo the newest alloc() (head always ends the newest live block).
    void release_back() {
        live.pop_back();
        head = live.empty() ? 0 : live.back().second;
    }
};

}  // namespace pg
