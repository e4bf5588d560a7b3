// pg_plan.hpp -- which queued requests form the next served batch (host
// only, no HIP calls, so the CPU tests can drive it directly).
//
// Unordered: the longest prefix of the FIFO queue that fits the staging
// space, the key cap and the request cap.
//
// Per-table ordered: the queue is scanned in order.  A request that
// conflicts with the batch's access to its table (a write with anything, a
// read with a write) blocks that table for the rest of the scan, so no
// later request of that table can overtake it; requests of other tables may
// still join.  Skipped requests keep their queue positions and lead the
// next scan.  With the batch-to-batch waits of the launcher (a batch waits
// for every earlier in-flight batch it conflicts with), every pair of
// conflicting requests of one table then takes effect in submission order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <utility>
#include <vector>

#include "../../include/poggers_amd.h"

namespace pg {

// Per-table access of a batch: bit 0 reads (query), bit 1 writes (insert,
// remove).  Two accesses to one table conflict unless both only read.
constexpr uint8_t kRead = 1, kWrite = 2;
inline uint8_t access_of(uint32_t op) { return op == PG_OP_QUERY ? kRead : kWrite; }
inline bool conflicts(uint8_t a, uint8_t b) { return ((a | b) & kWrite) ? (a && b) : false; }

struct PlanLimits {
    uint64_t free_bytes;   // largest staging block the arena can give now
    uint64_t max_keys;     // a batch's key cap (a larger single request still goes alone)
    size_t max_reqs;       // a batch's request cap
    bool ordered;          // per-table ordering
};

struct Plan {
    std::vector<size_t> pick;                          // queue positions, ascending
    std::vector<std::pair<uint32_t, uint8_t>> access;  // tables touched (ordered only)
    uint64_t keys = 0, need = 0;
};

// `queue[i]` has fields table, op, n (keys) and need (staging bytes);
// meta_bytes(k, nk) is the batch's fixed metadata for k requests over nk
// keys.  `acc` and `blocked` are zeroed per-table scratch of at least the
// number of tables (ordered mode only) and are zero again on return.
template <class Queue, class MetaBytes>
void plan_batch(const Queue &queue, const PlanLimits &lim, MetaBytes meta_bytes, std::vector<uint8_t> &acc,
                std::vector<uint8_t> &blocked, Plan &plan) {
    plan.pick.clear();
    plan.access.clear();
    plan.keys = plan.need = 0;
    size_t take = 0;
    if (!lim.ordered) {
        while (take < queue.size() && take < lim.max_reqs) {
            const auto &q = queue[take];
            if (meta_bytes(take + 1, plan.keys + q.n) + plan.need + q.need > lim.free_bytes) break;
            if (take > 0 && plan.keys + q.n > lim.max_keys) break;
            plan.need += q.need;
            plan.keys += q.n;
            plan.pick.push_back(take++);
        }
        return;
    }
    std::vector<uint32_t> marked;
    const size_t window = 4 * lim.max_reqs;
    for (size_t i = 0; i < queue.size() && i < window && take < lim.max_reqs; i++) {
        const auto &q = queue[i];
        if (blocked[q.table]) continue;
        const uint8_t a = access_of(q.op);
        if (conflicts(acc[q.table], a)) {
            blocked[q.table] = 1;
            marked.push_back(q.table);
            continue;
        }
        if (meta_bytes(take + 1, plan.keys + q.n) + plan.need + q.need > lim.free_bytes) break;
        if (take > 0 && plan.keys + q.n > lim.max_keys) break;
        if (!acc[q.table]) plan.access.push_back({q.table, 0});
        acc[q.table] |= a;
        plan.need += q.need;
        plan.keys += q.n;
        plan.pick.push_back(i);
        take++;
    }
    for (uint32_t t : marked) blocked[t] = 0;
    for (auto &tu : plan.access) {
        tu.second = acc[tu.first];
        acc[tu.first] = 0;
    }
}

}  // namespace pg
