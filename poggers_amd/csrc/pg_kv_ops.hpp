// pg_kv_ops.hpp -- per-key device ops for key_val_pair<u64,u64>
// tables (representations/key_val_pair.cuh:63-171), gfx950 wave64.
//
// Layout in HBM: 16 B slots {key, val}, B contiguous slots per bucket, bucket
// b at slot b*B.  Empty key 0, tombstone key ~0.
//
// Two tile shapes, chosen per table:
//   * stripe mode (B > 1): lane r of a P-lane tile owns slot stripe*P + r of
//     the probe bucket and loads it with one 16 B global_load_dwordx4; this is
//     the reference's stripe scan (insert_schemes/bucket_insert.cuh:127-357)
//     with the same early exit (a stripe holding an empty slot ends a query).
//   * probe-parallel mode (B == 1, P > 1): lane r evaluates probe j0 + r, so
//     one tile step examines P consecutive probes (for linear probing, P
//     adjacent slots in one or two 64 B sectors).  The lowest decisive lane
//     (match, empty, or end of the probe sequence) decides, which is exactly
//     the reference's probe-by-probe order.
#pragma once

#include "pg_internal.hpp"
#include "pg_tile.hpp"

namespace pg {

constexpr uint64_t kKvEmpty = 0ULL;
constexpr uint64_t kKvTomb = ~0ULL;

__device__ inline ulonglong2 load_slot(const ulonglong2 *p) { return *p; }

// Lanes strictly below the lowest set bit of `endm` (all lanes when 0).
__device__ inline uint64_t below_first(uint64_t endm) {
    return endm ? ((1ull << ctz64(endm)) - 1ull) : ~0ull;
}

// bucket_insert::insert (bucket_insert.cuh:127-203) / single_slot_insert::
// insert (single_slot_insert.cuh:125-166): lowest empty slot in probe order;
// key_val_pair::atomic_swap = CAS key 0 -> key, then store the value
// (key_val_pair.cuh:84-95).
template <int P>
__device__ bool kv_level_insert(const LevelDev &L, uint64_t key, uint64_t val, const Tile<P> &t) {
    Probe pr;
    pr.begin(L, key, key);
    pr.max_probes = pr.useful_probes(L.nb.d);
    ulonglong2 *slots = (ulonglong2 *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            uint64_t si = 0;
            ulonglong2 s = make_ulonglong2(kKvTomb, 0);
            if (!end) {
                si = fastmod(p, L.nb);
                s = load_slot(slots + si);
            }
            const uint64_t endm = t.bits(end);
            bool cand = !end && s.x == kKvEmpty;
            uint64_t cm = t.bits(cand) & below_first(endm);
            t.acc[0] += cm ? (uint32_t)ctz64(cm) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            while (cm) {
                const uint32_t l = (uint32_t)ctz64(cm);
                bool success = false;
                if (t.rank == l) {
                    success = atomicCAS((unsigned long long *)&slots[si].x, kKvEmpty, key) == kKvEmpty;
                    if (success) slots[si].y = val;
                }
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                cm &= cm - 1ull;
            }
            if (endm) return false;
        }
    }
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t base = fastmod(p, L.nb) * L.B;
        for (uint32_t s0 = 0; s0 < L.B; s0 += P) {
            const uint32_t i = s0 + t.rank;
            const bool ok = i < L.B;
            const ulonglong2 s = ok ? load_slot(slots + base + i) : make_ulonglong2(kKvTomb, 0);
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            uint64_t cm = t.bits(ok && s.x == kKvEmpty);
            while (cm) {
                const uint32_t l = (uint32_t)ctz64(cm);
                bool success = false;
                if (t.rank == l) {
                    success = atomicCAS((unsigned long long *)&slots[base + i].x, kKvEmpty, key) == kKvEmpty;
                    if (success) slots[base + i].y = val;
                }
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                cm &= cm - 1ull;
            }
        }
    }
}

// bucket_insert::query (bucket_insert.cuh:295-357), single_slot_insert::query
// (single_slot_insert.cuh:168-229).
template <int P>
__device__ bool kv_level_query(const LevelDev &L, uint64_t key, const Tile<P> &t, uint64_t &val,
                               bool &owner) {
    Probe pr;
    pr.begin(L, key, key);
    pr.max_probes = pr.useful_probes(L.nb.d);
    const ulonglong2 *slots = (const ulonglong2 *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            ulonglong2 s = make_ulonglong2(kKvTomb, 0);
            if (!end) s = load_slot(slots + fastmod(p, L.nb));
            const uint64_t endm = t.bits(end);
            const uint64_t lim = below_first(endm);
            const uint64_t mm = t.bits(!end && s.x == key) & lim;
            const uint64_t em = t.bits(!end && s.x == kKvEmpty) & lim;
            const uint64_t dec = mm | em;
            t.acc[0] += dec ? (uint32_t)ctz64(dec) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            if (dec) {
                const uint32_t l = (uint32_t)ctz64(dec);
                if ((mm >> l) & 1ull) {
                    owner = (t.rank == l);
                    if (owner) val = s.y;
                    return true;
                }
                return false;
            }
            if (endm) return false;
        }
    }
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t base = fastmod(p, L.nb) * L.B;
        for (uint32_t s0 = 0; s0 < L.B; s0 += P) {
            const uint32_t i = s0 + t.rank;
            const bool ok = i < L.B;
            const ulonglong2 s = ok ? load_slot(slots + base + i) : make_ulonglong2(kKvTomb, 0);
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            const uint64_t mm = t.bits(ok && s.x == key);
            if (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                owner = (t.rank == l);
                if (owner) val = s.y;
                return true;
            }
            if (t.any(ok && s.x == kKvEmpty)) return false;
        }
    }
}

// bucket_insert::remove (bucket_insert.cuh:205-293): CAS the lowest match to
// the tombstone key (key_val_pair.cuh:143-151); a stripe with an empty slot
// and no successful removal ends the search.
template <int P>
__device__ bool kv_level_remove(const LevelDev &L, uint64_t key, const Tile<P> &t) {
    Probe pr;
    pr.begin(L, key, key);
    pr.max_probes = pr.useful_probes(L.nb.d);
    ulonglong2 *slots = (ulonglong2 *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            uint64_t si = 0;
            ulonglong2 s = make_ulonglong2(kKvTomb, 0);
            if (!end) {
                si = fastmod(p, L.nb);
                s = load_slot(slots + si);
            }
            const uint64_t endm = t.bits(end);
            const uint64_t lim = below_first(endm);
            uint64_t mm = t.bits(!end && s.x == key) & lim;
            const uint64_t em = t.bits(!end && s.x == kKvEmpty) & lim;
            // Walk decisive lanes in probe order: matches are tried, the first
            // empty ends the search.
            uint64_t dec = mm | em;
            t.acc[0] += dec ? (uint32_t)ctz64(dec) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            while (dec) {
                const uint32_t l = (uint32_t)ctz64(dec);
                if (!((mm >> l) & 1ull)) return false;
                bool success = false;
                if (t.rank == l)
                    success = atomicCAS((unsigned long long *)&slots[si].x, key, kKvTomb) == key;
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                dec &= dec - 1ull;
            }
            if (endm) return false;
        }
    }
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t base = fastmod(p, L.nb) * L.B;
        for (uint32_t s0 = 0; s0 < L.B; s0 += P) {
            const uint32_t i = s0 + t.rank;
            const bool ok = i < L.B;
            const ulonglong2 s = ok ? load_slot(slots + base + i) : make_ulonglong2(kKvTomb, 0);
            const bool match = ok && s.x == key;
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            uint64_t mm = t.bits(match);
            while (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                bool success = false;
                if (t.rank == l)
                    success = atomicCAS((unsigned long long *)&slots[base + i].x, key, kKvTomb) == key;
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                mm &= mm - 1ull;
            }
            if (t.any(ok && !match && s.x == kKvEmpty)) return false;
        }
    }
}

// Whole-table ops: level 0, then the backing level on false
// (tables/base_table.cuh:167-308).
template <int P>
__device__ int kv_insert_key(const TableDev &T, uint64_t key, uint64_t val, const Tile<P> &t) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_insert<P>(T.lv[l], key, val, t)) return 1 + (int)l;
    return 0;
}
template <int P>
__device__ bool kv_query_key(const TableDev &T, uint64_t key, const Tile<P> &t, uint64_t &val, bool &owner) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_query<P>(T.lv[l], key, t, val, owner)) return true;
    return false;
}
template <int P>
__device__ bool kv_remove_key(const TableDev &T, uint64_t key, const Tile<P> &t) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_remove<P>(T.lv[l], key, t)) return true;
    return false;
}

}  // namespace pg
