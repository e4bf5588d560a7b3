// pg_kv.hip -- bulk insert / query / remove kernels for key_val_pair<u64,u64>
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

constexpr int kBlock = 256;

template <int P, bool ACC>
__global__ __launch_bounds__(kBlock) void k_kv_insert(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    const bool act = tid < a.n;
    int st = 1;
    if (act) {
        const uint64_t val = a.vals ? ((const uint64_t *)a.vals)[tid] : 0ull;
        st = kv_insert_key<P>(T, a.keys[tid], val, t);
        if (a.flags && t.leader()) a.flags[tid] = (uint8_t)st;
    }
    wave_count(a.fail_count, act && t.leader() && st == 0);
    if constexpr (ACC) flush_acc(a.stats, t, act);
}

template <int P, bool ACC>
__global__ __launch_bounds__(kBlock) void k_kv_query(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    if (tid < a.n) {
        uint64_t val = 0;
        bool own = false;
        const bool found = kv_query_key<P>(T, a.keys[tid], t, val, own);
        const bool writer = found ? own : t.leader();
        if (writer) {
            if (a.flags) a.flags[tid] = found ? 1 : 0;
            if (a.vals_out) ((uint64_t *)a.vals_out)[tid] = found ? val : 0ull;
        }
    }
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P, bool ACC>
__global__ __launch_bounds__(kBlock) void k_kv_remove(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    if (tid < a.n) {
        const bool ok = kv_remove_key<P>(T, a.keys[tid], t);
        if (a.flags && t.leader()) a.flags[tid] = ok ? 1 : 0;
    }
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P>
__global__ __launch_bounds__(64) void k_kv_insert_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    uint64_t fails = 0;
    for (uint64_t i = 0; i < a.n; i++) {
        const uint64_t val = a.vals ? ((const uint64_t *)a.vals)[i] : 0ull;
        const int st = kv_insert_key<P>(T, a.keys[i], val, t);
        if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
        fails += (st == 0);
        fence_agent();
    }
    if (a.fail_count && t.leader() && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
    flush_acc(a.stats, t, true);
}

template <int P>
__global__ __launch_bounds__(64) void k_kv_remove_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    for (uint64_t i = 0; i < a.n; i++) {
        const bool ok = kv_remove_key<P>(T, a.keys[i], t);
        if (a.flags && t.leader()) a.flags[i] = ok ? 1 : 0;
        fence_agent();
    }
    flush_acc(a.stats, t, true);
}

template <int P, bool ACC>
static void launch_kv_conc(Op op, unsigned grid, const TableDev &T, const OpArgs &a, hipStream_t s) {
    if (op == Op::Query) hipLaunchKernelGGL((k_kv_query<P, ACC>), dim3(grid), dim3(kBlock), 0, s, T, a);
    else if (op == Op::Insert) hipLaunchKernelGGL((k_kv_insert<P, ACC>), dim3(grid), dim3(kBlock), 0, s, T, a);
    else hipLaunchKernelGGL((k_kv_remove<P, ACC>), dim3(grid), dim3(kBlock), 0, s, T, a);
}

template <int P>
static int launch_kv_p(const TableDev &T, Op op, const OpArgs &a, hipStream_t s) {
    if (a.n == 0) return PG_OK;
    const unsigned grid = (unsigned)((a.n * (uint64_t)P + kBlock - 1) / kBlock);
    if (op != Op::Query && a.deterministic) {
        if (op == Op::Insert) hipLaunchKernelGGL((k_kv_insert_seq<P>), dim3(1), dim3(P), 0, s, T, a);
        else hipLaunchKernelGGL((k_kv_remove_seq<P>), dim3(1), dim3(P), 0, s, T, a);
    } else if (a.stats) {
        launch_kv_conc<P, true>(op, grid, T, a, s);
    } else {
        launch_kv_conc<P, false>(op, grid, T, a, s);
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

int launch_kv(const TableDev &T, uint32_t tile, Op op, const OpArgs &a, hipStream_t s) {
    switch (tile) {
        case 1: return launch_kv_p<1>(T, op, a, s);
        case 2: return launch_kv_p<2>(T, op, a, s);
        case 4: return launch_kv_p<4>(T, op, a, s);
        case 8: return launch_kv_p<8>(T, op, a, s);
        case 16: return launch_kv_p<16>(T, op, a, s);
        case 32: return launch_kv_p<32>(T, op, a, s);
        case 64: return launch_kv_p<64>(T, op, a, s);
        default: return PG_ERR_INVALID;
    }
}

}  // namespace pg
