// pg_tile.hpp -- wave64 sub-wavefront tiles (device only).
//
// The reference runs one cooperative_groups::thread_block_tile<P> per key and
// scans a bucket in stripes `for (i = rank; i < B; i += P)` with one
// ballot per stripe (representations/packed_bucket.cuh:40-271).  Here a
// tile is P consecutive lanes of a 64-wide wavefront.  Every lane of a tile
// holds V consecutive slot words per row (one global_load_dwordx{1,2,4}), a
// bucket is S rows, and a tile-local ballot (the wave's 64-bit __ballot
// shifted to the tile) gives the tile-uniform masks every decision is taken
// from, so control flow stays uniform inside a tile.  Only the owner lane of
// the chosen slot touches it (CAS or result write): no shuffles are needed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pg {

template <int P>
struct Tile {
    static_assert(P == 1 || P == 2 || P == 4 || P == 8 || P == 16 || P == 32 || P == 64,
                  "tile width must be a power of two <= 64");
    uint32_t rank;  // lane within the tile
    uint32_t base;  // first lane of the tile within the wave
    // Algorithmic accounting (tile-uniform): [0] level-0 buckets (or KV
    // stripes) examined, [1] backing-level buckets examined, [2] slots written.
    mutable uint32_t acc[3];
    // Set by kernels whose keys are in flight together: the shortcut insert
    // then bounds its first pass by slot index (see fast_shortcut_insert).
    bool concurrent;

    __device__ inline Tile() {
        const uint32_t lane = __lane_id();
        rank = lane & (P - 1);
        base = lane & ~uint32_t(P - 1);
        acc[0] = acc[1] = acc[2] = 0;
        concurrent = false;
    }
    // Tile-local ballot: bit r set when lane `rank == r` of this tile has pred.
    __device__ inline uint64_t bits(bool pred) const {
        const uint64_t m = __ballot(pred);
        if constexpr (P == 64) return m;
        else return (m >> base) & ((1ull << P) - 1ull);
    }
    __device__ inline bool any(bool pred) const { return bits(pred) != 0; }
    __device__ inline int popc(bool pred) const { return __popcll(bits(pred)); }
    __device__ inline bool leader() const { return rank == 0; }
};

__device__ inline int ctz64(uint64_t x) { return __builtin_ctzll(x); }

// Compile-time bucket shape for packed 32-bit slot words.
template <int P, int B>
struct Shape {
    static constexpr int V = (B >= 4 * P) ? 4 : ((B >= 2 * P) ? 2 : 1);  // words per lane per row
    static constexpr int ROW = P * V;                                     // words per row
    static constexpr int S = (B + ROW - 1) / ROW;                         // rows
    static_assert(V == 1 || B % ROW == 0, "rows must tile the bucket");
};

// Load V consecutive 32-bit words (V in {1,2,4}) with one vector load.
template <int V>
__device__ inline void load_words(const uint32_t *p, uint32_t (&w)[V]) {
    if constexpr (V == 4) {
        const uint4 v = *reinterpret_cast<const uint4 *>(p);
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else if constexpr (V == 2) {
        const uint2 v = *reinterpret_cast<const uint2 *>(p);
        w[0] = v.x; w[1] = v.y;
    } else {
        w[0] = *p;
    }
}

// Agent-scope acquire/release fence: waits for this wave's memory operations
// and invalidates the CU's L1, so the next load observes every earlier atomic.
__device__ inline void fence_agent() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "agent"); }

// Adds a tile leader's accounting counters to stats[0..2].  Only accounting
// runs pass stats (bench.py's byte counting), so one atomic per leader is fine.
template <int P>
__device__ inline void flush_acc(uint64_t *stats, const Tile<P> &t, bool act) {
    if (!stats || !act || t.rank != 0) return;
#pragma unroll
    for (int c = 0; c < 3; c++)
        if (t.acc[c]) atomicAdd((unsigned long long *)&stats[c], (unsigned long long)t.acc[c]);
}

// Per-wave aggregated atomic increment of a 64-bit device counter.
__device__ inline void wave_count(uint64_t *counter, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m && counter) {
        const uint32_t lane = __lane_id();
        const uint32_t first = (uint32_t)ctz64(__ballot(1));
        if (lane == first) atomicAdd((unsigned long long *)counter, (unsigned long long)__popcll(m));
    }
}

}  // namespace pg
