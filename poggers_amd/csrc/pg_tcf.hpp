// pg_tcf.hpp -- bulk insert / query / remove kernels for packed 16/16 filter
// slots (the Two-Choice Filter and its linear backing table), gfx950 wave64.
//
// Semantics follow the reference's device ops:
//   level 0 POWER_OF_N_SHORTCUT  insert_schemes/power_of_n_shortcut_buckets.cuh
//   level 1 LINEAR_BUCKETS       insert_schemes/linear_insert_buckets.cuh
//   buckets                      representations/packed_bucket.cuh
//   slot words                   representations/grouped_storage_sub_bits.cuh
// Every bucket routine acts on the LOWEST-index qualifying slot, which is what
// the reference's ascending stripes + lowest-lane leader give for any tile
// width, so results do not depend on P.
//
// Layout in HBM: one uint32 word per slot (val16 << 16 | fp16), buckets of B
// contiguous words (64 B at B = 16), bucket b at word b*B.
#pragma once

#include <type_traits>

#include "pg_internal.hpp"
#include "pg_tile.hpp"

namespace pg {

__device__ inline bool w_empty(uint32_t w) { return (w & 0xFFFFu) == 0u; }
__device__ inline bool w_tomb(uint32_t w) { return (w & 0xFFFFu) == 0xFFFFu; }
__device__ inline bool w_filled(uint32_t w) { return !(w_empty(w) || w_tomb(w)); }

// The same predicates for any slot format (the key bits are the low kbits).
template <class F>
__device__ inline bool f_empty(uint32_t w) { return (w & F::kmask) == 0u; }
template <class F>
__device__ inline bool f_tomb(uint32_t w) { return (w & F::kmask) == F::kmask; }
template <class F>
__device__ inline bool f_filled(uint32_t w) { return !(f_empty<F>(w) || f_tomb<F>(w)); }

// Compare-and-swap of one slot word.  16-bit words are swapped through the
// aligned 32-bit word that holds them (global atomics are 32/64-bit): the
// neighbour half is carried over unchanged, and the loop only repeats while
// that neighbour changes under us, so it ends as soon as either our half
// differs from `expect` or the swap lands.  Slot arrays are 256 B aligned and
// padded to 256 B, so the containing word is always inside the allocation.
__device__ inline bool cas_word(uint32_t *p, uint32_t expect, uint32_t desired) {
    return atomicCAS(p, expect, desired) == expect;
}
// Replace the `mask` bits of a 16-bit word if they equal expect's
// (replace_uint16, representation_helpers.cuh:451-489, on the aligned 32-bit
// word that holds it).
__device__ inline bool cas_bits16(uint16_t *p, uint32_t mask, uint32_t expect, uint32_t desired) {
    uint32_t *w32 = (uint32_t *)((uintptr_t)p & ~uintptr_t(3));
    const uint32_t sh = ((uintptr_t)p & 2u) ? 16u : 0u;
    const uint32_t m32 = (mask & 0xFFFFu) << sh;
    const uint32_t e32 = (expect & mask & 0xFFFFu) << sh, d32 = (desired & mask & 0xFFFFu) << sh;
    uint32_t old = __hip_atomic_load(w32, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if ((old & m32) != e32) return false;
        const uint32_t prev = atomicCAS(w32, old, (old & ~m32) | d32);
        if (prev == old) return true;
        old = prev;
    }
}
__device__ inline bool cas_word(uint16_t *p, uint32_t expect, uint32_t desired) {
    return cas_bits16(p, 0xFFFFu, expect, desired);
}

template <int V>
__device__ inline uint32_t select_word(const uint32_t (&w)[V], int v) {
    uint32_t x = w[0];
#pragma unroll
    for (int i = 1; i < V; i++) x = (v == i) ? w[i] : x;
    return x;
}

// Outcome of a bounded slot claim.
constexpr int kClaimed = 1;  // the CAS landed
constexpr int kNoSlot = 0;   // no empty (or reusable) slot left: the bucket cannot take the key
constexpr int kBudget = -1;  // the allowed number of CAS attempts ran out

// ---------------------------------------------------------------------------
// A bucket held in registers: compile-time B and P.
// ---------------------------------------------------------------------------
template <int P, int B>
struct Bucket {
    using Sh = Shape<P, B>;
    static constexpr int S = Sh::S, V = Sh::V, ROW = Sh::ROW;
    uint32_t w[S][V];

    __device__ static inline bool ok(int s, const Tile<P> &t) {
        return (uint32_t)(s * ROW) + t.rank * V < (uint32_t)B;
    }

    __device__ inline void load(const uint32_t *slots, uint64_t b, const Tile<P> &t) {
        const uint32_t *base = slots + b * (uint64_t)B;
        t.acc[0]++;
#pragma unroll
        for (int s = 0; s < S; s++) {
            if (ok(s, t)) {
                load_words<V>(base + s * ROW + t.rank * V, w[s]);
            } else {
#pragma unroll
                for (int v = 0; v < V; v++) w[s][v] = 0;
            }
        }
    }

    // get_fill: packed_bucket.cuh:40-62
    __device__ inline int fill(const Tile<P> &t) const {
        int f = 0;
#pragma unroll
        for (int s = 0; s < S; s++)
#pragma unroll
            for (int v = 0; v < V; v++) f += t.popc(ok(s, t) && w_filled(w[s][v]));
        return f;
    }

    // insert / insert_delete: claim the lowest empty (or, with wd, tombstone)
    // slot (packed_bucket.cuh:64-180; CAS words grouped_storage_sub_bits.cuh:
    // 127-157) with at most `tries` CAS attempts, passing over the first
    // `taken` candidates of the view: slots an earlier attempt on this bucket
    // found taken since the load (candidates are tried lowest first and a
    // taken slot stays taken, so they are exactly the lowest ones).  Every
    // failed attempt adds one to `taken`, so fill() + taken is the bucket's
    // fill as far as this key knows.  Returns kClaimed, kNoSlot (no candidate
    // left in the view: the bucket cannot take the key) or kBudget (tries ran
    // out).  `taken` and `tries` are tile-uniform.
    __device__ inline int claim(uint32_t *slots, uint64_t b, uint32_t word, bool wd, const Tile<P> &t,
                                uint32_t &taken, uint32_t tries = ~0u) const {
        uint32_t *base = slots + b * (uint64_t)B;
        uint32_t skip = taken;
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint32_t cand = 0;
            if (ok(s, t)) {
#pragma unroll
                for (int v = 0; v < V; v++)
                    if (w_empty(w[s][v]) || (wd && w_tomb(w[s][v]))) cand |= 1u << v;
            }
            uint64_t lanes = t.bits(cand != 0);
            while (lanes) {
                const uint32_t l = (uint32_t)ctz64(lanes);
                if (skip) {
                    skip--;
                    if (t.rank == l) cand &= cand - 1u;
                    lanes = t.bits(cand != 0);
                    continue;
                }
                if (tries == 0) return kBudget;
                tries--;
                bool success = false;
                if (t.rank == l) {
                    const int v = __builtin_ctz(cand);
                    const uint32_t cur = select_word<V>(w[s], v);
                    const uint32_t expect = (wd && w_tomb(cur)) ? kTomb16 : 0u;
                    success = atomicCAS(base + s * ROW + t.rank * V + v, expect, word) == expect;
                    cand &= cand - 1u;
                }
                if (t.any(success)) {
                    t.acc[2]++;
                    return kClaimed;
                }
                taken++;
                lanes = t.bits(cand != 0);
            }
        }
        return kNoSlot;
    }

    // query: lowest matching slot, packed_bucket.cuh:182-219.  The owner lane
    // (the one holding the match) receives the value.
    __device__ inline bool find(uint32_t fp, const Tile<P> &t, uint32_t &val, bool &owner) const {
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint32_t m = 0;
            if (ok(s, t)) {
#pragma unroll
                for (int v = 0; v < V; v++)
                    if ((w[s][v] & 0xFFFFu) == fp) m |= 1u << v;
            }
            const uint64_t lanes = t.bits(m != 0);
            if (lanes) {
                const uint32_t l = (uint32_t)ctz64(lanes);
                owner = (t.rank == l);
                if (owner) val = select_word<V>(w[s], __builtin_ctz(m)) >> 16;
                return true;
            }
        }
        return false;
    }

    // remove: CAS the lowest match to the tombstone word, packed_bucket.cuh:
    // 227-271, grouped_storage_sub_bits.cuh:217-234.
    __device__ inline bool remove(uint32_t *slots, uint64_t b, uint32_t fp, const Tile<P> &t) const {
        uint32_t *base = slots + b * (uint64_t)B;
#pragma unroll
        for (int s = 0; s < S; s++) {
            uint32_t m = 0;
            if (ok(s, t)) {
#pragma unroll
                for (int v = 0; v < V; v++)
                    if ((w[s][v] & 0xFFFFu) == fp) m |= 1u << v;
            }
            uint64_t lanes = t.bits(m != 0);
            while (lanes) {
                const uint32_t l = (uint32_t)ctz64(lanes);
                bool success = false;
                if (t.rank == l) {
                    const int v = __builtin_ctz(m);
                    const uint32_t cur = select_word<V>(w[s], v);
                    success = atomicCAS(base + s * ROW + t.rank * V + v, cur, kTomb16) == cur;
                    m &= m - 1u;
                }
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                lanes = t.bits(m != 0);
            }
        }
        return false;
    }
};

// ---------------------------------------------------------------------------
// Slot access for the generic routines, per slot format.  A slot reads as a
// "composite": the stored word (packed formats) or, for the 12-bit bucket,
// the slot's 12 bits gathered from its three-word group
// (representation_helpers.cuh:600-682 read_*).
// ---------------------------------------------------------------------------
template <class F>
__device__ inline uint32_t s_get(const void *bk, uint32_t i) {
    if constexpr (is_twelve<F>) {
        const uint16_t *w = (const uint16_t *)bk + 3 * (i >> 2);
        switch (i & 3u) {
            case 0: return w[0] & 0x0FFFu;
            case 1: return (w[0] & 0xF000u) | (w[1] & 0x00FFu);
            case 2: return (w[1] & 0xFF00u) | (w[2] & 0x000Fu);
            default: return w[2] & 0xFFF0u;
        }
    } else {
        return (uint32_t)((const typename F::W *)bk)[i];
    }
}
template <class F>
__device__ inline bool s_empty(uint32_t c) {
    if constexpr (is_twelve<F>) return c == 0u;
    else return f_empty<F>(c);
}
template <class F>
__device__ inline bool s_tomb(uint32_t c) {
    if constexpr (is_twelve<F>) return false;  // no tombstones
    else return f_tomb<F>(c);
}
template <class F>
__device__ inline bool s_filled(uint32_t c) { return !(s_empty<F>(c) || s_tomb<F>(c)); }
// A slot a claim can take: the empty word or, with wd, the tombstone word
// itself (the CAS expects exactly that word).  A 10/6 key whose stored bits
// are 0 or 0x3FF leaves a word that only looks empty or buried, its value
// bits being set; the reference's CAS fails on it and moves on
// (grouped_storage_sub_bits.cuh:127-157), so it is no candidate.  Counting
// it as one would also make a bounded concurrent claim read its certain
// failure as a slot taken by another key.  For 16/16 and 8/8 slots these
// are exactly the empty and tombstone slots.
template <class F>
__device__ inline bool s_claimable(uint32_t c, bool wd) {
    if constexpr (is_twelve<F>) return c == 0u;
    else return c == 0u || (wd && c == F::kmask);
}
// contains(): the key bits equal fp's; for the 12-bit bucket the composite
// equals slot i's projection of the item (sub_byte_match).
template <class F>
__device__ inline bool s_match(uint32_t c, uint32_t fp, uint32_t i) {
    if constexpr (is_twelve<F>) return c == twelve_proj(fp, i);
    else return (c & F::kmask) == (fp & F::kmask);
}
template <class F>
__device__ inline uint32_t s_value(uint32_t c) {
    if constexpr (is_twelve<F>) return 0u;  // twelve_bucket::query reports 0
    else return (c >> F::kbits) & F::vmask;
}
// Claim slot i, seen empty (or, with from_tomb, seen as a tombstone), for
// `word`.  Packed formats CAS the whole word from the empty / tombstone word
// (grouped_storage_sub_bits.cuh:127-157).  The 12-bit bucket replaces the
// slot's first part only if it is still zero, then writes the second part
// (representation_helpers.cuh:779-903 replace_*: two masked swaps, the
// second asserted to succeed).
template <class F>
__device__ inline bool s_claim(void *bk, uint32_t i, bool from_tomb, uint32_t word) {
    if constexpr (is_twelve<F>) {
        uint16_t *w = (uint16_t *)bk + 3 * (i >> 2);
        const uint32_t m = twelve_proj(word, i);
        switch (i & 3u) {
            case 0: return cas_bits16(w, 0x0FFFu, 0u, m);
            case 1:
                if (!cas_bits16(w, 0xF000u, 0u, m)) return false;
                cas_bits16(w + 1, 0x00FFu, 0u, m);
                return true;
            case 2:
                if (!cas_bits16(w + 1, 0xFF00u, 0u, m)) return false;
                cas_bits16(w + 2, 0x000Fu, 0u, m);
                return true;
            default: return cas_bits16(w + 2, 0xFFF0u, 0u, m);
        }
    } else {
        return cas_word((typename F::W *)bk + i, from_tomb ? F::kmask : 0u, word);
    }
}
// Tombstone slot i seen as c (grouped_storage_sub_bits.cuh:217-234).
template <class F>
__device__ inline bool s_bury(void *bk, uint32_t i, uint32_t c) {
    if constexpr (is_twelve<F>) return false;  // removal is not offered
    else return cas_word((typename F::W *)bk + i, c, F::kmask);
}

// ---------------------------------------------------------------------------
// Generic bucket routines: runtime B, one slot per lane per row, re-read per
// pass (the reference reads the bucket once for its fill and again for the
// claim).  Used by the backing level, grouped levels, the 8/8, 10/6 and
// 12-bit slots and the shapes without a fast path.
// ---------------------------------------------------------------------------
template <int P, class F = Fmt16>
__device__ int gen_fill(const void *base, uint32_t B, const Tile<P> &t) {
    int f = 0;
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const bool ok = i < B;
        const uint32_t w = ok ? s_get<F>(base, i) : 0u;
        f += t.popc(ok && s_filled<F>(w));
    }
    return f;
}

// Occupancy index of a linear level (LevelDev::occ): set bucket b's bit.
// A relaxed device-scope OR with no return value; it is issued before the
// claim's CAS, and a kernel boundary or a release fence publishes both.
template <int P>
__device__ inline void occ_mark(uint32_t *occ, uint64_t b, const Tile<P> &t) {
    if (t.leader())
        __hip_atomic_fetch_or(occ + (b >> 5), 1u << (uint32_t)(b & 31u), __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline bool occ_clear(const uint32_t *occ, uint64_t b) {
    return ((occ[b >> 5] >> (uint32_t)(b & 31u)) & 1u) == 0u;
}

// Claim the lowest empty slot (with `wd`, the lowest empty-or-tombstone one)
// with at most `tries` CAS attempts.  packed_bucket.cuh:64-180,
// 12_bit_bucket.cuh:94-144.  Returns kClaimed, kNoSlot or kBudget.  With an
// occupancy index (occ, bucket b), the bucket's bit is set once the tile
// sees a candidate slot, before its first CAS; walks past full buckets
// set nothing.
template <int P, class F = Fmt16>
__device__ int gen_claim(void *base, uint32_t B, uint32_t word, bool wd, const Tile<P> &t,
                         uint32_t tries = ~0u, uint32_t *occ = nullptr, uint64_t b = 0) {
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const bool ok = i < B;
        const uint32_t w = ok ? s_get<F>(base, i) : 0u;
        bool cand = ok && s_claimable<F>(w, wd);
        uint64_t lanes = t.bits(cand);
        if (lanes && occ) {
            occ_mark(occ, b, t);
            occ = nullptr;
        }
        while (lanes) {
            if (tries == 0) return kBudget;
            tries--;
            const uint32_t l = (uint32_t)ctz64(lanes);
            bool success = false;
            if (t.rank == l) {
                success = s_claim<F>(base, i, wd && s_tomb<F>(w), word);
                cand = false;
            }
            if (t.any(success)) {
                t.acc[2]++;
                return kClaimed;
            }
            lanes = t.bits(cand);
        }
    }
    return kNoSlot;
}

// Lowest matching slot; when none, *fill receives the bucket's fill (same
// pass).  packed_bucket.cuh:182-219, 12_bit_bucket.cuh:259-300.
template <int P, class F = Fmt16>
__device__ bool gen_find(const void *base, uint32_t B, uint32_t fp, const Tile<P> &t,
                         uint32_t &val, bool &owner, int &fill) {
    fill = 0;
    if constexpr (std::is_same<F, Fmt16>::value) {
        if ((B & 3u) == 0 && B <= 4u * P) {
            // The whole bucket in one row: lane r holds words 4r..4r+3, read
            // with one 16 B load (B % 4 == 0 keeps every bucket 16 B aligned,
            // grouped or not).
            const uint32_t *b32 = (const uint32_t *)base;
            const bool ok = 4u * t.rank < B;
            uint32_t w[4] = {0u, 0u, 0u, 0u};
            if (ok) load_words<4>(b32 + 4 * t.rank, w);
            uint32_t m = 0;
#pragma unroll
            for (int v = 0; v < 4; v++)
                if (ok && (w[v] & 0xFFFFu) == fp) m |= 1u << v;
            const uint64_t lanes = t.bits(m != 0);
            if (lanes) {
                const uint32_t l = (uint32_t)ctz64(lanes);
                owner = (t.rank == l);
                if (owner) val = select_word<4>(w, __builtin_ctz(m)) >> 16;
                return true;
            }
#pragma unroll
            for (int v = 0; v < 4; v++) fill += t.popc(ok && w_filled(w[v]));
            return false;
        }
    }
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const bool ok = i < B;
        const uint32_t w = ok ? s_get<F>(base, i) : 0u;
        const uint64_t lanes = t.bits(ok && s_match<F>(w, fp, i));
        if (lanes) {
            const uint32_t l = (uint32_t)ctz64(lanes);
            owner = (t.rank == l);
            if (owner) val = s_value<F>(w);
            return true;
        }
        fill += t.popc(ok && s_filled<F>(w));
    }
    return false;
}

// Tombstone the lowest matching slot.  packed_bucket.cuh:227-271.
template <int P, class F = Fmt16>
__device__ bool gen_remove(void *base, uint32_t B, uint32_t fp, const Tile<P> &t, int &fill) {
    fill = 0;
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const bool ok = i < B;
        const uint32_t w = ok ? s_get<F>(base, i) : 0u;
        bool m = ok && s_match<F>(w, fp, i);
        uint64_t lanes = t.bits(m);
        while (lanes) {
            const uint32_t l = (uint32_t)ctz64(lanes);
            bool success = false;
            if (t.rank == l) {
                success = s_bury<F>(base, i, w);
                m = false;
            }
            if (t.any(success)) {
                t.acc[2]++;
                return true;
            }
            lanes = t.bits(m);
        }
        fill += t.popc(ok && s_filled<F>(w));
    }
    return false;
}

// ---------------------------------------------------------------------------
// Level operations, generic shapes.
// ---------------------------------------------------------------------------

// Concurrent kernels and the two passes of the shortcut insert.
//
// The first pass claims a slot of a probe bucket only while the bucket's fill
// is below 3B/4 (power_of_n_shortcut_buckets.cuh:223); the second pass claims
// in the least-filled bucket (:254-300).  Applied in key order the fills read
// just before each claim are exact.  With many keys in flight they are read
// one or more memory round trips before the claim lands, and other keys land
// in between, so a rule checked on the read fills can be broken when the claim
// lands.  The checks are therefore made on the fill at the moment of the
// claim: slots go empty -> filled (and, for a tombstone reuse, tombstone ->
// filled) and claims take the lowest candidate slot first, so when a claim
// lands, the bucket's fill is the filled count of the view read plus the
// number of CAS attempts that failed on slots the view showed free.  Bounding
// the attempts bounds the fill at the moment of the claim:
//   * first pass: at most ceil(3B/4) - f attempts (f = filled count read);
//   * second pass: claim in the least-filled bucket (lowest probe index on
//     ties) with at most (next-least fill) - f + 1 attempts, so the bucket
//     taken is never fuller than the other choice; a bucket that runs out of
//     attempts has its known fill raised by that many and the choice is made
//     again.
// Without these bounds, concurrent keys pile into buckets that looked
// emptier than they had become, the two choices fill unevenly, and keys spill
// to the backing level faster than in key order (at 2^22 slots with four-lane
// tiles: 1.5x the backing keys after a 90 % fill, and under repeated 10 %
// insert/delete churn a full backing level after seven rounds).  Key-order
// (deterministic) kernels keep the unbounded reference loops, whose layouts
// are bit-exact.
template <int P>
__device__ inline uint32_t first_pass_tries(const Tile<P> &t, int B, int f) {
    return t.concurrent ? (uint32_t)((3 * B + 3) / 4 - f) : ~0u;
}

// Index j of the least-filled of np probe buckets (lowest index on ties), the
// next-least fill, and how many claim attempts keep j's fill at the moment of
// the claim within the key-order rule: up to the next-least fill when j comes
// first in probe order among the buckets at that fill (ties go to the first
// probe), else strictly below it.
__device__ inline void least_filled(const int (&fill)[kMaxShortcutProbes], int np, int B, int &j, int &next,
                                    uint32_t &tries) {
    j = 0;
#pragma unroll
    for (int q = 1; q < (int)kMaxShortcutProbes; q++)
        if (q < np && fill[q] < fill[j]) j = q;
    next = B;
    int first_at_next = (int)kMaxShortcutProbes;
#pragma unroll
    for (int q = 0; q < (int)kMaxShortcutProbes; q++) {
        if (q < np && q != j) {
            if (fill[q] < next) next = fill[q], first_at_next = q;
            else if (fill[q] == next && q < first_at_next) first_at_next = q;
        }
    }
    int fj = B;
#pragma unroll
    for (int q = 0; q < (int)kMaxShortcutProbes; q++)
        if (q == j) fj = fill[q];
    tries = (uint32_t)(next - fj + (j < first_at_next ? 1 : 0));
}

// power_of_n_shortcut_buckets.cuh:198-306 (insert) and :309-417
// (insert_with_delete), runtime B.
template <int P, class F = Fmt16>
__device__ bool gen_shortcut_insert(const LevelDev &L, uint64_t key, uint32_t fp, uint32_t word,
                                    bool wd, const Tile<P> &t, int ai) {
    Probe pr;
    pr.begin(L, key, fp);
    const int B = (int)L.B;
    uint64_t bidx[kMaxShortcutProbes];
    int fill[kMaxShortcutProbes];
    int np = 0, min_fill = B;
#pragma unroll
    for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) break;
        const uint64_t b = fastmod(p, L.nb);
        bidx[j] = b;
        t.acc[ai]++;
        int f = gen_fill<P, F>(bucket_ptr(L, b), L.B, t);
        if (4 * f < 3 * B) {
            const int r = gen_claim<P, F>(bucket_ptr(L, b), L.B, word, wd, t, first_pass_tries(t, B, f));
            if (r == kClaimed) return true;
            f = (r == kNoSlot) ? B : (3 * B + 3) / 4;
        }
        fill[j] = f;
        if (f < min_fill) min_fill = f;
        np = j + 1;
    }
    if (!t.concurrent) {
        int lvl = min_fill;
        while (lvl < B) {
            int next = B;
#pragma unroll
            for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
                if (j < np) {
                    if (fill[j] == lvl && gen_claim<P, F>(bucket_ptr(L, bidx[j]), L.B, word, wd, t) == kClaimed)
                        return true;
                    if (fill[j] > lvl && fill[j] < next) next = fill[j];
                }
            }
            lvl = next;
        }
        return false;
    }
    // The generic claim reads the bucket again, so the fill is read again just
    // before it: the bound then counts from that fill.  Every round raises one
    // known fill, so the loop ends within np * B rounds.
    if (np == 0) return false;
    for (int round = 0; round < (int)kMaxShortcutProbes * (B + 1); round++) {
        int j, next;
        uint32_t tries;
        least_filled(fill, np, B, j, next, tries);
        int fj = B;
        uint64_t bj = 0;
#pragma unroll
        for (int q = 0; q < (int)kMaxShortcutProbes; q++)
            if (q == j) fj = fill[q], bj = bidx[q];
        if (fj >= B) return false;
        void *base = bucket_ptr(L, bj);
        const int now = gen_fill<P, F>(base, L.B, t);
        int nf;
        if (now > fj) {
            nf = now;  // it filled since: choose again
        } else {
            const int r = gen_claim<P, F>(base, L.B, word, wd, t, tries);
            if (r == kClaimed) return true;
            nf = (r == kNoSlot) ? B : fj + (int)tries;
        }
#pragma unroll
        for (int q = 0; q < (int)kMaxShortcutProbes; q++)
            if (q == j) fill[q] = nf < B ? nf : B;
    }
    return false;
}

// linear_insert_buckets.cuh:200-244: first probe bucket with a free slot.
template <int P, class F = Fmt16>
__device__ bool gen_linear_insert(const LevelDev &L, uint64_t key, uint32_t fp, uint32_t word,
                                  bool wd, const Tile<P> &t, int ai) {
    Probe pr;
    pr.begin(L, key, fp);
    pr.max_probes = pr.useful_probes(L.nb.d);
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t b = fastmod(p, L.nb);
        t.acc[ai]++;
        if (gen_claim<P, F>(bucket_ptr(L, b), L.B, word, wd, t, ~0u, L.occ, b) == kClaimed) return true;
    }
}

template <int P, class F = Fmt16>
__device__ bool gen_level_insert(const LevelDev &L, uint64_t key, uint32_t fp, uint32_t word, bool wd,
                                 const Tile<P> &t, int ai) {
    if (L.scheme != PG_SCHEME_LINEAR_BUCKETS) return gen_shortcut_insert<P, F>(L, key, fp, word, wd, t, ai);
    return gen_linear_insert<P, F>(L, key, fp, word, wd, t, ai);
}

// Shortcut query (power_of_n_shortcut_buckets.cuh:676-707): every probe
// bucket, first match wins.  Linear query (linear_insert_buckets.cuh:362-398):
// stop at the first probe bucket whose fill != B; with an occupancy index, a
// bucket whose bit is clear is empty and stops the walk unread.
template <int P, class F = Fmt16>
__device__ bool gen_level_query(const LevelDev &L, uint64_t key, uint32_t fp, const Tile<P> &t,
                                uint32_t &val, bool &owner, int ai) {
    Probe pr;
    pr.begin(L, key, fp);
    const bool linear = L.scheme == PG_SCHEME_LINEAR_BUCKETS;
    if (linear) pr.max_probes = pr.useful_probes(L.nb.d);
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t b = fastmod(p, L.nb);
        if (L.occ && occ_clear(L.occ, b)) return false;
        int f;
        t.acc[ai]++;
        if (gen_find<P, F>(bucket_ptr(L, b), L.B, fp, t, val, owner, f)) return true;
        if (linear && f != (int)L.B) return false;
    }
}

// Shortcut remove (power_of_n_shortcut_buckets.cuh:709-740); linear remove
// (linear_insert_buckets.cuh:323-360) stops at the first bucket with fill != B
// (or, as the query does, at a bucket whose occupancy bit is clear).
template <int P, class F = Fmt16>
__device__ bool gen_level_remove(const LevelDev &L, uint64_t key, uint32_t fp, const Tile<P> &t, int ai) {
    Probe pr;
    pr.begin(L, key, fp);
    const bool linear = L.scheme == PG_SCHEME_LINEAR_BUCKETS;
    if (linear) pr.max_probes = pr.useful_probes(L.nb.d);
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t b = fastmod(p, L.nb);
        if (L.occ && occ_clear(L.occ, b)) return false;
        int f;
        t.acc[ai]++;
        if (gen_remove<P, F>(bucket_ptr(L, b), L.B, fp, t, f)) return true;
        if (linear && f != (int)L.B) return false;
    }
}


// ---------------------------------------------------------------------------
// Level-0 fast path: POWER_OF_N_SHORTCUT with compile-time B and at most two
// probes (every two-choice prober, and double hashing with Max_Probes 2; the
// launchers send levels of more probes to the generic routines).  The caller
// has already hashed the key (pr) and loaded its first probe bucket (bk0,
// b0), so a tile can issue the first-choice loads of several keys at once;
// both probe buckets stay in registers for both passes.
// power_of_n_shortcut_buckets.cuh:198-306 (insert), :309-417 (with delete).
// ---------------------------------------------------------------------------
template <int P, int B>
__device__ bool fast_shortcut_insert(const LevelDev &L, const Probe &pr, uint32_t word, bool wd,
                                     const Tile<P> &t, const Bucket<P, B> &bk0, uint64_t b0) {
    uint32_t *slots = (uint32_t *)L.slots;
    if (pr.at(0) == kEnd) return false;
    constexpr int kLim = (3 * B + 3) / 4;
    uint32_t taken0 = 0, taken1 = 0;  // view candidates found taken, per bucket
    int f0 = bk0.fill(t);
    if (4 * f0 < 3 * B) {
        const int r = bk0.claim(slots, b0, word, wd, t, taken0, first_pass_tries(t, B, f0));
        if (r == kClaimed) return true;
        f0 = (r == kNoSlot) ? B : kLim;
    }
    const uint64_t p1 = pr.at(1);
    const bool two = p1 != kEnd;
    Bucket<P, B> bk1;
    uint64_t b1 = 0;
    int f1 = B;  // a missing second probe counts as full
    if (two) {
        b1 = fastmod(p1, L.nb);
        bk1.load(slots, b1, t);
        f1 = bk1.fill(t);
        if (4 * f1 < 3 * B) {
            const int r = bk1.claim(slots, b1, word, wd, t, taken1, first_pass_tries(t, B, f1));
            if (r == kClaimed) return true;
            f1 = (r == kNoSlot) ? B : kLim;
        }
    }
    if (!t.concurrent) {
        // key order: ascending fill levels, probe order within a level
        int lvl = f0 < f1 ? f0 : f1;
        while (lvl < B) {
            int next = B;
            if (f0 == lvl && bk0.claim(slots, b0, word, wd, t, taken0) == kClaimed) return true;
            if (f0 > lvl && f0 < next) next = f0;
            if (two) {
                if (f1 == lvl && bk1.claim(slots, b1, word, wd, t, taken1) == kClaimed) return true;
                if (f1 > lvl && f1 < next) next = f1;
            }
            lvl = next;
        }
        return false;
    }
    // Bounded fill-ordered pass on the register views: the first probe may
    // land while no fuller than the second, the second only while less full
    // than the first (ties go to the first probe, as in key order).  Every
    // round raises one known fill, so the loop ends within 2 * B rounds.
    for (int round = 0; round < 2 * (B + 1); round++) {
        const bool pick1 = f1 < f0;  // the least filled, first probe on ties
        const int fj = pick1 ? f1 : f0;
        if (fj >= B) return false;
        const uint32_t tries = pick1 ? (uint32_t)(f0 - f1) : (uint32_t)(f1 - f0 + 1);
        const int r = pick1 ? bk1.claim(slots, b1, word, wd, t, taken1, tries)
                            : bk0.claim(slots, b0, word, wd, t, taken0, tries);
        if (r == kClaimed) return true;
        int nf = (r == kNoSlot) ? B : fj + (int)tries;
        nf = nf < B ? nf : B;
        if (pick1) f1 = nf;
        else f0 = nf;
    }
    return false;
}

template <int P, int B>
__device__ bool fast_shortcut_remove(const LevelDev &L, uint64_t key, uint32_t fp, const Tile<P> &t) {
    Probe pr;
    pr.begin(L, key, fp);
    uint32_t *slots = (uint32_t *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t b = fastmod(p, L.nb);
        Bucket<P, B> bk;
        bk.load(slots, b, t);
        if (bk.remove(slots, b, fp, t)) return true;
    }
}

// Insert one key whose level-0 probe state is ready (bucketed_table.cuh:
// 271-291: level 0, then the backing levels on false).  Returns 0 on failure,
// else 1 + level.
template <int P, int B>
__device__ int tcf_insert_prepared(const TableDev &T, uint64_t key, uint32_t fp, uint32_t word, bool wd,
                                   const Tile<P> &t, const Probe &pr, const Bucket<P, B> &bk0, uint64_t b0) {
    if (fast_shortcut_insert<P, B>(T.lv[0], pr, word, wd, t, bk0, b0)) return 1;
    for (uint32_t l = 1; l < T.nlevels; l++)
        if (gen_level_insert<P>(T.lv[l], key, fp, word, wd, t, 1)) return 1 + (int)l;
    return 0;
}

// Whole-table insert of one key; B == 0 selects the generic level-0 path (the
// only one for the 8/8 and 10/6 formats).
template <int P, int B, class F = Fmt16>
__device__ int tcf_insert_key(const TableDev &T, uint64_t key, uint32_t val, bool wd, const Tile<P> &t) {
    const uint32_t fp = fp_of<F>(key);
    const uint32_t word = pack_word<F>(fp, val);
    if constexpr (B != 0) {
        static_assert(F::kbits == 16, "register-bucket path is for 16/16 slots");
        const LevelDev &L = T.lv[0];
        Probe pr;
        pr.begin(L, key, fp);
        Bucket<P, B> bk0;
        const uint64_t p0 = pr.at(0);
        const uint64_t b0 = (p0 == kEnd) ? 0 : fastmod(p0, L.nb);
        if (p0 != kEnd) bk0.load((const uint32_t *)L.slots, b0, t);
        return tcf_insert_prepared<P, B>(T, key, fp, word, wd, t, pr, bk0, b0);
    } else {
        if (gen_level_insert<P, F>(T.lv[0], key, fp, word, wd, t, 0)) return 1;
        for (uint32_t l = 1; l < T.nlevels; l++)
            if (gen_level_insert<P, F>(T.lv[l], key, fp, word, wd, t, 1)) return 1 + (int)l;
        return 0;
    }
}

// ---------------------------------------------------------------------------
// Upserts: insert_if_not_exists / insert_if_not_exists_delete.
// Shortcut level (power_of_n_shortcut_buckets.cuh:421-674): look the key up in
// every probe bucket; a match is returned (and, for _delete, tombstoned);
// otherwise the shortcut insert (with tombstone reuse for _delete).
// Linear level (linear_insert_buckets.cuh:249-321): per probe bucket, a match
// is returned (and removed for _delete), else a free slot takes the key
// (plain insert in both forms).
// ---------------------------------------------------------------------------
template <int P, class F = Fmt16>
__device__ bool gen_level_upsert(const LevelDev &L, uint64_t key, uint32_t fp, uint32_t word, bool del,
                                 const Tile<P> &t, int ai, uint32_t &val, bool &owner, bool &found) {
    Probe pr;
    pr.begin(L, key, fp);
    const bool linear = L.scheme == PG_SCHEME_LINEAR_BUCKETS;
    if (linear) pr.max_probes = pr.useful_probes(L.nb.d);
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) break;
        const uint64_t b = fastmod(p, L.nb);
        void *base = bucket_ptr(L, b);
        int f;
        t.acc[ai]++;
        if (gen_find<P, F>(base, L.B, fp, t, val, owner, f)) {
            found = true;
            if (del) gen_remove<P, F>(base, L.B, fp, t, f);
            return true;
        }
        if (linear && gen_claim<P, F>(base, L.B, word, false, t, ~0u, L.occ, b) == kClaimed) return true;
    }
    if (linear) return false;
    return gen_shortcut_insert<P, F>(L, key, fp, word, del, t, ai);
}

template <int P, int B, class F = Fmt16>
__device__ int tcf_upsert_key(const TableDev &T, uint64_t key, uint32_t v_in, bool del, const Tile<P> &t,
                              uint32_t &val, bool &owner, bool &found) {
    const uint32_t fp = fp_of<F>(key);
    const uint32_t word = pack_word<F>(fp, v_in);
    found = false;
    bool ok;
    if constexpr (B != 0) {
        static_assert(F::kbits == 16, "register-bucket path is for 16/16 slots");
        const LevelDev &L = T.lv[0];
        uint32_t *slots = (uint32_t *)L.slots;
        Probe pr;
        pr.begin(L, key, fp);
        Bucket<P, B> bk0;
        const uint64_t p0 = pr.at(0);
        const uint64_t b0 = (p0 == kEnd) ? 0 : fastmod(p0, L.nb);
        ok = false;
        if (p0 != kEnd) {
            bk0.load(slots, b0, t);
            if (bk0.find(fp, t, val, owner)) {
                found = true;
                if (del) bk0.remove(slots, b0, fp, t);
                ok = true;
            }
            for (uint32_t j = 1; !ok; j++) {
                const uint64_t p = pr.at(j);
                if (p == kEnd) break;
                const uint64_t b = fastmod(p, L.nb);
                Bucket<P, B> bk;
                bk.load(slots, b, t);
                if (bk.find(fp, t, val, owner)) {
                    found = true;
                    if (del) bk.remove(slots, b, fp, t);
                    ok = true;
                }
            }
        }
        if (!ok) ok = fast_shortcut_insert<P, B>(L, pr, word, del, t, bk0, b0);
    } else {
        ok = gen_level_upsert<P, F>(T.lv[0], key, fp, word, del, t, 0, val, owner, found);
    }
    if (ok) return 1;
    for (uint32_t l = 1; l < T.nlevels; l++)
        if (gen_level_upsert<P, F>(T.lv[l], key, fp, word, del, t, 1, val, owner, found)) return 1 + (int)l;
    return 0;
}

template <int P, int B, class F = Fmt16>
__device__ bool tcf_remove_key(const TableDev &T, uint64_t key, const Tile<P> &t) {
    const uint32_t fp = fp_of<F>(key);
    bool ok;
    if constexpr (B != 0) ok = fast_shortcut_remove<P, B>(T.lv[0], key, fp, t);
    else ok = gen_level_remove<P, F>(T.lv[0], key, fp, t, 0);
    if (ok) return true;
    for (uint32_t l = 1; l < T.nlevels; l++)
        if (gen_level_remove<P, F>(T.lv[l], key, fp, t, 1)) return true;
    return false;
}

// ---------------------------------------------------------------------------
// Kernels.  I = keys per tile (tile_id + i * tiles_in_grid, so the tiles of a
// wave read consecutive keys); ACC = the accounting variant (a distinct
// kernel, so profiles of timed launches are not mixed with counting runs).
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;

// W8: plain inserts of the one-key register path ask the compiler for 8
// waves per SIMD (it then trims the scalar registers that otherwise hold the
// kernel at 7; the same 55 VGPRs, no scratch).  insert_with_delete keeps 7:
// with more keys in flight the fills they read are staler, and under
// sustained churn with tombstone reuse that fills the backing level faster.
template <int P, int B, int I, bool ACC, class F = Fmt16, bool W8 = false>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(W8 ? 8 : 1, 8)))
void k_tcf_insert(TableDev T, OpArgs a) {
    Tile<P> t;
    t.concurrent = true;
    const uint64_t tile_id = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    const uint64_t tiles = (uint64_t(gridDim.x) * kBlock) / P;
    const bool wd = a.with_delete != 0;
    uint64_t idx[I];
    bool live[I];
    int st[I];
    if constexpr (B != 0) {
        const LevelDev &L = T.lv[0];
        uint64_t key[I], b0[I];
        uint32_t fp[I], word[I];
        Probe pr[I];
        Bucket<P, B> bk0[I];
#pragma unroll
        for (int i = 0; i < I; i++) {
            idx[i] = tile_id + (uint64_t)i * tiles;
            live[i] = idx[i] < a.n;
            key[i] = live[i] ? a.keys[idx[i]] : 0;
            const uint32_t val = (live[i] && a.vals) ? ((const uint16_t *)a.vals)[idx[i]] : 0u;
            fp[i] = fp16_of(key[i]);  // B != 0: 16/16 slots only
            word[i] = (val << 16) | fp[i];
            pr[i].begin(L, key[i], fp[i]);
            const uint64_t p0 = pr[i].at(0);
            b0[i] = (p0 == kEnd) ? 0 : fastmod(p0, L.nb);
            if (live[i] && p0 != kEnd) bk0[i].load((const uint32_t *)L.slots, b0[i], t);
        }
#pragma unroll
        for (int i = 0; i < I; i++) {
            st[i] = 1;
            if (live[i]) st[i] = tcf_insert_prepared<P, B>(T, key[i], fp[i], word[i], wd, t, pr[i], bk0[i], b0[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < I; i++) {
            idx[i] = tile_id + (uint64_t)i * tiles;
            live[i] = idx[i] < a.n;
            st[i] = 1;
            if (live[i]) {
                const uint32_t val = a.vals ? ((const typename F::V *)a.vals)[idx[i]] : 0u;
                st[i] = tcf_insert_key<P, 0, F>(T, a.keys[idx[i]], val, wd, t);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < I; i++) {
        if (live[i] && a.flags && t.leader()) a.flags[idx[i]] = (uint8_t)st[i];
        wave_count(a.fail_count, live[i] && t.leader() && st[i] == 0);
    }
    if constexpr (ACC) flush_acc(a.stats, t, live[0]);
}

template <int P, int B, bool ACC, class F = Fmt16>
__global__ __launch_bounds__(kBlock) void k_tcf_remove(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    if (tid < a.n) {
        const bool ok = tcf_remove_key<P, B, F>(T, a.keys[tid], t);
        if (a.flags && t.leader()) a.flags[tid] = ok ? 1 : 0;
    }
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P, int B, class F = Fmt16>
__device__ inline void upsert_one(const TableDev &T, const OpArgs &a, uint64_t i, const Tile<P> &t,
                                  uint64_t &fails) {
    const uint32_t v_in = a.vals ? ((const typename F::V *)a.vals)[i] : 0u;
    uint32_t val = 0;
    bool owner = false, found = false;
    const int st = tcf_upsert_key<P, B, F>(T, a.keys[i], v_in, a.with_delete != 0, t, val, owner, found);
    if (t.leader()) {
        if (a.flags) a.flags[i] = (uint8_t)st;
        if (a.found) a.found[i] = found ? 1 : 0;
        if (a.vals_out && !found) ((typename F::V *)a.vals_out)[i] = 0;
    }
    if (found && owner && a.vals_out) ((typename F::V *)a.vals_out)[i] = (typename F::V)val;
    fails += (st == 0);
}

template <int P, int B, bool ACC, class F = Fmt16>
__global__ __launch_bounds__(kBlock) void k_tcf_upsert(TableDev T, OpArgs a) {
    Tile<P> t;
    t.concurrent = true;
    const uint64_t tid = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    uint64_t fails = 0;
    if (tid < a.n) upsert_one<P, B, F>(T, a, tid, t, fails);
    wave_count(a.fail_count, tid < a.n && t.leader() && fails);
    if constexpr (ACC) flush_acc(a.stats, t, tid < a.n);
}

template <int P, int B, class F = Fmt16>
__global__ __launch_bounds__(64) void k_tcf_upsert_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    uint64_t fails = 0;
    for (uint64_t i = 0; i < a.n; i++) {
        upsert_one<P, B, F>(T, a, i, t, fails);
        fence_agent();
    }
    if (a.fail_count && t.leader() && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
    flush_acc(a.stats, t, true);
}

// Deterministic mode: one tile applies the keys in array order.  The fence
// after every key makes the next key's loads observe this key's CAS.
template <int P, int B, class F = Fmt16>
__global__ __launch_bounds__(64) void k_tcf_insert_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    uint64_t fails = 0;
    for (uint64_t i = 0; i < a.n; i++) {
        const uint64_t key = a.keys[i];
        const uint32_t val = a.vals ? ((const typename F::V *)a.vals)[i] : 0u;
        const int st = tcf_insert_key<P, B, F>(T, key, val, a.with_delete != 0, t);
        if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
        fails += (st == 0);
        fence_agent();
    }
    if (a.fail_count && t.leader() && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
    flush_acc(a.stats, t, true);
}

template <int P, int B, class F = Fmt16>
__global__ __launch_bounds__(64) void k_tcf_remove_seq(TableDev T, OpArgs a) {
    Tile<P> t;
    for (uint64_t i = 0; i < a.n; i++) {
        const bool ok = tcf_remove_key<P, B, F>(T, a.keys[i], t);
        if (a.flags && t.leader()) a.flags[i] = ok ? 1 : 0;
        fence_agent();
    }
    flush_acc(a.stats, t, true);
}

// Query (power_of_n_shortcut_buckets.cuh:676-707, then the backing level,
// bucketed_table.cuh:332-349).  The level-0 bucket loads of all I keys are
// issued before any is examined, so a wave keeps 64/P * I random 64 B reads
// in flight.
template <int P, int B, int I, bool ACC, class F = Fmt16>
__global__ __launch_bounds__(kBlock) void k_tcf_query(TableDev T, OpArgs a) {
    Tile<P> t;
    const uint64_t tile_id = (uint64_t(blockIdx.x) * kBlock + threadIdx.x) / P;
    const uint64_t tiles = (uint64_t(gridDim.x) * kBlock) / P;
    const LevelDev &L = T.lv[0];
    const uint32_t *slots = (const uint32_t *)L.slots;

    uint64_t idx[I], key[I];
    uint32_t fp[I], val[I];
    bool live[I], pend[I], found[I], own[I];
    Probe pr[I];
#pragma unroll
    for (int i = 0; i < I; i++) {
        idx[i] = tile_id + (uint64_t)i * tiles;
        live[i] = idx[i] < a.n;
        key[i] = live[i] ? a.keys[idx[i]] : 0;
        fp[i] = fp_of<F>(key[i]);
        pr[i].begin(L, key[i], fp[i]);
        pend[i] = live[i];
        found[i] = false;
        own[i] = false;
        val[i] = 0;
    }
    if constexpr (B != 0) {
        for (uint32_t j = 0; j < L.max_probes; j++) {
            Bucket<P, B> bk[I];
#pragma unroll
            for (int i = 0; i < I; i++) {
                if (pend[i]) {
                    const uint64_t p = pr[i].at(j);
                    if (p == kEnd) pend[i] = false;
                    else bk[i].load(slots, fastmod(p, L.nb), t);
                }
            }
            bool more = false;
#pragma unroll
            for (int i = 0; i < I; i++) {
                if (pend[i] && bk[i].find(fp[i], t, val[i], own[i])) {
                    found[i] = true;
                    pend[i] = false;
                }
                more |= pend[i];
            }
            if (!__ballot(more)) break;
        }
    } else {
#pragma unroll
        for (int i = 0; i < I; i++)
            if (live[i]) found[i] = gen_level_query<P, F>(L, key[i], fp[i], t, val[i], own[i], 0);
    }
#pragma unroll
    for (int i = 0; i < I; i++) {
        if (live[i] && !found[i]) {
            for (uint32_t l = 1; l < T.nlevels && !found[i]; l++)
                found[i] = gen_level_query<P, F>(T.lv[l], key[i], fp[i], t, val[i], own[i], 1);
        }
        if (live[i]) {
            const bool writer = found[i] ? own[i] : t.leader();
            if (writer) {
                if (a.flags) a.flags[idx[i]] = found[i] ? 1 : 0;
                if (a.vals_out) ((typename F::V *)a.vals_out)[idx[i]] = found[i] ? (typename F::V)val[i] : 0;
            }
        }
    }
    if constexpr (ACC) flush_acc(a.stats, t, live[0]);
}

// ---------------------------------------------------------------------------
// Dispatch for one tile width (instantiated once per P in pg_tcf_p<P>.hip).
// ---------------------------------------------------------------------------
inline unsigned tcf_grid(uint64_t tiles, int P) {
    const uint64_t threads = tiles * (uint64_t)P;
    return (unsigned)((threads + kBlock - 1) / kBlock);
}

template <int P, int B, int I, bool ACC, class F = Fmt16>
void launch_tcf_kernel(Op op, const TableDev &T, const OpArgs &a, hipStream_t s) {
    const uint64_t tiles = (a.n + I - 1) / I;
    constexpr bool w8 = !ACC && I == 1 && B != 0 && B <= 16 && P <= 16;
    if (op == Op::Query)
        hipLaunchKernelGGL((k_tcf_query<P, B, I, ACC, F>), dim3(tcf_grid(tiles, P)), dim3(kBlock), 0, s, T, a);
    else if (w8 && !a.with_delete)
        hipLaunchKernelGGL((k_tcf_insert<P, B, I, ACC, F, w8>), dim3(tcf_grid(tiles, P)), dim3(kBlock), 0, s, T, a);
    else
        hipLaunchKernelGGL((k_tcf_insert<P, B, I, ACC, F>), dim3(tcf_grid(tiles, P)), dim3(kBlock), 0, s, T, a);
}

template <int P, int B, class F = Fmt16>
int launch_tcf_pb(const TableDev &T, Op op, const OpArgs &a, hipStream_t s) {
    if (a.n == 0) return PG_OK;
    const bool acc = a.stats != nullptr;
    if (op != Op::Query && a.deterministic) {
        if (op == Op::Insert) hipLaunchKernelGGL((k_tcf_insert_seq<P, B, F>), dim3(1), dim3(P), 0, s, T, a);
        else if (op == Op::Upsert) hipLaunchKernelGGL((k_tcf_upsert_seq<P, B, F>), dim3(1), dim3(P), 0, s, T, a);
        else hipLaunchKernelGGL((k_tcf_remove_seq<P, B, F>), dim3(1), dim3(P), 0, s, T, a);
    } else if (op == Op::Upsert) {
        if (acc) hipLaunchKernelGGL((k_tcf_upsert<P, B, true, F>), dim3(tcf_grid(a.n, P)), dim3(kBlock), 0, s, T, a);
        else hipLaunchKernelGGL((k_tcf_upsert<P, B, false, F>), dim3(tcf_grid(a.n, P)), dim3(kBlock), 0, s, T, a);
    } else if (op == Op::Remove) {
        if (acc) hipLaunchKernelGGL((k_tcf_remove<P, B, true, F>), dim3(tcf_grid(a.n, P)), dim3(kBlock), 0, s, T, a);
        else hipLaunchKernelGGL((k_tcf_remove<P, B, false, F>), dim3(tcf_grid(a.n, P)), dim3(kBlock), 0, s, T, a);
    } else if (acc) {
        launch_tcf_kernel<P, B, 1, true, F>(op, T, a, s);
    } else if constexpr (B == 0) {
        launch_tcf_kernel<P, B, 1, false, F>(op, T, a, s);
    } else {
        switch (a.items) {
            case 2: launch_tcf_kernel<P, B, 2, false, F>(op, T, a, s); break;
            case 4: launch_tcf_kernel<P, B, 4, false, F>(op, T, a, s); break;
            default: launch_tcf_kernel<P, B, 1, false, F>(op, T, a, s); break;
        }
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

template <int P>
int launch_tcf_p(const TableDev &T, Op op, const OpArgs &a, hipStream_t s) {
    const LevelDev &L = T.lv[0];
    if (T.rep == PG_REP_FP8_V8) return launch_tcf_pb<P, 0, Fmt8>(T, op, a, s);
    if (T.rep == PG_REP_FP10_V6) return launch_tcf_pb<P, 0, Fmt10>(T, op, a, s);
    if (T.rep == PG_REP_FP12) return launch_tcf_pb<P, 0, Fmt12>(T, op, a, s);
    if (L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT && L.max_probes <= 2) {  // fast_shortcut_insert's shape
        switch (L.B) {
            case 8: return launch_tcf_pb<P, 8>(T, op, a, s);
            case 16: return launch_tcf_pb<P, 16>(T, op, a, s);
            case 32: return launch_tcf_pb<P, 32>(T, op, a, s);
            default: break;
        }
    }
    return launch_tcf_pb<P, 0>(T, op, a, s);
}

}  // namespace pg
