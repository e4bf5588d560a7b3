// pg_kv_ops.hpp -- per-key device ops for key_val_pair<u64,u64>
// tables (representations/key_val_pair.cuh:63-171), for the key-only
// dynamic_container<key_container, uint64_t> tables (the CMS allocator's
// free table, allocators/cms.cuh:203) and for the 4-byte
// dynamic_container<key_val_pair<u64,u16>, u16> slots of the flat TCF,
// gfx950 wave64.
//
// Layout in HBM: 16 B slots {key, val} (SlotKV), 8 B keys (SlotKey) or
// 32-bit words val16 << 16 | key16 (SlotKV16), B contiguous slots per
// bucket, bucket b at slot b*B.  Empty key 0, tombstone key all ones.
//
// Two tile shapes, chosen per table:
//   * stripe mode (B > 1): lane r of a P-lane tile owns slot stripe*P + r of
//     the probe bucket and loads it with one 16 B global_load_dwordx4; this is
//     the reference's stripe scan (insert_schemes/bucket_insert.cuh:127-357)
//     with the same early exit (a stripe holding an empty slot ends a query).
//     When B is 2P or 4P the whole bucket is loaded at once (KvBucket: each
//     lane issues its 2 or 4 slot loads together) and the stripes are then
//     evaluated in the same order, so a walk over several stripes costs one
//     round trip instead of one per stripe.
//   * probe-parallel mode (B == 1, P > 1): lane r evaluates probe j0 + r, so
//     one tile step examines P consecutive probes (for linear probing, P
//     adjacent slots in one or two 64 B sectors).  The lowest decisive lane
//     (match, empty, or end of the probe sequence) decides, which is exactly
//     the reference's probe-by-probe order.
#pragma once

#include <type_traits>

#include "pg_internal.hpp"
#include "pg_tile.hpp"

namespace pg {

constexpr uint64_t kKvEmpty = 0ULL;
constexpr uint64_t kKvTomb = ~0ULL;

// Slot layouts.  Probing always hashes the key as given; the stored key is
// stored(key): the key itself for key_val_pair, get_smallKey(key) for the
// dynamic_container slots (0 -> 1, all-ones -> all-ones - 1,
// dynamic_container.cuh:35-50).  tag(key) is rep_type::tag, the probe tag of
// the xor prober: the key for key_val_pair, the small key otherwise.
// claim(p, cur, sk, v) takes slot p, seen as `cur` with an empty key;
// bury(p, cur, sk) turns the key of slot p, seen as `cur` holding sk, into
// the tombstone.  Both return whether their swap landed.
struct SlotKV {
    using S = ulonglong2;
    using V = uint64_t;  // value buffer type
    __device__ static uint64_t stored(uint64_t key) { return key; }
    __device__ static uint64_t tag(uint64_t key) { return key; }
    __device__ static S load(const S *p) { return *p; }
    __device__ static S none() { return make_ulonglong2(kKvTomb, 0); }
    __device__ static uint64_t key(const S &s) { return s.x; }
    __device__ static uint64_t val(const S &s) { return s.y; }
    // key_val_pair::atomic_swap (key_val_pair.cuh:84-95) CASes the key from
    // empty and then stores the value, so a reader can see the key before its
    // value.  Here the slot is claimed with the tombstone key instead, which
    // every op already reads as occupied and never matching: inserts do not
    // claim it (key-value slots never reuse tombstones), queries and removes
    // walk past it.  Key and value are then written by one 16 B store, so a
    // reader's 16 B slot load sees the slot empty, busy, or holding the key
    // with its value, never the key alone.  The final slot bytes are the
    // reference's.
    __device__ static bool claim(S *p, const S &, uint64_t sk, uint64_t v) {
        if (atomicCAS((unsigned long long *)&p->x, kKvEmpty, kKvTomb) != kKvEmpty) return false;
        *p = make_ulonglong2(sk, v);
        return true;
    }
    // key_val_pair::atomic_reset: CAS key -> tombstone key, value kept
    // (key_val_pair.cuh:143-151)
    __device__ static bool bury(S *p, const S &, uint64_t sk) {
        return atomicCAS((unsigned long long *)&p->x, sk, kKvTomb) == sk;
    }
    // The replace scheme's swap on a slot seen holding sk: the value word is
    // stored alone, so a reader's 16 B load sees the old or the new value.
    __device__ static bool replace(S *p, const S &, uint64_t, uint64_t v) {
        __hip_atomic_store(&p->y, (unsigned long long)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
};
struct SlotKey {
    using S = uint64_t;
    using V = uint16_t;  // the CMS table's Val; queries report 0
    __device__ static uint64_t stored(uint64_t key) {
        key += (key == kKvEmpty);
        key -= (key == kKvTomb);
        return key;
    }
    __device__ static uint64_t tag(uint64_t key) { return stored(key); }
    __device__ static S load(const S *p) { return *p; }
    __device__ static S none() { return kKvTomb; }
    __device__ static uint64_t key(const S &s) { return s; }
    __device__ static uint64_t val(const S &) { return 0; }
    __device__ static bool claim(S *p, const S &, uint64_t sk, uint64_t) {
        return atomicCAS((unsigned long long *)p, kKvEmpty, sk) == kKvEmpty;
    }
    __device__ static bool bury(S *p, const S &, uint64_t sk) {
        return atomicCAS((unsigned long long *)p, sk, kKvTomb) == sk;
    }
    // key only: a slot holding the key already is the result
    __device__ static bool replace(S *, const S &, uint64_t, uint64_t) { return true; }
};
// dynamic_container<key_val_pair<uint64_t, uint16_t>, uint16_t>: one 32-bit
// word val16 << 16 | key16 (the flat TCF's slot, data_structs/tcf.cuh:29-30).
// The reference swaps the 16-bit key and then writes the value; here one
// 32-bit CAS writes both, from the word seen empty, so a reader never sees
// the key without its value.  A tombstone keeps the value bits
// (key_val_pair::atomic_reset sets the key only).
struct SlotKV16 {
    using S = uint32_t;
    using V = uint16_t;
    __device__ static uint64_t stored(uint64_t key) { return fp16_of(key); }
    __device__ static uint64_t tag(uint64_t key) { return fp16_of(key); }
    __device__ static S load(const S *p) { return *p; }
    __device__ static S none() { return 0xFFFFu; }
    __device__ static uint64_t key(const S &s) { return s & 0xFFFFu; }
    __device__ static uint64_t val(const S &s) { return s >> 16; }
    __device__ static bool claim(S *p, const S &cur, uint64_t sk, uint64_t v) {
        return atomicCAS(p, cur, ((uint32_t)v << 16) | (uint32_t)sk) == cur;
    }
    __device__ static bool bury(S *p, const S &cur, uint64_t) {
        return atomicCAS(p, cur, (cur & 0xFFFF0000u) | 0xFFFFu) == cur;
    }
    // New value bits under the same key16; repeats only while another thread
    // changes the word and the key stays, so it ends once the swap lands or
    // the key is gone (then the slot is no longer a candidate).
    __device__ static bool replace(S *p, const S &cur, uint64_t sk, uint64_t v) {
        S w = cur;
        while ((w & 0xFFFFu) == (uint32_t)sk) {
            const S prev = atomicCAS(p, w, ((uint32_t)v << 16) | (uint32_t)sk);
            if (prev == w) return true;
            w = prev;
        }
        return false;
    }
};

// Lanes strictly below the lowest set bit of `endm` (all lanes when 0).
__device__ inline uint64_t below_first(uint64_t endm) {
    return endm ? ((1ull << ctz64(endm)) - 1ull) : ~0ull;
}

// A probe bucket of NS stripes of P slots held in registers: lane r holds
// slot k*P + r of stripe k.  The NS loads are independent, so they are all in
// flight together.
template <int P, int NS, class SL>
struct KvBucket {
    typename SL::S s[NS];
    __device__ inline void load(const typename SL::S *slots, uint64_t base, const Tile<P> &t) {
#pragma unroll
        for (int k = 0; k < NS; k++) s[k] = SL::load(slots + base + k * P + t.rank);
    }
};

// The stripe scans below over a register bucket: the same decisions in the
// same order as the stripe-by-stripe walks of kv_level_insert / _query /
// _remove (bucket_insert.cuh:127-357), on a bucket loaded in one round trip.
template <int P, int NS, class SL>
__device__ bool kv_reg_insert(const LevelDev &L, uint64_t key, uint64_t val, const Tile<P> &t) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t base = fastmod(p, L.nb) * (uint64_t)(NS * P);
        KvBucket<P, NS, SL> bk;
        bk.load(slots, base, t);
#pragma unroll
        for (int k = 0; k < NS; k++) {
            t.acc[0] += P;
            uint64_t cm = t.bits(SL::key(bk.s[k]) == kKvEmpty);
            while (cm) {
                const uint32_t l = (uint32_t)ctz64(cm);
                bool success = false;
                if (t.rank == l) success = SL::claim(slots + base + k * P + t.rank, bk.s[k], sk, val);
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                cm &= cm - 1ull;
            }
        }
    }
}

template <int P, int NS, class SL>
__device__ bool kv_reg_query(const LevelDev &L, uint64_t key, const Tile<P> &t, uint64_t &val, bool &owner) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    const S *slots = (const S *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        KvBucket<P, NS, SL> bk;
        bk.load(slots, fastmod(p, L.nb) * (uint64_t)(NS * P), t);
#pragma unroll
        for (int k = 0; k < NS; k++) {
            t.acc[0] += P;
            const uint64_t mm = t.bits(SL::key(bk.s[k]) == sk);
            if (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                owner = (t.rank == l);
                if (owner) val = SL::val(bk.s[k]);
                return true;
            }
            if (t.any(SL::key(bk.s[k]) == kKvEmpty)) return false;
        }
    }
}

template <int P, int NS, class SL>
__device__ bool kv_reg_remove(const LevelDev &L, uint64_t key, const Tile<P> &t) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const uint64_t base = fastmod(p, L.nb) * (uint64_t)(NS * P);
        KvBucket<P, NS, SL> bk;
        bk.load(slots, base, t);
#pragma unroll
        for (int k = 0; k < NS; k++) {
            const bool match = SL::key(bk.s[k]) == sk;
            t.acc[0] += P;
            uint64_t mm = t.bits(match);
            while (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                bool success = false;
                if (t.rank == l) success = SL::bury(slots + base + k * P + t.rank, bk.s[k], sk);
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                mm &= mm - 1ull;
            }
            if (t.any(!match && SL::key(bk.s[k]) == kKvEmpty)) return false;
        }
    }
}

// ---------------------------------------------------------------------------
// Flat power-of-n schemes over key-value slots: power_of_n_insert_scheme
// (insert_schemes/power_of_n.cuh:39-492) and power_of_n_insert_shortcut_scheme
// (power_of_n_shortcut.cuh:39-505).  A bucket's fill counts every non-empty
// slot, tombstones included (check_fill_bucket, power_of_n_shortcut.cuh:
// 234-261); a claim takes the lowest empty slot (insert_into_bucket,
// :134-191); queries and removes check every probe bucket in probe order and
// act on the lowest matching slot (:193-232, :263-325, :437-505).  Runtime B, stripes of P slots,
// each bucket read again for the claim as in the reference.
// ---------------------------------------------------------------------------
template <int P, class SL>
__device__ int kv_pow_fill(const typename SL::S *bk, uint32_t B, const Tile<P> &t) {
    int f = 0;
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        f += t.popc(i < B && SL::key(SL::load(bk + i)) != kKvEmpty);
    }
    t.acc[0] += B;
    return f;
}

// The fill and whether a slot holds sk (the replace scheme's
// check_fill_bucket, power_of_n_shortcut_key_replace.cuh:233-270).
template <int P, class SL>
__device__ int kv_pow_fill_find(const typename SL::S *bk, uint32_t B, uint64_t sk, const Tile<P> &t, bool &found) {
    int f = 0;
    bool hit = false;
    for (uint32_t s0 = 0; s0 < B; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const uint64_t k = i < B ? SL::key(SL::load(bk + i)) : kKvEmpty;
        f += t.popc(k != kKvEmpty);
        hit |= t.any(i < B && k == sk);
    }
    t.acc[0] += B;
    found = hit;
    return f;
}

// The lowest empty slot among [0, lim); with RP (the replace scheme), the
// lowest slot that is empty or holds sk, whose value is then replaced
// (insert_into_bucket, power_of_n_shortcut_key_replace.cuh:134-190).
template <int P, class SL, bool RP = false>
__device__ bool kv_pow_claim(typename SL::S *bk, uint32_t B, uint32_t lim, uint64_t sk, uint64_t val,
                             const Tile<P> &t) {
    using S = typename SL::S;
    const uint32_t end = lim < B ? lim : B;
    for (uint32_t s0 = 0; s0 < end; s0 += P) {
        const uint32_t i = s0 + t.rank;
        const bool ok = i < end;
        const S s = ok ? SL::load(bk + i) : SL::none();
        t.acc[0] += (end - s0 < P) ? end - s0 : P;
        uint64_t cm = t.bits(ok && (SL::key(s) == kKvEmpty || (RP && SL::key(s) == sk)));
        while (cm) {
            const uint32_t l = (uint32_t)ctz64(cm);
            bool success = false;
            if (t.rank == l) {
                if constexpr (RP) {
                    success = SL::key(s) == kKvEmpty ? SL::claim(bk + i, s, sk, val) : SL::replace(bk + i, s, sk, val);
                } else {
                    success = SL::claim(bk + i, s, sk, val);
                }
            }
            if (t.any(success)) {
                t.acc[2]++;
                return true;
            }
            cm &= cm - 1ull;
        }
    }
    return false;
}

// power_of_n.cuh:326-422 / power_of_n_shortcut.cuh:327-435.  The shortcut
// claims in a probe bucket whose fill is below 3B/4.  Since empties are
// claimed lowest first and slots never return to empty, the non-empty slots
// of a bucket are always a prefix, so a claim at slot i lands at fill i:
// concurrent kernels bound the shortcut claim to slots below ceil(3B/4),
// which applies the rule at the moment of the claim, and a bucket whose
// bounded claim fails enters the fill-ordered pass at fill ceil(3B/4).  In
// key order the bound is B, and a failed claim means the bucket is full.
//
// The replace scheme (power_of_n_shortcut_key_replace.cuh:336-444) is the
// shortcut scheme whose first pass also enters a probe bucket that holds the
// key, whatever its fill, and whose claims take the lowest slot that is empty
// or holds the key.  By the prefix property no empty slot lies below a held
// key, so such a claim replaces the value; the bound is then B, as the key's
// slot may lie past 3B/4.  RP selects it at compile time, so the other
// schemes' kernels carry none of its code.
template <int P, class SL, bool RP = false>
__device__ bool kv_pow_insert(const LevelDev &L, uint64_t key, uint64_t val, const Tile<P> &t) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    const bool shortcut = RP || L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS;
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    const int B = (int)L.B;
    const uint32_t lim1 = t.concurrent ? (3u * L.B + 3u) / 4u : L.B;
    uint64_t bidx[kMaxShortcutProbes];
    int fill[kMaxShortcutProbes];
    int np = 0, min_fill = B;
#pragma unroll
    for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) break;
        const uint64_t b = fastmod(p, L.nb);
        bidx[j] = b;
        bool found = false;
        int f;
        if constexpr (RP) f = kv_pow_fill_find<P, SL>(slots + b * L.B, L.B, sk, t, found);
        else f = kv_pow_fill<P, SL>(slots + b * L.B, L.B, t);
        if (shortcut && (4 * f < 3 * B || found)) {
            if (kv_pow_claim<P, SL, RP>(slots + b * L.B, L.B, found ? L.B : lim1, sk, val, t)) return true;
            f = (int)lim1;
        }
        fill[j] = f;
        if (f < min_fill) min_fill = f;
        np = j + 1;
    }
    if (t.concurrent) {
        // Fills read while many keys are in flight are stale: every key of a
        // fresh table reads zeros and would pile into its first probe bucket.
        // So the least-filled bucket (first in probe order on ties) is
        // claimed only at a slot index no larger than the next-least fill,
        // which by the prefix property keeps it no fuller than the others at
        // the claim.  A bounded claim that fails has found every slot below
        // its bound taken, so that bucket's known fill rises past the bound;
        // the fills read again each round only raise the known fills (a read
        // can come from a stale cache line).  Every round raises the sum of
        // the known fills, so np*B rounds suffice.
        for (int round = 0; round <= np * B; round++) {
            if (round) {
#pragma unroll
                for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
                    if (j < np) {
                        const int f = kv_pow_fill<P, SL>(slots + bidx[j] * L.B, L.B, t);
                        if (f > fill[j]) fill[j] = f;
                    }
                }
            }
            int best = -1, bf = B, second = B;
#pragma unroll
            for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
                if (j < np) {
                    if (fill[j] < bf) {
                        second = bf;
                        bf = fill[j];
                        best = j;
                    } else if (fill[j] < second) {
                        second = fill[j];
                    }
                }
            }
            if (best < 0) return false;  // every probe bucket is full
            uint64_t bb = bidx[0];
#pragma unroll
            for (int j = 1; j < (int)kMaxShortcutProbes; j++) bb = (j == best) ? bidx[j] : bb;
            if (kv_pow_claim<P, SL, RP>(slots + bb * L.B, L.B, (uint32_t)second + 1u, sk, val, t)) return true;
#pragma unroll
            for (int j = 0; j < (int)kMaxShortcutProbes; j++)
                if (j == best) fill[j] = second + 1 < B ? second + 1 : B;
        }
        return false;
    }
    int lvl = min_fill;
    while (lvl < B) {
        int next = B;
#pragma unroll
        for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
            if (j < np) {
                if (fill[j] == lvl && kv_pow_claim<P, SL, RP>(slots + bidx[j] * L.B, L.B, L.B, sk, val, t))
                    return true;
                if (fill[j] > lvl && fill[j] < next) next = fill[j];
            }
        }
        lvl = next;
    }
    return false;
}

template <int P, class SL>
__device__ bool kv_pow_query(const LevelDev &L, uint64_t key, const Tile<P> &t, uint64_t &val, bool &owner) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    const uint64_t sk = SL::stored(key);
    const S *slots = (const S *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        const S *bk = slots + fastmod(p, L.nb) * L.B;
        for (uint32_t s0 = 0; s0 < L.B; s0 += P) {
            const uint32_t i = s0 + t.rank;
            const bool ok = i < L.B;
            const S s = ok ? SL::load(bk + i) : SL::none();
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            const uint64_t mm = t.bits(ok && SL::key(s) == sk);
            if (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                owner = (t.rank == l);
                if (owner) val = SL::val(s);
                return true;
            }
        }
    }
}

template <int P, class SL>
__device__ bool kv_pow_remove(const LevelDev &L, uint64_t key, const Tile<P> &t) {
    using S = typename SL::S;
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        S *bk = slots + fastmod(p, L.nb) * L.B;
        for (uint32_t s0 = 0; s0 < L.B; s0 += P) {
            const uint32_t i = s0 + t.rank;
            const bool ok = i < L.B;
            const S s = ok ? SL::load(bk + i) : SL::none();
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            uint64_t mm = t.bits(ok && SL::key(s) == sk);
            while (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                bool success = false;
                if (t.rank == l) success = SL::bury(bk + i, s, sk);
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                mm &= mm - 1ull;
            }
        }
    }
}

// The flat schemes over 4-byte slots with B % V == 0 and B <= V*P, V = 4 or
// 8 (the flat TCF's 16-slot buckets at tile 4 or 2): the whole bucket is one
// row, lane r holding slots V*r..V*r+V-1 from V/4 16 B loads.  A row read once serves the fill
// and the claim: slots only fill up, so a row can only show more empties
// than the table holds, and trying them lowest first with a CAS each claims
// the lowest slot that is still empty.
template <int V>
__device__ inline uint32_t row_word(const uint32_t (&w)[V], int v) {
    uint32_t x = w[0];
#pragma unroll
    for (int i = 1; i < V; i++) x = (v == i) ? w[i] : x;
    return x;
}

template <int P, int V>
struct Row16 {
    static_assert(V == 4 || V == 8, "rows of 4 or 8 words per lane");
    uint32_t w[V];
    bool ok;
    __device__ inline void load(const uint32_t *bk, uint32_t B, const Tile<P> &t) {
        ok = (uint32_t)V * t.rank < B;
#pragma unroll
        for (int c = 0; c < V / 4; c++) {
            uint32_t q[4] = {0u, 0u, 0u, 0u};
            if (ok) load_words<4>(bk + V * t.rank + 4 * c, q);
#pragma unroll
            for (int v = 0; v < 4; v++) w[4 * c + v] = q[v];
        }
        t.acc[0] += B;
    }
    __device__ inline int fill(const Tile<P> &t) const {
        int f = 0;
#pragma unroll
        for (int v = 0; v < V; v++) f += t.popc(ok && (w[v] & 0xFFFFu) != 0u);
        return f;
    }
    // The lowest slot below `lim` seen empty.
    __device__ inline bool claim(uint32_t *bk, uint32_t lim, uint32_t word, const Tile<P> &t) const {
        uint32_t cand = 0;
        if (ok) {
#pragma unroll
            for (int v = 0; v < V; v++)
                if ((w[v] & 0xFFFFu) == 0u && (uint32_t)V * t.rank + (uint32_t)v < lim) cand |= 1u << v;
        }
        uint64_t lanes = t.bits(cand != 0);
        while (lanes) {
            const uint32_t l = (uint32_t)ctz64(lanes);
            bool success = false;
            if (t.rank == l) {
                const int v = __builtin_ctz(cand);
                const uint32_t cur = row_word<V>(w, v);
                success = atomicCAS(bk + V * t.rank + v, cur, word) == cur;
                cand &= cand - 1u;
            }
            if (t.any(success)) {
                t.acc[2]++;
                return true;
            }
            lanes = t.bits(cand != 0);
        }
        return false;
    }
    // The lowest slot holding key16 `sk`; its lane receives the value.
    __device__ inline bool find(uint32_t sk, const Tile<P> &t, uint64_t &val, bool &owner) const {
        uint32_t m = 0;
        if (ok) {
#pragma unroll
            for (int v = 0; v < V; v++)
                if ((w[v] & 0xFFFFu) == sk) m |= 1u << v;
        }
        const uint64_t lanes = t.bits(m != 0);
        if (!lanes) return false;
        const uint32_t l = (uint32_t)ctz64(lanes);
        owner = (t.rank == l);
        if (owner) val = row_word<V>(w, __builtin_ctz(m)) >> 16;
        return true;
    }
    // Tombstone the lowest slot holding `sk`, keeping its value bits.
    __device__ inline bool bury(uint32_t *bk, uint32_t sk, const Tile<P> &t) const {
        uint32_t m = 0;
        if (ok) {
#pragma unroll
            for (int v = 0; v < V; v++)
                if ((w[v] & 0xFFFFu) == sk) m |= 1u << v;
        }
        uint64_t lanes = t.bits(m != 0);
        while (lanes) {
            const uint32_t l = (uint32_t)ctz64(lanes);
            bool success = false;
            if (t.rank == l) {
                const int v = __builtin_ctz(m);
                const uint32_t cur = row_word<V>(w, v);
                success = atomicCAS(bk + V * t.rank + v, cur, (cur & 0xFFFF0000u) | 0xFFFFu) == cur;
                m &= m - 1u;
            }
            if (t.any(success)) {
                t.acc[2]++;
                return true;
            }
            lanes = t.bits(m != 0);
        }
        return false;
    }
};

// kv_pow_insert over rows: the same decisions.  Only the fills are kept
// across probes: the shortcut claims from the row it has just read, and the
// fill-ordered pass (rare under the shortcut) reads the chosen bucket's row
// again before its claim, so one row of registers serves any probe count.
template <int P, int V>
__device__ bool kv16_pow_insert(const LevelDev &L, uint64_t key, uint64_t val, const Tile<P> &t) {
    const uint32_t sk = fp16_of(key);
    Probe pr;
    pr.begin(L, key, sk);
    const bool shortcut = L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS;
    const uint32_t word = ((uint32_t)val << 16) | sk;
    uint32_t *slots = (uint32_t *)L.slots;
    const int B = (int)L.B;
    const uint32_t lim1 = t.concurrent ? (3u * L.B + 3u) / 4u : L.B;
    uint64_t bidx[kMaxShortcutProbes];
    int fill[kMaxShortcutProbes];
    int np = 0, min_fill = B;
#pragma unroll
    for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) break;
        const uint64_t b = fastmod(p, L.nb);
        bidx[j] = b;
        Row16<P, V> row;
        row.load(slots + b * L.B, L.B, t);
        int f = row.fill(t);
        if (shortcut && 4 * f < 3 * B) {
            if (row.claim(slots + b * L.B, lim1, word, t)) return true;
            f = (int)lim1;
        }
        fill[j] = f;
        if (f < min_fill) min_fill = f;
        np = j + 1;
    }
    if (t.concurrent) {
        // the bounded fill-ordered pass of kv_pow_insert
        for (int round = 0; round <= np * B; round++) {
            if (round) {
#pragma unroll
                for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
                    if (j < np) {
                        Row16<P, V> row;
                        row.load(slots + bidx[j] * L.B, L.B, t);
                        const int f = row.fill(t);
                        if (f > fill[j]) fill[j] = f;
                    }
                }
            }
            int best = -1, bf = B, second = B;
#pragma unroll
            for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
                if (j < np) {
                    if (fill[j] < bf) {
                        second = bf;
                        bf = fill[j];
                        best = j;
                    } else if (fill[j] < second) {
                        second = fill[j];
                    }
                }
            }
            if (best < 0) return false;
            uint64_t bb = bidx[0];
#pragma unroll
            for (int j = 1; j < (int)kMaxShortcutProbes; j++) bb = (j == best) ? bidx[j] : bb;
            Row16<P, V> row;
            row.load(slots + bb * L.B, L.B, t);
            if (row.claim(slots + bb * L.B, (uint32_t)second + 1u, word, t)) return true;
#pragma unroll
            for (int j = 0; j < (int)kMaxShortcutProbes; j++)
                if (j == best) fill[j] = second + 1 < B ? second + 1 : B;
        }
        return false;
    }
    int lvl = min_fill;
    while (lvl < B) {
        int next = B;
#pragma unroll
        for (int j = 0; j < (int)kMaxShortcutProbes; j++) {
            if (j < np) {
                if (fill[j] == lvl) {
                    Row16<P, V> row;
                    row.load(slots + bidx[j] * L.B, L.B, t);
                    if (row.claim(slots + bidx[j] * L.B, L.B, word, t)) return true;
                }
                if (fill[j] > lvl && fill[j] < next) next = fill[j];
            }
        }
        lvl = next;
    }
    return false;
}

template <int P, int V>
__device__ bool kv16_pow_query(const LevelDev &L, uint64_t key, const Tile<P> &t, uint64_t &val, bool &owner) {
    const uint32_t sk = fp16_of(key);
    Probe pr;
    pr.begin(L, key, sk);
    const uint32_t *slots = (const uint32_t *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        Row16<P, V> row;
        row.load(slots + fastmod(p, L.nb) * L.B, L.B, t);
        if (row.find(sk, t, val, owner)) return true;
    }
}

template <int P, int V>
__device__ bool kv16_pow_remove(const LevelDev &L, uint64_t key, const Tile<P> &t) {
    const uint32_t sk = fp16_of(key);
    Probe pr;
    pr.begin(L, key, sk);
    uint32_t *slots = (uint32_t *)L.slots;
    for (uint32_t j = 0;; j++) {
        const uint64_t p = pr.at(j);
        if (p == kEnd) return false;
        uint32_t *bk = slots + fastmod(p, L.nb) * L.B;
        Row16<P, V> row;
        row.load(bk, L.B, t);
        if (row.bury(bk, sk, t)) return true;
    }
}

// Words per lane of the row shape of a level (0: no row shape).  Rows of
// 8 words are offered to tiles of one or two lanes only: in a wider tile's
// kernel their registers (up to four probe rows) would lower the occupancy
// of the 4-word path, measured 12 % slower at tile 4.
template <int P>
__device__ inline int row16_words(const LevelDev &L) {
    if ((L.B & 3u) == 0 && L.B <= 4u * P) return 4;
    if (P <= 2 && (L.B & 7u) == 0 && L.B <= 8u * P) return 8;
    return 0;
}

__device__ inline bool pow_scheme(const LevelDev &L) {
    return L.scheme == PG_SCHEME_POWER_OF_N || L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS ||
           L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
}

// bucket_insert::insert (bucket_insert.cuh:127-203) / single_slot_insert::
// insert (single_slot_insert.cuh:125-166): lowest empty slot in probe order;
// key_val_pair::atomic_swap = CAS key 0 -> key, then store the value
// (key_val_pair.cuh:84-95).
template <int P, class SL = SlotKV, bool RP = false>
__device__ bool kv_level_insert(const LevelDev &L, uint64_t key, uint64_t val, const Tile<P> &t) {
    using S = typename SL::S;
    if (pow_scheme(L)) {
        if constexpr (RP)
            if (L.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE) return kv_pow_insert<P, SL, true>(L, key, val, t);
        if constexpr (std::is_same<SL, SlotKV16>::value) {
            const int rw = row16_words<P>(L);
            if (rw == 4) return kv16_pow_insert<P, 4>(L, key, val, t);
            if constexpr (P <= 2)
                if (rw == 8) return kv16_pow_insert<P, 8>(L, key, val, t);
        }
        return kv_pow_insert<P, SL>(L, key, val, t);
    }
    if constexpr (P <= 16) {
        if (L.B == 4 * P) return kv_reg_insert<P, 4, SL>(L, key, val, t);
        if (L.B == 2 * P) return kv_reg_insert<P, 2, SL>(L, key, val, t);
    }
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            uint64_t si = 0;
            S s = SL::none();
            if (!end) {
                si = fastmod(p, L.nb);
                s = SL::load(slots + si);
            }
            const uint64_t endm = t.bits(end);
            bool cand = !end && SL::key(s) == kKvEmpty;
            uint64_t cm = t.bits(cand) & below_first(endm);
            t.acc[0] += cm ? (uint32_t)ctz64(cm) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            while (cm) {
                const uint32_t l = (uint32_t)ctz64(cm);
                bool success = false;
                if (t.rank == l) success = SL::claim(slots + si, s, sk, val);
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
            const S s = ok ? SL::load(slots + base + i) : SL::none();
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            uint64_t cm = t.bits(ok && SL::key(s) == kKvEmpty);
            while (cm) {
                const uint32_t l = (uint32_t)ctz64(cm);
                bool success = false;
                if (t.rank == l) success = SL::claim(slots + base + i, s, sk, val);
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
template <int P, class SL = SlotKV>
__device__ bool kv_level_query(const LevelDev &L, uint64_t key, const Tile<P> &t, uint64_t &val,
                               bool &owner) {
    using S = typename SL::S;
    if (pow_scheme(L)) {
        if constexpr (std::is_same<SL, SlotKV16>::value) {
            const int rw = row16_words<P>(L);
            if (rw == 4) return kv16_pow_query<P, 4>(L, key, t, val, owner);
            if constexpr (P <= 2)
                if (rw == 8) return kv16_pow_query<P, 8>(L, key, t, val, owner);
        }
        return kv_pow_query<P, SL>(L, key, t, val, owner);
    }
    if constexpr (P <= 16) {
        if (L.B == 4 * P) return kv_reg_query<P, 4, SL>(L, key, t, val, owner);
        if (L.B == 2 * P) return kv_reg_query<P, 2, SL>(L, key, t, val, owner);
    }
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    const S *slots = (const S *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            S s = SL::none();
            if (!end) s = SL::load(slots + fastmod(p, L.nb));
            const uint64_t endm = t.bits(end);
            const uint64_t lim = below_first(endm);
            const uint64_t mm = t.bits(!end && SL::key(s) == sk) & lim;
            const uint64_t em = t.bits(!end && SL::key(s) == kKvEmpty) & lim;
            const uint64_t dec = mm | em;
            t.acc[0] += dec ? (uint32_t)ctz64(dec) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            if (dec) {
                const uint32_t l = (uint32_t)ctz64(dec);
                if ((mm >> l) & 1ull) {
                    owner = (t.rank == l);
                    if (owner) val = SL::val(s);
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
            const S s = ok ? SL::load(slots + base + i) : SL::none();
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            const uint64_t mm = t.bits(ok && SL::key(s) == sk);
            if (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                owner = (t.rank == l);
                if (owner) val = SL::val(s);
                return true;
            }
            if (t.any(ok && SL::key(s) == kKvEmpty)) return false;
        }
    }
}

// bucket_insert::remove (bucket_insert.cuh:205-293): CAS the lowest match to
// the tombstone key (key_val_pair.cuh:143-151); a stripe with an empty slot
// and no successful removal ends the search.
template <int P, class SL = SlotKV>
__device__ bool kv_level_remove(const LevelDev &L, uint64_t key, const Tile<P> &t) {
    using S = typename SL::S;
    if (pow_scheme(L)) {
        if constexpr (std::is_same<SL, SlotKV16>::value) {
            const int rw = row16_words<P>(L);
            if (rw == 4) return kv16_pow_remove<P, 4>(L, key, t);
            if constexpr (P <= 2)
                if (rw == 8) return kv16_pow_remove<P, 8>(L, key, t);
        }
        return kv_pow_remove<P, SL>(L, key, t);
    }
    if constexpr (P <= 16) {
        if (L.B == 4 * P) return kv_reg_remove<P, 4, SL>(L, key, t);
        if (L.B == 2 * P) return kv_reg_remove<P, 2, SL>(L, key, t);
    }
    Probe pr;
    pr.begin(L, key, SL::tag(key));
    pr.max_probes = pr.useful_probes(L.nb.d);
    const uint64_t sk = SL::stored(key);
    S *slots = (S *)L.slots;
    if (P > 1 && L.B == 1) {
        for (uint32_t j0 = 0;; j0 += P) {
            const uint64_t p = pr.at(j0 + t.rank);
            const bool end = (p == kEnd);
            uint64_t si = 0;
            S s = SL::none();
            if (!end) {
                si = fastmod(p, L.nb);
                s = SL::load(slots + si);
            }
            const uint64_t endm = t.bits(end);
            const uint64_t lim = below_first(endm);
            uint64_t mm = t.bits(!end && SL::key(s) == sk) & lim;
            const uint64_t em = t.bits(!end && SL::key(s) == kKvEmpty) & lim;
            // Walk decisive lanes in probe order: matches are tried, the first
            // empty ends the search.
            uint64_t dec = mm | em;
            t.acc[0] += dec ? (uint32_t)ctz64(dec) + 1 : (endm ? (uint32_t)ctz64(endm) : P);
            while (dec) {
                const uint32_t l = (uint32_t)ctz64(dec);
                if (!((mm >> l) & 1ull)) return false;
                bool success = false;
                if (t.rank == l) success = SL::bury(slots + si, s, sk);
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
            const S s = ok ? SL::load(slots + base + i) : SL::none();
            const bool match = ok && SL::key(s) == sk;
            t.acc[0] += (L.B - s0 < P) ? L.B - s0 : P;
            uint64_t mm = t.bits(match);
            while (mm) {
                const uint32_t l = (uint32_t)ctz64(mm);
                bool success = false;
                if (t.rank == l) success = SL::bury(slots + base + i, s, sk);
                if (t.any(success)) {
                    t.acc[2]++;
                    return true;
                }
                mm &= mm - 1ull;
            }
            if (t.any(ok && !match && SL::key(s) == kKvEmpty)) return false;
        }
    }
}

// Whole-table ops: level 0, then the backing level on false
// (tables/base_table.cuh:167-308).
// RP: the table has a level of the replace scheme (a separate kernel, so the
// other tables' kernels carry none of that scheme's code).
template <int P, class SL = SlotKV, bool RP = false>
__device__ int kv_insert_key(const TableDev &T, uint64_t key, uint64_t val, const Tile<P> &t) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_insert<P, SL, RP>(T.lv[l], key, val, t)) return 1 + (int)l;
    return 0;
}
template <int P, class SL = SlotKV>
__device__ bool kv_query_key(const TableDev &T, uint64_t key, const Tile<P> &t, uint64_t &val, bool &owner) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_query<P, SL>(T.lv[l], key, t, val, owner)) return true;
    return false;
}
template <int P, class SL = SlotKV>
__device__ bool kv_remove_key(const TableDev &T, uint64_t key, const Tile<P> &t) {
    for (uint32_t l = 0; l < T.nlevels; l++)
        if (kv_level_remove<P, SL>(T.lv[l], key, t)) return true;
    return false;
}

}  // namespace pg
