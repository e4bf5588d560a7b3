// pg_internal.hpp -- structures and helpers shared by the HIP kernels and the
// host side of the C ABI.  Host + device.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/poggers_amd.h"

namespace pg {

constexpr uint64_t kEnd = ~uint64_t(0);  // probing end() sentinel (probing_schemes/*.cuh end())
constexpr uint32_t kMaxShortcutProbes = 4;

// Range reduction n % d with an invariant divisor (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1):
//   l = ceil(log2 d), m = floor(2^64 (2^l - d) / d) + 1,
//   t = mulhi(m, n), q = (t + ((n - t) >> 1)) >> (l - 1),  r = n - q d.
// Powers of two use a mask.  Exact for every 64-bit n and d >= 1; the
// reference computes "probe % num_buckets" (power_of_n_shortcut_buckets.cuh:216).
struct FastMod {
    uint64_t d;
    uint64_t magic;
    uint64_t mask;   // d - 1 when d is a power of two
    uint32_t sh2;
    uint32_t pow2;
};

__host__ __device__ inline uint64_t mulhi64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

inline FastMod make_fastmod(uint64_t d) {
    FastMod f{};
    f.d = d;
    if ((d & (d - 1)) == 0) {
        f.pow2 = 1;
        f.mask = d - 1;
        return f;
    }
    uint32_t l = 64 - (uint32_t)__builtin_clzll(d - 1);  // ceil(log2 d), d >= 3
    unsigned __int128 two_l = (unsigned __int128)1 << l;
    unsigned __int128 num = (two_l - d) << 64;
    f.magic = (uint64_t)(num / d) + 1;
    f.sh2 = l - 1;
    f.pow2 = 0;
    return f;
}

__host__ __device__ inline uint64_t fastmod(uint64_t n, const FastMod &f) {
    if (f.pow2) return n & f.mask;
    uint64_t t = mulhi64(f.magic, n);
    uint64_t q = (t + ((n - t) >> 1)) >> f.sh2;
    return n - q * f.d;
}

// ---- hashers (hash_schemes/murmurhash.cuh:20-61 and MurmurHash3 fmix64) ----
__host__ __device__ inline uint64_t fmix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

// MurmurHash64A specialised to one 8-byte block (len = 8, no tail).
__host__ __device__ inline uint64_t murmur64a_u64(uint64_t key, uint64_t seed) {
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    uint64_t h = seed ^ (8ULL * m);
    uint64_t k = key;
    k *= m;
    k ^= k >> 47;
    k *= m;
    h ^= k;
    h *= m;
    h ^= h >> 47;
    h *= m;
    h ^= h >> 47;
    return h;
}

__host__ __device__ inline uint64_t hash64(uint32_t hasher, uint64_t key, uint64_t seed) {
    return hasher == PG_HASH_MURMUR64A ? murmur64a_u64(key, seed) : fmix64(key ^ seed);
}

// splitmix64 key stream (the seeded generator of every benchmark input).
__host__ __device__ inline uint64_t key_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    if (z == 0) z = 0x8000000000000000ULL;
    if (z == ~0ULL) z = 0x7FFFFFFFFFFFFFFFULL;
    return z;
}

// One level as the kernels see it (passed by value in the kernarg segment:
// no device-side struct pointer chase, unlike the reference's
// table -> scheme -> slots chain, tables/bucketed_table.cuh:93-95).
struct LevelDev {
    void *slots;
    FastMod nb;          // number of buckets and its divider
    uint64_t seed;       // H seed
    uint64_t seed2;      // seed^2, second hash of DOUBLE / P2
    uint32_t B;          // slots per bucket
    uint32_t scheme;
    uint32_t prober;
    uint32_t hasher;
    uint32_t max_probes;
    uint32_t group;      // GROUPED_SHORTCUT: buckets per 128-byte line, else 0
    uint32_t bucket_bytes;
    uint32_t pad;
    // Occupancy index of a backing level (has_occ_index), else null: bit b
    // of occ[b / 32] is set before any claim in bucket b and cleared only
    // with the slots, so a clear bit proves the bucket empty.
    uint32_t *occ = nullptr;
};

// A linear-bucket backing level (level >= 1, which only keys that spill
// reach) of 16/16 or 8/8 filter slots keeps an occupancy index.  Its query
// and remove stop at the first probe bucket whose fill is not B
// (linear_insert_buckets.cuh:323-398); an empty bucket has fill 0 and, since
// these fingerprints are never 0, no match, so a clear bit answers for the
// bucket without reading it.  A 10/6 fingerprint can match an empty word, so
// 10/6 levels read every bucket.  A primary level keeps none: every insert
// there would pay the extra atomic.
__host__ __device__ inline bool has_occ_index(uint32_t rep, uint32_t scheme, uint32_t level) {
    return level > 0 && scheme == PG_SCHEME_LINEAR_BUCKETS && (rep == PG_REP_FP16_V16 || rep == PG_REP_FP8_V8);
}
__host__ __device__ inline uint64_t occ_bytes(uint64_t nb) { return ((nb + 31) / 32 * 4 + 255) & ~uint64_t(255); }

// Bytes of one bucket of B slots; the 12-bit bucket packs four slots in
// three uint16 words: (12B - 1)/16 + 1 words (12_bit_bucket.cuh:46).
__host__ __device__ inline uint64_t bucket_bytes(uint32_t rep, uint32_t B) {
    if (rep == PG_REP_FP12) return 2ull * ((12ull * B - 1) / 16 + 1);
    if (rep == PG_REP_KV64) return 16ull * B;
    if (rep == PG_REP_KEY64) return 8ull * B;
    return (uint64_t)B * (rep == PG_REP_FP16_V16 || rep == PG_REP_KV16 ? 4u : 2u);
}
// Key-value slot tables (the bucket_insert / single_slot_insert / flat
// power-of-n family); the others are the packed filter slots.
__host__ __device__ inline bool is_kv_rep(uint32_t rep) {
    return rep == PG_REP_KV64 || rep == PG_REP_KEY64 || rep == PG_REP_KV16;
}

// Start of bucket b: plain levels are arrays of buckets; a grouped level puts
// `group` buckets in each 128-byte line and pads the rest
// (grouped_power_buckets.cuh:27-37 bucket_wrapper, :150-160 b / K, b % K).
__host__ __device__ inline void *bucket_ptr(const LevelDev &L, uint64_t b) {
    uint64_t off;
    if (L.group) {
        const uint64_t line = b / L.group;
        off = line * 128 + (b - line * L.group) * L.bucket_bytes;
    } else {
        off = b * L.bucket_bytes;
    }
    return (char *)L.slots + off;
}

struct TableDev {
    LevelDev lv[PG_MAX_LEVELS];
    uint32_t nlevels;
    uint32_t rep;
};

// Raw probe sequence of a key (probing_schemes/*.cuh begin()/next()).
// Probe i is returned by at(i); a value of kEnd ends the sequence, exactly
// like the reference's `p != end()` loop condition.
struct Probe {
    uint64_t h1, h2, tag;
    uint32_t prober, max_probes;

    __host__ __device__ inline void begin(const LevelDev &L, uint64_t key, uint64_t tagv) {
        prober = L.prober;
        max_probes = L.max_probes;
        tag = tagv;
        h1 = hash64(L.hasher, key, L.seed);
        h2 = (prober == PG_PROBE_DOUBLE || prober == PG_PROBE_P2) ? hash64(L.hasher, key, L.seed2) : 0;
    }
    // Probes worth visiting at a level of nb buckets.  Linear probes h + i
    // revisit the same buckets with period nb unless h + i wraps past 2^64
    // (which also covers the ~0 end marker), so when no wrap can occur
    // within max_probes, probes beyond the first nb repeat buckets already
    // seen and cannot change any result: the walk stops after nb probes.
    __host__ __device__ inline uint32_t useful_probes(uint64_t nb) const {
        if (prober == PG_PROBE_LINEAR && h1 <= kEnd - (uint64_t)max_probes && nb < max_probes)
            return (uint32_t)nb;
        return max_probes;
    }
    __host__ __device__ inline uint64_t at(uint32_t i) const {
        if (i >= max_probes) return kEnd;
        if (i == 0) return prober == PG_PROBE_SPLIT64 ? (h1 & 0xffffffffULL) : h1;
        switch (prober) {
            case PG_PROBE_LINEAR: return h1 + (uint64_t)i;
            case PG_PROBE_DOUBLE: return h1 + (uint64_t)i * h2;
            case PG_PROBE_P2: return h2;
            case PG_PROBE_XOR_P2: return h1 ^ (tag * 0x5bd1e995ULL);
            default: return h1 >> 32;  // SPLIT64
        }
    }
};

// Fingerprint of the FP16_V16 slot: low 16 bits of the raw key, 0 -> 1,
// 0xFFFF -> 0xFFFE (representations/dynamic_container.cuh:35-50).
__host__ __device__ inline uint32_t fp16_of(uint64_t key) {
    uint32_t f = (uint32_t)(key & 0xFFFFu);
    f += (f == 0u);
    f -= (f == 0xFFFFu);
    return f;
}
constexpr uint32_t kTomb16 = 0x0000FFFFu;  // join(tombstone fp 0xFFFF, val 0)

// Packed filter slot formats: one W-bit word = (val & vmask) << kbits |
// (fp & kmask) (grouped_storage_sub_bits.cuh:77-95 join, :35-74 retrieve).
// The fingerprint is the SmallKey-wide remap of the raw key's low bits
// (dynamic_container.cuh:35-50): 16 bits for the 16/16 and 10/6 slots, 8 bits
// for 8/8.  The probe tag is that remapped SmallKey; only the stored and
// compared key bits are masked to kbits (grouped_storage_sub_bits.cuh:85-87,
// :186-195), so a 10/6 key whose masked bits are 0 or 0x3FF is stored as an
// empty- or tombstone-looking word, exactly as the reference stores it.
// Empty word 0; tombstone word = kmask (join(get_empty()-1, 0)).
struct Fmt16 {  // bit_grouped_container<16,16>, SmallKey uint16, 32-bit word
    using W = uint32_t;
    using V = uint16_t;  // value type of the bulk value buffers
    static constexpr uint32_t small_bits = 16, kbits = 16, kmask = 0xFFFFu, vmask = 0xFFFFu;
};
struct Fmt8 {  // bit_grouped_container<8,8>, SmallKey uint8, 16-bit word
    using W = uint16_t;
    using V = uint8_t;
    static constexpr uint32_t small_bits = 8, kbits = 8, kmask = 0xFFu, vmask = 0xFFu;
};
struct Fmt10 {  // bit_grouped_container<10,6>, SmallKey uint16, 16-bit word
    using W = uint16_t;
    using V = uint8_t;
    static constexpr uint32_t small_bits = 16, kbits = 10, kmask = 0x3FFu, vmask = 0x3Fu;
};
// The 12-bit sub-word bucket (12_bit_bucket.cuh:33-342): key-only, no
// tombstones; the "fingerprint" is the key's low 16 bits with no remap, and
// slot 4g+o keeps and remaps its own 12 of them (twelve_proj).
struct Fmt12 {
    using W = uint16_t;
    using V = uint16_t;
    static constexpr uint32_t small_bits = 16, kbits = 16, kmask = 0xFFFFu, vmask = 0u;
};
template <class F> inline constexpr bool is_twelve = false;
template <> inline constexpr bool is_twelve<Fmt12> = true;

template <class F>
__host__ __device__ inline uint32_t fp_of(uint64_t key) {
    if constexpr (is_twelve<F>) {
        return (uint32_t)key & 0xFFFFu;
    } else {
        constexpr uint32_t all = (1u << F::small_bits) - 1u;
        uint32_t f = (uint32_t)key & all;
        f += (f == 0u);
        f -= (f == all);
        return f;
    }
}
template <class F>
__host__ __device__ inline uint32_t pack_word(uint32_t fp, uint32_t val) {
    if constexpr (is_twelve<F>) return fp;  // the item; each slot stores its projection
    else return ((val & F::vmask) << F::kbits) | (fp & F::kmask);
}

// 12-bit slot o's stored bits of item, remapped so a filled slot never reads
// as empty (representation_helpers.cuh:684-777 match_*, :779-903 replace_*):
//   o=0 item & 0x0FFF (0 -> 1)        o=1 item & 0xF0FF (| 0x1000 if bits 12-15 are 0)
//   o=2 item & 0xFF0F (| 0x0100 if bits 8-15 are 0)   o=3 item & 0xFFF0 (0 -> 0x0010)
__host__ __device__ inline uint32_t twelve_proj(uint32_t item, uint32_t o) {
    item &= 0xFFFFu;
    switch (o & 3u) {
        case 0: { const uint32_t m = item & 0x0FFFu; return m ? m : 1u; }
        case 1: { const uint32_t m = item & 0xF0FFu; return (item & 0xF000u) ? m : (m | 0x1000u); }
        case 2: { const uint32_t m = item & 0xFF0Fu; return (item & 0xFF00u) ? m : (m | 0x0100u); }
        default: { const uint32_t m = item & 0xFFF0u; return m ? m : 0x0010u; }
    }
}

// ---- launchers (defined in the .hip files) ----
struct OpArgs {
    const uint64_t *keys;
    const void *vals;
    uint64_t n;
    void *vals_out;
    uint8_t *flags;        // status (insert, upsert) or found (query/remove)
    uint8_t *found;        // upsert: key was already present
    uint64_t *fail_count;
    uint64_t *stats;       // optional accounting counters (pg_set_stats)
    uint32_t with_delete;
    uint32_t deterministic;
    uint32_t items;        // keys per tile (pg_set_items), 0 = default
};

enum class Op { Insert, Query, Remove, Upsert };

// The device view of a table and its tile width (pg_capi.hip); NULL for a
// bad handle.
const TableDev *table_dev(pg_handle h, uint32_t *tile);

int launch_tcf(const TableDev &T, uint32_t tile, Op op, const OpArgs &a, hipStream_t s);
int launch_kv(const TableDev &T, uint32_t tile, Op op, const OpArgs &a, hipStream_t s);
int launch_fill(const TableDev &T, uint64_t *d_counter, hipStream_t s);
int launch_zero(void *p, uint64_t bytes, hipStream_t s);
int launch_occ_rebuild(const LevelDev &L, uint32_t rep, hipStream_t s);
int launch_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out, hipStream_t s);
int launch_random_sector(void *buf, uint64_t bytes, uint64_t n, uint32_t op, uint64_t seed,
                         uint64_t *sink, hipStream_t s);

}  // namespace pg
