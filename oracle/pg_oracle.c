/*
 * pg_oracle.c -- single-threaded CPU restatement of poggers' bulk
 * insert / query / remove path for the modular hash tables and the
 * Two-Choice Filter (TCF), plus a multi-threaded run of the same filter
 * algorithm (atomic slot claims) for the all-cores CPU baseline.
 *
 * THIS IS TEST INFRASTRUCTURE.  It is the checker that the HIP path is
 * compared against (and the CPU baseline that bench.py times).  Nothing in
 * the product library links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference's include/poggers/ directory unless marked tests/).
 *
 * Pinning (what checks each function, see tests/test_oracle_*.py):
 *   or_murmur64a  -- SMHasher verification code 0x1F0D3804 and the recorded
 *                    value MurmurHash64A(1, seed 42) = 0x63cc585a9ca03f57.
 *   or_fmix64     -- SMHasher verification code of MurmurHash3_x64_128
 *                    (0x6384BA69, built on this finalizer) and inversion by
 *                    an independently derived inverse.
 *   probes        -- recorded double-hash probes of key 1 (seed 42) and the
 *                    closed forms of each scheme.
 *   slot formats  -- exhaustive pack/unpack round trip (as the reference's
 *                    tests/src/kmer_speed_tests.cu:664-684) and the
 *                    known answers of tests/src/test_key_val_pair.cu:50-62.
 *   TCF insert    -- hand-derived placement scenarios (shortcut threshold,
 *                    fill-ordered retry, tie order, backing spill), zero
 *                    false negatives, FP-rate closed form.
 *   KV insert     -- tests/src/test_components.cu:88-127 and :345-366 known
 *                    answers, linear-probing closed-form probe lengths,
 *                    brute force on tiny tables.
 *   flat power-of-n (kv_pow_*) and the 16/16 key-value slot -- the
 *                    tests/src/test_components.cu:88-200 p2_table checks;
 *                    hand-derived slot words, kept tombstone values, fills
 *                    that count tombstones, least-filled choice and tie
 *                    order (tests/test_oracle_flat.py).
 *   replace scheme -- hand-derived scenarios (value replaced in place, in
 *                    a full bucket, a copy stored in an earlier probe
 *                    bucket with room and what a remove then uncovers, the
 *                    16/16 word and the key-only slot) and a dictionary
 *                    model of insert-or-assign (tests/test_oracle_replace.py).
 *   multi-thread  -- with one thread, byte-identical layouts to the
 *                    sequential functions; with several, zero false
 *                    negatives, fill = stored keys, queries identical to
 *                    the sequential query of the same table.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#define OR_END (~(uint64_t)0)
#define OR_MAX_LEVELS 4   /* a primary level and up to three backing levels */

/* ---- enums (values mirror include/poggers_amd.h so tests can share one
 *      descriptor layout; the oracle does not include that header) ---- */
enum { OR_REP_KV64 = 0, OR_REP_FP16_V16 = 1, OR_REP_FP8_V8 = 2, OR_REP_FP10_V6 = 3, OR_REP_FP12 = 4,
       OR_REP_KEY64 = 5, OR_REP_KV16 = 6 };
enum { OR_SCHEME_BUCKET = 0, OR_SCHEME_SINGLE_SLOT = 1,
       OR_SCHEME_POWER_OF_N_SHORTCUT = 2, OR_SCHEME_LINEAR_BUCKETS = 3,
       OR_SCHEME_GROUPED = 4, OR_SCHEME_POWER_OF_N = 5, OR_SCHEME_POWER_OF_N_SHORTCUT_SLOTS = 6,
       OR_SCHEME_POWER_OF_N_SHORTCUT_REPLACE = 7 };
enum { OR_PROBE_LINEAR = 0, OR_PROBE_DOUBLE = 1, OR_PROBE_P2 = 2,
       OR_PROBE_XOR_P2 = 3, OR_PROBE_SPLIT64 = 4 };
enum { OR_HASH_FMIX64 = 0, OR_HASH_MURMUR64A = 1 };

/* ------------------------------------------------------------------ */
/* Hashers                                                             */
/* ------------------------------------------------------------------ */

/* Austin Appleby's MurmurHash64A over len bytes.
 * Follows hash_schemes/murmurhash.cuh:20-61. */
uint64_t or_murmur64a(const void *key, int len, uint64_t seed)
{
    const uint64_t m = 0xc6a4a7935bd1e995ULL;
    const int r = 47;
    uint64_t h = seed ^ ((uint64_t)len * m);
    const unsigned char *p = (const unsigned char *)key;
    int nblocks = len / 8;
    for (int b = 0; b < nblocks; b++) {
        uint64_t k;
        memcpy(&k, p + 8 * b, 8);
        k *= m;
        k ^= k >> r;
        k *= m;
        h ^= k;
        h *= m;
    }
    const unsigned char *tail = p + 8 * nblocks;
    switch (len & 7) {
    case 7: h ^= (uint64_t)tail[6] << 48; /* fall through */
    case 6: h ^= (uint64_t)tail[5] << 40; /* fall through */
    case 5: h ^= (uint64_t)tail[4] << 32; /* fall through */
    case 4: h ^= (uint64_t)tail[3] << 24; /* fall through */
    case 3: h ^= (uint64_t)tail[2] << 16; /* fall through */
    case 2: h ^= (uint64_t)tail[1] << 8;  /* fall through */
    case 1: h ^= (uint64_t)tail[0];
            h *= m;
    }
    h ^= h >> r;
    h *= m;
    h ^= h >> r;
    return h;
}

/* MurmurHash3's 64-bit finalizer (SMHasher, MurmurHash3.cpp fmix64).
 * The reference has no counterpart (its hasher is MurmurHash64A); this is
 * the hasher the MI355X build uses by default. */
uint64_t or_fmix64(uint64_t k)
{
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ULL;
    k ^= k >> 33;
    return k;
}

/* MurmurHash3_x64_128 (SMHasher), only used to pin or_fmix64 through the
 * published verification code of that function. */
static uint64_t rotl64(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }

void or_murmur3_x64_128(const void *key, int len, uint32_t seed, void *out)
{
    const unsigned char *data = (const unsigned char *)key;
    const int nblocks = len / 16;
    uint64_t h1 = seed, h2 = seed;
    const uint64_t c1 = 0x87c37b91114253d5ULL;
    const uint64_t c2 = 0x4cf5ad432745937fULL;
    for (int i = 0; i < nblocks; i++) {
        uint64_t k1, k2;
        memcpy(&k1, data + 16 * i, 8);
        memcpy(&k2, data + 16 * i + 8, 8);
        k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
        h1 = rotl64(h1, 27); h1 += h2; h1 = h1 * 5 + 0x52dce729;
        k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
        h2 = rotl64(h2, 31); h2 += h1; h2 = h2 * 5 + 0x38495ab5;
    }
    const unsigned char *tail = data + nblocks * 16;
    uint64_t k1 = 0, k2 = 0;
    switch (len & 15) {
    case 15: k2 ^= (uint64_t)tail[14] << 48; /* fall through */
    case 14: k2 ^= (uint64_t)tail[13] << 40; /* fall through */
    case 13: k2 ^= (uint64_t)tail[12] << 32; /* fall through */
    case 12: k2 ^= (uint64_t)tail[11] << 24; /* fall through */
    case 11: k2 ^= (uint64_t)tail[10] << 16; /* fall through */
    case 10: k2 ^= (uint64_t)tail[9] << 8;   /* fall through */
    case 9:  k2 ^= (uint64_t)tail[8];
             k2 *= c2; k2 = rotl64(k2, 33); k2 *= c1; h2 ^= k2;
             /* fall through */
    case 8:  k1 ^= (uint64_t)tail[7] << 56; /* fall through */
    case 7:  k1 ^= (uint64_t)tail[6] << 48; /* fall through */
    case 6:  k1 ^= (uint64_t)tail[5] << 40; /* fall through */
    case 5:  k1 ^= (uint64_t)tail[4] << 32; /* fall through */
    case 4:  k1 ^= (uint64_t)tail[3] << 24; /* fall through */
    case 3:  k1 ^= (uint64_t)tail[2] << 16; /* fall through */
    case 2:  k1 ^= (uint64_t)tail[1] << 8;  /* fall through */
    case 1:  k1 ^= (uint64_t)tail[0];
             k1 *= c1; k1 = rotl64(k1, 31); k1 *= c2; h1 ^= k1;
    }
    h1 ^= (uint64_t)len; h2 ^= (uint64_t)len;
    h1 += h2; h2 += h1;
    h1 = or_fmix64(h1); h2 = or_fmix64(h2);
    h1 += h2; h2 += h1;
    memcpy(out, &h1, 8);
    memcpy((unsigned char *)out + 8, &h2, 8);
}

/* Seeded 64-bit key hash used by every probing scheme.
 *   MURMUR64A: murmurHasher<uint64_t>::hash, hash_schemes/murmurhash.cuh:93-100
 *              (MurmurHash64A over the 8 key bytes with the stored seed).
 *   FMIX64:    fmix64(key XOR seed). */
uint64_t or_hash(uint32_t hasher, uint64_t key, uint64_t seed)
{
    if (hasher == OR_HASH_MURMUR64A)
        return or_murmur64a(&key, 8, seed);
    return or_fmix64(key ^ seed);
}

/* ------------------------------------------------------------------ */
/* Probing schemes: produce the raw 64-bit probe values (before the    */
/* "% num_buckets" the consumer applies).  The reference iterates     */
/*   for (p = begin(key); p != end(); p = next(tag))                   */
/* with end() == ~0, so a probe value equal to ~0 ends the sequence.   */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t prober, max_probes;
    uint32_t i;
    uint64_t h1, h2, tag;
} or_probe_it;

/* begin(): probing_schemes/double_hashing.cuh:72-81 (h1 = H_seed, h2 =
 * H_seed^2), xor_power_of_two.cuh:69-76, linear_probing.cuh:64-70,
 * power_of_two.cuh:74-83.  SPLIT64 has no reference counterpart: one
 * 64-bit hash split into its low and high 32-bit halves. */
static uint64_t probe_begin(or_probe_it *it, uint32_t prober, uint32_t hasher,
                            uint32_t max_probes, uint64_t seed, uint64_t key,
                            uint64_t tag)
{
    it->prober = prober;
    it->max_probes = max_probes;
    it->i = 0;
    it->tag = tag;
    it->h1 = or_hash(hasher, key, seed);
    it->h2 = 0;
    if (prober == OR_PROBE_DOUBLE || prober == OR_PROBE_P2)
        it->h2 = or_hash(hasher, key, seed * seed);
    if (max_probes == 0)
        return OR_END;
    if (prober == OR_PROBE_SPLIT64)
        return it->h1 & 0xffffffffULL;
    return it->h1;
}

/* next(): double_hashing.cuh:83-89 (h1 + i*h2), xor_power_of_two.cuh:78-84
 * (h1 ^ (tag * 0x5bd1e995)), linear_probing.cuh:72-78 (h + i),
 * power_of_two.cuh:85-91 (h2). */
static uint64_t probe_next(or_probe_it *it)
{
    it->i += 1;
    if (it->i >= it->max_probes)
        return OR_END;
    switch (it->prober) {
    case OR_PROBE_LINEAR: return it->h1 + (uint64_t)it->i;
    case OR_PROBE_DOUBLE: return it->h1 + (uint64_t)it->i * it->h2;
    case OR_PROBE_P2:     return it->h2;
    case OR_PROBE_XOR_P2: return it->h1 ^ (it->tag * 0x5bd1e995ULL);
    case OR_PROBE_SPLIT64: return it->h1 >> 32;
    }
    return OR_END;
}

/* Test hook: write up to cap raw probe values, return their count. */
int or_probes(uint32_t prober, uint32_t hasher, uint32_t max_probes,
              uint64_t seed, uint64_t key, uint64_t tag, uint64_t *out, int cap)
{
    or_probe_it it;
    int n = 0;
    for (uint64_t p = probe_begin(&it, prober, hasher, max_probes, seed, key, tag);
         p != OR_END && n < cap; p = probe_next(&it))
        out[n++] = p;
    return n;
}

/* ------------------------------------------------------------------ */
/* Slot encodings                                                      */
/* ------------------------------------------------------------------ */

/* Fingerprint: the low bits of the RAW key, 0 -> 1, all-ones -> all-ones-1,
 * at the width of the SmallKey type (representations/dynamic_container.cuh:
 * 35-50): uint8 for the 8/8 slot, uint16 for 16/16 and for 10/6.  The 10/6
 * slot keeps only the low 10 bits of that (grouped_storage_sub_bits.cuh:
 * 85-87), so the remap happens before the masking; the probe tag is the
 * unmasked SmallKey. */
static uint32_t fp_of(uint32_t rep, uint64_t key)
{
    /* twelve_bucket takes (uint16_t)key with no remap (12_bit_bucket.cuh:
     * 104-122, representation_helpers.cuh:930-962); each slot remaps its
     * own 12 bits (or_fp12_proj). */
    if (rep == OR_REP_FP12) return (uint16_t)key;
    if (rep == OR_REP_FP8_V8) {
        uint32_t f = (uint8_t)key;
        if (f == 0) f += 1;
        if (f == 0xFF) f -= 1;
        return f;
    }
    uint32_t f = (uint16_t)key;
    if (f == 0) f += 1;
    if (f == 0xFFFF) f -= 1;
    return f;
}

/* grouped_bits_pair<K,V,kb,vb>: word = (val & vmask) << kb | (fp & kmask).
 * representations/grouped_storage_sub_bits.cuh:77-95 (join_in_storage),
 * :35-74 (retrieve_key/val_from_storage). */
uint32_t or_pack(uint32_t kbits, uint32_t vbits, uint32_t fp, uint32_t val)
{
    uint32_t vm = (vbits >= 32) ? 0xffffffffu : ((1u << vbits) - 1u);
    uint32_t km = (kbits >= 32) ? 0xffffffffu : ((1u << kbits) - 1u);
    return ((val & vm) << kbits) | (fp & km);
}
uint32_t or_unpack_fp(uint32_t kbits, uint32_t w) { return w & ((1u << kbits) - 1u); }
uint32_t or_unpack_val(uint32_t kbits, uint32_t vbits, uint32_t w)
{
    return (w >> kbits) & ((1u << vbits) - 1u);
}

/* ------------------------------------------------------------------ */
/* Table state                                                         */
/* ------------------------------------------------------------------ */
typedef struct {
    uint64_t nslots;
    uint32_t scheme, bucket_size, tile, prober, hasher, max_probes;
    uint32_t group;      /* GROUPED: buckets per 128-byte line; else 0 */
    uint32_t reserved;
} or_level_desc;

typedef struct {
    or_level_desc d;
    uint64_t seed;
    uint64_t nb;        /* buckets (for SINGLE_SLOT: slots) */
    uint32_t B;         /* slots per bucket */
    void *slots;        /* uint32_t[] (FP16_V16), uint16_t[] (FP8_V8, FP10_V6, FP12), uint64_t[2*] (KV64) */
    uint64_t nslot_alloc;
    uint64_t bucket_bytes;
    uint64_t bytes;     /* size of the slot array */
} or_level;

typedef struct {
    uint32_t rep;
    uint32_t nlevels;
    or_level lv[OR_MAX_LEVELS];
    /* algorithmic accounting, reset by or_stats_reset */
    uint64_t sectors_read;     /* distinct 64 B sectors examined */
    uint64_t sectors_written;  /* 64 B sectors dirtied */
    uint64_t buckets_read[2];  /* bucket examinations: [0] level 0, [1] all backing levels */
} or_table;

static inline int is_kv_rep(uint32_t rep)
{
    return rep == OR_REP_KV64 || rep == OR_REP_KEY64 || rep == OR_REP_KV16;
}
static inline int narrow(const or_table *t)
{
    return t->rep == OR_REP_FP8_V8 || t->rep == OR_REP_FP10_V6 || t->rep == OR_REP_FP12;
}

/* Bucket b of a level.  Plain levels are arrays of buckets; GROUPED levels
 * put `group` buckets in each 128-byte line and pad the rest
 * (grouped_power_buckets.cuh:27-37 bucket_wrapper, :150-160 slot b / K, b % K). */
static inline uint8_t *bucket_at(const or_level *L, uint64_t b)
{
    uint64_t off = L->d.group ? (b / L->d.group) * 128 + (b % L->d.group) * L->bucket_bytes
                              : b * L->bucket_bytes;
    return (uint8_t *)L->slots + off;
}

/* The 12-bit sub-word bucket (12_bit_bucket.cuh:33-313): four slots in three
 * uint16 words; slot 4g+o keeps these bits of the 16-bit item, and remaps a
 * zero first part so a filled slot never reads as empty
 * (representation_helpers.cuh:600-777 read / match, :779-903 replace):
 *   o=0  w0[0:12]              m = item & 0x0FFF,  0 -> 1
 *   o=1  w0[12:16] + w1[0:8]   m = item & 0xF0FF,  (item & 0xF000) == 0 -> |= 0x1000
 *   o=2  w1[8:16] + w2[0:4]    m = item & 0xFF0F,  (item & 0xFF00) == 0 -> |= 0x0100
 *   o=3  w2[4:16]              m = item & 0xFFF0,  0 -> 0x0010              */
uint32_t or_fp12_proj(uint32_t item, uint32_t o)
{
    item &= 0xFFFFu;
    switch (o & 3u) {
    case 0: { uint32_t m = item & 0x0FFFu; return m ? m : 1u; }
    case 1: { uint32_t m = item & 0xF0FFu; return (item & 0xF000u) ? m : (m | 0x1000u); }
    case 2: { uint32_t m = item & 0xFF0Fu; return (item & 0xFF00u) ? m : (m | 0x0100u); }
    default: { uint32_t m = item & 0xFFF0u; return m ? m : 0x0010u; }
    }
}
static uint32_t fp12_get(const uint16_t *w, uint32_t i)
{
    w += 3 * (i / 4);
    switch (i & 3u) {
    case 0: return w[0] & 0x0FFFu;
    case 1: return (w[0] & 0xF000u) | (w[1] & 0x00FFu);
    case 2: return (w[1] & 0xFF00u) | (w[2] & 0x000Fu);
    default: return w[2] & 0xFFF0u;
    }
}
static void fp12_put(uint16_t *w, uint32_t i, uint32_t m)
{
    w += 3 * (i / 4);
    switch (i & 3u) {
    case 0: w[0] = (uint16_t)((w[0] & 0xF000u) | (m & 0x0FFFu)); break;
    case 1: w[0] = (uint16_t)((w[0] & 0x0FFFu) | (m & 0xF000u));
            w[1] = (uint16_t)((w[1] & 0xFF00u) | (m & 0x00FFu)); break;
    case 2: w[1] = (uint16_t)((w[1] & 0x00FFu) | (m & 0xFF00u));
            w[2] = (uint16_t)((w[2] & 0xFFF0u) | (m & 0x000Fu)); break;
    default: w[2] = (uint16_t)((w[2] & 0x000Fu) | (m & 0xFFF0u)); break;
    }
}

/* Slot i of bucket b: the stored word (packed slots) or the slot's 12-bit
 * composite (FP12). */
static inline uint32_t sget(const or_table *t, const or_level *L, uint64_t b, uint32_t i)
{
    const uint8_t *p = bucket_at(L, b);
    if (t->rep == OR_REP_FP12) return fp12_get((const uint16_t *)p, i);
    if (narrow(t)) return ((const uint16_t *)p)[i];
    return ((const uint32_t *)p)[i];
}
/* Store: a packed word, or for FP12 the item's projection for slot i. */
static inline void sset(const or_table *t, or_level *L, uint64_t b, uint32_t i, uint32_t w)
{
    uint8_t *p = bucket_at(L, b);
    if (t->rep == OR_REP_FP12) fp12_put((uint16_t *)p, i, or_fp12_proj(w, i & 3u));
    else if (narrow(t)) ((uint16_t *)p)[i] = (uint16_t)w;
    else ((uint32_t *)p)[i] = w;
}
static inline uint32_t kbits_of(const or_table *t)
{
    return t->rep == OR_REP_FP8_V8 ? 8 : (t->rep == OR_REP_FP10_V6 ? 10 : 16);
}
static inline uint32_t vbits_of(const or_table *t)
{
    return t->rep == OR_REP_FP8_V8 ? 8 : (t->rep == OR_REP_FP10_V6 ? 6 : 16);
}
static inline uint32_t kmask_of(const or_table *t) { return (1u << kbits_of(t)) - 1u; }
static inline uint32_t slot_bytes(const or_table *t)
{
    if (t->rep == OR_REP_KEY64) return 8u;
    if (t->rep == OR_REP_KV16) return 4u;
    return t->rep == OR_REP_KV64 ? 16u : (narrow(t) ? 2u : 4u);
}
/* Bytes of one bucket: B slots, or (12*B - 1)/16 + 1 uint16 words for the
 * 12-bit bucket (12_bit_bucket.cuh:46). */
static inline uint64_t bucket_bytes_of(uint32_t rep, uint32_t B)
{
    if (rep == OR_REP_FP12) return 2ull * ((12ull * B - 1) / 16 + 1);
    if (rep == OR_REP_KV64) return 16ull * B;
    if (rep == OR_REP_KEY64) return 8ull * B;
    if (rep == OR_REP_KV16) return 4ull * B;
    return (uint64_t)B * ((rep == OR_REP_FP16_V16) ? 4u : 2u);
}

/* Packed-slot predicates: grouped_storage_sub_bits.cuh:160-195.
 * empty == fp bits 0; tombstone == fp bits all-ones (get_empty()-1). */
static inline int p_empty(const or_table *t, uint32_t w)
{
    return t->rep == OR_REP_FP12 ? w == 0 : (w & kmask_of(t)) == 0;
}
/* The 12-bit bucket has no tombstones (its remove does not compile in the
 * reference, 12_bit_bucket.cuh:207-246). */
static inline int p_tomb(const or_table *t, uint32_t w)
{
    return t->rep == OR_REP_FP12 ? 0 : (w & kmask_of(t)) == kmask_of(t);
}
/* Query match of slot i (contains / sub_byte_match) and its value. */
static inline int p_match(const or_table *t, uint32_t w, uint32_t fp, uint32_t i)
{
    if (t->rep == OR_REP_FP12) return w == or_fp12_proj(fp, i & 3u);
    return (w & kmask_of(t)) == (fp & kmask_of(t));
}
static inline uint32_t p_val(const or_table *t, uint32_t w)
{
    return t->rep == OR_REP_FP12 ? 0u : w >> kbits_of(t);  /* 12-bit query returns 0 */
}
static inline uint32_t tomb_word(const or_table *t) { return kmask_of(t); } /* join(tombstone, 0) */

/* KV predicates: representations/key_val_pair.cuh:70,111-125 (empty key 0,
 * tombstone key ~0). */
#define KV_EMPTY 0ULL
#define KV_TOMB (~0ULL)

static void account_bucket(or_table *t, int level, uint64_t b)
{
    (void)b;
    t->buckets_read[level ? 1 : 0] += 1;
    t->sectors_read += (t->lv[level].bucket_bytes + 63) / 64;
}

/* ------------------------------------------------------------------ */
/* Bucket routines for packed slots.                                   */
/* representations/packed_bucket.cuh: the tile scans slot stripes in    */
/* ascending order and the leader is the lowest lane, so every routine */
/* acts on the LOWEST-index qualifying slot, whatever the tile width.  */
/* ------------------------------------------------------------------ */

/* get_fill: slots neither empty nor tombstone. packed_bucket.cuh:40-62 */
static int pb_fill(const or_table *t, const or_level *L, uint64_t b)
{
    int f = 0;
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = sget(t, L, b, i);
        if (!(p_empty(t, w) || p_tomb(t, w))) f++;
    }
    return f;
}

/* insert: claim the lowest EMPTY slot (CAS from the all-zero word).
 * packed_bucket.cuh:64-114, grouped_storage_sub_bits.cuh:127-141. */
static int pb_insert(or_table *t, or_level *L, uint64_t b, uint32_t word)
{
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = sget(t, L, b, i);
        if (p_empty(t, w)) {
            if (w == 0) { /* CAS(0 -> word) */
                sset(t, L, b, i, word);
                t->sectors_written += 1;
                return 1;
            }
        }
    }
    return 0;
}

/* insert_delete: lowest empty-or-tombstone slot; CAS from the empty word or
 * from the tombstone word.  packed_bucket.cuh:116-180,
 * grouped_storage_sub_bits.cuh:143-157. */
static int pb_insert_delete(or_table *t, or_level *L, uint64_t b, uint32_t word)
{
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = sget(t, L, b, i);
        if (p_empty(t, w) || p_tomb(t, w)) {
            uint32_t expect = p_empty(t, w) ? 0u : tomb_word(t);
            if (w == expect) {
                sset(t, L, b, i, word);
                t->sectors_written += 1;
                return 1;
            }
        }
    }
    return 0;
}

/* query: value of the lowest slot whose fp matches.  packed_bucket.cuh:182-219 */
static int pb_query(const or_table *t, const or_level *L, uint64_t b, uint32_t fp,
                    uint32_t *val)
{
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = sget(t, L, b, i);
        if (p_match(t, w, fp, i)) {
            *val = p_val(t, w);
            return 1;
        }
    }
    return 0;
}

/* remove: CAS the lowest matching slot's word to the tombstone word.
 * packed_bucket.cuh:227-271, grouped_storage_sub_bits.cuh:217-234. */
static int pb_remove(or_table *t, or_level *L, uint64_t b, uint32_t fp)
{
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = sget(t, L, b, i);
        if (p_match(t, w, fp, i)) {
            sset(t, L, b, i, tomb_word(t));
            t->sectors_written += 1;
            return 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Insert schemes for packed slots                                     */
/* ------------------------------------------------------------------ */
#define MAXP 8

/* power_of_n_insert_shortcut_bucket_scheme::insert / insert_with_delete,
 * insert_schemes/power_of_n_shortcut_buckets.cuh:198-306 and :309-417.
 *   pass 1: for each probe j (in probe order) b_j = p_j % nb, f_j = fill;
 *           if f_j < 0.75*B try to insert into b_j (return on success; on
 *           failure f_j := B).  "fill < B*.75" is evaluated exactly as
 *           4*fill < 3*B.
 *   pass 2: for level = min f, ascending through the distinct fills < B,
 *           try every probe j with f_j == level, in probe order. */
static int shortcut_insert(or_table *t, or_level *L, uint64_t key, uint32_t word,
                           int with_delete)
{
    uint64_t buckets[MAXP];
    int fill[MAXP];
    int np = 0;
    int B = (int)L->B;
    int min_fill = B;
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END && np < MAXP; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        buckets[np] = b;
        account_bucket(t, (int)(L - t->lv), b);
        int f = pb_fill(t, L, b);
        if (4 * f < 3 * B) {
            int ok = with_delete ? pb_insert_delete(t, L, b, word) : pb_insert(t, L, b, word);
            if (ok) return 1;
            f = B;
        }
        fill[np] = f;
        if (f < min_fill) min_fill = f;
        np++;
    }
    int lvl = min_fill;
    while (lvl < B) {
        int next = B;
        for (int j = 0; j < np; j++) {
            if (fill[j] == lvl) {
                int ok = with_delete ? pb_insert_delete(t, L, buckets[j], word)
                                     : pb_insert(t, L, buckets[j], word);
                if (ok) return 1;
            }
            if (fill[j] > lvl && fill[j] < next) next = fill[j];
        }
        lvl = next;
    }
    return 0;
}

/* power_of_n_shortcut_buckets.cuh:676-707: check every probe bucket in
 * order, the first match wins (no early exit on empty slots). */
static int shortcut_query(or_table *t, or_level *L, uint64_t key, uint32_t *val)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        account_bucket(t, (int)(L - t->lv), b);
        if (pb_query(t, L, b, fp, val)) return 1;
    }
    return 0;
}

/* power_of_n_shortcut_buckets.cuh:709-740 */
static int shortcut_remove(or_table *t, or_level *L, uint64_t key)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        account_bucket(t, (int)(L - t->lv), b);
        if (pb_remove(t, L, b, fp)) return 1;
    }
    return 0;
}

/* linear_insert_bucket_scheme, insert_schemes/linear_insert_buckets.cuh:
 * insert :200-221 (first probe bucket with an empty slot), insert_with_delete
 * :223-244, query :362-398 and remove :323-360 (stop at the first probe
 * bucket whose fill != B). */
static int linear_insert(or_table *t, or_level *L, int lvl, uint64_t key, uint32_t word,
                         int with_delete)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        account_bucket(t, lvl, b);
        int ok = with_delete ? pb_insert_delete(t, L, b, word) : pb_insert(t, L, b, word);
        if (ok) return 1;
    }
    return 0;
}

static int linear_query(or_table *t, or_level *L, int lvl, uint64_t key, uint32_t *val)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        account_bucket(t, lvl, b);
        if (pb_query(t, L, b, fp, val)) return 1;
        if (pb_fill(t, L, b) != (int)L->B) return 0;
    }
    return 0;
}

static int linear_remove(or_table *t, or_level *L, int lvl, uint64_t key)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        account_bucket(t, lvl, b);
        if (pb_remove(t, L, b, fp)) return 1;
        if (pb_fill(t, L, b) != (int)L->B) return 0;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* KV (key_val_pair<uint64,uint64>) schemes, and the key-only variant    */
/* dynamic_container<key_container, uint64_t> (KEY64) the CMS allocator   */
/* keeps its free list in (allocators/cms.cuh:203).                     */
/* ------------------------------------------------------------------ */
/* Key-value slots.  KV64 = key_val_pair<u64,u64> {key, val} (16 B);
 * KEY64 = dynamic_container<key_container, u64> (8 B key); KV16 =
 * dynamic_container<key_val_pair, u16> (4 B {u16 key, u16 val}, as one
 * little-endian word val << 16 | key).  The dynamic containers store
 * get_smallKey(key): KEY64 remaps 0 -> 1 and ~0 -> ~0 - 1, KV16 keeps the
 * low 16 bits with 0 -> 1 and 0xFFFF -> 0xFFFE
 * (representations/dynamic_container.cuh:35-50); its tag() is that small key
 * (:52-54), key_val_pair's is the key (key_val_pair.cuh:124-126).  Empty key
 * 0, tombstone key get_empty() - 1 (~0 or 0xFFFF, key_val_pair.cuh:70,
 * 115-117); atomic_reset changes the key only (:143-151). */
static inline int kv_keyonly(const or_table *t) { return t->rep == OR_REP_KEY64; }
static inline uint64_t kv_tomb_key(const or_table *t) { return t->rep == OR_REP_KV16 ? 0xFFFFull : KV_TOMB; }
static inline uint64_t kv_key_at(const or_table *t, const or_level *L, uint64_t s)
{
    if (t->rep == OR_REP_KV16) return ((const uint32_t *)L->slots)[s] & 0xFFFFu;
    return ((const uint64_t *)L->slots)[(kv_keyonly(t) ? 1 : 2) * s];
}
static inline uint64_t kv_val_at(const or_table *t, const or_level *L, uint64_t s)
{
    if (t->rep == OR_REP_KV16) return ((const uint32_t *)L->slots)[s] >> 16;
    return kv_keyonly(t) ? 0 : ((const uint64_t *)L->slots)[2 * s + 1];
}
static inline void kv_put(const or_table *t, or_level *L, uint64_t s, uint64_t skey, uint64_t val)
{
    if (t->rep == OR_REP_KV16) {
        ((uint32_t *)L->slots)[s] = (uint32_t)((val & 0xFFFFu) << 16 | (skey & 0xFFFFu));
        return;
    }
    uint64_t *p = (uint64_t *)L->slots + (kv_keyonly(t) ? 1 : 2) * s;
    p[0] = skey;
    if (!kv_keyonly(t)) p[1] = val;
}
static inline void kv_bury(const or_table *t, or_level *L, uint64_t s)
{
    if (t->rep == OR_REP_KV16) {
        uint32_t *w = (uint32_t *)L->slots + s;
        *w = (*w & 0xFFFF0000u) | 0xFFFFu;
        return;
    }
    ((uint64_t *)L->slots)[(kv_keyonly(t) ? 1 : 2) * s] = KV_TOMB;
}
/* The stored key: the key itself (key_val_pair) or get_smallKey(key). */
static inline uint64_t kv_skey(const or_table *t, uint64_t key)
{
    if (t->rep == OR_REP_KV16) return fp_of(OR_REP_FP16_V16, key);
    if (!kv_keyonly(t)) return key;
    if (key == KV_EMPTY) key += 1;
    if (key == KV_TOMB) key -= 1;
    return key;
}
/* The probe tag (XOR power-of-two probing). */
static inline uint64_t kv_tag(const or_table *t, uint64_t key)
{
    return t->rep == OR_REP_KV64 ? key : kv_skey(t, key);
}

/* Sector accounting for KV probes: count distinct 64 B sectors in the
 * examined slot range [s0, s1) of the slot array, skipping the sector the
 * previous probe already covered (linear probing walks adjacent buckets). */
static void account_kv_range(or_table *t, int lvl, uint64_t s0, uint64_t s1, uint64_t *last_sector)
{
    t->buckets_read[lvl ? 1 : 0] += 1;
    uint64_t sb = slot_bytes(t);
    uint64_t a = (s0 * sb) / 64, z = (s1 * sb - 1) / 64;
    for (uint64_t s = a; s <= z; s++) {
        if (s != *last_sector) t->sectors_read += 1;
        *last_sector = s;
    }
}

/* bucket_insert::insert, insert_schemes/bucket_insert.cuh:127-203: the
 * lowest empty slot of the first probe bucket that has one.
 * key_val_pair::atomic_swap (key_val_pair.cuh:84-95): CAS key 0 -> key,
 * then store the value. */
static int kv_insert(or_table *t, or_level *L, int lvl, uint64_t key, uint64_t val)
{
    or_probe_it it;
    uint64_t last = OR_END;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, kv_tag(t, key));
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        uint64_t base = b * L->B;
        for (uint32_t i = 0; i < L->B; i++) {
            if (kv_key_at(t, L, base + i) == KV_EMPTY) {
                account_kv_range(t, lvl, base, base + i + 1, &last);
                kv_put(t, L, base + i, kv_skey(t, key), val);
                t->sectors_written += 1;
                return 1;
            }
        }
        account_kv_range(t, lvl, base, base + L->B, &last);
    }
    return 0;
}

/* bucket_insert::query, bucket_insert.cuh:295-357.  The tile scans stripes
 * of P slots (lane i owns slot stripe*P + i); a stripe with a match returns
 * the lowest matching slot; otherwise a stripe holding an empty slot ends
 * the search.  single_slot_insert::query (single_slot_insert.cuh:168-229)
 * is the P=B=1 case. */
static int kv_query(or_table *t, or_level *L, int lvl, uint64_t key, uint64_t *val)
{
    uint32_t P = L->d.tile ? L->d.tile : 1;
    if (L->d.scheme == OR_SCHEME_SINGLE_SLOT) P = 1;
    or_probe_it it;
    uint64_t last = OR_END;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, kv_tag(t, key));
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        uint64_t base = b * L->B;
        for (uint32_t s0 = 0; s0 < L->B; s0 += P) {
            uint32_t s1 = s0 + P < L->B ? s0 + P : L->B;
            account_kv_range(t, lvl, base + s0, base + s1, &last);
            for (uint32_t i = s0; i < s1; i++)
                if (kv_key_at(t, L, base + i) == kv_skey(t, key)) { *val = kv_val_at(t, L, base + i); return 1; }
            for (uint32_t i = s0; i < s1; i++)
                if (kv_key_at(t, L, base + i) == KV_EMPTY) return 0;
        }
    }
    return 0;
}

/* bucket_insert::remove, bucket_insert.cuh:205-293: CAS the lowest match in
 * the stripe to the tombstone key (key_val_pair.cuh:143-151); a stripe with
 * no match but an empty slot ends the search. */
static int kv_remove(or_table *t, or_level *L, int lvl, uint64_t key)
{
    uint32_t P = L->d.tile ? L->d.tile : 1;
    or_probe_it it;
    uint64_t last = OR_END;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, kv_tag(t, key));
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        uint64_t base = b * L->B;
        for (uint32_t s0 = 0; s0 < L->B; s0 += P) {
            uint32_t s1 = s0 + P < L->B ? s0 + P : L->B;
            account_kv_range(t, lvl, base + s0, base + s1, &last);
            for (uint32_t i = s0; i < s1; i++) {
                if (kv_key_at(t, L, base + i) == kv_skey(t, key)) {
                    kv_bury(t, L, base + i);
                    t->sectors_written += 1;
                    return 1;
                }
            }
            for (uint32_t i = s0; i < s1; i++)
                if (kv_key_at(t, L, base + i) == KV_EMPTY) return 0;
        }
    }
    return 0;
}

/* The flat power-of-n schemes over key-value slots: power_of_n_insert_scheme
 * (insert_schemes/power_of_n.cuh:326-422) and power_of_n_insert_shortcut_scheme
 * (power_of_n_shortcut.cuh:327-435, data_structs/tcf.cuh:30), each "bucket"
 * being B consecutive slots.  Unlike the packed-bucket scheme:
 *   fill      = slots whose key is not empty, tombstones included
 *               (check_fill_bucket, power_of_n_shortcut.cuh:234-261);
 *   claim     = the lowest EMPTY slot only (insert_into_bucket, :134-191);
 *   query     = every probe bucket in order, lowest match (:193-232, :437-468);
 *   remove    = the lowest match's key -> tombstone, value kept (:263-325, :470-505).
 * insert: the shortcut scheme first tries each probe bucket whose fill is
 * below 0.75 B (4*fill < 3*B); then, for both schemes, the fill levels in
 * ascending order, each probe with that fill in probe order. */
static int kv_pow_fill(or_table *t, or_level *L, uint64_t b)
{
    int f = 0;
    for (uint32_t i = 0; i < L->B; i++) f += kv_key_at(t, L, b * L->B + i) != KV_EMPTY;
    return f;
}

static int kv_pow_claim(or_table *t, or_level *L, uint64_t b, uint64_t key, uint64_t val)
{
    for (uint32_t i = 0; i < L->B; i++) {
        if (kv_key_at(t, L, b * L->B + i) == KV_EMPTY) {
            kv_put(t, L, b * L->B + i, kv_skey(t, key), val);
            t->sectors_written += 1;
            return 1;
        }
    }
    return 0;
}

/* power_of_n_insert_shortcut_replace_scheme (power_of_n_shortcut_key_replace.cuh:
 * 39-523), read with three slips fixed: insert()'s `if` at :361 closes after
 * `current_fill = Bucket_Size;` as in power_of_n_shortcut.cuh; found_ballot
 * (:235) starts false, so a bucket reads Bucket_Size + 1 only when a slot holds
 * the key; the swap on a slot holding the key replaces its value.
 *   fill  = as above, or B + 1 when a slot holds the key (:233-270);
 *   claim = the lowest slot that is empty or holds the key (:134-190): an
 *           empty slot takes key and value, a held key gets the new value.
 * Query and remove are the shortcut scheme's (:446-523). */
static int kv_pow_find(or_table *t, or_level *L, uint64_t b, uint64_t key)
{
    for (uint32_t i = 0; i < L->B; i++)
        if (kv_key_at(t, L, b * L->B + i) == kv_skey(t, key)) return 1;
    return 0;
}

static int kv_pow_claim_replace(or_table *t, or_level *L, uint64_t b, uint64_t key, uint64_t val)
{
    for (uint32_t i = 0; i < L->B; i++) {
        const uint64_t k = kv_key_at(t, L, b * L->B + i);
        if (k == KV_EMPTY || k == kv_skey(t, key)) {
            kv_put(t, L, b * L->B + i, kv_skey(t, key), val);
            t->sectors_written += 1;
            return 1;
        }
    }
    return 0;
}

static int kv_pow_insert(or_table *t, or_level *L, int lvl, uint64_t key, uint64_t val)
{
    const int replace = L->d.scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
    const int shortcut = L->d.scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_SLOTS || replace;
    int (*claim)(or_table *, or_level *, uint64_t, uint64_t, uint64_t) =
        replace ? kv_pow_claim_replace : kv_pow_claim;
    const int B = (int)L->B;
    uint64_t buckets[MAXP];
    int fill[MAXP];
    int np = 0, min_fill = B;
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes, L->seed, key, kv_tag(t, key));
         p != OR_END && np < MAXP; p = probe_next(&it)) {
        const uint64_t b = p % L->nb;
        buckets[np] = b;
        account_bucket(t, lvl, b);
        int f = kv_pow_fill(t, L, b);
        if (replace && kv_pow_find(t, L, b, key)) f = B + 1;
        if (shortcut && (4 * f < 3 * B || f == B + 1)) {
            if (claim(t, L, b, key, val)) return 1;
            f = B;
        }
        fill[np] = f;
        if (f < min_fill) min_fill = f;
        np++;
    }
    for (int lvl_fill = min_fill; lvl_fill < B;) {
        int next = B;
        for (int j = 0; j < np; j++) {
            if (fill[j] == lvl_fill && claim(t, L, buckets[j], key, val)) return 1;
            if (fill[j] > lvl_fill && fill[j] < next) next = fill[j];
        }
        lvl_fill = next;
    }
    return 0;
}

static int kv_pow_query(or_table *t, or_level *L, int lvl, uint64_t key, uint64_t *val)
{
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes, L->seed, key, kv_tag(t, key));
         p != OR_END; p = probe_next(&it)) {
        const uint64_t b = p % L->nb;
        account_bucket(t, lvl, b);
        for (uint32_t i = 0; i < L->B; i++)
            if (kv_key_at(t, L, b * L->B + i) == kv_skey(t, key)) {
                *val = kv_val_at(t, L, b * L->B + i);
                return 1;
            }
    }
    return 0;
}

static int kv_pow_remove(or_table *t, or_level *L, int lvl, uint64_t key)
{
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes, L->seed, key, kv_tag(t, key));
         p != OR_END; p = probe_next(&it)) {
        const uint64_t b = p % L->nb;
        account_bucket(t, lvl, b);
        for (uint32_t i = 0; i < L->B; i++)
            if (kv_key_at(t, L, b * L->B + i) == kv_skey(t, key)) {
                kv_bury(t, L, b * L->B + i);
                t->sectors_written += 1;
                return 1;
            }
    }
    return 0;
}

static inline int kv_pow_scheme(const or_level *L)
{
    return L->d.scheme == OR_SCHEME_POWER_OF_N || L->d.scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_SLOTS ||
           L->d.scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
}

/* ------------------------------------------------------------------ */
/* Table lifecycle                                                     */
/* ------------------------------------------------------------------ */

/* bucketed_table::generate_on_device / static_table::generate_on_device
 * (tables/bucketed_table.cuh:132-192, tables/base_table.cuh:110-145): level
 * i takes the next size of the sizing (sizing/default_sizing.cuh:54-73) and
 * seed, the backing level seed*seed.  Buckets = ceil(nslots / B)
 * (power_of_n_shortcut_buckets.cuh:94, bucket_insert.cuh:90-92); the single
 * slot scheme allocates nslots slots (single_slot_insert.cuh:89-96). */
or_table *or_create(uint32_t rep, uint32_t nlevels, const or_level_desc *lv, uint64_t seed)
{
    if (nlevels < 1 || nlevels > OR_MAX_LEVELS || rep > OR_REP_KV16) return NULL;
    for (uint32_t l = 0; l < nlevels; l++) {
        /* GROUPED levels hold `group` packed buckets per 128-byte line
         * (grouped_power_buckets.cuh:58: sizeof(bucket) * K <= 128). */
        int grouped = lv[l].scheme == OR_SCHEME_GROUPED;
        if (grouped != (lv[l].group != 0)) return NULL;
        if (grouped && (is_kv_rep(rep) || lv[l].bucket_size == 0 ||
                        lv[l].group * bucket_bytes_of(rep, lv[l].bucket_size) > 128))
            return NULL;
        const int flat = lv[l].scheme == OR_SCHEME_POWER_OF_N || lv[l].scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_SLOTS ||
                         lv[l].scheme == OR_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
        const int kv_scheme = lv[l].scheme <= OR_SCHEME_SINGLE_SLOT || flat;
        if (is_kv_rep(rep) != kv_scheme) return NULL;
        /* the flat power-of-n schemes keep one fill per probe (MAXP) */
        if (flat &&
            (lv[l].max_probes == 0 || lv[l].max_probes > MAXP))
            return NULL;
    }
    or_table *t = (or_table *)calloc(1, sizeof(or_table));
    if (!t) return NULL;
    t->rep = rep;
    t->nlevels = nlevels;
    uint64_t s = seed;
    for (uint32_t l = 0; l < nlevels; l++) {
        or_level *L = &t->lv[l];
        L->d = lv[l];
        L->seed = s;
        s = s * s;
        if (lv[l].nslots == 0) { free(t); return NULL; }
        if (lv[l].scheme == OR_SCHEME_SINGLE_SLOT) {
            L->B = 1;
            L->nb = lv[l].nslots;
        } else {
            L->B = lv[l].bucket_size;
            L->nb = (lv[l].nslots - 1) / L->B + 1;
        }
        L->nslot_alloc = L->nb * L->B;
        L->bucket_bytes = bucket_bytes_of(rep, L->B);
        L->bytes = lv[l].group ? ((L->nb - 1) / lv[l].group + 1) * 128 : L->nb * L->bucket_bytes;
        L->slots = calloc(L->bytes, 1);
        if (!L->slots) { for (uint32_t q = 0; q < l; q++) free(t->lv[q].slots); free(t); return NULL; }
    }
    return t;
}

void or_destroy(or_table *t)
{
    if (!t) return;
    for (uint32_t l = 0; l < t->nlevels; l++) free(t->lv[l].slots);
    free(t);
}

void or_clear(or_table *t)
{
    for (uint32_t l = 0; l < t->nlevels; l++)
        memset(t->lv[l].slots, 0, t->lv[l].bytes);
}

/* Per-level dispatch for the packed representations.  Every table op
 * tries level 0 and, on false, the backing level (bucketed_table.cuh:271-381).
 * GROUPED levels run the shortcut scheme's ops: power_of_n_insert_grouped_buckets
 * (grouped_power_buckets.cuh:164-384) repeats them over grouped buckets. */
static int lvl_insert(or_table *t, int l, uint64_t key, uint32_t word, int with_delete)
{
    or_level *L = &t->lv[l];
    if (L->d.scheme != OR_SCHEME_LINEAR_BUCKETS) return shortcut_insert(t, L, key, word, with_delete);
    return linear_insert(t, L, l, key, word, with_delete);
}
static int lvl_query(or_table *t, int l, uint64_t key, uint32_t *val)
{
    or_level *L = &t->lv[l];
    if (L->d.scheme != OR_SCHEME_LINEAR_BUCKETS) return shortcut_query(t, L, key, val);
    return linear_query(t, L, l, key, val);
}
static int lvl_remove(or_table *t, int l, uint64_t key)
{
    or_level *L = &t->lv[l];
    if (L->d.scheme != OR_SCHEME_LINEAR_BUCKETS) return shortcut_remove(t, L, key);
    return linear_remove(t, L, l, key);
}

/* Insert one key; returns 0 on failure, else 1 + level index. */
int or_insert(or_table *t, uint64_t key, uint64_t val, int with_delete)
{
    if (is_kv_rep(t->rep)) {
        for (uint32_t l = 0; l < t->nlevels; l++) {
            or_level *L = &t->lv[l];
            if (kv_pow_scheme(L) ? kv_pow_insert(t, L, (int)l, key, val) : kv_insert(t, L, (int)l, key, val))
                return 1 + (int)l;
        }
        return 0;
    }
    /* twelve_bucket has no insert_delete that compiles (12_bit_bucket.cuh:
     * 148-203 calls a sub_byte_match overload that does not exist) */
    if (t->rep == OR_REP_FP12 && with_delete) return -1;
    uint32_t word = t->rep == OR_REP_FP12 ? fp_of(t->rep, key)
                                          : or_pack(kbits_of(t), vbits_of(t), fp_of(t->rep, key), (uint32_t)val);
    for (uint32_t l = 0; l < t->nlevels; l++)
        if (lvl_insert(t, (int)l, key, word, with_delete)) return 1 + (int)l;
    return 0;
}

int or_query(or_table *t, uint64_t key, uint64_t *val)
{
    if (is_kv_rep(t->rep)) {
        for (uint32_t l = 0; l < t->nlevels; l++) {
            or_level *L = &t->lv[l];
            if (kv_pow_scheme(L) ? kv_pow_query(t, L, (int)l, key, val) : kv_query(t, L, (int)l, key, val))
                return 1;
        }
        return 0;
    }
    for (uint32_t l = 0; l < t->nlevels; l++) {
        uint32_t v;
        if (lvl_query(t, (int)l, key, &v)) { *val = v; return 1; }
    }
    return 0;
}

int or_remove(or_table *t, uint64_t key)
{
    if (is_kv_rep(t->rep)) {
        for (uint32_t l = 0; l < t->nlevels; l++)
            if (t->lv[l].d.scheme == OR_SCHEME_SINGLE_SLOT) return -1; /* no remove in the reference */
        for (uint32_t l = 0; l < t->nlevels; l++) {
            or_level *L = &t->lv[l];
            if (kv_pow_scheme(L) ? kv_pow_remove(t, L, (int)l, key) : kv_remove(t, L, (int)l, key)) return 1;
        }
        return 0;
    }
    if (t->rep == OR_REP_FP12) return -1; /* twelve_bucket::remove does not compile (12_bit_bucket.cuh:207-246) */
    for (uint32_t l = 0; l < t->nlevels; l++)
        if (lvl_remove(t, (int)l, key)) return 1;
    return 0;
}

/* ---- upserts ------------------------------------------------------ */

/* power_of_n_insert_shortcut_bucket_scheme::insert_if_not_exists
 * (insert_schemes/power_of_n_shortcut_buckets.cuh:421-544): query every probe
 * bucket first (a match returns its value, found = true); otherwise the plain
 * shortcut insert.  The _delete form (:546-674) removes the match instead of
 * returning it, and inserts with tombstone reuse. */
static int shortcut_upsert(or_table *t, or_level *L, uint64_t key, uint32_t word, int del,
                           int *found, uint32_t *val)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        if (pb_query(t, L, b, fp, val)) {
            *found = 1;
            if (del) pb_remove(t, L, b, fp);
            return 1;
        }
    }
    return shortcut_insert(t, L, key, word, del);
}

/* linear_insert_bucket_scheme::insert_if_not_exists (linear_insert_buckets.cuh:
 * 249-282): per probe bucket, a match returns it, else a free slot takes the
 * key.  The _delete form (:285-321) removes the match; its insert does not
 * reuse tombstones. */
static int linear_upsert(or_table *t, or_level *L, uint64_t key, uint32_t word, int del,
                         int *found, uint32_t *val)
{
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes,
                                  L->seed, key, fp);
         p != OR_END; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        if (pb_query(t, L, b, fp, val)) {
            *found = 1;
            if (del) pb_remove(t, L, b, fp);
            return 1;
        }
        if (pb_insert(t, L, b, word)) return 1;
    }
    return 0;
}

/* Table-level upsert (bucketed_table.cuh:226-268): level 0, then the backing
 * level when level 0 returns false.  Returns 0 on failure, else 1 + level;
 * *found tells whether the key was already present (value in *val). */
int or_insert_if_not_exists(or_table *t, uint64_t key, uint64_t val, int del, int *found, uint64_t *val_out)
{
    *found = 0;
    *val_out = 0;
    if (is_kv_rep(t->rep)) return -1; /* bucket_insert has no upsert */
    if (t->rep == OR_REP_FP12) return -1; /* not offered for the 12-bit bucket */
    uint32_t word = or_pack(kbits_of(t), vbits_of(t), fp_of(t->rep, key), (uint32_t)val);
    for (uint32_t l = 0; l < t->nlevels; l++) {
        or_level *L = &t->lv[l];
        uint32_t v = 0;
        int r = (L->d.scheme != OR_SCHEME_LINEAR_BUCKETS)
                    ? shortcut_upsert(t, L, key, word, del, found, &v)
                    : linear_upsert(t, L, key, word, del, found, &v);
        if (r) {
            *val_out = *found ? v : 0;
            return 1 + (int)l;
        }
    }
    return 0;
}

void or_bulk_insert_if_not_exists(or_table *t, const uint64_t *keys, const uint16_t *vals, uint64_t n,
                                  int del, uint16_t *vals_out, uint8_t *found, uint8_t *status)
{
    for (uint64_t i = 0; i < n; i++) {
        int f = 0;
        uint64_t v = 0;
        int r = or_insert_if_not_exists(t, keys[i], vals ? vals[i] : 0, del, &f, &v);
        if (status) status[i] = (uint8_t)(r < 0 ? 0 : r);
        if (found) found[i] = (uint8_t)f;
        if (vals_out) vals_out[i] = (uint16_t)v;
    }
}

/* ---- bulk wrappers: the harness kernels of tests/src/test_mhm_tcf.cu:
 * 318-487 apply one op per key; sequential order here. ---- */
void or_bulk_insert(or_table *t, const uint64_t *keys, const void *vals, uint32_t val_bytes,
                    uint64_t n, uint8_t *status, int with_delete)
{
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v = 0;
        if (vals) {
            if (val_bytes == 8) v = ((const uint64_t *)vals)[i];
            else if (val_bytes == 2) v = ((const uint16_t *)vals)[i];
            else if (val_bytes == 1) v = ((const uint8_t *)vals)[i];
        }
        int r = or_insert(t, keys[i], v, with_delete);
        if (status) status[i] = (uint8_t)r;
    }
}

void or_bulk_query(or_table *t, const uint64_t *keys, uint64_t n, void *vals_out,
                   uint32_t val_bytes, uint8_t *found)
{
    for (uint64_t i = 0; i < n; i++) {
        uint64_t v = 0;
        int r = or_query(t, keys[i], &v);
        if (found) found[i] = (uint8_t)r;
        if (vals_out) {
            if (!r) v = 0;
            if (val_bytes == 8) ((uint64_t *)vals_out)[i] = v;
            else if (val_bytes == 2) ((uint16_t *)vals_out)[i] = (uint16_t)v;
            else if (val_bytes == 1) ((uint8_t *)vals_out)[i] = (uint8_t)v;
        }
    }
}

void or_bulk_remove(or_table *t, const uint64_t *keys, uint64_t n, uint8_t *found)
{
    for (uint64_t i = 0; i < n; i++) {
        int r = or_remove(t, keys[i]);
        if (found) found[i] = (uint8_t)(r > 0);
    }
}

/* bucketed_table::get_fill (tables/bucketed_table.cuh:498-548): filled
 * slots (neither empty nor tombstone) summed over levels. */
uint64_t or_fill(const or_table *t)
{
    uint64_t f = 0;
    for (uint32_t l = 0; l < t->nlevels; l++) {
        const or_level *L = &t->lv[l];
        if (is_kv_rep(t->rep)) {
            for (uint64_t s = 0; s < L->nslot_alloc; s++) {
                uint64_t k = kv_key_at(t, L, s);
                if (k != KV_EMPTY && k != kv_tomb_key(t)) f++;
            }
            continue;
        }
        for (uint64_t b = 0; b < L->nb; b++)
            for (uint32_t i = 0; i < L->B; i++) {
                uint32_t w = sget(t, L, b, i);
                if (!(p_empty(t, w) || p_tomb(t, w))) f++;
            }
    }
    return f;
}

uint64_t or_level_bytes(const or_table *t, uint32_t l)
{
    if (l >= t->nlevels) return 0;
    return t->lv[l].bytes;
}
const void *or_level_data(const or_table *t, uint32_t l)
{
    if (l >= t->nlevels) return NULL;
    return t->lv[l].slots;
}
uint64_t or_level_buckets(const or_table *t, uint32_t l)
{
    if (l >= t->nlevels) return 0;
    return t->lv[l].nb;
}

void or_stats(const or_table *t, uint64_t *out4)
{
    out4[0] = t->sectors_read;
    out4[1] = t->sectors_written;
    out4[2] = t->buckets_read[0];
    out4[3] = t->buckets_read[1];
}
void or_stats_reset(or_table *t)
{
    t->sectors_read = t->sectors_written = 0;
    t->buckets_read[0] = t->buckets_read[1] = 0;
}

/* ------------------------------------------------------------------ */
/* Seeded key streams: splitmix64 (Vigna) output of counter i of a      */
/* stream; a bijection of the counter, so keys of one seed are distinct.*/
/* 0 and ~0 are reserved slot words; they map to fixed other values.   */
/* ------------------------------------------------------------------ */
uint64_t or_key_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    z ^= z >> 31;
    if (z == 0) z = 0x8000000000000000ULL;
    if (z == ~0ULL) z = 0x7FFFFFFFFFFFFFFFULL;
    return z;
}

void or_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *out)
{
    for (uint64_t i = 0; i < n; i++) out[i] = or_key_at(seed, start + i);
}

/* ------------------------------------------------------------------ */
/* Multi-threaded run of the same algorithm (the CPU baseline's         */
/* all-cores figure).  16/16 slots, a shortcut level 0 and an optional  */
/* linear backing level: the table shape the filter benchmark uses.    */
/* nthreads host threads take contiguous slices of the keys; slots are  */
/* read with relaxed atomic loads and claimed with an atomic compare-   */
/* and-swap from the empty word (the reference's atomicCAS,             */
/* grouped_storage_sub_bits.cuh:127-141); a lost CAS moves on to the    */
/* next empty slot as the reference's leader loop does                   */
/* (packed_bucket.cuh:64-114).  With one thread the slot layout is the   */
/* sequential oracle's; with more, placement depends on thread timing   */
/* (as on the GPU) while membership does not.  No accounting counters.  */
/* ------------------------------------------------------------------ */
static inline uint32_t *mt_slots(const or_level *L, uint64_t b) { return (uint32_t *)bucket_at(L, b); }

static int mt_fill(const or_level *L, uint64_t b)
{
    const uint32_t *s = mt_slots(L, b);
    int f = 0;
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t fp = __atomic_load_n(s + i, __ATOMIC_RELAXED) & 0xFFFFu;
        f += (fp != 0 && fp != 0xFFFFu);
    }
    return f;
}

static int mt_claim(const or_level *L, uint64_t b, uint32_t word)
{
    uint32_t *s = mt_slots(L, b);
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t expect = 0;
        if (__atomic_load_n(s + i, __ATOMIC_RELAXED) == 0 &&
            __atomic_compare_exchange_n(s + i, &expect, word, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED))
            return 1;
    }
    return 0;
}

static int mt_find(const or_level *L, uint64_t b, uint32_t fp, uint32_t *val)
{
    const uint32_t *s = mt_slots(L, b);
    for (uint32_t i = 0; i < L->B; i++) {
        uint32_t w = __atomic_load_n(s + i, __ATOMIC_RELAXED);
        if ((w & 0xFFFFu) == fp) {
            *val = w >> 16;
            return 1;
        }
    }
    return 0;
}

/* shortcut_insert / linear_insert above, over atomic slots */
static int mt_insert(const or_table *t, uint64_t key, uint32_t word)
{
    const or_level *L = &t->lv[0];
    uint32_t fp = fp_of(t->rep, key);
    uint64_t buckets[MAXP];
    int fill[MAXP];
    int np = 0, B = (int)L->B, min_fill = B;
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes, L->seed, key, fp);
         p != OR_END && np < MAXP; p = probe_next(&it)) {
        uint64_t b = p % L->nb;
        buckets[np] = b;
        int f = mt_fill(L, b);
        if (4 * f < 3 * B) {
            if (mt_claim(L, b, word)) return 1;
            f = B;
        }
        fill[np] = f;
        if (f < min_fill) min_fill = f;
        np++;
    }
    for (int lvl = min_fill; lvl < B;) {
        int next = B;
        for (int j = 0; j < np; j++) {
            if (fill[j] == lvl && mt_claim(L, buckets[j], word)) return 1;
            if (fill[j] > lvl && fill[j] < next) next = fill[j];
        }
        lvl = next;
    }
    if (t->nlevels > 1) {
        const or_level *K = &t->lv[1];
        for (uint64_t p = probe_begin(&it, K->d.prober, K->d.hasher, K->d.max_probes, K->seed, key, fp);
             p != OR_END; p = probe_next(&it))
            if (mt_claim(K, p % K->nb, word)) return 2;
    }
    return 0;
}

/* shortcut_query, then linear_query's walk to the first bucket not full */
static int mt_query(const or_table *t, uint64_t key, uint32_t *val)
{
    const or_level *L = &t->lv[0];
    uint32_t fp = fp_of(t->rep, key);
    or_probe_it it;
    for (uint64_t p = probe_begin(&it, L->d.prober, L->d.hasher, L->d.max_probes, L->seed, key, fp);
         p != OR_END; p = probe_next(&it))
        if (mt_find(L, p % L->nb, fp, val)) return 1;
    if (t->nlevels > 1) {
        const or_level *K = &t->lv[1];
        for (uint64_t p = probe_begin(&it, K->d.prober, K->d.hasher, K->d.max_probes, K->seed, key, fp);
             p != OR_END; p = probe_next(&it)) {
            uint64_t b = p % K->nb;
            if (mt_find(K, b, fp, val)) return 1;
            if (mt_fill(K, b) != (int)K->B) return 0;
        }
    }
    return 0;
}

typedef struct {
    const or_table *t;
    const uint64_t *keys;
    const uint16_t *vals;
    uint64_t lo, hi;
    int query;
    uint8_t *out;        /* insert: status; query: found */
    uint16_t *vals_out;  /* query values */
} mt_job;

static void *mt_worker(void *arg)
{
    const mt_job *j = (const mt_job *)arg;
    for (uint64_t i = j->lo; i < j->hi; i++) {
        if (j->query) {
            uint32_t v = 0;
            int r = mt_query(j->t, j->keys[i], &v);
            if (j->out) j->out[i] = (uint8_t)r;
            if (j->vals_out) j->vals_out[i] = r ? (uint16_t)v : 0;
        } else {
            uint32_t word = or_pack(16, 16, fp_of(j->t->rep, j->keys[i]), j->vals ? j->vals[i] : 0u);
            int r = mt_insert(j->t, j->keys[i], word);
            if (j->out) j->out[i] = (uint8_t)r;
        }
    }
    return NULL;
}

/* Host threads the process may run on. */
uint32_t or_hw_threads(void)
{
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (uint32_t)n : 1u;
}

static int mt_shape_ok(const or_table *t)
{
    if (t->rep != OR_REP_FP16_V16 || t->lv[0].d.scheme != OR_SCHEME_POWER_OF_N_SHORTCUT) return 0;
    return t->nlevels == 1 || (t->nlevels == 2 && t->lv[1].d.scheme == OR_SCHEME_LINEAR_BUCKETS);
}

static int mt_run(const or_table *t, const uint64_t *keys, const uint16_t *vals, uint64_t n, int query,
                  uint8_t *out, uint16_t *vals_out, uint32_t nthreads)
{
    if (!mt_shape_ok(t)) return -1;
    if (nthreads == 0) nthreads = or_hw_threads();
    if (nthreads > 256) nthreads = 256;
    if ((uint64_t)nthreads > n) nthreads = n ? (uint32_t)n : 1u;
    pthread_t th[256];
    mt_job jobs[256];
    int created[256];
    for (uint32_t k = 0; k < nthreads; k++) {
        jobs[k] = (mt_job){t, keys, vals, n * k / nthreads, n * (k + 1) / nthreads, query, out, vals_out};
        /* the last slice runs on this thread; a slice whose thread cannot be
         * started runs here too */
        created[k] = k + 1 < nthreads && pthread_create(&th[k], NULL, mt_worker, &jobs[k]) == 0;
        if (!created[k] && k + 1 < nthreads) mt_worker(&jobs[k]);
    }
    mt_worker(&jobs[nthreads - 1]);
    for (uint32_t k = 0; k + 1 < nthreads; k++)
        if (created[k]) pthread_join(th[k], NULL);
    return (int)nthreads;
}

/* Returns the number of threads used, or -1 for an unsupported shape. */
int or_bulk_insert_mt(or_table *t, const uint64_t *keys, const uint16_t *vals, uint64_t n, uint8_t *status,
                      uint32_t nthreads)
{
    return mt_run(t, keys, vals, n, 0, status, NULL, nthreads);
}

int or_bulk_query_mt(or_table *t, const uint64_t *keys, uint64_t n, uint16_t *vals_out, uint8_t *found,
                     uint32_t nthreads)
{
    return mt_run(t, keys, NULL, n, 1, found, vals_out, nthreads);
}
