// poggers_amd/tables.hpp -- C++ templated front-end over the C ABI.
//
// Mirrors the reference's composition API: a table type is
//   bucketed_table<Key, Val, Internal_Rep, Partition_Size, Bucket_Size,
//                  Insert_Scheme, Max_Probes, Probing_Scheme, Hasher,
//                  Is_Recursive, Recursive_Type>
// with the reference's template-parameter order
// (tables/bucketed_table.cuh:72, tables/base_table.cuh:52).  The template
// arguments are checked at compile time and turned into a pg_desc; every
// operation is a bulk, stream-ordered call of the C ABI (include/poggers_amd.h)
// that runs hand-written gfx950 kernels.  Device ops that take a
// cooperative-groups tile in the reference (insert(tile, key, val), ...)
// become bulk_insert / bulk_query / bulk_remove over device arrays.
#pragma once

#include <cstdarg>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../poggers_amd.h"

namespace poggers_amd {

// ---- representations (representations/*.cuh) ----
struct key_val_pair {  // key_val_pair<Key, Val>: 16 B {u64 key, u64 val}
    static constexpr uint32_t rep = PG_REP_KV64;
};
template <std::size_t KeyBits, std::size_t ValBits>
struct bit_grouped_container {  // dynamic_container<bit_grouped_container<K,V>, SmallKey>
    static_assert((KeyBits == 16 && ValBits == 16) || (KeyBits == 8 && ValBits == 8) ||
                      (KeyBits == 10 && ValBits == 6),
                  "packed slots are 16/16, 8/8 or 10/6 (the reference's compositions)");
    static constexpr uint32_t rep =
        KeyBits == 16 ? PG_REP_FP16_V16 : (KeyBits == 8 ? PG_REP_FP8_V8 : PG_REP_FP10_V6);
};

// wrapper_half_bucket<uint16_t> (representations/12_bit_bucket.cuh:319-323):
// the 12-bit sub-word bucket, key-only (Val = uint16_t, queries report 0).
template <typename Internal_Storage>
struct wrapper_half_bucket {
    static_assert(std::is_same<Internal_Storage, uint16_t>::value, "the 12-bit bucket stores uint16_t words");
    static constexpr uint32_t rep = PG_REP_FP12;
};

// dynamic_container<key_container, uint64_t> (representations/key_only.cuh
// under dynamic_container.cuh): key-only 8-byte slots, the CMS allocator's
// free table (allocators/cms.cuh:203).
struct key_container {
    static constexpr uint32_t rep = PG_REP_KEY64;
};

// dynamic_container<Rep, SmallKey> (representations/dynamic_container.cuh:
// 35-120): the slot stores get_smallKey(key), the key's low bits with 0 and
// all-ones remapped.  Over key_val_pair with a uint16_t small key it is the
// flat TCF's 4-byte slot (data_structs/tcf.cuh:29-30): one word
// val16 << 16 | key16; over key_container with uint64_t it is the 8-byte
// key-only slot above.
template <typename Rep, typename SmallKey>
struct dynamic_container {
    static_assert((std::is_same<Rep, key_val_pair>::value && std::is_same<SmallKey, uint16_t>::value) ||
                      (std::is_same<Rep, key_container>::value && std::is_same<SmallKey, uint64_t>::value),
                  "dynamic_container<key_val_pair, uint16_t> or dynamic_container<key_container, uint64_t>");
    static constexpr uint32_t rep = std::is_same<Rep, key_val_pair>::value ? PG_REP_KV16 : PG_REP_KEY64;
};

// ---- insert schemes (insert_schemes/*.cuh) ----
struct bucket_insert { static constexpr uint32_t scheme = PG_SCHEME_BUCKET; };
struct single_slot_insert { static constexpr uint32_t scheme = PG_SCHEME_SINGLE_SLOT; };
struct power_of_n_insert_shortcut_bucket_scheme {
    static constexpr uint32_t scheme = PG_SCHEME_POWER_OF_N_SHORTCUT;
};
struct linear_insert_bucket_scheme { static constexpr uint32_t scheme = PG_SCHEME_LINEAR_BUCKETS; };
// The flat power-of-n schemes over key-value slots (power_of_n.cuh:39-492,
// power_of_n_shortcut.cuh:39-505).
struct power_of_n_insert_scheme { static constexpr uint32_t scheme = PG_SCHEME_POWER_OF_N; };
struct power_of_n_insert_shortcut_scheme {
    static constexpr uint32_t scheme = PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS;
};
// power_of_n_insert_shortcut_replace_scheme
// (power_of_n_shortcut_key_replace.cuh:39-523): the shortcut scheme whose
// insert also replaces the value of a slot holding the key.
struct power_of_n_insert_shortcut_replace_scheme {
    static constexpr uint32_t scheme = PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
};
// blocked_bucket_insert<K> (insert_schemes/grouped_power_buckets.cuh:388-393):
// the shortcut scheme over K buckets per 128-byte line.
template <std::size_t Buckets_per_cache_line>
struct blocked_bucket_insert {
    static_assert(Buckets_per_cache_line >= 1, "at least one bucket per line");
    static constexpr uint32_t scheme = PG_SCHEME_GROUPED_SHORTCUT;
    static constexpr uint32_t group = (uint32_t)Buckets_per_cache_line;
};

// ---- probing schemes (probing_schemes/*.cuh) ----
struct linearProber { static constexpr uint32_t prober = PG_PROBE_LINEAR; };
struct doubleHasher { static constexpr uint32_t prober = PG_PROBE_DOUBLE; };
struct powerOfTwoHasher { static constexpr uint32_t prober = PG_PROBE_P2; };
struct XORPowerOfTwoHasher { static constexpr uint32_t prober = PG_PROBE_XOR_P2; };
struct split64Hasher { static constexpr uint32_t prober = PG_PROBE_SPLIT64; };  // no reference counterpart

// ---- hashers (hash_schemes/murmurhash.cuh) ----
struct murmurHasher { static constexpr uint32_t hasher = PG_HASH_MURMUR64A; };
struct fmix64Hasher { static constexpr uint32_t hasher = PG_HASH_FMIX64; };  // MurmurHash3 finalizer

// Recursion terminator (tables/recursive_end_table.cuh:50-172).
struct recursive_end_table {
    static constexpr uint32_t levels = 0;
};

// size_in_num_slots<N> (sizing/default_sizing.cuh:30-95): per-level slot
// counts, consumed level by level at construction.
template <std::size_t Num_Tables>
class size_in_num_slots {
   public:
    explicit size_in_num_slots(uint64_t start, ...) : i_(0) {
        slots_[0] = start;
        std::va_list args;
        va_start(args, start);
        for (std::size_t j = 1; j < Num_Tables; j++) slots_[j] = va_arg(args, uint64_t);
        va_end(args);
    }
    uint64_t next() {
        i_ += 1;
        return i_ <= Num_Tables ? slots_[i_ - 1] : end();
    }
    void reset() { i_ = 0; }
    static constexpr uint64_t end() { return ~uint64_t(0); }
    uint64_t total() const {
        uint64_t t = 0;
        for (std::size_t j = 0; j < Num_Tables; j++) t += slots_[j];
        return t;
    }

   private:
    uint64_t slots_[Num_Tables];
    std::size_t i_;
};

class pg_error : public std::runtime_error {
   public:
    pg_error(const std::string &what, int status)
        : std::runtime_error(what + ": " + pg_error_string(status)), status_(status) {}
    int status() const { return status_; }

   private:
    int status_;
};

inline void check(int rc, const char *what) {
    if (rc != PG_OK) throw pg_error(what, rc);
}

// Buckets per 128-byte line of a grouped scheme (0 for the others).
template <typename S, typename = void>
struct group_trait { static constexpr uint32_t value = 0; };
template <typename S>
struct group_trait<S, std::void_t<decltype(S::group)>> { static constexpr uint32_t value = S::group; };
template <typename S>
constexpr uint32_t group_of() { return group_trait<S>::value; }

template <typename Key, typename Val, typename Internal_Rep, std::size_t Partition_Size, std::size_t Bucket_Size,
          typename Insert_Scheme, std::size_t Max_Probes, typename Probing_Scheme, typename Hasher,
          bool Is_Recursive = false, typename Recursive_Type = recursive_end_table>
class bucketed_table {
   public:
    using key_type = Key;
    using val_type = Val;
    using my_type = bucketed_table<Key, Val, Internal_Rep, Partition_Size, Bucket_Size, Insert_Scheme, Max_Probes,
                                   Probing_Scheme, Hasher, Is_Recursive, Recursive_Type>;
    static constexpr uint32_t rep = Internal_Rep::rep;
    static constexpr uint32_t levels = 1 + (Is_Recursive ? Recursive_Type::levels : 0);

    static_assert(std::is_same<Key, uint64_t>::value, "keys are uint64_t");
    static_assert(rep != PG_REP_KV64 || std::is_same<Val, uint64_t>::value, "key_val_pair values are uint64_t");
    static_assert(rep != PG_REP_KEY64 || std::is_same<Val, uint16_t>::value,
                  "key-only tables take uint16_t value buffers (queries report 0)");
    static_assert(rep != PG_REP_FP12 || std::is_same<Val, uint16_t>::value, "12-bit bucket values are uint16_t");
    static_assert(rep != PG_REP_FP12 || !std::is_same<Probing_Scheme, XORPowerOfTwoHasher>::value,
                  "twelve_bucket::tag() returns nothing (12_bit_bucket.cuh:313): no XOR probing");
    static_assert(rep != PG_REP_FP16_V16 || std::is_same<Val, uint16_t>::value,
                  "16/16 filter values are uint16_t");
    static_assert(!(rep == PG_REP_FP8_V8 || rep == PG_REP_FP10_V6) || std::is_same<Val, uint8_t>::value,
                  "8/8 and 10/6 filter values are uint8_t");
    static_assert(Partition_Size >= 1 && Partition_Size <= 64 && (Partition_Size & (Partition_Size - 1)) == 0,
                  "Partition_Size is a power of two of at most one wavefront (64 lanes)");
    static_assert(Bucket_Size >= 1 && Bucket_Size <= 64, "Bucket_Size in 1..64");
    static_assert(Max_Probes >= 1, "at least one probe");
    static_assert(!(std::is_same<Probing_Scheme, powerOfTwoHasher>::value ||
                    std::is_same<Probing_Scheme, XORPowerOfTwoHasher>::value ||
                    std::is_same<Probing_Scheme, split64Hasher>::value) ||
                      Max_Probes == 2,
                  "two-choice probing schemes take exactly two probes (power_of_two.cuh:32)");
    static_assert(rep != PG_REP_KV16 || std::is_same<Val, uint16_t>::value,
                  "16/16 key_val_pair values are uint16_t");
    static constexpr bool flat_pow = Insert_Scheme::scheme == PG_SCHEME_POWER_OF_N ||
                                     Insert_Scheme::scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS ||
                                     Insert_Scheme::scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
    static_assert(!flat_pow || Max_Probes <= 4, "power-of-n schemes take at most 4 probes");
    static_assert((rep == PG_REP_KV64 || rep == PG_REP_KEY64 || rep == PG_REP_KV16)
                      ? (Insert_Scheme::scheme == PG_SCHEME_BUCKET || Insert_Scheme::scheme == PG_SCHEME_SINGLE_SLOT ||
                         flat_pow)
                      : (Insert_Scheme::scheme == PG_SCHEME_POWER_OF_N_SHORTCUT ||
                         Insert_Scheme::scheme == PG_SCHEME_LINEAR_BUCKETS ||
                         Insert_Scheme::scheme == PG_SCHEME_GROUPED_SHORTCUT),
                  "insert scheme does not apply to this representation");
    static_assert(!(Insert_Scheme::scheme == PG_SCHEME_SINGLE_SLOT) || Partition_Size == Bucket_Size,
                  "single_slot_insert needs Partition_Size == Bucket_Size (single_slot_insert.cuh:43)");
    static_assert(levels <= PG_MAX_LEVELS, "at most PG_MAX_LEVELS levels (a primary and three backing levels)");

    // Fill the descriptor levels from `first`, consuming the sizing.
    template <typename Sizing>
    static void describe(pg_desc &d, uint32_t first, Sizing *sizing) {
        const uint64_t nslots = sizing->next();
        if (nslots == sizing->end())
            throw std::runtime_error("Hash Table expects a num_slots from the sizing, but the size container is empty");
        pg_level_desc &l = d.level[first];
        l.nslots = nslots;
        l.scheme = Insert_Scheme::scheme;
        l.bucket_size = (uint32_t)Bucket_Size;
        l.tile = (uint32_t)Partition_Size;
        l.prober = Probing_Scheme::prober;
        l.hasher = Hasher::hasher;
        l.max_probes = (uint32_t)Max_Probes;
        l.group = group_of<Insert_Scheme>();
        l.reserved = 0;
        if constexpr (Is_Recursive) {
            static_assert(Recursive_Type::rep == rep, "levels share one representation");
            Recursive_Type::describe(d, first + 1, sizing);
        }
    }

    // bucketed_table::generate_on_device(Sizing*, seed)
    // (tables/bucketed_table.cuh:132-192): throws on a sizing mismatch.
    template <typename Sizing>
    static my_type *generate_on_device(Sizing *sizing, uint64_t seed) {
        pg_desc d{};
        d.rep = rep;
        d.num_levels = levels;
        describe(d, 0, sizing);
        if (sizing->next() != sizing->end())
            throw std::runtime_error("Hash Table Does not expect recursion but the size container is not empty.");
        pg_handle h = nullptr;
        check(pg_create(&d, seed, &h), "pg_create");
        return new my_type(h);
    }

    static void free_on_device(my_type *t) {
        if (!t) return;
        check(pg_destroy(t->h_), "pg_destroy");
        t->h_ = nullptr;
        delete t;
    }

    // Device arrays in, stream-ordered.
    void bulk_insert(const Key *d_keys, const Val *d_vals, uint64_t n, uint8_t *d_status = nullptr,
                     uint64_t *d_fail_count = nullptr, void *stream = nullptr) {
        check(pg_bulk_insert(h_, d_keys, d_vals, n, d_status, d_fail_count, stream), "bulk_insert");
    }
    void bulk_insert_with_delete(const Key *d_keys, const Val *d_vals, uint64_t n, uint8_t *d_status = nullptr,
                                 uint64_t *d_fail_count = nullptr, void *stream = nullptr) {
        check(pg_bulk_insert_with_delete(h_, d_keys, d_vals, n, d_status, d_fail_count, stream),
              "bulk_insert_with_delete");
    }
    void bulk_query(const Key *d_keys, uint64_t n, Val *d_vals_out, uint8_t *d_found, void *stream = nullptr) {
        check(pg_bulk_query(h_, d_keys, n, d_vals_out, d_found, stream), "bulk_query");
    }
    void bulk_remove(const Key *d_keys, uint64_t n, uint8_t *d_found = nullptr, void *stream = nullptr) {
        check(pg_bulk_remove(h_, d_keys, n, d_found, stream), "bulk_remove");
    }
    void clear(void *stream = nullptr) { check(pg_clear(h_, stream), "clear"); }
    void set_mode(uint32_t mode) { check(pg_set_mode(h_, mode), "set_mode"); }
    void set_items(uint32_t insert_items, uint32_t query_items) {
        check(pg_set_items(h_, insert_items, query_items), "set_items");
    }

    uint64_t get_fill(void *stream = nullptr) {
        uint64_t v = 0;
        check(pg_fill(h_, &v, stream), "get_fill");
        return v;
    }
    uint64_t get_num_slots() {
        uint64_t v = 0;
        check(pg_num_slots(h_, &v), "get_num_slots");
        return v;
    }
    uint64_t host_bytes_in_use() {
        uint64_t v = 0;
        check(pg_bytes(h_, &v), "host_bytes_in_use");
        return v;
    }
    pg_handle handle() const { return h_; }

   private:
    explicit bucketed_table(pg_handle h) : h_(h) {}
    pg_handle h_;
};

// The reference's static_table (tables/base_table.cuh:52) composes the same
// way; key_val_pair tables use it.
template <typename Key, typename Val, typename Internal_Rep, std::size_t Partition_Size, std::size_t Bucket_Size,
          typename Insert_Scheme, std::size_t Max_Probes, typename Probing_Scheme, typename Hasher,
          bool Is_Recursive = false, typename Recursive_Type = recursive_end_table>
using static_table = bucketed_table<Key, Val, Internal_Rep, Partition_Size, Bucket_Size, Insert_Scheme, Max_Probes,
                                    Probing_Scheme, Hasher, Is_Recursive, Recursive_Type>;

namespace data_structs {

// The reference's pre-baked TCF (data_structs/tcf.cuh:29-30): flat 16/16
// key_val_pair slots, P=4, B=16, power_of_n_insert_shortcut_scheme with two
// double-hashing probes, backed by tiny_static_table_4 (B=4,
// bucket_insert, 20 probes).  Size with size_in_num_slots<2>.
using tiny_static_table_4 = static_table<uint64_t, uint16_t, dynamic_container<key_val_pair, uint16_t>, 4, 4,
                                         bucket_insert, 20, doubleHasher, murmurHasher>;
using tcf = static_table<uint64_t, uint16_t, dynamic_container<key_val_pair, uint16_t>, 4, 16,
                         power_of_n_insert_shortcut_scheme, 2, doubleHasher, murmurHasher, true, tiny_static_table_4>;
// The same composition at tile 2, the fastest measured shape at 2^28 slots
// (each lane holds 8 slots of a bucket).
using fast_flat_tcf_backing = static_table<uint64_t, uint16_t, dynamic_container<key_val_pair, uint16_t>, 2, 4,
                                           bucket_insert, 20, doubleHasher, murmurHasher>;
using fast_flat_tcf = static_table<uint64_t, uint16_t, dynamic_container<key_val_pair, uint16_t>, 2, 16,
                                   power_of_n_insert_shortcut_scheme, 2, doubleHasher, murmurHasher, true,
                                   fast_flat_tcf_backing>;

// The TCF as the reference's delete tests compose it, del_TCF
// (tests/src/delete_tests.cu:138-152): 16/16 grouped slots, P=4, B=16,
// shortcut two-choice with XOR power-of-two probing, backed by a linear
// bucket table (B=8, 4000 probes).  Size with size_in_num_slots<2>.
using del_tcf_backing = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 4, 8,
                                       linear_insert_bucket_scheme, 4000, linearProber, murmurHasher>;
using del_tcf = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 4, 16,
                               power_of_n_insert_shortcut_bucket_scheme, 2, XORPowerOfTwoHasher, murmurHasher, true,
                               del_tcf_backing>;

// The reference's 8/8 and 10/6 compositions: del_TCF_small
// (tests/src/delete_tests.cu:165-179) and the TCF of
// tests/src/test_mhm_tcf.cu:105-116.
using small_tcf_backing = bucketed_table<uint64_t, uint8_t, bit_grouped_container<8, 8>, 4, 8,
                                         linear_insert_bucket_scheme, 400, linearProber, murmurHasher>;
using small_tcf = bucketed_table<uint64_t, uint8_t, bit_grouped_container<8, 8>, 4, 16,
                                 power_of_n_insert_shortcut_bucket_scheme, 2, XORPowerOfTwoHasher, murmurHasher,
                                 true, small_tcf_backing>;
using tcf_10_6_backing = bucketed_table<uint64_t, uint8_t, bit_grouped_container<10, 6>, 1, 8,
                                        linear_insert_bucket_scheme, 20, doubleHasher, murmurHasher>;
using tcf_10_6 = bucketed_table<uint64_t, uint8_t, bit_grouped_container<10, 6>, 1, 8,
                                power_of_n_insert_shortcut_bucket_scheme, 2, doubleHasher, murmurHasher, true,
                                tcf_10_6_backing>;

// The reference's 12-bit bucket benchmarks (tests/src/bucket_speed_tests.cu:
// 82-102): plain and cache-line-grouped (blocked_bucket_insert<K>).
using tcqf_twelve = bucketed_table<uint64_t, uint16_t, wrapper_half_bucket<uint16_t>, 4, 16,
                                   power_of_n_insert_shortcut_bucket_scheme, 2, doubleHasher, murmurHasher>;
using tcqf_twelve_16 = bucketed_table<uint64_t, uint16_t, wrapper_half_bucket<uint16_t>, 4, 16,
                                      blocked_bucket_insert<5>, 2, doubleHasher, murmurHasher>;
using tcqf_twelve_8 = bucketed_table<uint64_t, uint16_t, wrapper_half_bucket<uint16_t>, 2, 8,
                                     blocked_bucket_insert<10>, 2, doubleHasher, murmurHasher>;

// The MI355X default: two choices from one fmix64 hash split into 32-bit
// halves, tile 2 (the fastest measured shape at 2^28 slots).
using fast_tcf_backing = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 2, 8,
                                        linear_insert_bucket_scheme, 4000, linearProber, fmix64Hasher>;
using fast_tcf = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 2, 16,
                                power_of_n_insert_shortcut_bucket_scheme, 2, split64Hasher, fmix64Hasher, true,
                                fast_tcf_backing>;

// The CMS allocator's free table (allocators/cms.cuh:203): key-only,
// tile 1, 4-slot buckets, double hashing with 20 probes.
using cms_free_table = static_table<uint64_t, uint16_t, key_container, 1, 4, bucket_insert, 20, doubleHasher,
                                    murmurHasher>;

// hash_table_64 (data_structs/uint64_hash_table.cuh:35): u64 -> u64, tile 4,
// 16-slot buckets, double hashing with 20 probes.
using hash_table_64 = static_table<uint64_t, uint64_t, key_val_pair, 4, 16, bucket_insert, 20, doubleHasher,
                                   murmurHasher>;

}  // namespace data_structs
}  // namespace poggers_amd
