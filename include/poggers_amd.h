/*
 * poggers_amd.h -- C ABI of the MI355X (gfx950) bulk hash-table / Two-Choice
 * Filter library.
 *
 * The reference (huntermBerkeley/poggers) is a header-only CUDA template
 * library with no host bulk API: callers compose
 *     bucketed_table<Key, Val, Rep, P, B, InsertScheme, MaxProbes, Prober,
 *                    Hasher, IsRecursive, Backing>
 * (tables/bucketed_table.cuh:72) and write their own __global__ kernels
 * that call the per-tile device ops (tests/src/test_mhm_tcf.cu:318-487).
 * This ABI replaces that composition with a runtime descriptor and bulk,
 * stream-ordered entry points.  Each entry point cites the reference
 * interface it replaces.
 *
 * Conventions
 *   - Every function returns an int status (PG_OK == 0); no exception
 *     crosses the ABI.  pg_error_string() names a status.
 *   - Key, value and result buffers are caller-owned DEVICE pointers
 *     (the reference tests pass cudaMalloc'ed arrays the same way,
 *     tests/src/test_mhm_tcf.cu:616-621).  The handle owns table memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls are asynchronous and stream-ordered, except pg_create,
 *     pg_destroy and pg_fill, which synchronise.
 *   - A handle is not safe for concurrent host calls without external
 *     ordering.  Device-side ops within one call are lock-free per slot.
 *   - A bulk call takes any n: calls needing more than 2^31 work-items
 *     (n * tile) are issued as several launches, in key order, on the
 *     same stream.
 *   - Keys 0 and ~0 are reserved (empty / tombstone of the KV slot,
 *     representations/key_val_pair.cuh:70,115-117); filters remap their
 *     fingerprints so any key is accepted there.
 */
#ifndef POGGERS_AMD_H
#define POGGERS_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PG_MAX_LEVELS 4  /* a primary level and up to three backing levels */

/* status codes */
enum pg_status {
    PG_OK = 0,
    PG_ERR_INVALID = 1,      /* bad descriptor / argument */
    PG_ERR_UNSUPPORTED = 2,  /* op or shape not offered by this scheme */
    PG_ERR_NOMEM = 3,        /* device allocation failed */
    PG_ERR_HIP = 4,          /* a HIP runtime call failed */
    PG_ERR_NODEVICE = 5      /* no usable GPU */
};

/* Slot representation.
 *   KV64      key_val_pair<uint64_t,uint64_t>: 16 B {key, val}
 *             (representations/key_val_pair.cuh:63-171).  Schemes BUCKET,
 *             SINGLE_SLOT, POWER_OF_N, POWER_OF_N_SHORTCUT_SLOTS and
 *             POWER_OF_N_SHORTCUT_REPLACE.
 *   FP16_V16  dynamic_container<bit_grouped_container<16,16>, uint16_t>:
 *             one uint32 word = val16 << 16 | fp16
 *             (representations/dynamic_container.cuh:23-141,
 *              grouped_storage_sub_bits.cuh:99-260)
 *   FP8_V8    dynamic_container<bit_grouped_container<8,8>, uint8_t>:
 *             one uint16 word = val8 << 8 | fp8; fp8 = low byte of the key,
 *             0 -> 1, 0xFF -> 0xFE (tests/src/delete_tests.cu:165-179)
 *   FP10_V6   dynamic_container<bit_grouped_container<10,6>, uint16_t>:
 *             one uint16 word = val6 << 10 | (fp16 & 0x3FF); fp16 is
 *             remapped BEFORE the masking, as in the reference, so a key
 *             whose low 10 bits are 0 (or 0x3FF) stores as an empty-looking
 *             (or tombstone-looking) word (tests/src/test_mhm_tcf.cu:105-116).
 *             The probe tag (XOR_P2) is the unmasked fp16.
 *   FP12      twelve_bucket<uint64_t, uint16_t, uint16_t, P, B> through
 *             wrapper_half_bucket<uint16_t> (representations/12_bit_bucket.cuh:
 *             33-342): key-only, four slots in three uint16 words, slot 4g+o
 *             keeping its own 12 bits of the key's low 16 bits
 *             (representation_helpers.cuh:600-903); a bucket is
 *             2*((12B-1)/16+1) bytes.  Insert and query only: the
 *             reference's remove and insert_delete for this bucket do not
 *             compile, so remove, insert_with_delete and the upserts return
 *             PG_ERR_UNSUPPORTED; queries report value 0 (uint16 buffers).
 *   KEY64     dynamic_container<key_container, uint64_t>: one 8 B key per
 *             slot, the key-only table the CMS allocator keeps its free
 *             list in (allocators/cms.cuh:203; representations/key_only.cuh).
 *             The stored key is get_smallKey(key): 0 -> 1, ~0 -> ~0 - 1
 *             (dynamic_container.cuh:35-50), so every key is accepted and
 *             0/1 (and ~0/~0-1) share a slot; probes hash the key as
 *             given.  The schemes of KV64; queries
 *             report value 0 (uint16 buffers, the CMS table's Val).
 *   KV16      dynamic_container<key_val_pair<uint64_t, uint16_t>, uint16_t>
 *             (representations/dynamic_container.cuh:35-120 over
 *             key_val_pair.cuh:63-171): one uint32 word per slot =
 *             val16 << 16 | key16, key16 = the key's low 16 bits with
 *             0 -> 1 and 0xFFFF -> 0xFFFE; empty key 0, tombstone key 0xFFFF.
 *             A remove turns the key into the tombstone and keeps the value
 *             bits (key_val_pair::atomic_reset).  The slot of the
 *             reference's flat TCF (data_structs/tcf.cuh:29-30).  Schemes
 *             BUCKET, SINGLE_SLOT, POWER_OF_N, POWER_OF_N_SHORTCUT_SLOTS and
 *             POWER_OF_N_SHORTCUT_REPLACE, as KV64 and KEY64.
 *   Value buffers hold the reference's Val type: uint64 (KV64), uint16
 *   (FP16_V16, FP12, KEY64, KV16), uint8 (FP8_V8 and FP10_V6, which keep the low
 *   8 / 6 bits).  The 8/8, 10/6 and 12-bit formats and grouped levels run the
 *   generic (runtime bucket size) kernels; batches take FP16_V16 tables, the
 *   server KV64, FP16_V16 and KV16 tables.                                             */
enum pg_rep {
    PG_REP_KV64 = 0,
    PG_REP_FP16_V16 = 1,
    PG_REP_FP8_V8 = 2,
    PG_REP_FP10_V6 = 3,
    PG_REP_FP12 = 4,
    PG_REP_KEY64 = 5,
    PG_REP_KV16 = 6
};

/* Insert scheme of a level.
 *   BUCKET        insert_schemes/bucket_insert.cuh:38-361
 *   SINGLE_SLOT   insert_schemes/single_slot_insert.cuh:38-233 (no remove;
 *                 bucket_size must equal tile, as its static_assert at :43)
 *   POWER_OF_N_SHORTCUT  insert_schemes/power_of_n_shortcut_buckets.cuh:35-783
 *                 (the TCF primary: two choices + 75 % shortcut)
 *   LINEAR_BUCKETS insert_schemes/linear_insert_buckets.cuh:35-445
 *                 (the TCF backing table)
 *   GROUPED_SHORTCUT  blocked_bucket_insert<K> =
 *                 power_of_n_insert_grouped_buckets
 *                 (insert_schemes/grouped_power_buckets.cuh:46-388): the
 *                 shortcut scheme over buckets packed `group` = K to a
 *                 128-byte line (bucket b at byte 128*(b/K) + size*(b%K)),
 *                 packed filter slots only
 *   POWER_OF_N    insert_schemes/power_of_n.cuh:39-492: each key goes to
 *                 the least-filled of its max_probes (at most 4) probe
 *                 buckets, ties in probe order; a bucket's fill counts every
 *                 non-empty slot, tombstones included; inserts claim empty
 *                 slots only; queries and removes check every probe bucket.
 *                 Key-value slots (KV64, KEY64, KV16) only.
 *   POWER_OF_N_SHORTCUT_SLOTS  insert_schemes/power_of_n_shortcut.cuh:
 *                 39-505: POWER_OF_N with the 75 % shortcut (a probe bucket
 *                 whose fill is below 3B/4 takes the key at once), over
 *                 key-value slots (the reference's flat TCF primary)
 *   POWER_OF_N_SHORTCUT_REPLACE  power_of_n_insert_shortcut_replace_scheme
 *                 (insert_schemes/power_of_n_shortcut_key_replace.cuh:39-523):
 *                 POWER_OF_N_SHORTCUT_SLOTS whose insert also takes a slot
 *                 that already holds the key and replaces its value: a probe
 *                 bucket holding the key is tried at once whatever its fill
 *                 (check_fill_bucket's Bucket_Size + 1, :233-270), and a claim
 *                 takes the lowest slot that is empty or holds the key
 *                 (insert_into_bucket, :134-190).  Probe buckets are visited
 *                 in order, so a key held in a later probe bucket is stored
 *                 again in an earlier one with room, as in the reference;
 *                 queries find the earlier copy.  The reference header does
 *                 not compile (a brace of insert() is missing after :366,
 *                 found_ballot starts true at :235, atomic_swap only swaps
 *                 from the empty key); this is the scheme its code describes
 *                 with those three slips fixed.  Key-value slots only.   */
enum pg_scheme {
    PG_SCHEME_BUCKET = 0,
    PG_SCHEME_SINGLE_SLOT = 1,
    PG_SCHEME_POWER_OF_N_SHORTCUT = 2,
    PG_SCHEME_LINEAR_BUCKETS = 3,
    PG_SCHEME_GROUPED_SHORTCUT = 4,
    PG_SCHEME_POWER_OF_N = 5,
    PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS = 6,
    PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE = 7
};

/* Probing scheme (the probing_schemes directory).  Probe i of a key, before the
 * consumer's "% num_buckets":
 *   LINEAR   h + i                       (linear_probing.cuh:72-78)
 *   DOUBLE   h1 + i*h2, h2 seeded seed^2 (double_hashing.cuh:66-89)
 *   P2       {h1, h2}                    (power_of_two.cuh:68-91)
 *   XOR_P2   {h1, h1 ^ (tag*0x5bd1e995)} (xor_power_of_two.cuh:64-84)
 *   SPLIT64  {lo32(h), hi32(h)}: one 64-bit hash split into halves
 *            (no reference counterpart); its probes are 32-bit, so levels
 *            of more than 2^32 buckets are refused (PG_ERR_UNSUPPORTED)  */
enum pg_prober {
    PG_PROBE_LINEAR = 0,
    PG_PROBE_DOUBLE = 1,
    PG_PROBE_P2 = 2,
    PG_PROBE_XOR_P2 = 3,
    PG_PROBE_SPLIT64 = 4
};

/* Hasher: FMIX64 = MurmurHash3 fmix64(key ^ seed);
 *         MURMUR64A = MurmurHash64A(&key, 8, seed)
 *         (hash_schemes/murmurhash.cuh:20-100)                          */
enum pg_hasher { PG_HASH_FMIX64 = 0, PG_HASH_MURMUR64A = 1 };

/* Execution mode (pg_set_mode). */
enum pg_mode {
    PG_CONCURRENT = 0,    /* one tile per key, all keys in flight */
    PG_DETERMINISTIC = 1  /* keys applied in array order by one tile: the
                             slot layout equals the sequential oracle's */
};

/* One level of a (possibly recursive) table.  Mirrors one instantiation of
 * the reference's Insert_Scheme<Key, Val, Partition_Size, Bucket_Size,
 * Internal_Rep, Max_Probes, Hasher, Probing_Scheme>.                     */
typedef struct pg_level_desc {
    uint64_t nslots;       /* size_in_num_slots entry (sizing/default_sizing.cuh:30-95) */
    uint32_t scheme;       /* enum pg_scheme */
    uint32_t bucket_size;  /* Bucket_Size (slots per bucket) */
    uint32_t tile;         /* Partition_Size: lanes per key, 1..64, power of two */
    uint32_t prober;       /* enum pg_prober */
    uint32_t hasher;       /* enum pg_hasher */
    uint32_t max_probes;   /* Max_Probes */
    uint32_t group;        /* GROUPED_SHORTCUT: buckets per 128-byte line (K of
                              blocked_bucket_insert<K>, K * bucket bytes <= 128);
                              0 for every other scheme */
    uint32_t reserved;     /* 0 */
} pg_level_desc;

/* A table: level 0 plus up to three backing levels (Is_Recursive /
 * Recursive_Type, tables/bucketed_table.cuh:72,166-176).  Each backing level
 * is seeded with the square of the level above's seed
 * (bucketed_table.cuh:170).                                               */
typedef struct pg_desc {
    uint32_t rep;          /* enum pg_rep */
    uint32_t num_levels;   /* 1 .. PG_MAX_LEVELS */
    pg_level_desc level[PG_MAX_LEVELS];
} pg_desc;

typedef struct pg_table_s *pg_handle;

/* Replaces bucketed_table::generate_on_device(Sizing*, seed)
 * (tables/bucketed_table.cuh:132-192) and static_table::generate_on_device
 * (tables/base_table.cuh:110-145).  Allocates and zeroes the slot arrays. */
int pg_create(const pg_desc *desc, uint64_t seed, pg_handle *out);

/* Replaces free_on_device (tables/bucketed_table.cuh:194-224). */
int pg_destroy(pg_handle h);

/* Zero every slot (a fill kernel, safe to record in a hipGraph). */
int pg_clear(pg_handle h, void *stream);

/* Select PG_CONCURRENT or PG_DETERMINISTIC for later mutating calls.
 * The reference's only ordered mode is a single-tile launch
 * (tests/src/test_mhm_tcf.cu:265-314,848). */
int pg_set_mode(pg_handle h, uint32_t mode);

/* Bulk insert: replaces the per-key device op insert(tile, key, val)
 * (tables/bucketed_table.cuh:271-291) driven by speed_insert_kernel
 * (tests/src/test_mhm_tcf.cu:318-345).
 *   d_vals       one Val per key: uint64 (KV64), uint16 (FP16_V16) or
 *                uint8 (FP8_V8, FP10_V6);
 *                NULL inserts value 0 (bucketed_table.cuh:324-329).
 *   d_status     optional, one byte per key: 0 failed, 1 + level index.
 *   d_fail_count optional uint64 device counter, incremented per failure
 *                (the reference's `misses` counter). */
int pg_bulk_insert(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                   uint8_t *d_status, uint64_t *d_fail_count, void *stream);

/* Bulk insert that may reuse tombstones: replaces insert_with_delete
 * (tables/bucketed_table.cuh:293-321; power_of_n_shortcut_buckets.cuh:
 * 309-417; linear_insert_buckets.cuh:223-244).  Filters only. */
int pg_bulk_insert_with_delete(pg_handle h, const uint64_t *d_keys, const void *d_vals,
                               uint64_t n, uint8_t *d_status, uint64_t *d_fail_count,
                               void *stream);

/* Bulk upsert: replaces insert_if_not_exists(tile, key, val, Val& ext,
 * bool& found) (tables/bucketed_table.cuh:226-246;
 * power_of_n_shortcut_buckets.cuh:421-544; linear_insert_buckets.cuh:249-282).
 * A key already present is left in place: d_found = 1 and d_vals_out gets
 * its stored value; otherwise it is inserted.  d_status as for insert (0 =
 * the op returned false).  Filters only. */
int pg_bulk_insert_if_not_exists(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                                 void *d_vals_out, uint8_t *d_found, uint8_t *d_status,
                                 uint64_t *d_fail_count, void *stream);

/* Toggle upsert: replaces insert_if_not_exists_delete
 * (bucketed_table.cuh:248-268; power_of_n_shortcut_buckets.cuh:546-674;
 * linear_insert_buckets.cuh:285-321).  A key already present is removed
 * (d_found = 1, d_vals_out = its value); otherwise it is inserted, reusing
 * tombstones in the shortcut level.  Filters only. */
int pg_bulk_insert_if_not_exists_delete(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                                        void *d_vals_out, uint8_t *d_found, uint8_t *d_status,
                                        uint64_t *d_fail_count, void *stream);

/* Bulk query: replaces query(tile, key, Val&) (bucketed_table.cuh:332-349)
 * driven by speed_query_kernel (tests/src/test_mhm_tcf.cu:456-487).
 *   d_vals_out optional; value per key (0 when absent).
 *   d_found    one byte per key. */
int pg_bulk_query(pg_handle h, const uint64_t *d_keys, uint64_t n, void *d_vals_out,
                  uint8_t *d_found, void *stream);

/* Bulk remove: replaces remove(tile, key) (bucketed_table.cuh:351-381)
 * driven by speed_delete_kernel (tests/src/delete_tests.cu:335-374).
 *   d_found optional, one byte per key: 1 if a slot was tombstoned. */
int pg_bulk_remove(pg_handle h, const uint64_t *d_keys, uint64_t n, uint8_t *d_found,
                   void *stream);

/* Filled slots over all levels: replaces get_fill() and count_fill_kernel
 * (tables/bucketed_table.cuh:36-52,498-548).  Synchronises `stream`;
 * *out is a HOST pointer. */
int pg_fill(pg_handle h, uint64_t *out, void *stream);

/* pg_fill without any allocation: no counterpart in the reference.  For
 * callers that cannot allocate device memory (tight serving loops that
 * report fill between batches, harnesses that forbid hipMalloc).  Counts into
 * a 32-byte file-scope device array (one 8-byte counter per level, for the
 * PG_MAX_LEVELS levels a table can have; a handle claiming more is refused
 * with PG_ERR_UNSUPPORTED) instead of the table's own counter, calls no
 * hipMalloc / hipFree, and returns the same total as pg_fill.  The table's device must be current (else
 * PG_ERR_INVALID).  All calls on one device share that device's array, so
 * calls that overlap on one device race: order them (one stream, or
 * external synchronisation).  Synchronises `stream`; *out is a HOST pointer. */
int pg_fill_noalloc(pg_handle h, uint64_t *out, void *stream);

/* Bytes of slot memory over all levels: replaces host_bytes_in_use
 * (tables/bucketed_table.cuh:405-425). */
int pg_bytes(pg_handle h, uint64_t *out);

/* Slots over all levels: replaces get_num_slots (bucketed_table.cuh:462-489). */
int pg_num_slots(pg_handle h, uint64_t *out);

/* Device pointer and byte size of one level's slot array (raw bucket dump
 * for layout checks).  No reference counterpart: the reference never
 * exposes its slot arrays outside the device structs.  After writing slots
 * through the pointer, call pg_level_sync. */
int pg_level_ptr(pg_handle h, uint32_t level, void **d_ptr, uint64_t *bytes);

/* Rebuild what a level derives from its slots, after they were written
 * through pg_level_ptr (pg_load does this itself).  A linear-bucket backing
 * level (level >= 1) of 16/16 or 8/8 filter slots keeps an occupancy index:
 * one bit per bucket, set before any slot of the bucket is claimed and
 * cleared only by pg_clear, so a clear bit proves the bucket empty and its
 * queries and removes stop there without reading it (the reference's linear
 * walk stops at the first bucket that is not full,
 * linear_insert_buckets.cuh:323-398).
 * The rebuild sets the bit of every bucket holding a non-empty word.  PG_OK
 * and nothing to do for other levels.  Ordered on `stream`.  No reference
 * counterpart. */
int pg_level_sync(pg_handle h, uint32_t level, void *stream);

/* ---- batches: many independent filters, one persistent launch ----
 * No reference counterpart: the reference launches one kernel per table
 * (tests/src/test_mhm_tcf.cu:318-487).  A batch holds `ntables` filters of
 * uneven sizes sharing the prototype's shape (levels, bucket size, tile,
 * probing, hashing; FP16_V16 with a POWER_OF_N_SHORTCUT level 0).  Keys of all
 * tables come in one flat device array with CSR device offsets: table t owns
 * keys[offsets[t] .. offsets[t+1]), offsets[ntables] == total. */
typedef struct pg_batch_s *pg_batch;

/* nslots[t] = level-0 slots of table t (host array); backing_slots[t] the
 * backing level's (host array, NULL = 1 % of level 0, at least one bucket). */
int pg_batch_create(const pg_desc *proto, uint32_t ntables, const uint64_t *nslots,
                    const uint64_t *backing_slots, uint64_t seed, pg_batch *out);
int pg_batch_destroy(pg_batch b);
/* PG_DETERMINISTIC makes pg_batch_insert apply each table's keys in order
 * (one tile per table, tables in parallel): layouts equal the oracle's. */
int pg_batch_set_mode(pg_batch b, uint32_t mode);
/* Persistent launch shape: keys per work chunk and blocks (0 = fill the GPU). */
int pg_batch_set_launch(pg_batch b, uint32_t chunk, uint32_t blocks);
/* Keys each tile has in flight in the persistent kernels (1, 2 or 4; 0 = 1):
 * the first-choice bucket loads of that many keys are issued before any is
 * examined.  The reference has one key per tile (test_mhm_tcf.cu:318-487);
 * like the tile width, a launch-shape choice that leaves results unchanged. */
int pg_batch_set_items(pg_batch b, uint32_t items);
int pg_batch_clear(pg_batch b, void *stream);
int pg_batch_insert(pg_batch b, const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total,
                    const uint16_t *d_vals, uint8_t *d_status, uint64_t *d_fail_count, void *stream);
int pg_batch_query(pg_batch b, const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total,
                   uint16_t *d_vals_out, uint8_t *d_found, void *stream);
/* Record [clear] + insert + query into a hipGraph held by the batch (kernels
 * only, no memset nodes); pg_batch_replay launches it.  The graph holds the
 * buffers' addresses: they must stay allocated while it can be replayed. */
int pg_batch_capture(pg_batch b, const uint64_t *d_ins_keys, const uint64_t *d_ins_offsets, uint64_t ins_total,
                     const uint64_t *d_q_keys, const uint64_t *d_q_offsets, uint64_t q_total, uint8_t *d_status,
                     uint8_t *d_found, uint64_t *d_fail_count, uint32_t clear_first, void *stream);
int pg_batch_replay(pg_batch b, void *stream);
/* Filled slots of one table (synchronises; *out is a host pointer). */
int pg_batch_fill(pg_batch b, uint32_t table, uint64_t *out, void *stream);
/* Device pointer and byte size of one table's level, for reading layouts.
 * The backing levels keep the occupancy index described at pg_level_sync;
 * a batch has no sync call, so slots written through this pointer would
 * leave the index stale.  Write a batch's tables only through its ops. */
int pg_batch_level_ptr(pg_batch b, uint32_t table, uint32_t level, void **d_ptr, uint64_t *bytes);
/* Bytes of slot memory over every table and level. */
int pg_batch_bytes(pg_batch b, uint64_t *out);

/* ---- serving: continuous batching over resident tables ----
 * No reference counterpart (the reference has no host request path).
 * Requests carry host or device buffers.  Buffers the GPU can address
 * (pinned with hipHostMalloc, or device memory) are used in place: the
 * kernel reads keys and writes results through them, with no copy calls and
 * no staging bytes.  Pageable buffers are staged through the device budget
 * with copies in and out.  pg_server_progress never blocks: it reports
 * finished requests, admits queued requests in FIFO order while their
 * staging bytes fit the budget and a stream is free, and launches each
 * admitted group as one batch (one copy of the request table, one kernel
 * with a tile per key over all its requests, plus the staged requests'
 * copies).  Up to `nstreams` batches overlap on the device.  Every served table
 * must use the same tile width (at most 16) and hold KV64, FP16_V16 or KV16
 * slots (PG_ERR_UNSUPPORTED otherwise).  Requests that touch the same
 * keys of one table from different in-flight batches are not ordered with
 * each other (as concurrent bulk calls on two streams would not be). */
enum pg_op { PG_OP_INSERT = 0, PG_OP_QUERY = 1, PG_OP_REMOVE = 2 };
/* Request ordering (pg_server_set_order).
 *   PG_ORDER_NONE   requests in flight together run concurrently, in one
 *                   batch or in overlapping batches (the default).
 *   PG_ORDER_TABLE  on each table, a request that writes (insert, remove)
 *                   takes effect after every request submitted before it on
 *                   that table, and before every request submitted after
 *                   it; queries of a table between two writes still run
 *                   together.  Batches on other tables still overlap. */
enum pg_order { PG_ORDER_NONE = 0, PG_ORDER_TABLE = 1 };
typedef struct pg_server_s *pg_server;

int pg_server_create(const pg_handle *tables, uint32_t ntables, uint64_t budget_bytes, uint32_t nstreams,
                     pg_server *out);
int pg_server_destroy(pg_server s);
/* Queue a request; PG_ERR_NOMEM when it could never fit the budget.
 * h_vals: insert values (uint64 KV / uint16 filter) or NULL; h_vals_out:
 * query values or NULL; h_flags: per-key status / found or NULL.  In-place
 * buffers must stay valid and untouched until the request is reported. */
int pg_server_submit(pg_server s, uint32_t table, uint32_t op, const uint64_t *h_keys, const void *h_vals,
                     uint64_t n, void *h_vals_out, uint8_t *h_flags, uint64_t *ticket);
/* One request of pg_server_submit_many (the arguments of pg_server_submit). */
typedef struct pg_request {
    uint32_t table;
    uint32_t op;            /* enum pg_op */
    const uint64_t *keys;
    const void *vals;
    uint64_t n;
    void *vals_out;
    uint8_t *flags;
} pg_request;
/* Queue `count` requests in order, as that many pg_server_submit calls:
 * tickets[i] (if not NULL) receives request i's ticket.  On an invalid
 * request it stops and returns that request's error; *nqueued tells how
 * many were queued before it. */
int pg_server_submit_many(pg_server s, const pg_request *reqs, uint32_t count, uint64_t *tickets,
                          uint32_t *nqueued);
/* Engine step: writes up to `cap` finished tickets to done[]; finished
 * tickets that do not fit are written by later calls. */
int pg_server_progress(pg_server s, uint64_t *done, uint32_t cap, uint32_t *ndone);
int pg_server_pending(pg_server s, uint64_t *queued, uint64_t *inflight);
/* Batch formation: at most max_batch_keys keys (a larger single request
 * still forms its own batch) and max_batch_requests (1..4096) requests per
 * batch. */
int pg_server_set_batch(pg_server s, uint64_t max_batch_keys, uint32_t max_batch_requests);
/* Ordering of requests (enum pg_order).  Under PG_ORDER_TABLE a request
 * that conflicts with one already in the forming batch on the same table
 * waits for a later batch (other tables' requests may still join), and a
 * batch waits on the device for earlier in-flight batches it conflicts
 * with; batches launched before the switch count as conflicting. */
int pg_server_set_order(pg_server s, uint32_t order);
/* One request alone: copies in, the table's bulk kernel, copies out, wait.
 * The per-request baseline; only when nothing is queued or in flight. */
int pg_server_run_plain(pg_server s, uint32_t table, uint32_t op, const uint64_t *h_keys, const void *h_vals,
                        uint64_t n, void *h_vals_out, uint8_t *h_flags);

/* ---- raw bucket-array files (layout checks, checkpoint / resume) ----
 * No reference counterpart: the reference never serializes a table
 * (tables/bucketed_table.cuh:194-224 only frees).  Format, little endian:
 *   "PGAMDTB1" | u32 version (1) | u32 header bytes | pg_desc | u64 seed |
 *   u64 bytes of level 0 | ... | raw slot array of level 0 | ...
 * The slot arrays are the device layout byte for byte (see pg_level_ptr). */
int pg_save(pg_handle h, const char *path);        /* synchronises */
int pg_load(const char *path, pg_handle *out);     /* pg_create + copy in */
/* Header only, no device work: descriptor and seed of a saved table. */
int pg_file_info(const char *path, pg_desc *desc, uint64_t *seed);

/* Human-readable name of a status code. */
const char *pg_error_string(int status);

/* ---- measurement helpers (no reference counterpart) ---- */

/* Fill d_out[i] = splitmix64 key i+start of stream `seed` (the seeded key
 * streams every benchmark input is drawn from). */
int pg_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *d_out, void *stream);

/* Random 64 B-sector microbenchmark (the measured random-access roofline):
 * n_reads independent 64 B reads at hashed sector indices of d_buf (bytes, a
 * multiple of 64).  `op` bits 0-1: 0 read only, 1 add a CAS on one word of
 * each sector (the insert pattern: read, then dirty the line), 2 add a plain
 * store of that word instead; bits 4-7 = v, 1 << (v - 1) sectors in flight
 * per tile for v in 1..4 (0 -> 4); bit 8 = nontemporal loads; bits 9-10 =
 * tile shape (0: 4 lanes x 16 B, 1: 1 lane x 64 B, 2: 16 lanes x 4 B); bits
 * 11-14 = w, 0 for a grid covering the work, else w blocks of 256 threads
 * per CU walking it grid-stride (caps the requests in flight).  Writes a
 * checksum word to d_sink so nothing is optimised away. */
int pg_random_sector_bench(void *d_buf, uint64_t bytes, uint64_t n_reads, uint32_t op,
                           uint64_t seed, uint64_t *d_sink, void *stream);

/* Accounting: while d_stats (3 device uint64 counters) is set, bulk ops add
 * [0] level-0 buckets examined (KV: slots examined), [1] backing-level
 * buckets examined, [2] slots written.  These give the algorithmic bytes
 * of a call (each bucket examination reads B slot words; each write dirties
 * one 64 B sector).  NULL disables it.  No reference counterpart. */
int pg_set_stats(pg_handle h, uint64_t *d_stats);

/* Launch shape: keys handled per tile by concurrent bulk inserts and queries
 * (1, 2 or 4; 0 = library default).  A tile issues the first-choice bucket
 * loads of all its keys before examining any, trading registers for memory
 * requests in flight.  Results do not depend on it.  The reference launches
 * exactly one key per tile (tables/bucketed_table.cuh:395-403); this is the
 * knob its manual tile/bucket sweeps tuned by recompiling. */
int pg_set_items(pg_handle h, uint32_t insert_items, uint32_t query_items);

/* Host-side range reduction used by every probe (exported for tests):
 * returns n % d computed with the library's invariant-divisor method. */
uint64_t pg_fastmod_host(uint64_t n, uint64_t d);

#ifdef __cplusplus
}
#endif

#endif /* POGGERS_AMD_H */
