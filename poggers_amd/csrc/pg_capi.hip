// pg_capi.hip -- host side of the C ABI (include/poggers_amd.h).
#include <mutex>
#include <new>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "pg_internal.hpp"

struct pg_table_s {
    pg_desc desc;
    pg::TableDev dev;
    uint64_t seed;
    uint32_t mode;
    int device;
    uint64_t bytes[PG_MAX_LEVELS];
    uint64_t *d_counter;  // 8 B device scratch for pg_fill
    uint64_t *d_stats;    // caller's accounting counters, or NULL
    uint32_t items_insert, items_query;  // keys per tile (pg_set_items)
};

namespace {

bool pow2_tile(uint32_t p) { return p >= 1 && p <= 64 && (p & (p - 1)) == 0; }

// Validate one level; returns PG_OK or an error.
int check_level(uint32_t rep, const pg_level_desc &l, uint32_t index) {
    if (l.nslots == 0) return PG_ERR_INVALID;
    if (!pow2_tile(l.tile)) return PG_ERR_INVALID;
    if (l.prober > PG_PROBE_SPLIT64 || l.hasher > PG_HASH_MURMUR64A) return PG_ERR_INVALID;
    if (l.max_probes == 0) return PG_ERR_INVALID;
    // SPLIT64 probes are the two 32-bit halves of one hash: a level of more
    // than 2^32 buckets could not reach its upper buckets.
    if (l.prober == PG_PROBE_SPLIT64 && l.bucket_size &&
        (l.nslots - 1) / (l.scheme == PG_SCHEME_SINGLE_SLOT ? 1u : l.bucket_size) + 1 > (1ull << 32))
        return PG_ERR_UNSUPPORTED;
    // power_of_two.cuh:32, xor_power_of_two.cuh:32: exactly two probes.
    if ((l.prober == PG_PROBE_P2 || l.prober == PG_PROBE_XOR_P2 || l.prober == PG_PROBE_SPLIT64) &&
        l.max_probes != 2)
        return PG_ERR_INVALID;
    if (pg::is_kv_rep(rep)) {
        const bool pow = l.scheme == PG_SCHEME_POWER_OF_N || l.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS ||
                         l.scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE;
        if (l.scheme != PG_SCHEME_BUCKET && l.scheme != PG_SCHEME_SINGLE_SLOT && !pow) return PG_ERR_UNSUPPORTED;
        if (l.scheme != PG_SCHEME_SINGLE_SLOT && (l.bucket_size == 0 || l.bucket_size > 64)) return PG_ERR_INVALID;
        // single_slot_insert static_asserts Partition_Size == Bucket_Size
        // (single_slot_insert.cuh:43); its slots are probed one by one
        if (l.scheme == PG_SCHEME_SINGLE_SLOT && l.bucket_size != l.tile) return PG_ERR_INVALID;
        // the per-probe fills live in registers (power_of_n.cuh:336-337
        // keeps Max_Probes of them)
        if (pow && l.max_probes > pg::kMaxShortcutProbes) return PG_ERR_UNSUPPORTED;
    } else if (rep <= PG_REP_FP12) {  // the packed filter slots
        if (l.scheme != PG_SCHEME_POWER_OF_N_SHORTCUT && l.scheme != PG_SCHEME_LINEAR_BUCKETS &&
            l.scheme != PG_SCHEME_GROUPED_SHORTCUT)
            return PG_ERR_UNSUPPORTED;
        if (l.bucket_size == 0 || l.bucket_size > 64) return PG_ERR_INVALID;
        if (l.scheme != PG_SCHEME_LINEAR_BUCKETS && l.max_probes > pg::kMaxShortcutProbes)
            return PG_ERR_UNSUPPORTED;
        // twelve_bucket::tag() returns nothing (12_bit_bucket.cuh:313-315), so
        // a tag-driven probe is undefined in the reference: not offered.
        if (rep == PG_REP_FP12 && l.prober == PG_PROBE_XOR_P2) return PG_ERR_UNSUPPORTED;
    } else {
        return PG_ERR_INVALID;
    }
    // blocked_bucket_insert<K>: K buckets in one 128-byte line
    // (grouped_power_buckets.cuh:58).
    if ((l.scheme == PG_SCHEME_GROUPED_SHORTCUT) != (l.group != 0)) return PG_ERR_INVALID;
    if (l.group && l.group * pg::bucket_bytes(rep, l.bucket_size) > 128) return PG_ERR_INVALID;
    if (l.reserved != 0) return PG_ERR_INVALID;
    (void)index;
    return PG_OK;
}

// Bytes of one level's slot array: nb buckets, or ceil(nb / group) 128-byte
// lines for a grouped level.
uint64_t level_bytes(uint32_t rep, const pg_level_desc &l, uint32_t B, uint64_t nb) {
    if (l.group) return ((nb - 1) / l.group + 1) * 128;
    return nb * pg::bucket_bytes(rep, B);
}

}  // namespace

namespace pg {
const TableDev *table_dev(pg_handle h, uint32_t *tile) {
    if (!h) return nullptr;
    if (tile) *tile = h->desc.level[0].tile;
    return &h->dev;
}
}  // namespace pg

extern "C" {

const char *pg_error_string(int status) {
    switch (status) {
        case PG_OK: return "ok";
        case PG_ERR_INVALID: return "invalid argument";
        case PG_ERR_UNSUPPORTED: return "unsupported by this scheme";
        case PG_ERR_NOMEM: return "device allocation failed";
        case PG_ERR_HIP: return "HIP runtime error";
        case PG_ERR_NODEVICE: return "no usable GPU";
        default: return "unknown status";
    }
}

uint64_t pg_fastmod_host(uint64_t n, uint64_t d) {
    if (d == 0) return 0;
    return pg::fastmod(n, pg::make_fastmod(d));
}

int pg_create(const pg_desc *desc, uint64_t seed, pg_handle *out) {
    if (!desc || !out) return PG_ERR_INVALID;
    *out = nullptr;
    if (desc->num_levels < 1 || desc->num_levels > PG_MAX_LEVELS) return PG_ERR_INVALID;
    // The tile is shared by all levels: the reference passes the level-0 tile
    // to the backing table's ops (bucketed_table.cuh:280-282).
    for (uint32_t i = 0; i < desc->num_levels; i++) {
        const int rc = check_level(desc->rep, desc->level[i], i);
        if (rc != PG_OK) return rc;
        if (desc->level[i].tile != desc->level[0].tile) return PG_ERR_INVALID;
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PG_ERR_NODEVICE;

    pg_table_s *t = new (std::nothrow) pg_table_s();
    if (!t) return PG_ERR_NOMEM;
    t->desc = *desc;
    t->seed = seed;
    t->mode = PG_CONCURRENT;
    if (hipGetDevice(&t->device) != hipSuccess) {
        delete t;
        return PG_ERR_HIP;
    }
    t->dev.nlevels = desc->num_levels;
    t->dev.rep = desc->rep;
    uint64_t s = seed;
    int rc = PG_OK;
    for (uint32_t i = 0; i < desc->num_levels && rc == PG_OK; i++) {
        const pg_level_desc &l = desc->level[i];
        pg::LevelDev &L = t->dev.lv[i];
        // Buckets = ceil(nslots / B) (power_of_n_shortcut_buckets.cuh:94,
        // bucket_insert.cuh:90-92); single slot allocates nslots slots
        // (single_slot_insert.cuh:89-96).
        const uint32_t B = (l.scheme == PG_SCHEME_SINGLE_SLOT) ? 1u : l.bucket_size;
        const uint64_t nb = (l.nslots - 1) / B + 1;
        L.nb = pg::make_fastmod(nb);
        L.B = B;
        L.seed = s;
        L.seed2 = s * s;
        L.scheme = l.scheme;
        L.prober = l.prober;
        L.hasher = l.hasher;
        L.max_probes = l.max_probes;
        L.group = l.group;
        L.bucket_bytes = (uint32_t)pg::bucket_bytes(desc->rep, B);
        s = s * s;  // recursive level seeded seed*seed (bucketed_table.cuh:170)
        t->bytes[i] = level_bytes(desc->rep, l, B, nb);
        // padded to 256 B: 16-bit slots are swapped through their 32-bit word
        const uint64_t bytes = (t->bytes[i] + 255) & ~uint64_t(255);
        if (hipMalloc(&L.slots, bytes) != hipSuccess) {
            L.slots = nullptr;
            rc = PG_ERR_NOMEM;
            break;
        }
        rc = pg::launch_zero(L.slots, bytes, nullptr);
        if (rc == PG_OK && pg::has_occ_index(desc->rep, l.scheme, i)) {
            if (hipMalloc((void **)&L.occ, pg::occ_bytes(nb)) != hipSuccess) {
                L.occ = nullptr;
                rc = PG_ERR_NOMEM;
                break;
            }
            rc = pg::launch_zero(L.occ, pg::occ_bytes(nb), nullptr);
        }
    }
    if (rc == PG_OK && hipMalloc((void **)&t->d_counter, 256) != hipSuccess) rc = PG_ERR_NOMEM;
    if (rc == PG_OK && hipDeviceSynchronize() != hipSuccess) rc = PG_ERR_HIP;
    if (rc != PG_OK) {
        pg_destroy(t);
        return rc;
    }
    *out = t;
    return PG_OK;
}

int pg_destroy(pg_handle h) {
    if (!h) return PG_ERR_INVALID;
    // release everything even if a call fails; report the first failure
    int rc = PG_OK;
    auto note = [&rc](hipError_t e) {
        if (e != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    };
    note(hipDeviceSynchronize());
    for (uint32_t i = 0; i < h->dev.nlevels; i++) {
        if (h->dev.lv[i].slots) note(hipFree(h->dev.lv[i].slots));
        if (h->dev.lv[i].occ) note(hipFree(h->dev.lv[i].occ));
    }
    if (h->d_counter) note(hipFree(h->d_counter));
    delete h;
    return rc;
}

int pg_clear(pg_handle h, void *stream) {
    if (!h) return PG_ERR_INVALID;
    for (uint32_t i = 0; i < h->dev.nlevels; i++) {
        const pg::LevelDev &L = h->dev.lv[i];
        int rc = pg::launch_zero(L.slots, (h->bytes[i] + 255) & ~uint64_t(255), (hipStream_t)stream);
        if (rc == PG_OK && L.occ) rc = pg::launch_zero(L.occ, pg::occ_bytes(L.nb.d), (hipStream_t)stream);
        if (rc != PG_OK) return rc;
    }
    return PG_OK;
}

int pg_set_mode(pg_handle h, uint32_t mode) {
    if (!h || mode > PG_DETERMINISTIC) return PG_ERR_INVALID;
    h->mode = mode;
    return PG_OK;
}

// Bytes per element of the value buffers (vals, vals_out) of a representation.
static uint64_t value_bytes(uint32_t rep) {
    if (rep == PG_REP_KV64) return 8;
    if (rep == PG_REP_FP8_V8 || rep == PG_REP_FP10_V6) return 1;
    return 2;
}

// One tile per key: a launch covers n * tile work-items, and HIP caps the
// work-items of one grid dimension below 2^32.  Bulk calls are cut into
// launches of at most 2^31 work-items, issued in order on the stream, each
// with its key, value and output pointers advanced.
static int run_op(pg_handle h, pg::Op op, const pg::OpArgs &a, hipStream_t s) {
    const uint32_t tile = h->desc.level[0].tile;
    const bool kv = pg::is_kv_rep(h->dev.rep);
    const uint64_t max_keys = (1ull << 31) / tile;
    const uint64_t vb = value_bytes(h->dev.rep);
    uint64_t done = 0;
    do {
        pg::OpArgs c = a;
        c.n = a.n - done < max_keys ? a.n - done : max_keys;
        if (done) {
            c.keys = a.keys + done;
            if (a.vals) c.vals = (const uint8_t *)a.vals + done * vb;
            if (a.vals_out) c.vals_out = (uint8_t *)a.vals_out + done * vb;
            if (a.flags) c.flags = a.flags + done;
            if (a.found) c.found = a.found + done;
        }
        const int rc = kv ? pg::launch_kv(h->dev, tile, op, c, s) : pg::launch_tcf(h->dev, tile, op, c, s);
        if (rc != PG_OK) return rc;
        done += c.n;
    } while (done < a.n);
    return PG_OK;
}

static int do_insert(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n, uint8_t *d_status,
                     uint64_t *d_fail_count, void *stream, uint32_t with_delete) {
    if (!h || (n && !d_keys)) return PG_ERR_INVALID;
    pg::OpArgs a{};
    a.keys = d_keys;
    a.vals = d_vals;
    a.n = n;
    a.flags = d_status;
    a.fail_count = d_fail_count;
    a.with_delete = with_delete;
    a.deterministic = h->mode == PG_DETERMINISTIC;
    a.stats = h->d_stats;
    a.items = h->items_insert;
    return run_op(h, pg::Op::Insert, a, (hipStream_t)stream);
}

int pg_bulk_insert(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n, uint8_t *d_status,
                   uint64_t *d_fail_count, void *stream) {
    return do_insert(h, d_keys, d_vals, n, d_status, d_fail_count, stream, 0);
}

int pg_bulk_insert_with_delete(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                               uint8_t *d_status, uint64_t *d_fail_count, void *stream) {
    if (!h) return PG_ERR_INVALID;
    // bucket_insert / single_slot_insert have no insert_with_delete, and
    // twelve_bucket's does not compile (12_bit_bucket.cuh:148-203).
    if (pg::is_kv_rep(h->dev.rep) || h->dev.rep == PG_REP_FP12)
        return PG_ERR_UNSUPPORTED;
    return do_insert(h, d_keys, d_vals, n, d_status, d_fail_count, stream, 1);
}

static int do_upsert(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n, void *d_vals_out,
                     uint8_t *d_found, uint8_t *d_status, uint64_t *d_fail_count, void *stream, uint32_t del) {
    if (!h || (n && !d_keys)) return PG_ERR_INVALID;
    // bucket_insert / single_slot_insert have no insert_if_not_exists; not
    // offered for the 12-bit bucket (its _delete form needs remove).
    if (pg::is_kv_rep(h->dev.rep) || h->dev.rep == PG_REP_FP12)
        return PG_ERR_UNSUPPORTED;
    pg::OpArgs a{};
    a.keys = d_keys;
    a.vals = d_vals;
    a.n = n;
    a.vals_out = d_vals_out;
    a.found = d_found;
    a.flags = d_status;
    a.fail_count = d_fail_count;
    a.with_delete = del;
    a.deterministic = h->mode == PG_DETERMINISTIC;
    a.stats = h->d_stats;
    return run_op(h, pg::Op::Upsert, a, (hipStream_t)stream);
}

int pg_bulk_insert_if_not_exists(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                                 void *d_vals_out, uint8_t *d_found, uint8_t *d_status, uint64_t *d_fail_count,
                                 void *stream) {
    return do_upsert(h, d_keys, d_vals, n, d_vals_out, d_found, d_status, d_fail_count, stream, 0);
}

int pg_bulk_insert_if_not_exists_delete(pg_handle h, const uint64_t *d_keys, const void *d_vals, uint64_t n,
                                        void *d_vals_out, uint8_t *d_found, uint8_t *d_status,
                                        uint64_t *d_fail_count, void *stream) {
    return do_upsert(h, d_keys, d_vals, n, d_vals_out, d_found, d_status, d_fail_count, stream, 1);
}

int pg_bulk_query(pg_handle h, const uint64_t *d_keys, uint64_t n, void *d_vals_out, uint8_t *d_found,
                  void *stream) {
    if (!h || (n && !d_keys)) return PG_ERR_INVALID;
    pg::OpArgs a{};
    a.keys = d_keys;
    a.n = n;
    a.vals_out = d_vals_out;
    a.flags = d_found;
    a.stats = h->d_stats;
    a.items = h->items_query;
    return run_op(h, pg::Op::Query, a, (hipStream_t)stream);
}

int pg_bulk_remove(pg_handle h, const uint64_t *d_keys, uint64_t n, uint8_t *d_found, void *stream) {
    if (!h || (n && !d_keys)) return PG_ERR_INVALID;
    for (uint32_t i = 0; i < h->dev.nlevels; i++)
        if (h->dev.lv[i].scheme == PG_SCHEME_SINGLE_SLOT) return PG_ERR_UNSUPPORTED;
    // twelve_bucket::remove does not compile (12_bit_bucket.cuh:207-246)
    if (h->dev.rep == PG_REP_FP12) return PG_ERR_UNSUPPORTED;
    pg::OpArgs a{};
    a.keys = d_keys;
    a.n = n;
    a.flags = d_found;
    a.deterministic = h->mode == PG_DETERMINISTIC;
    a.stats = h->d_stats;
    return run_op(h, pg::Op::Remove, a, (hipStream_t)stream);
}

int pg_set_stats(pg_handle h, uint64_t *d_stats) {
    if (!h) return PG_ERR_INVALID;
    h->d_stats = d_stats;
    return PG_OK;
}

int pg_set_items(pg_handle h, uint32_t insert_items, uint32_t query_items) {
    auto ok = [](uint32_t v) { return v == 0 || v == 1 || v == 2 || v == 4; };
    if (!h || !ok(insert_items) || !ok(query_items)) return PG_ERR_INVALID;
    h->items_insert = insert_items;
    h->items_query = query_items;
    return PG_OK;
}

int pg_fill(pg_handle h, uint64_t *out, void *stream) {
    if (!h || !out) return PG_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int rc = pg::launch_fill(h->dev, h->d_counter, s);
    if (rc != PG_OK) return rc;
    if (hipMemcpyAsync(out, h->d_counter, sizeof(uint64_t), hipMemcpyDeviceToHost, s) != hipSuccess)
        return PG_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? PG_OK : PG_ERR_HIP;
}

// Scratch of pg_fill_noalloc: one 8-byte fill counter per level, in a
// file-scope device array instead of the table's hipMalloc'd counter.  Each
// device holds its own copy at its own address, looked up once per device
// with hipGetSymbolAddress; host code uses only that address (the array's
// name is a host-side stand-in, not a device pointer).  Every call on a
// device shares that device's copy: calls that overlap on one device race.
constexpr uint32_t kNoAllocCounters = PG_MAX_LEVELS;  // 32 B: one per level a table can have
static __device__ __align__(256) uint64_t scratch[kNoAllocCounters];
constexpr int kMaxScratchDevices = 64;
static std::mutex g_scratch_mu;
static uint64_t *g_scratch_ptr[kMaxScratchDevices];

static int noalloc_scratch(int dev, uint64_t **out) {
    if (dev < 0 || dev >= kMaxScratchDevices) return PG_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PG_ERR_HIP;
        g_scratch_ptr[dev] = (uint64_t *)p;
    }
    *out = g_scratch_ptr[dev];
    return PG_OK;
}

int pg_fill_noalloc(pg_handle h, uint64_t *out, void *stream) {
    if (!h || !out) return PG_ERR_INVALID;
    if (h->dev.nlevels > kNoAllocCounters) return PG_ERR_UNSUPPORTED;  // more levels than counters (a guard)
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess) return PG_ERR_HIP;
    if (dev != h->device) return PG_ERR_INVALID;  // the scratch of the table's device must be current
    uint64_t *ctr = nullptr;
    int rc = noalloc_scratch(dev, &ctr);
    if (rc != PG_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    for (uint32_t l = 0; l < h->dev.nlevels; l++) {  // level l counts into ctr[l]
        pg::TableDev one = h->dev;
        one.nlevels = 1;
        one.lv[0] = h->dev.lv[l];
        rc = pg::launch_fill(one, ctr + l, s);
        if (rc != PG_OK) return rc;
    }
    uint64_t per_level[PG_MAX_LEVELS] = {0, 0, 0, 0};
    if (hipMemcpyAsync(per_level, ctr, sizeof(uint64_t) * h->dev.nlevels, hipMemcpyDeviceToHost, s) != hipSuccess)
        return PG_ERR_HIP;
    if (hipStreamSynchronize(s) != hipSuccess) return PG_ERR_HIP;
    uint64_t total = 0;
    for (uint32_t l = 0; l < h->dev.nlevels; l++) total += per_level[l];
    *out = total;
    return PG_OK;
}

int pg_bytes(pg_handle h, uint64_t *out) {
    if (!h || !out) return PG_ERR_INVALID;
    uint64_t b = 0;
    for (uint32_t i = 0; i < h->dev.nlevels; i++) b += h->bytes[i];
    *out = b;
    return PG_OK;
}

int pg_num_slots(pg_handle h, uint64_t *out) {
    if (!h || !out) return PG_ERR_INVALID;
    uint64_t n = 0;
    for (uint32_t i = 0; i < h->dev.nlevels; i++) n += h->dev.lv[i].nb.d * h->dev.lv[i].B;
    *out = n;
    return PG_OK;
}

int pg_level_ptr(pg_handle h, uint32_t level, void **d_ptr, uint64_t *bytes) {
    if (!h || level >= h->dev.nlevels) return PG_ERR_INVALID;
    if (d_ptr) *d_ptr = h->dev.lv[level].slots;
    if (bytes) *bytes = h->bytes[level];
    return PG_OK;
}

int pg_level_sync(pg_handle h, uint32_t level, void *stream) {
    if (!h || level >= h->dev.nlevels) return PG_ERR_INVALID;
    if (!h->dev.lv[level].occ) return PG_OK;
    return pg::launch_occ_rebuild(h->dev.lv[level], h->dev.rep, (hipStream_t)stream);
}

// ---- raw bucket-array files ----
static const char kMagic[8] = {'P', 'G', 'A', 'M', 'D', 'T', 'B', '1'};
static const uint32_t kFileVersion = 3;  // 2: pg_level_desc carries group + reserved; 3: four levels

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    pg_desc desc;
    uint64_t seed;
    uint64_t level_bytes[PG_MAX_LEVELS];
};

static int read_header(FILE *f, FileHeader *h) {
    if (fread(h, sizeof(FileHeader), 1, f) != 1) return PG_ERR_INVALID;
    if (memcmp(h->magic, kMagic, 8) != 0 || h->version != kFileVersion || h->header_bytes != sizeof(FileHeader))
        return PG_ERR_INVALID;
    if (h->desc.num_levels < 1 || h->desc.num_levels > PG_MAX_LEVELS) return PG_ERR_INVALID;
    return PG_OK;
}

int pg_file_info(const char *path, pg_desc *desc, uint64_t *seed) {
    if (!path) return PG_ERR_INVALID;
    FILE *f = fopen(path, "rb");
    if (!f) return PG_ERR_INVALID;
    FileHeader h;
    const int rc = read_header(f, &h);
    fclose(f);
    if (rc != PG_OK) return rc;
    if (desc) *desc = h.desc;
    if (seed) *seed = h.seed;
    return PG_OK;
}

int pg_save(pg_handle h, const char *path) {
    if (!h || !path) return PG_ERR_INVALID;
    FileHeader fh{};
    memcpy(fh.magic, kMagic, 8);
    fh.version = kFileVersion;
    fh.header_bytes = sizeof(FileHeader);
    fh.desc = h->desc;
    fh.seed = h->seed;
    for (uint32_t l = 0; l < h->dev.nlevels; l++) fh.level_bytes[l] = h->bytes[l];
    if (hipDeviceSynchronize() != hipSuccess) return PG_ERR_HIP;
    FILE *f = fopen(path, "wb");
    if (!f) return PG_ERR_INVALID;
    int rc = fwrite(&fh, sizeof(fh), 1, f) == 1 ? PG_OK : PG_ERR_INVALID;
    const uint64_t chunk = 64ull << 20;
    void *buf = nullptr;
    if (rc == PG_OK && hipHostMalloc(&buf, chunk) != hipSuccess) rc = PG_ERR_NOMEM;
    for (uint32_t l = 0; l < h->dev.nlevels && rc == PG_OK; l++) {
        for (uint64_t off = 0; off < h->bytes[l] && rc == PG_OK; off += chunk) {
            const uint64_t n = h->bytes[l] - off < chunk ? h->bytes[l] - off : chunk;
            if (hipMemcpy(buf, (char *)h->dev.lv[l].slots + off, n, hipMemcpyDeviceToHost) != hipSuccess)
                rc = PG_ERR_HIP;
            else if (fwrite(buf, 1, n, f) != n)
                rc = PG_ERR_INVALID;
        }
    }
    if (buf && hipHostFree(buf) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    if (fclose(f) != 0 && rc == PG_OK) rc = PG_ERR_INVALID;
    return rc;
}

int pg_load(const char *path, pg_handle *out) {
    if (!path || !out) return PG_ERR_INVALID;
    *out = nullptr;
    FILE *f = fopen(path, "rb");
    if (!f) return PG_ERR_INVALID;
    FileHeader fh;
    int rc = read_header(f, &fh);
    pg_handle h = nullptr;
    if (rc == PG_OK) rc = pg_create(&fh.desc, fh.seed, &h);
    for (uint32_t l = 0; l < fh.desc.num_levels && rc == PG_OK; l++)
        if (fh.level_bytes[l] != h->bytes[l]) rc = PG_ERR_INVALID;
    const uint64_t chunk = 64ull << 20;
    void *buf = nullptr;
    if (rc == PG_OK && hipHostMalloc(&buf, chunk) != hipSuccess) rc = PG_ERR_NOMEM;
    for (uint32_t l = 0; l < fh.desc.num_levels && rc == PG_OK; l++) {
        for (uint64_t off = 0; off < h->bytes[l] && rc == PG_OK; off += chunk) {
            const uint64_t n = h->bytes[l] - off < chunk ? h->bytes[l] - off : chunk;
            if (fread(buf, 1, n, f) != n)
                rc = PG_ERR_INVALID;
            else if (hipMemcpy((char *)h->dev.lv[l].slots + off, buf, n, hipMemcpyHostToDevice) != hipSuccess)
                rc = PG_ERR_HIP;
        }
    }
    if (buf && hipHostFree(buf) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    fclose(f);
    for (uint32_t l = 0; l < fh.desc.num_levels && rc == PG_OK; l++) rc = pg_level_sync(h, l, nullptr);
    if (rc == PG_OK && hipDeviceSynchronize() != hipSuccess) rc = PG_ERR_HIP;
    if (rc != PG_OK) {
        if (h) pg_destroy(h);
        return rc;
    }
    *out = h;
    return PG_OK;
}

int pg_gen_keys(uint64_t seed, uint64_t start, uint64_t n, uint64_t *d_out, void *stream) {
    if (n && !d_out) return PG_ERR_INVALID;
    return pg::launch_gen_keys(seed, start, n, d_out, (hipStream_t)stream);
}

int pg_random_sector_bench(void *d_buf, uint64_t bytes, uint64_t n_reads, uint32_t op, uint64_t seed,
                           uint64_t *d_sink, void *stream) {
    if (!d_buf || !d_sink || bytes < 64 || (bytes % 64) || op >= (1u << 15)) return PG_ERR_INVALID;
    return pg::launch_random_sector(d_buf, bytes, n_reads, op, seed, d_sink, (hipStream_t)stream);
}

}  // extern "C"
