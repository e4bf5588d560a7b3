// pg_batch.hip -- many independent filters updated by one persistent kernel.
//
// A batch is N Two-Choice Filters of uneven sizes that share one shape
// (representation, bucket size, tile width, probing, hashing) and live in one
// arena allocation.  A bulk call takes one flat key array and a CSR offset
// array (keys of table t are keys[off[t] .. off[t+1])).  The persistent
// kernel is launched with a grid sized to fill the GPU once; its waves pull
// fixed-size chunks of consecutive keys from a device work counter, find the
// chunk's first table by one binary search over the offsets (L2-resident),
// walk forward across table boundaries inside the chunk, and run the same
// per-key device code as the single-table kernels (pg_tcf.hpp).  Every wave exits when the counter passes the total, so the
// grid drains.  The counter is reset by a kernel, not a memset, so the whole
// sequence can be recorded in a hipGraph (pg_batch_capture).
//
// The reference runs one kernel per table with one tile per key
// (tests/src/test_mhm_tcf.cu:318-487); it has no multi-table launch.
#include <new>
#include <vector>

#include "pg_tcf.hpp"

namespace pg {

constexpr int kBatchBlock = 256;

// Table of key index `i`: largest t with off[t] <= i (binary search).
__device__ inline uint32_t find_table(const uint64_t *off, uint32_t ntables, uint64_t i) {
    uint32_t lo = 0, hi = ntables;  // invariant: off[lo] <= i < off[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

struct BatchArgs {
    const TableDev *tables;   // device array, one TableDev per table
    uint32_t ntables;
    const uint64_t *keys;
    const uint64_t *off;      // ntables + 1 offsets (device)
    uint64_t total;           // off[ntables]
    const uint16_t *vals;     // optional
    uint8_t *flags;           // insert: status; query: found
    uint16_t *vals_out;       // query values, optional
    uint64_t *fail_count;     // optional
    uint64_t *counter;        // work counter (device, reset before launch)
    uint32_t chunk;           // keys per chunk
};

__global__ __launch_bounds__(64) void k_batch_reset(uint64_t *counter) {
    if (threadIdx.x == 0) counter[0] = 0;
}

// One key of a batch: the single-table per-key code on the key's table.
template <int P, int B, bool QUERY>
__device__ inline void batch_key(const BatchArgs &a, const TableDev &T, uint64_t i, const Tile<P> &t,
                                 uint64_t &fails) {
    const uint64_t key = a.keys[i];
    if constexpr (QUERY) {
        const uint32_t fp = fp16_of(key);
        uint32_t val = 0;
        bool own = false, found = false;
        if constexpr (B != 0) {
            const LevelDev &L = T.lv[0];
            Probe pr;
            pr.begin(L, key, fp);
            for (uint32_t j = 0; j < L.max_probes && !found; j++) {
                const uint64_t p = pr.at(j);
                if (p == kEnd) break;
                Bucket<P, B> bk;
                bk.load((const uint32_t *)L.slots, fastmod(p, L.nb), t);
                found = bk.find(fp, t, val, own);
            }
        } else {
            found = gen_level_query<P>(T.lv[0], key, fp, t, val, own, 0);
        }
        for (uint32_t l = 1; l < T.nlevels && !found; l++)
            found = gen_level_query<P>(T.lv[l], key, fp, t, val, own, 1);
        const bool writer = found ? own : t.leader();
        if (writer) {
            if (a.flags) a.flags[i] = found ? 1 : 0;
            if (a.vals_out) a.vals_out[i] = found ? (uint16_t)val : 0;
        }
    } else {
        const uint32_t v = a.vals ? a.vals[i] : 0u;
        const int st = tcf_insert_key<P, B>(T, key, v, false, t);
        if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
        fails += (st == 0 && t.leader());
    }
}

template <int P, int B, bool QUERY>
__device__ inline void batch_one(const BatchArgs &a, uint64_t i, uint32_t &tb, const Tile<P> &t, uint64_t &fails) {
    while (a.off[tb + 1] <= i) tb++;
    batch_key<P, B, QUERY>(a, a.tables[tb], i, t, fails);
}

// I keys of one tile at once (register-bucket shapes, B != 0): every key's
// table is found and its first probe bucket loaded before any is examined,
// so a tile has I first-choice reads in flight; then each key runs the
// single-table code from that loaded bucket.
template <int P, int B, bool QUERY, int I>
__device__ inline void batch_many(const BatchArgs &a, uint64_t i0, uint64_t end, uint32_t stride, uint32_t &tb,
                                  const Tile<P> &t, uint64_t &fails) {
    const TableDev *T[I];
    uint64_t idx[I], key[I], b0[I];
    uint32_t fp[I];
    bool live[I];
    Probe pr[I];
    Bucket<P, B> bk[I];
#pragma unroll
    for (int k = 0; k < I; k++) {
        idx[k] = i0 + (uint64_t)k * stride;
        live[k] = idx[k] < end;
        if (live[k])
            while (a.off[tb + 1] <= idx[k]) tb++;
        T[k] = &a.tables[tb];
        key[k] = live[k] ? a.keys[idx[k]] : 0;
        fp[k] = fp16_of(key[k]);
        const LevelDev &L = T[k]->lv[0];
        pr[k].begin(L, key[k], fp[k]);
        const uint64_t p0 = pr[k].at(0);
        b0[k] = (p0 == kEnd) ? 0 : fastmod(p0, L.nb);
        if (live[k] && p0 != kEnd) bk[k].load((const uint32_t *)L.slots, b0[k], t);
    }
#pragma unroll
    for (int k = 0; k < I; k++) {
        if (!live[k]) continue;
        const TableDev &Tk = *T[k];
        const uint64_t i = idx[k];
        if constexpr (QUERY) {
            const LevelDev &L = Tk.lv[0];
            uint32_t val = 0;
            bool own = false;
            bool found = pr[k].at(0) != kEnd && bk[k].find(fp[k], t, val, own);
            for (uint32_t j = 1; j < L.max_probes && !found; j++) {
                const uint64_t p = pr[k].at(j);
                if (p == kEnd) break;
                Bucket<P, B> bj;
                bj.load((const uint32_t *)L.slots, fastmod(p, L.nb), t);
                found = bj.find(fp[k], t, val, own);
            }
            for (uint32_t l = 1; l < Tk.nlevels && !found; l++)
                found = gen_level_query<P>(Tk.lv[l], key[k], fp[k], t, val, own, 1);
            const bool writer = found ? own : t.leader();
            if (writer) {
                if (a.flags) a.flags[i] = found ? 1 : 0;
                if (a.vals_out) a.vals_out[i] = found ? (uint16_t)val : 0;
            }
        } else {
            const uint32_t v = a.vals ? a.vals[i] : 0u;
            const int st = tcf_insert_prepared<P, B>(Tk, key[k], fp[k], (v << 16) | fp[k], false, t, pr[k],
                                                     bk[k], b0[k]);
            if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
            fails += (st == 0 && t.leader());
        }
    }
}

// With one key per tile the compiler is asked for 7 waves per SIMD (72
// VGPRs, 12 B of scratch) instead of its own 6 (75 VGPRs): 5-9 % faster on
// 4096 uneven filters.  8 waves spill 48 B per lane and are slower.
template <int P, int B, bool QUERY, int I>
__global__ __launch_bounds__(kBatchBlock) __attribute__((amdgpu_waves_per_eu(I == 1 ? 7 : 1, 8)))
void k_batch_persistent(BatchArgs a) {
    const uint64_t *off = a.off;
    Tile<P> t;
    t.concurrent = true;
    constexpr uint32_t kTilesPerWave = 64 / P;
    const uint32_t tile_in_wave = (threadIdx.x & 63) / P;
    uint64_t fails = 0;
    while (true) {
        uint64_t base = 0;
        uint32_t tb0 = 0;
        if ((threadIdx.x & 63) == 0) {
            base = atomicAdd((unsigned long long *)a.counter, (unsigned long long)a.chunk);
            if (base < a.total) tb0 = find_table(off, a.ntables, base);
        }
        base = __shfl(base, 0, 64);
        tb0 = __shfl(tb0, 0, 64);
        if (base >= a.total) break;
        const uint64_t end = (base + a.chunk < a.total) ? base + a.chunk : a.total;
        uint32_t tb = tb0;
        if constexpr (I == 1 || B == 0) {
            // Most chunks lie inside one table: its descriptor is then read
            // once per wave into scalar registers instead of per key.
            const uint32_t tu = __builtin_amdgcn_readfirstlane(tb0);
            if (off[tu + 1] >= end) {
                const TableDev &T = a.tables[tu];
                for (uint64_t i = base + tile_in_wave; i < end; i += kTilesPerWave)
                    batch_key<P, B, QUERY>(a, T, i, t, fails);
            } else {
                for (uint64_t i = base + tile_in_wave; i < end; i += kTilesPerWave)
                    batch_one<P, B, QUERY>(a, i, tb, t, fails);
            }
        } else {
            for (uint64_t i0 = base + tile_in_wave; i0 < end; i0 += (uint64_t)kTilesPerWave * I)
                batch_many<P, B, QUERY, I>(a, i0, end, kTilesPerWave, tb, t, fails);
        }
    }
    if (!QUERY && a.fail_count && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
}

// Deterministic batched insert: tile t applies table t's keys in order (the
// tables are independent, so every table's layout equals the sequential
// oracle's while all tables proceed in parallel).
template <int P, int B>
__global__ __launch_bounds__(kBatchBlock) void k_batch_insert_ordered(BatchArgs a) {
    Tile<P> t;
    const uint64_t tb = (uint64_t(blockIdx.x) * kBatchBlock + threadIdx.x) / P;
    if (tb >= a.ntables) return;
    const TableDev &T = a.tables[tb];
    uint64_t fails = 0;
    for (uint64_t i = a.off[tb]; i < a.off[tb + 1]; i++) {
        const uint32_t v = a.vals ? a.vals[i] : 0u;
        const int st = tcf_insert_key<P, B>(T, a.keys[i], v, false, t);
        if (a.flags && t.leader()) a.flags[i] = (uint8_t)st;
        fails += (st == 0);
        fence_agent();
    }
    if (a.fail_count && t.leader() && fails) atomicAdd((unsigned long long *)a.fail_count, fails);
}

template <int P, int B, bool QUERY>
static void launch_persistent(const BatchArgs &a, unsigned blocks, uint32_t items, hipStream_t s) {
    if constexpr (B != 0) {
        if (items == 4) {
            hipLaunchKernelGGL((k_batch_persistent<P, B, QUERY, 4>), dim3(blocks), dim3(kBatchBlock), 0, s, a);
            return;
        }
        if (items == 2) {
            hipLaunchKernelGGL((k_batch_persistent<P, B, QUERY, 2>), dim3(blocks), dim3(kBatchBlock), 0, s, a);
            return;
        }
    }
    hipLaunchKernelGGL((k_batch_persistent<P, B, QUERY, 1>), dim3(blocks), dim3(kBatchBlock), 0, s, a);
}

template <int P, int B>
static int launch_batch_pb(const BatchArgs &a, bool query, bool ordered, unsigned persistent_blocks,
                           uint32_t items, hipStream_t s) {
    hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, a.counter);
    if (ordered && !query) {
        const unsigned grid = (unsigned)(((uint64_t)a.ntables * P + kBatchBlock - 1) / kBatchBlock);
        hipLaunchKernelGGL((k_batch_insert_ordered<P, B>), dim3(grid), dim3(kBatchBlock), 0, s, a);
    } else if (query) {
        launch_persistent<P, B, true>(a, persistent_blocks, items, s);
    } else {
        launch_persistent<P, B, false>(a, persistent_blocks, items, s);
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

template <int P>
static int launch_batch_p(uint32_t B, const BatchArgs &a, bool query, bool ordered, unsigned blocks,
                          uint32_t items, hipStream_t s) {
    switch (B) {
        case 8: return launch_batch_pb<P, 8>(a, query, ordered, blocks, items, s);
        case 16: return launch_batch_pb<P, 16>(a, query, ordered, blocks, items, s);
        case 32: return launch_batch_pb<P, 32>(a, query, ordered, blocks, items, s);
        default: return launch_batch_pb<P, 0>(a, query, ordered, blocks, items, s);
    }
}

static int launch_batch(uint32_t P, uint32_t B, const BatchArgs &a, bool query, bool ordered, unsigned blocks,
                        uint32_t items, hipStream_t s) {
    switch (P) {
        case 4: return launch_batch_p<4>(B, a, query, ordered, blocks, items, s);
        case 8: return launch_batch_p<8>(B, a, query, ordered, blocks, items, s);
        case 16: return launch_batch_p<16>(B, a, query, ordered, blocks, items, s);
        case 32: return launch_batch_p<32>(B, a, query, ordered, blocks, items, s);
        case 64: return launch_batch_p<64>(B, a, query, ordered, blocks, items, s);
        case 2: return launch_batch_p<2>(B, a, query, ordered, blocks, items, s);
        case 1: return launch_batch_p<1>(B, a, query, ordered, blocks, items, s);
        default: return PG_ERR_INVALID;
    }
}

}  // namespace pg

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct pg_batch_s {
    pg_desc proto;
    uint32_t ntables;
    uint32_t mode;
    uint32_t chunk;
    uint32_t items;            // keys per tile in flight (1, 2, 4)
    unsigned blocks;           // persistent grid
    void *arena;               // every table's slot arrays
    uint64_t arena_bytes;
    pg::TableDev *d_tables;
    std::vector<pg::TableDev> h_tables;
    std::vector<uint64_t> level_bytes;  // per table, per level
    uint64_t *d_counter;
    uint64_t *d_fill;
    hipGraphExec_t graph;
    hipGraph_t graph_def;
};

namespace {

unsigned persistent_grid() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    // 8 blocks of 4 waves per CU: one resident wave per SIMD slot pair;
    // enough waves to keep many random reads in flight.
    return (unsigned)(cus > 0 ? cus : 256) * 8u;
}

uint64_t align256(uint64_t x) { return (x + 255) & ~uint64_t(255); }

}  // namespace

extern "C" {

int pg_batch_create(const pg_desc *proto, uint32_t ntables, const uint64_t *nslots, const uint64_t *backing_slots,
                    uint64_t seed, pg_batch *out) {
    if (!proto || !nslots || !out || ntables == 0) return PG_ERR_INVALID;
    *out = nullptr;
    if (proto->rep != PG_REP_FP16_V16) return PG_ERR_UNSUPPORTED;
    if (proto->num_levels < 1 || proto->num_levels > PG_MAX_LEVELS) return PG_ERR_INVALID;
    const pg_level_desc &l0 = proto->level[0];
    if (l0.scheme != PG_SCHEME_POWER_OF_N_SHORTCUT) return PG_ERR_UNSUPPORTED;
    for (uint32_t i = 0; i < proto->num_levels; i++) {
        if (proto->level[i].tile != l0.tile) return PG_ERR_INVALID;
        if (i > 0 && proto->level[i].scheme != PG_SCHEME_LINEAR_BUCKETS) return PG_ERR_UNSUPPORTED;
        if (proto->level[i].group != 0) return PG_ERR_UNSUPPORTED;
        if (proto->level[i].bucket_size == 0 || proto->level[i].bucket_size > 64) return PG_ERR_INVALID;
    }
    const uint32_t P = l0.tile;
    if (P == 0 || P > 64 || (P & (P - 1))) return PG_ERR_INVALID;
    if ((l0.prober == PG_PROBE_P2 || l0.prober == PG_PROBE_XOR_P2 || l0.prober == PG_PROBE_SPLIT64) &&
        l0.max_probes != 2)
        return PG_ERR_INVALID;
    if (l0.max_probes == 0 || l0.max_probes > pg::kMaxShortcutProbes) return PG_ERR_UNSUPPORTED;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PG_ERR_NODEVICE;

    pg_batch_s *b = new (std::nothrow) pg_batch_s();
    if (!b) return PG_ERR_NOMEM;
    b->proto = *proto;
    b->ntables = ntables;
    b->mode = PG_CONCURRENT;
    b->chunk = 1024;
    b->items = 1;
    b->blocks = persistent_grid();
    b->h_tables.resize(ntables);
    b->level_bytes.resize((size_t)ntables * PG_MAX_LEVELS, 0);
    // lay out every table's levels in one arena
    uint64_t total = 0;
    std::vector<uint64_t> offs((size_t)ntables * PG_MAX_LEVELS, 0);
    // occupancy index of each linear backing level, after its slots (~0: none)
    std::vector<uint64_t> occ_offs((size_t)ntables * PG_MAX_LEVELS, ~uint64_t(0));
    for (uint32_t t = 0; t < ntables; t++) {
        if (nslots[t] == 0) {
            delete b;
            return PG_ERR_INVALID;
        }
        for (uint32_t l = 0; l < proto->num_levels; l++) {
            uint64_t ns = l == 0 ? nslots[t] : (backing_slots ? backing_slots[t] : nslots[t] / 100);
            if (ns == 0) ns = proto->level[l].bucket_size;  // tiny tables keep one backing bucket
            const uint32_t B = proto->level[l].bucket_size;
            const uint64_t nb = (ns - 1) / B + 1;
            const uint64_t bytes = nb * B * 4;
            offs[(size_t)t * PG_MAX_LEVELS + l] = total;
            b->level_bytes[(size_t)t * PG_MAX_LEVELS + l] = bytes;
            total += align256(bytes);
            if (pg::has_occ_index(proto->rep, proto->level[l].scheme, l)) {
                occ_offs[(size_t)t * PG_MAX_LEVELS + l] = total;
                total += pg::occ_bytes(nb);
            }
            pg::TableDev &T = b->h_tables[t];
            pg::LevelDev &L = T.lv[l];
            L.nb = pg::make_fastmod(nb);
            L.B = B;
            L.group = 0;
            L.bucket_bytes = B * 4;
            L.scheme = proto->level[l].scheme;
            L.prober = proto->level[l].prober;
            L.hasher = proto->level[l].hasher;
            L.max_probes = proto->level[l].max_probes;
            uint64_t s = seed;
            for (uint32_t q = 0; q < l; q++) s = s * s;
            L.seed = s;
            L.seed2 = s * s;
            T.nlevels = proto->num_levels;
            T.rep = proto->rep;
        }
    }
    b->arena_bytes = total;
    int rc = PG_OK;
    if (hipMalloc(&b->arena, total) != hipSuccess) rc = PG_ERR_NOMEM;
    if (rc == PG_OK) {
        for (uint32_t t = 0; t < ntables; t++)
            for (uint32_t l = 0; l < proto->num_levels; l++) {
                const size_t i = (size_t)t * PG_MAX_LEVELS + l;
                b->h_tables[t].lv[l].slots = (char *)b->arena + offs[i];
                if (occ_offs[i] != ~uint64_t(0)) b->h_tables[t].lv[l].occ = (uint32_t *)((char *)b->arena + occ_offs[i]);
            }
        if (hipMalloc((void **)&b->d_tables, sizeof(pg::TableDev) * ntables) != hipSuccess ||
            hipMalloc((void **)&b->d_counter, 256) != hipSuccess ||
            hipMalloc((void **)&b->d_fill, 256) != hipSuccess)
            rc = PG_ERR_NOMEM;
    }
    if (rc == PG_OK &&
        hipMemcpy(b->d_tables, b->h_tables.data(), sizeof(pg::TableDev) * ntables, hipMemcpyHostToDevice) != hipSuccess)
        rc = PG_ERR_HIP;
    if (rc == PG_OK) rc = pg::launch_zero(b->arena, total, nullptr);
    if (rc == PG_OK && hipDeviceSynchronize() != hipSuccess) rc = PG_ERR_HIP;
    if (rc != PG_OK) {
        pg_batch_destroy(b);
        return rc;
    }
    *out = b;
    return PG_OK;
}

int pg_batch_destroy(pg_batch b) {
    if (!b) return PG_ERR_INVALID;
    // release everything even if a call fails; report the first failure
    int rc = PG_OK;
    auto note = [&rc](hipError_t e) {
        if (e != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    };
    note(hipDeviceSynchronize());
    if (b->graph) note(hipGraphExecDestroy(b->graph));
    if (b->graph_def) note(hipGraphDestroy(b->graph_def));
    if (b->arena) note(hipFree(b->arena));
    if (b->d_tables) note(hipFree(b->d_tables));
    if (b->d_counter) note(hipFree(b->d_counter));
    if (b->d_fill) note(hipFree(b->d_fill));
    delete b;
    return rc;
}

int pg_batch_set_mode(pg_batch b, uint32_t mode) {
    if (!b || mode > PG_DETERMINISTIC) return PG_ERR_INVALID;
    b->mode = mode;
    return PG_OK;
}

int pg_batch_set_launch(pg_batch b, uint32_t chunk, uint32_t blocks) {
    if (!b || chunk == 0) return PG_ERR_INVALID;
    b->chunk = chunk;
    b->blocks = blocks ? blocks : persistent_grid();
    return PG_OK;
}

int pg_batch_set_items(pg_batch b, uint32_t items) {
    if (!b || !(items == 0 || items == 1 || items == 2 || items == 4)) return PG_ERR_INVALID;
    b->items = items ? items : 1;
    return PG_OK;
}

int pg_batch_clear(pg_batch b, void *stream) {
    if (!b) return PG_ERR_INVALID;
    return pg::launch_zero(b->arena, b->arena_bytes, (hipStream_t)stream);
}

static int batch_op(pg_batch b, bool query, const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total,
                    const uint16_t *d_vals, uint8_t *d_flags, uint16_t *d_vals_out, uint64_t *d_fail, void *stream) {
    if (!b || (total && (!d_keys || !d_offsets))) return PG_ERR_INVALID;
    if (total == 0) return PG_OK;
    pg::BatchArgs a{};
    a.tables = b->d_tables;
    a.ntables = b->ntables;
    a.keys = d_keys;
    a.off = d_offsets;
    a.total = total;
    a.vals = d_vals;
    a.flags = d_flags;
    a.vals_out = d_vals_out;
    a.fail_count = d_fail;
    a.counter = b->d_counter;
    a.chunk = b->chunk;
    // the register fast path takes levels of at most two probes
    const uint32_t fastB = b->proto.level[0].max_probes <= 2 ? b->proto.level[0].bucket_size : 0u;
    return pg::launch_batch(b->proto.level[0].tile, fastB, a, query,
                            b->mode == PG_DETERMINISTIC, b->blocks, b->items, (hipStream_t)stream);
}

int pg_batch_insert(pg_batch b, const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total,
                    const uint16_t *d_vals, uint8_t *d_status, uint64_t *d_fail_count, void *stream) {
    return batch_op(b, false, d_keys, d_offsets, total, d_vals, d_status, nullptr, d_fail_count, stream);
}

int pg_batch_query(pg_batch b, const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total,
                   uint16_t *d_vals_out, uint8_t *d_found, void *stream) {
    return batch_op(b, true, d_keys, d_offsets, total, nullptr, d_found, d_vals_out, nullptr, stream);
}

int pg_batch_capture(pg_batch b, const uint64_t *d_ins_keys, const uint64_t *d_ins_offsets, uint64_t ins_total,
                     const uint64_t *d_q_keys, const uint64_t *d_q_offsets, uint64_t q_total, uint8_t *d_status,
                     uint8_t *d_found, uint64_t *d_fail_count, uint32_t clear_first, void *stream) {
    if (!b) return PG_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    bool own_stream = false;
    if (!s) {
        if (hipStreamCreate(&s) != hipSuccess) return PG_ERR_HIP;
        own_stream = true;
    }
    int rc = PG_OK;
    if (b->graph) {
        if (hipGraphExecDestroy(b->graph) != hipSuccess) rc = PG_ERR_HIP;
        b->graph = nullptr;
    }
    if (b->graph_def) {
        if (hipGraphDestroy(b->graph_def) != hipSuccess) rc = PG_ERR_HIP;
        b->graph_def = nullptr;
    }
    const bool capturing =
        rc == PG_OK && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (!capturing) rc = PG_ERR_HIP;
    if (rc == PG_OK && clear_first) rc = pg_batch_clear(b, s);
    if (rc == PG_OK) rc = pg_batch_insert(b, d_ins_keys, d_ins_offsets, ins_total, nullptr, d_status, d_fail_count, s);
    if (rc == PG_OK) rc = pg_batch_query(b, d_q_keys, d_q_offsets, q_total, nullptr, d_found, s);
    hipGraph_t g = nullptr;
    if (capturing && hipStreamEndCapture(s, &g) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    if (rc == PG_OK && hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0) != hipSuccess) rc = PG_ERR_HIP;
    b->graph_def = g;
    if (own_stream && hipStreamDestroy(s) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    return rc;
}

int pg_batch_replay(pg_batch b, void *stream) {
    if (!b || !b->graph) return PG_ERR_INVALID;
    return hipGraphLaunch(b->graph, (hipStream_t)stream) == hipSuccess ? PG_OK : PG_ERR_HIP;
}

int pg_batch_fill(pg_batch b, uint32_t table, uint64_t *out, void *stream) {
    if (!b || !out || table >= b->ntables) return PG_ERR_INVALID;
    hipStream_t s = (hipStream_t)stream;
    int rc = pg::launch_fill(b->h_tables[table], b->d_fill, s);
    if (rc != PG_OK) return rc;
    if (hipMemcpyAsync(out, b->d_fill, 8, hipMemcpyDeviceToHost, s) != hipSuccess) return PG_ERR_HIP;
    return hipStreamSynchronize(s) == hipSuccess ? PG_OK : PG_ERR_HIP;
}

int pg_batch_level_ptr(pg_batch b, uint32_t table, uint32_t level, void **d_ptr, uint64_t *bytes) {
    if (!b || table >= b->ntables || level >= b->proto.num_levels) return PG_ERR_INVALID;
    if (d_ptr) *d_ptr = b->h_tables[table].lv[level].slots;
    if (bytes) *bytes = b->level_bytes[(size_t)table * PG_MAX_LEVELS + level];
    return PG_OK;
}

int pg_batch_bytes(pg_batch b, uint64_t *out) {
    if (!b || !out) return PG_ERR_INVALID;
    uint64_t s = 0;
    for (uint64_t v : b->level_bytes) s += v;
    *out = s;
    return PG_OK;
}

}  // extern "C"
