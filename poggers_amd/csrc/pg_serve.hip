// pg_serve.hip -- continuous-batching request loop over resident tables.
//
// Requests (table, op, n keys in host or device memory) are queued by
// pg_server_submit.  pg_server_progress is the engine: it retires finished
// batches (event query, no blocking), admits queued requests in FIFO order
// while their staging bytes fit the device arena (the HBM budget) and a
// stream slot is free, and launches each admitted group as one batch: one
// copy of the batch's request table, one persistent kernel that applies
// every request of the batch to its own table (mixed tables, mixed ops), all
// on one stream; up to `nstreams` batches are in flight, so requests that
// arrive while others run overlap with them on the device.
//
// Buffers the GPU can address (pinned host memory from hipHostMalloc, or
// device memory) are used in place: the kernel reads the keys and writes the
// results through them, so a request costs no copy calls and no staging
// bytes.  Other (pageable) buffers are staged through the arena with
// host-to-device / device-to-host copies.  pg_server_run_plain processes one
// request with its own kernel launch (and copies, if staged) and waits for
// it: the per-request baseline.
//
// No reference counterpart: the reference has no host request path.
#include <deque>
#include <new>
#include <vector>

#include "pg_kv_ops.hpp"
#include "pg_plan.hpp"
#include "pg_ring.hpp"
#include "pg_tcf.hpp"

namespace pg {

struct ReqDev {
    const uint64_t *keys;
    const void *vals;
    uint8_t *flags;
    void *vals_out;
    uint64_t n;
    uint32_t table;
    uint32_t op;
};

constexpr uint32_t kIdxShift = 8;  // one index entry per 256 keys of a batch

struct ServeArgs {
    const TableDev *tables;
    const ReqDev *reqs;
    const uint64_t *off;   // nreq + 1 key offsets
    const uint32_t *idx;   // idx[c]: the request holding key c << kIdxShift (NULL: search)
    uint32_t nreq;
    uint64_t total;
};

__device__ inline uint32_t find_req(const uint64_t *off, uint32_t n, uint64_t i) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

// RP: the server holds a table with a replace-scheme level; its kernel then
// carries that insert (a separate instantiation, so other servers' kernels
// carry none of its code).
template <int P, bool RP>
__device__ inline void serve_one(const TableDev &T, const ReqDev &r, uint64_t j, const Tile<P> &t) {
    const uint64_t key = r.keys[j];
    if (T.rep == PG_REP_KV64) {
        if (r.op == PG_OP_INSERT) {
            const uint64_t v = r.vals ? ((const uint64_t *)r.vals)[j] : 0ull;
            const int st = kv_insert_key<P, SlotKV, RP>(T, key, v, t);
            if (r.flags && t.leader()) r.flags[j] = (uint8_t)st;
        } else if (r.op == PG_OP_QUERY) {
            uint64_t v = 0;
            bool own = false;
            const bool f = kv_query_key<P>(T, key, t, v, own);
            if (f ? own : t.leader()) {
                if (r.flags) r.flags[j] = f ? 1 : 0;
                if (r.vals_out) ((uint64_t *)r.vals_out)[j] = f ? v : 0ull;
            }
        } else {
            const bool f = kv_remove_key<P>(T, key, t);
            if (r.flags && t.leader()) r.flags[j] = f ? 1 : 0;
        }
        return;
    }
    if (T.rep == PG_REP_KV16) {  // the flat TCF's 16/16 key-value slots
        if (r.op == PG_OP_INSERT) {
            const uint64_t v = r.vals ? ((const uint16_t *)r.vals)[j] : 0u;
            const int st = kv_insert_key<P, SlotKV16, RP>(T, key, v, t);
            if (r.flags && t.leader()) r.flags[j] = (uint8_t)st;
        } else if (r.op == PG_OP_QUERY) {
            uint64_t v = 0;
            bool own = false;
            const bool f = kv_query_key<P, SlotKV16>(T, key, t, v, own);
            if (f ? own : t.leader()) {
                if (r.flags) r.flags[j] = f ? 1 : 0;
                if (r.vals_out) ((uint16_t *)r.vals_out)[j] = f ? (uint16_t)v : 0;
            }
        } else {
            const bool f = kv_remove_key<P, SlotKV16>(T, key, t);
            if (r.flags && t.leader()) r.flags[j] = f ? 1 : 0;
        }
        return;
    }
    const bool fast16 = T.lv[0].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT && T.lv[0].B == 16 && T.lv[0].max_probes <= 2;
    if (r.op == PG_OP_INSERT) {
        const uint32_t v = r.vals ? ((const uint16_t *)r.vals)[j] : 0u;
        const int st = fast16 ? tcf_insert_key<P, 16>(T, key, v, false, t) : tcf_insert_key<P, 0>(T, key, v, false, t);
        if (r.flags && t.leader()) r.flags[j] = (uint8_t)st;
    } else if (r.op == PG_OP_QUERY) {
        const uint32_t fp = fp16_of(key);
        uint32_t v = 0;
        bool own = false, f = false;
        if (fast16) {
            const LevelDev &L = T.lv[0];
            Probe pr;
            pr.begin(L, key, fp);
            for (uint32_t q = 0; q < L.max_probes && !f; q++) {
                const uint64_t p = pr.at(q);
                if (p == kEnd) break;
                Bucket<P, 16> bk;
                bk.load((const uint32_t *)L.slots, fastmod(p, L.nb), t);
                f = bk.find(fp, t, v, own);
            }
        } else {
            f = gen_level_query<P>(T.lv[0], key, fp, t, v, own, 0);
        }
        for (uint32_t l = 1; l < T.nlevels && !f; l++) f = gen_level_query<P>(T.lv[l], key, fp, t, v, own, 1);
        if (f ? own : t.leader()) {
            if (r.flags) r.flags[j] = f ? 1 : 0;
            if (r.vals_out) ((uint16_t *)r.vals_out)[j] = f ? (uint16_t)v : 0;
        }
    } else {
        const bool f = fast16 ? tcf_remove_key<P, 16>(T, key, t) : tcf_remove_key<P, 0>(T, key, t);
        if (r.flags && t.leader()) r.flags[j] = f ? 1 : 0;
    }
}

// One tile per key of the batch, like the single-table bulk kernels, so every
// key's loads (its key, its buckets) are in flight at once.  A tile finds its
// request from the batch's coarse index (one load, then a short walk over the
// key offsets) rather than a binary search, whose dependent loads would sit
// in front of every key.  The grid is capped at 2^31 work-items (HIP's limit
// for one dimension is below 2^32), and a tile takes key i + k * tiles until
// the batch is done.
template <int P, bool RP>
__global__ __launch_bounds__(256) void k_serve(ServeArgs a) {
    Tile<P> t;
    t.concurrent = true;
    const uint64_t tiles = (uint64_t(gridDim.x) * 256) / P;
    for (uint64_t i = (uint64_t(blockIdx.x) * 256 + threadIdx.x) / P; i < a.total; i += tiles) {
        uint32_t r;
        if (a.idx) {
            r = a.idx[i >> kIdxShift];
            while (a.off[r + 1] <= i) r++;
        } else {
            r = find_req(a.off, a.nreq, i);
        }
        const ReqDev &rq = a.reqs[r];
        serve_one<P, RP>(a.tables[rq.table], rq, i - a.off[r], t);
    }
}

template <bool RP>
static int launch_serve_rp(uint32_t P, const ServeArgs &a, hipStream_t s) {
    const uint64_t want = (a.total * P + 255) / 256;
    const unsigned blocks = (unsigned)(want < (1ull << 23) ? want : (1ull << 23));
    switch (P) {
        case 1: hipLaunchKernelGGL((k_serve<1, RP>), dim3(blocks), dim3(256), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_serve<2, RP>), dim3(blocks), dim3(256), 0, s, a); break;
        case 4: hipLaunchKernelGGL((k_serve<4, RP>), dim3(blocks), dim3(256), 0, s, a); break;
        case 8: hipLaunchKernelGGL((k_serve<8, RP>), dim3(blocks), dim3(256), 0, s, a); break;
        case 16: hipLaunchKernelGGL((k_serve<16, RP>), dim3(blocks), dim3(256), 0, s, a); break;
        default: return PG_ERR_INVALID;
    }
    return hipGetLastError() == hipSuccess ? PG_OK : PG_ERR_HIP;
}

static int launch_serve(uint32_t P, bool rp, const ServeArgs &a, hipStream_t s) {
    return rp ? launch_serve_rp<true>(P, a, s) : launch_serve_rp<false>(P, a, s);
}

}  // namespace pg

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

constexpr uint64_t kAlign = 256;
constexpr uint32_t kMaxBatchReqs = 4096;
constexpr uint64_t kMaxIdx = (1ull << 24) >> pg::kIdxShift;  // index capacity: batches up to 16 M keys

inline uint64_t al(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

struct ServeReq {
    uint64_t ticket;
    uint32_t table, op;
    const uint64_t *h_keys;
    const void *h_vals;
    uint64_t n;
    void *h_vals_out;
    uint8_t *h_flags;
    uint32_t vbytes;   // value width of the table (8 KV, 2 filter)
    uint64_t need;     // staging bytes (0 when used in place)
    bool in_place;     // every buffer is GPU-addressable: no staging
    pg::ReqDev dev;    // in place: the buffers as the kernel sees them
};

// The address the GPU uses for p: device memory as is, pinned host memory
// through its mapping; false for memory the GPU cannot address (pageable).
bool gpu_view(const void *p, void **out) {
    *out = nullptr;
    if (!p) return true;
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (a.type == hipMemoryTypeDevice) {
        *out = const_cast<void *>(p);
        return true;
    }
    if (a.type == hipMemoryTypeHost && a.devicePointer) {
        const char *hp = a.hostPointer ? (const char *)a.hostPointer : (const char *)p;
        *out = (char *)a.devicePointer + ((const char *)p - hp);
        return true;
    }
    return false;
}

// Classify a request's buffers: in place when all of them are GPU-addressable.
void classify(ServeReq &r) {
    void *k = nullptr, *v = nullptr, *vo = nullptr, *f = nullptr;
    r.in_place = gpu_view(r.h_keys, &k) && (r.op != PG_OP_INSERT || gpu_view(r.h_vals, &v)) &&
                 (r.op != PG_OP_QUERY || gpu_view(r.h_vals_out, &vo)) && gpu_view(r.h_flags, &f);
    r.dev = pg::ReqDev{};
    if (r.in_place) {
        r.dev.keys = (const uint64_t *)k;
        r.dev.vals = r.op == PG_OP_INSERT ? v : nullptr;
        r.dev.vals_out = r.op == PG_OP_QUERY ? vo : nullptr;
        r.dev.flags = (uint8_t *)f;
    }
}

struct ServeBatch {
    std::vector<uint64_t> tickets;
    std::vector<std::pair<uint32_t, uint8_t>> access;  // tables touched (ordered mode only)
    bool tracked = false;                              // launched in ordered mode
    hipEvent_t done = nullptr;
    uint32_t stream = 0;
    bool finished = false;
    size_t nreported = 0;                              // tickets handed to the caller so far
    bool reported() const { return nreported == tickets.size(); }
};

uint64_t req_bytes(const ServeReq &r) {
    uint64_t b = al(r.n * 8);                                     // keys
    if (r.op == PG_OP_INSERT && r.h_vals) b += al(r.n * r.vbytes);
    if (r.h_flags) b += al(r.n);
    if (r.op == PG_OP_QUERY && r.h_vals_out) b += al(r.n * r.vbytes);
    return b;
}

}  // namespace

struct pg_server_s {
    std::vector<pg_handle> handles;
    std::vector<pg::TableDev> h_tables;
    pg::TableDev *d_tables = nullptr;
    uint32_t P = 0;
    bool rp = false;  // a table has a replace-scheme level: serve with that kernel
    char *arena = nullptr;
    pg::Ring ring;
    std::vector<hipStream_t> streams;
    std::vector<int> busy;
    std::vector<char *> h_meta;         // pinned request table per stream: offsets, descriptors, index
    std::deque<ServeReq> queue;
    std::deque<ServeBatch> inflight;
    uint64_t next_ticket = 1;
    uint32_t max_batch_reqs = kMaxBatchReqs;
    uint64_t max_batch_keys = 1ull << 22;  // bounds a batch's time on the device
    uint32_t order = PG_ORDER_NONE;
    std::vector<uint8_t> scratch_access;   // per table, while a batch is formed
    std::vector<uint8_t> scratch_blocked;  // per table, while a batch is formed
    pg::Plan plan;                         // the batch being formed
    // Set when a HIP call failed after a batch's kernel was enqueued: that
    // batch may have been applied, so its requests can be neither retried
    // nor reported, and every later call but destroy returns PG_ERR_HIP.
    bool failed = false;
};

extern "C" {

int pg_server_create(const pg_handle *tables, uint32_t ntables, uint64_t budget_bytes, uint32_t nstreams,
                     pg_server *out) {
    if (!tables || !out || ntables == 0 || nstreams == 0 || nstreams > 16 || budget_bytes < (1u << 20))
        return PG_ERR_INVALID;
    *out = nullptr;
    pg_server_s *s = new (std::nothrow) pg_server_s();
    if (!s) return PG_ERR_NOMEM;
    s->handles.assign(tables, tables + ntables);
    s->h_tables.resize(ntables);
    for (uint32_t i = 0; i < ntables; i++) {
        uint32_t tile = 0;
        const pg::TableDev *T = pg::table_dev(tables[i], &tile);
        // one tile width (<= 16) for every served table; KV64, 16/16 filters
        // and 16/16 key-value slots
        if (!T || (i > 0 && tile != s->P) || tile > 16 ||
            (T->rep != PG_REP_KV64 && T->rep != PG_REP_FP16_V16 && T->rep != PG_REP_KV16)) {
            delete s;
            return PG_ERR_UNSUPPORTED;
        }
        for (uint32_t l = 0; l < T->nlevels; l++)
            if (T->lv[l].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE) s->rp = true;
        s->P = tile;
        s->h_tables[i] = *T;
    }
    int rc = PG_OK;
    if (hipMalloc((void **)&s->d_tables, sizeof(pg::TableDev) * ntables) != hipSuccess ||
        hipMalloc((void **)&s->arena, budget_bytes) != hipSuccess)
        rc = PG_ERR_NOMEM;
    if (rc == PG_OK &&
        hipMemcpy(s->d_tables, s->h_tables.data(), sizeof(pg::TableDev) * ntables, hipMemcpyHostToDevice) !=
            hipSuccess)
        rc = PG_ERR_HIP;
    s->ring.cap = budget_bytes;
    for (uint32_t i = 0; i < nstreams && rc == PG_OK; i++) {
        hipStream_t st;
        char *m = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
            hipHostMalloc((void **)&m, al(8 * (kMaxBatchReqs + 1)) + al(sizeof(pg::ReqDev) * kMaxBatchReqs) +
                                           al(4 * (kMaxIdx + 1))) != hipSuccess) {
            rc = PG_ERR_HIP;
            break;
        }
        s->streams.push_back(st);
        s->busy.push_back(0);
        s->h_meta.push_back(m);
    }
    if (rc != PG_OK) {
        pg_server_destroy(s);
        return rc;
    }
    *out = s;
    return PG_OK;
}

int pg_server_destroy(pg_server s) {
    if (!s) return PG_ERR_INVALID;
    // release everything even if a call fails; report the first failure
    int rc = PG_OK;
    auto note = [&rc](hipError_t e) {
        if (e != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    };
    for (hipStream_t st : s->streams) note(hipStreamSynchronize(st));
    for (auto &b : s->inflight)
        if (b.done) note(hipEventDestroy(b.done));
    for (hipStream_t st : s->streams) note(hipStreamDestroy(st));
    for (auto *m : s->h_meta) note(hipHostFree(m));
    if (s->arena) note(hipFree(s->arena));
    if (s->d_tables) note(hipFree(s->d_tables));
    delete s;
    return rc;
}

int pg_server_submit(pg_server s, uint32_t table, uint32_t op, const uint64_t *h_keys, const void *h_vals,
                     uint64_t n, void *h_vals_out, uint8_t *h_flags, uint64_t *ticket) {
    if (!s || table >= s->handles.size() || op > PG_OP_REMOVE || (n && !h_keys)) return PG_ERR_INVALID;
    if (s->failed) return PG_ERR_HIP;
    const pg::TableDev &T = s->h_tables[table];
    if (op == PG_OP_REMOVE)
        for (uint32_t l = 0; l < T.nlevels; l++)
            if (T.lv[l].scheme == PG_SCHEME_SINGLE_SLOT) return PG_ERR_UNSUPPORTED;
    ServeReq r{};
    r.ticket = s->next_ticket++;
    r.table = table;
    r.op = op;
    r.h_keys = h_keys;
    r.h_vals = h_vals;
    r.n = n;
    r.h_vals_out = h_vals_out;
    r.h_flags = h_flags;
    r.vbytes = T.rep == PG_REP_KV64 ? 8u : 2u;
    classify(r);
    r.need = r.in_place ? 0 : req_bytes(r);
    // a request that can never fit the budget is refused at submission
    const uint64_t overhead = al(sizeof(pg::ReqDev)) + al(16) + kAlign;
    if (r.need + overhead > s->ring.cap) return PG_ERR_NOMEM;
    s->queue.push_back(r);
    if (ticket) *ticket = r.ticket;
    return PG_OK;
}

int pg_server_submit_many(pg_server s, const pg_request *reqs, uint32_t count, uint64_t *tickets,
                          uint32_t *nqueued) {
    if (nqueued) *nqueued = 0;
    if (!s || (count && !reqs)) return PG_ERR_INVALID;
    for (uint32_t i = 0; i < count; i++) {
        const pg_request &r = reqs[i];
        uint64_t t = 0;
        const int rc = pg_server_submit(s, r.table, r.op, r.keys, r.vals, r.n, r.vals_out, r.flags, &t);
        if (rc != PG_OK) return rc;
        if (tickets) tickets[i] = t;
        if (nqueued) *nqueued = i + 1;
    }
    return PG_OK;
}

// Launch queued requests that fit into one batch on a free stream.  The
// batch's request table (key offsets, then descriptors) is built in pinned
// memory and copied with one call; staged requests add their own copies.
static int launch_batch(pg_server s, uint32_t si) {
    const uint64_t free_bytes = s->ring.largest_free();
    const uint64_t fixed = 0;
    // offsets, descriptors and the coarse index of k requests over nk keys
    auto meta_bytes = [](size_t k, uint64_t nk) {
        const uint64_t nidx = (nk >> pg::kIdxShift) + 2;
        return al(8 * (k + 1)) + al(sizeof(pg::ReqDev) * k) + al(4 * (nidx < kMaxIdx + 1 ? nidx : kMaxIdx + 1));
    };
    const bool ordered = s->order == PG_ORDER_TABLE;
    pg::Plan &plan = s->plan;
    pg::plan_batch(s->queue, pg::PlanLimits{free_bytes - fixed, s->max_batch_keys, s->max_batch_reqs, ordered},
                   meta_bytes, s->scratch_access, s->scratch_blocked, plan);
    const std::vector<size_t> &pick = plan.pick;
    std::vector<std::pair<uint32_t, uint8_t>> &touched = plan.access;
    const size_t take = pick.size();
    const uint64_t keys = plan.keys, need = plan.need;
    if (take == 0) return PG_OK;
    const uint64_t meta = meta_bytes(take, keys);
    uint64_t at = 0;
    if (!s->ring.alloc(fixed + meta + need, at)) return PG_OK;
    hipStream_t st = s->streams[si];
    // A failure before the kernel is enqueued applied nothing: let copies
    // already enqueued on the stream finish, return the staging block and
    // leave the requests queued.  Once the kernel is enqueued the batch may
    // have been applied: the server then fails for good (see `failed`).
    auto undo = [&](int rc) {
        (void)hipStreamSynchronize(st);
        s->ring.release_back();
        return rc;
    };
    auto fail = [&](int rc) {
        s->failed = true;
        return rc;
    };
    if (ordered) {
        // earlier batches still in flight that touch a table of this batch
        // in a conflicting way finish (on the device) before this one starts
        for (const ServeBatch &e : s->inflight) {
            if (e.finished) continue;
            bool dep = !e.tracked;  // launched before ordering was switched on: wait for it
            for (const auto &x : e.access) {
                for (const auto &y : touched)
                    if (x.first == y.first && pg::conflicts(x.second, y.second)) dep = true;
                if (dep) break;
            }
            if (dep && hipStreamWaitEvent(st, e.done, 0) != hipSuccess) return undo(PG_ERR_HIP);
        }
    }
    char *base = s->arena + at;
    uint64_t *d_off = (uint64_t *)(base + fixed);
    pg::ReqDev *d_desc = (pg::ReqDev *)(base + fixed + al(8 * (take + 1)));
    char *p = base + fixed + meta;
    char *hm = s->h_meta[si];
    uint64_t *ho = (uint64_t *)hm;
    pg::ReqDev *hd = (pg::ReqDev *)(hm + al(8 * (take + 1)));
    ServeBatch b;
    b.stream = si;
    if (ordered) {
        b.access = std::move(touched);
        b.tracked = true;
    }
    uint64_t total = 0;
    for (size_t i = 0; i < take; i++) {
        const ServeReq &r = s->queue[pick[i]];
        pg::ReqDev &d = hd[i];
        if (r.in_place) {
            d = r.dev;
        } else {
            d = pg::ReqDev{};
            d.keys = (const uint64_t *)p;
            if (r.n && hipMemcpyAsync(p, r.h_keys, r.n * 8, hipMemcpyHostToDevice, st) != hipSuccess)
                return undo(PG_ERR_HIP);
            p += al(r.n * 8);
            if (r.op == PG_OP_INSERT && r.h_vals) {
                d.vals = p;
                if (r.n && hipMemcpyAsync(p, r.h_vals, r.n * r.vbytes, hipMemcpyHostToDevice, st) != hipSuccess)
                    return undo(PG_ERR_HIP);
                p += al(r.n * r.vbytes);
            }
            if (r.h_flags) {
                d.flags = (uint8_t *)p;
                p += al(r.n);
            }
            if (r.op == PG_OP_QUERY && r.h_vals_out) {
                d.vals_out = p;
                p += al(r.n * r.vbytes);
            }
        }
        d.n = r.n;
        d.table = r.table;
        d.op = r.op;
        ho[i] = total;
        total += r.n;
        b.tickets.push_back(r.ticket);
    }
    ho[take] = total;
    // coarse index: the request holding key c << kIdxShift, for every c
    const uint64_t nidx = total ? ((total - 1) >> pg::kIdxShift) + 1 : 0;
    const bool use_idx = take > 1 && nidx <= kMaxIdx;
    const uint64_t idx_at = al(8 * (take + 1)) + al(sizeof(pg::ReqDev) * take);
    uint64_t meta_copy = al(8 * (take + 1)) + sizeof(pg::ReqDev) * take;
    if (use_idx) {
        uint32_t *hi = (uint32_t *)(hm + idx_at);
        uint32_t r = 0;
        for (uint64_t c = 0; c < nidx; c++) {
            const uint64_t k = c << pg::kIdxShift;
            while (ho[r + 1] <= k) r++;
            hi[c] = r;
        }
        meta_copy = idx_at + 4 * nidx;
    }
    if (hipMemcpyAsync(d_off, hm, meta_copy, hipMemcpyHostToDevice, st) != hipSuccess) return undo(PG_ERR_HIP);
    if (total) {
        pg::ServeArgs a{};
        a.tables = s->d_tables;
        a.reqs = d_desc;
        a.off = d_off;
        a.idx = use_idx ? (const uint32_t *)((char *)d_off + idx_at) : nullptr;
        a.nreq = (uint32_t)take;
        a.total = total;
        const int rc = pg::launch_serve(s->P, s->rp, a, st);
        if (rc != PG_OK) return fail(rc);
    }
    for (size_t i = 0; i < take; i++) {
        const ServeReq &r = s->queue[pick[i]];
        if (r.in_place) continue;
        const pg::ReqDev &d = hd[i];
        if (r.h_flags && r.n &&
            hipMemcpyAsync(r.h_flags, d.flags, r.n, hipMemcpyDeviceToHost, st) != hipSuccess)
            return fail(PG_ERR_HIP);
        if (d.vals_out && r.n &&
            hipMemcpyAsync(r.h_vals_out, d.vals_out, r.n * r.vbytes, hipMemcpyDeviceToHost, st) != hipSuccess)
            return fail(PG_ERR_HIP);
    }
    if (hipEventCreateWithFlags(&b.done, hipEventDisableTiming) != hipSuccess) return fail(PG_ERR_HIP);
    if (hipEventRecord(b.done, st) != hipSuccess) {
        (void)hipEventDestroy(b.done);
        return fail(PG_ERR_HIP);
    }
    if (!ordered) {
        s->queue.erase(s->queue.begin(), s->queue.begin() + (long)take);
    } else {  // remove the picked positions (ascending), keep the rest in order
        std::deque<ServeReq> rest;
        size_t k = 0;
        for (size_t i = 0; i < s->queue.size(); i++) {
            if (k < pick.size() && pick[k] == i) k++;
            else rest.push_back(s->queue[i]);
        }
        s->queue.swap(rest);
    }
    s->busy[si] = 1;
    s->inflight.push_back(std::move(b));
    return PG_OK;
}

int pg_server_progress(pg_server s, uint64_t *done, uint32_t cap, uint32_t *ndone) {
    if (ndone) *ndone = 0;
    if (!s) return PG_ERR_INVALID;
    if (s->failed) return PG_ERR_HIP;
    uint32_t nd = 0;
    // retire: report a batch as soon as it finishes (any order); free its
    // staging memory in launch order
    for (auto &b : s->inflight) {
        if (!b.finished) {
            const hipError_t q = hipEventQuery(b.done);
            if (q == hipSuccess) {
                b.finished = true;
                s->busy[b.stream] = 0;
            } else if (q != hipErrorNotReady) {
                return PG_ERR_HIP;
            }
        }
        // a batch with more tickets than fit is reported over several calls
        while (b.finished && !b.reported() && (nd < cap || !done)) {
            if (done) done[nd++] = b.tickets[b.nreported];
            b.nreported++;
        }
    }
    while (!s->inflight.empty() && s->inflight.front().finished && s->inflight.front().reported()) {
        if (hipEventDestroy(s->inflight.front().done) != hipSuccess) return PG_ERR_HIP;
        s->ring.release_front();
        s->inflight.pop_front();
    }
    // admit and launch on every idle stream
    for (uint32_t si = 0; si < s->streams.size() && !s->queue.empty(); si++) {
        if (s->busy[si]) continue;
        const int rc = launch_batch(s, si);
        if (rc != PG_OK) return rc;
    }
    if (ndone) *ndone = nd;
    return PG_OK;
}

int pg_server_set_order(pg_server s, uint32_t order) {
    if (!s || order > PG_ORDER_TABLE) return PG_ERR_INVALID;
    s->order = order;
    s->scratch_access.assign(s->handles.size(), 0);
    s->scratch_blocked.assign(s->handles.size(), 0);
    return PG_OK;
}

int pg_server_set_batch(pg_server s, uint64_t max_batch_keys, uint32_t max_batch_requests) {
    if (!s || max_batch_keys == 0 || max_batch_requests == 0 || max_batch_requests > kMaxBatchReqs)
        return PG_ERR_INVALID;
    s->max_batch_keys = max_batch_keys;
    s->max_batch_reqs = max_batch_requests;
    return PG_OK;
}

int pg_server_pending(pg_server s, uint64_t *queued, uint64_t *inflight) {
    if (!s) return PG_ERR_INVALID;
    if (queued) *queued = s->queue.size();
    if (inflight) {
        uint64_t n = 0;
        for (auto &b : s->inflight) n += b.tickets.size();
        *inflight = n;
    }
    return PG_OK;
}

int pg_server_run_plain(pg_server s, uint32_t table, uint32_t op, const uint64_t *h_keys, const void *h_vals,
                        uint64_t n, void *h_vals_out, uint8_t *h_flags) {
    if (!s || table >= s->handles.size() || op > PG_OP_REMOVE || (n && !h_keys)) return PG_ERR_INVALID;
    if (s->failed) return PG_ERR_HIP;
    if (!s->inflight.empty() || !s->queue.empty()) return PG_ERR_INVALID;  // not while batches run
    const pg::TableDev &T = s->h_tables[table];
    const uint32_t vb = T.rep == PG_REP_KV64 ? 8u : 2u;
    ServeReq r{};
    r.op = op;
    r.n = n;
    r.h_vals = h_vals;
    r.h_flags = h_flags;
    r.h_vals_out = h_vals_out;
    r.h_keys = h_keys;
    r.vbytes = vb;
    hipStream_t st = s->streams[0];
    classify(r);
    if (r.in_place) {  // one launch on the caller's buffers, then wait
        pg_handle h = s->handles[table];
        int rc;
        if (op == PG_OP_INSERT) rc = pg_bulk_insert(h, r.dev.keys, r.dev.vals, n, r.dev.flags, nullptr, st);
        else if (op == PG_OP_QUERY) rc = pg_bulk_query(h, r.dev.keys, n, r.dev.vals_out, r.dev.flags, st);
        else rc = pg_bulk_remove(h, r.dev.keys, n, r.dev.flags, st);
        if (hipStreamSynchronize(st) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
        return rc;
    }
    uint64_t at = 0;
    if (!s->ring.alloc(req_bytes(r), at)) return PG_ERR_NOMEM;
    char *p = s->arena + at;
    uint64_t *d_keys = (uint64_t *)p;
    p += al(n * 8);
    void *d_vals = nullptr, *d_out = nullptr;
    uint8_t *d_flags = nullptr;
    if (op == PG_OP_INSERT && h_vals) {
        d_vals = p;
        p += al(n * vb);
    }
    if (h_flags) {
        d_flags = (uint8_t *)p;
        p += al(n);
    }
    if (op == PG_OP_QUERY && h_vals_out) d_out = p;
    int rc = PG_OK;
    if (n && hipMemcpyAsync(d_keys, h_keys, n * 8, hipMemcpyHostToDevice, st) != hipSuccess) rc = PG_ERR_HIP;
    if (rc == PG_OK && d_vals && n && hipMemcpyAsync(d_vals, h_vals, n * vb, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = PG_ERR_HIP;
    if (rc == PG_OK) {
        pg_handle h = s->handles[table];
        if (op == PG_OP_INSERT) rc = pg_bulk_insert(h, d_keys, d_vals, n, d_flags, nullptr, st);
        else if (op == PG_OP_QUERY) rc = pg_bulk_query(h, d_keys, n, d_out, d_flags, st);
        else rc = pg_bulk_remove(h, d_keys, n, d_flags, st);
    }
    if (rc == PG_OK && d_flags && n && hipMemcpyAsync(h_flags, d_flags, n, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = PG_ERR_HIP;
    if (rc == PG_OK && d_out && n && hipMemcpyAsync(h_vals_out, d_out, n * vb, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = PG_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess && rc == PG_OK) rc = PG_ERR_HIP;
    s->ring.release_front();
    return rc;
}

}  // extern "C"
