// poggers_amd/serving.hpp -- C++ wrappers of the many-table entry points of
// the C ABI: batches of filters (pg_batch_*) and the request server
// (pg_server_*).  Both own their C handle (released by the destructor,
// movable, not copyable) and turn status codes into pg_error exceptions, as
// the table front-end in tables.hpp does.  Neither has a reference
// counterpart: the reference launches one kernel per table per operation
// (tests/src/test_mhm_tcf.cu:318-487) and has no host request path.
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "tables.hpp"

namespace poggers_amd {

// The descriptor of a table type with the given level sizes (the backing
// size is ignored by single-level types).
template <typename Table>
pg_desc table_desc(uint64_t primary_slots, uint64_t backing_slots = 0) {
    pg_desc d{};
    d.rep = Table::rep;
    d.num_levels = Table::levels;
    if constexpr (Table::levels == 2) {
        size_in_num_slots<2> sizing(primary_slots, backing_slots);
        Table::describe(d, 0, &sizing);
    } else {
        size_in_num_slots<1> sizing(primary_slots);
        Table::describe(d, 0, &sizing);
    }
    return d;
}

// Many filters of one type and uneven sizes in one arena, updated by one
// persistent kernel per call (pg_batch_*).  Keys of filter t are
// keys[offsets[t] .. offsets[t+1]) (device arrays, CSR style).
template <typename Filter>
class filter_batch {
   public:
    static_assert(Filter::rep == PG_REP_FP16_V16, "batches hold 16/16 filters");
    using val_type = typename Filter::val_type;

    // backing_slots empty: each filter's backing level holds 1 % of its slots.
    explicit filter_batch(const std::vector<uint64_t> &nslots, uint64_t seed = 42,
                          const std::vector<uint64_t> &backing_slots = {}) {
        if (nslots.empty()) throw std::invalid_argument("filter_batch: no filters");
        if (!backing_slots.empty() && backing_slots.size() != nslots.size())
            throw std::invalid_argument("filter_batch: one backing size per filter");
        const pg_desc proto = table_desc<Filter>(4096, 64);
        check(pg_batch_create(&proto, (uint32_t)nslots.size(), nslots.data(),
                              backing_slots.empty() ? nullptr : backing_slots.data(), seed, &b_),
              "pg_batch_create");
        n_ = (uint32_t)nslots.size();
    }
    ~filter_batch() {
        if (b_) pg_batch_destroy(b_);
    }
    filter_batch(const filter_batch &) = delete;
    filter_batch &operator=(const filter_batch &) = delete;
    filter_batch(filter_batch &&o) noexcept : b_(std::exchange(o.b_, nullptr)), n_(o.n_) {}
    filter_batch &operator=(filter_batch &&o) noexcept {
        if (this != &o) {
            if (b_) pg_batch_destroy(b_);
            b_ = std::exchange(o.b_, nullptr);
            n_ = o.n_;
        }
        return *this;
    }

    uint32_t size() const { return n_; }
    void set_mode(uint32_t mode) { check(pg_batch_set_mode(b_, mode), "pg_batch_set_mode"); }
    void set_launch(uint32_t chunk, uint32_t blocks = 0) {
        check(pg_batch_set_launch(b_, chunk, blocks), "pg_batch_set_launch");
    }
    void set_items(uint32_t items) { check(pg_batch_set_items(b_, items), "pg_batch_set_items"); }
    void clear(void *stream = nullptr) { check(pg_batch_clear(b_, stream), "pg_batch_clear"); }

    void insert(const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total, const val_type *d_vals = nullptr,
                uint8_t *d_status = nullptr, uint64_t *d_fail_count = nullptr, void *stream = nullptr) {
        check(pg_batch_insert(b_, d_keys, d_offsets, total, d_vals, d_status, d_fail_count, stream),
              "pg_batch_insert");
    }
    void query(const uint64_t *d_keys, const uint64_t *d_offsets, uint64_t total, val_type *d_vals_out = nullptr,
               uint8_t *d_found = nullptr, void *stream = nullptr) {
        check(pg_batch_query(b_, d_keys, d_offsets, total, d_vals_out, d_found, stream), "pg_batch_query");
    }
    // Record clear (optional) + insert + query in a hipGraph; replay() relaunches it.
    void capture(const uint64_t *d_ins_keys, const uint64_t *d_ins_offsets, uint64_t ins_total,
                 const uint64_t *d_q_keys, const uint64_t *d_q_offsets, uint64_t q_total,
                 uint8_t *d_status = nullptr, uint8_t *d_found = nullptr, uint64_t *d_fail_count = nullptr,
                 bool clear_first = true, void *stream = nullptr) {
        check(pg_batch_capture(b_, d_ins_keys, d_ins_offsets, ins_total, d_q_keys, d_q_offsets, q_total, d_status,
                               d_found, d_fail_count, clear_first ? 1u : 0u, stream),
              "pg_batch_capture");
    }
    void replay(void *stream = nullptr) { check(pg_batch_replay(b_, stream), "pg_batch_replay"); }

    uint64_t get_fill(uint32_t filter, void *stream = nullptr) {
        uint64_t v = 0;
        check(pg_batch_fill(b_, filter, &v, stream), "pg_batch_fill");
        return v;
    }
    uint64_t host_bytes_in_use() {
        uint64_t v = 0;
        check(pg_batch_bytes(b_, &v), "pg_batch_bytes");
        return v;
    }
    pg_batch handle() const { return b_; }

   private:
    pg_batch b_ = nullptr;
    uint32_t n_ = 0;
};

// The continuous-batching request loop (pg_server_*) over resident tables.
// Buffers the GPU can address (device memory, pinned host memory) are used
// in place; pageable ones are staged through the HBM budget.
class request_server {
   public:
    template <typename... Tables>
    static request_server over(uint64_t budget_bytes, uint32_t nstreams, Tables *...tables) {
        return request_server(std::vector<pg_handle>{tables->handle()...}, budget_bytes, nstreams);
    }
    request_server(const std::vector<pg_handle> &tables, uint64_t budget_bytes, uint32_t nstreams) {
        check(pg_server_create(tables.data(), (uint32_t)tables.size(), budget_bytes, nstreams, &s_),
              "pg_server_create");
    }
    ~request_server() {
        if (s_) pg_server_destroy(s_);
    }
    request_server(const request_server &) = delete;
    request_server &operator=(const request_server &) = delete;
    request_server(request_server &&o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
    request_server &operator=(request_server &&o) noexcept {
        if (this != &o) {
            if (s_) pg_server_destroy(s_);
            s_ = std::exchange(o.s_, nullptr);
        }
        return *this;
    }

    // Queue one request; returns its ticket.
    uint64_t submit(uint32_t table, uint32_t op, const uint64_t *keys, uint64_t n, const void *vals = nullptr,
                    void *vals_out = nullptr, uint8_t *flags = nullptr) {
        uint64_t ticket = 0;
        check(pg_server_submit(s_, table, op, keys, vals, n, vals_out, flags, &ticket), "pg_server_submit");
        return ticket;
    }
    // One engine step; appends the finished tickets to `done`, returns how many.
    uint32_t progress(std::vector<uint64_t> &done) {
        uint64_t buf[1024];
        uint32_t nd = 0;
        check(pg_server_progress(s_, buf, 1024, &nd), "pg_server_progress");
        done.insert(done.end(), buf, buf + nd);
        return nd;
    }
    // Steps until every queued and in-flight request has finished.
    void drain(std::vector<uint64_t> &done) {
        for (;;) {
            progress(done);
            uint64_t queued = 0, inflight = 0;
            check(pg_server_pending(s_, &queued, &inflight), "pg_server_pending");
            if (queued == 0 && inflight == 0) return;
        }
    }
    void set_batch(uint64_t max_batch_keys, uint32_t max_batch_requests = 4096) {
        check(pg_server_set_batch(s_, max_batch_keys, max_batch_requests), "pg_server_set_batch");
    }
    void set_order(uint32_t order) { check(pg_server_set_order(s_, order), "pg_server_set_order"); }
    // One request with its own launch, waited for (the per-request baseline).
    void run_plain(uint32_t table, uint32_t op, const uint64_t *keys, uint64_t n, const void *vals = nullptr,
                   void *vals_out = nullptr, uint8_t *flags = nullptr) {
        check(pg_server_run_plain(s_, table, op, keys, vals, n, vals_out, flags), "pg_server_run_plain");
    }
    pg_server handle() const { return s_; }

   private:
    pg_server s_ = nullptr;
};

}  // namespace poggers_amd
