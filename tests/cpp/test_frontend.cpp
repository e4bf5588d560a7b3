// Exercises the C++ templated front-end (include/poggers_amd/tables.hpp)
// and the many-table wrappers (include/poggers_amd/serving.hpp).
//   test_frontend cpu   host-only checks (sizing errors, descriptors)
//   test_frontend gpu   bulk ops, a filter batch and the request server
//                       through the C ABI on device 0
// Exit status 0 on success; prints the failing check otherwise.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "poggers_amd/serving.hpp"
#include "poggers_amd/tables.hpp"

using namespace poggers_amd;
using namespace poggers_amd::data_structs;

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                   \
        }                                                               \
    } while (0)

static int cpu_checks() {
    // A recursive table needs two sizes (bucketed_table.cuh:162-176).
    size_in_num_slots<1> one(1 << 12);
    bool threw = false;
    try {
        del_tcf::generate_on_device(&one, 42);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    size_in_num_slots<3> three(1 << 12, 64, 64);
    threw = false;
    try {
        del_tcf::generate_on_device(&three, 42);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    size_in_num_slots<2> two(1000, 10);
    CHECK(two.total() == 1010);
    pg_desc d{};
    del_tcf::describe(d, 0, &two);
    CHECK(d.level[0].nslots == 1000 && d.level[0].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT);
    CHECK(d.level[0].prober == PG_PROBE_XOR_P2 && d.level[0].tile == 4 && d.level[0].bucket_size == 16);
    CHECK(d.level[1].nslots == 10 && d.level[1].scheme == PG_SCHEME_LINEAR_BUCKETS && d.level[1].max_probes == 4000);
    static_assert(del_tcf::levels == 2 && hash_table_64::levels == 1, "levels");
    // the reference's flat TCF (data_structs/tcf.cuh:29-30)
    static_assert(tcf::levels == 2 && tcf::rep == PG_REP_KV16 && tiny_static_table_4::rep == PG_REP_KV16, "flat tcf");
    static_assert(fast_flat_tcf::levels == 2 && fast_flat_tcf::rep == PG_REP_KV16, "fast flat tcf");
    size_in_num_slots<2> two_f(1000, 10);
    tcf::describe(d, 0, &two_f);
    CHECK(d.level[0].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_SLOTS && d.level[0].prober == PG_PROBE_DOUBLE &&
          d.level[0].max_probes == 2 && d.level[0].tile == 4 && d.level[0].bucket_size == 16);
    CHECK(d.level[1].scheme == PG_SCHEME_BUCKET && d.level[1].bucket_size == 4 && d.level[1].max_probes == 20);
    static_assert(small_tcf::rep == PG_REP_FP8_V8 && tcf_10_6::rep == PG_REP_FP10_V6, "packed formats");
    static_assert(tcqf_twelve::rep == PG_REP_FP12 && tcqf_twelve_16::rep == PG_REP_FP12, "12-bit bucket");
    size_in_num_slots<1> one_12(4096);
    tcqf_twelve_16::describe(d, 0, &one_12);
    CHECK(d.level[0].scheme == PG_SCHEME_GROUPED_SHORTCUT && d.level[0].group == 5 && d.level[0].tile == 4);
    size_in_num_slots<1> one_b(4096);
    tcqf_twelve::describe(d, 0, &one_b);
    CHECK(d.level[0].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT && d.level[0].group == 0);
    // power_of_n_insert_shortcut_replace_scheme over key_val_pair slots
    using replace_table = static_table<uint64_t, uint64_t, key_val_pair, 4, 16,
                                       power_of_n_insert_shortcut_replace_scheme, 2, doubleHasher, murmurHasher>;
    size_in_num_slots<1> one_r(4096);
    replace_table::describe(d, 0, &one_r);
    CHECK(d.level[0].scheme == PG_SCHEME_POWER_OF_N_SHORTCUT_REPLACE && d.level[0].bucket_size == 16 &&
          d.level[0].tile == 4 && d.level[0].max_probes == 2);
    size_in_num_slots<2> two_b(1000, 10);  // a sizing is consumed as levels take their sizes
    tcf_10_6::describe(d, 0, &two_b);
    CHECK(d.level[0].tile == 1 && d.level[0].bucket_size == 8);
    CHECK(d.level[0].prober == PG_PROBE_DOUBLE && d.level[1].max_probes == 20);
    std::printf("cpu checks ok\n");
    return 0;
}

// Bulk insert / query / remove through the template front-end.  Keys that
// fail to insert are allowed up to 1 in 10 000 (an overfull probe path, e.g.
// double hashing whose step shares factors with a power-of-two bucket count
// cycles through few buckets); every stored key must then be found.
// tests/src/test_components.cu:54: u64 -> u64, tile 8, 8-slot buckets,
// power_of_n_insert_scheme with three double-hashing probes.
using p2_table = static_table<uint64_t, uint64_t, key_val_pair, 8, 8, power_of_n_insert_scheme, 3, doubleHasher,
                              murmurHasher>;

template <typename Table>
static int roundtrip(const char *name, uint64_t nslots, uint64_t backing, uint64_t n) {
    size_in_num_slots<2> sz2(nslots, backing);
    size_in_num_slots<1> sz1(nslots);
    Table *t = Table::levels == 2 ? Table::generate_on_device(&sz2, 42) : Table::generate_on_device(&sz1, 42);
    uint64_t *keys = nullptr, *absent = nullptr, *fails = nullptr;
    typename Table::val_type *vals = nullptr, *out = nullptr;
    uint8_t *found = nullptr, *status = nullptr;
    CHECK(hipMalloc(&keys, n * 8) == hipSuccess && hipMalloc(&absent, n * 8) == hipSuccess);
    CHECK(hipMalloc(&vals, n * sizeof(*vals)) == hipSuccess && hipMalloc(&out, n * sizeof(*out)) == hipSuccess);
    CHECK(hipMalloc(&found, n) == hipSuccess && hipMalloc(&status, n) == hipSuccess &&
          hipMalloc(&fails, 8) == hipSuccess);
    CHECK(hipMemset(fails, 0, 8) == hipSuccess);
    CHECK(pg_gen_keys(7, 0, n, keys, nullptr) == PG_OK);
    CHECK(pg_gen_keys(7, 1ull << 40, n, absent, nullptr) == PG_OK);
    // values = low bits of the key
    std::vector<uint64_t> hk(n);
    CHECK(hipMemcpy(hk.data(), keys, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<typename Table::val_type> hv(n);
    for (uint64_t i = 0; i < n; i++) hv[i] = (typename Table::val_type)(hk[i] * 3);
    CHECK(hipMemcpy(vals, hv.data(), n * sizeof(*vals), hipMemcpyHostToDevice) == hipSuccess);
    t->bulk_insert(keys, vals, n, status, fails);
    uint64_t nfail = ~0ull;
    CHECK(hipMemcpy(&nfail, fails, 8, hipMemcpyDeviceToHost) == hipSuccess);
    std::vector<uint8_t> hs(n);
    CHECK(hipMemcpy(hs.data(), status, n, hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t nok = 0, nok_lo = 0;
    for (uint64_t i = 0; i < n; i++) {
        nok += hs[i] != 0;
        if (i < n / 2) nok_lo += hs[i] != 0;
    }
    CHECK(nfail == n - nok && nfail <= n / 10000);
    CHECK(t->get_fill() == nok);
    t->bulk_query(keys, n, out, found);
    std::vector<uint8_t> hf(n);
    std::vector<typename Table::val_type> ho(n);
    CHECK(hipMemcpy(hf.data(), found, n, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(ho.data(), out, n * sizeof(*out), hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t fn = 0, wrong = 0;
    for (uint64_t i = 0; i < n; i++) {
        if (!hs[i]) continue;
        fn += hf[i] == 0;
        wrong += ho[i] != hv[i];
    }
    CHECK(fn == 0);
    if (Table::rep == PG_REP_KV64) CHECK(wrong == 0);
    if constexpr (Table::rep == PG_REP_FP12) {
        // 24-byte buckets of 16 slots, five to a 128-byte line; key-only, no remove
        CHECK(t->host_bytes_in_use() == ((t->get_num_slots() / 16 - 1) / 5 + 1) * 128);
    } else {
        const uint64_t slot_bytes = Table::rep == PG_REP_KV64 ? 16
                                    : Table::rep == PG_REP_KEY64  ? 8
                                    : (Table::rep == PG_REP_FP16_V16 || Table::rep == PG_REP_KV16) ? 4 : 2;
        CHECK(t->host_bytes_in_use() == t->get_num_slots() * slot_bytes);
    }
    if (Table::rep == PG_REP_FP8_V8 || Table::rep == PG_REP_FP12) {
        // 8-bit fingerprints collide often: a remove may take another key's
        // word, so the exact post-remove counts below do not apply.
        for (void *p : {(void *)keys, (void *)absent, (void *)vals, (void *)out, (void *)found, (void *)status,
                        (void *)fails})
            CHECK(hipFree(p) == hipSuccess);
        Table::free_on_device(t);
        std::printf("%s ok (%llu keys)\n", name, (unsigned long long)n);
        return 0;
    }
    // A filter remove takes the lowest slot holding the key's fingerprint in
    // its buckets, which may belong to another key with the same fingerprint;
    // that key's own remove (or query) can then miss.  Fill follows the
    // removes that succeeded; exact-key tables must remove exactly.
    const bool exact = Table::rep == PG_REP_KV64 || Table::rep == PG_REP_KEY64;
    t->bulk_remove(keys, n / 2, found);
    std::vector<uint8_t> hr(n / 2);
    CHECK(hipMemcpy(hr.data(), found, n / 2, hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t removed = 0;
    for (uint64_t i = 0; i < n / 2; i++) removed += hr[i] != 0;
    CHECK(t->get_fill() == nok - removed);
    CHECK(exact ? removed == nok_lo : removed + n / 10000 >= nok_lo);
    t->bulk_query(keys, n, out, found);
    CHECK(hipMemcpy(hf.data(), found, n, hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t live = 0;
    for (uint64_t i = n / 2; i < n; i++) live += hf[i] && hs[i];
    CHECK(exact ? live == nok - nok_lo : live + n / 10000 >= nok - nok_lo);
    if (exact) {
        for (uint64_t i = 0; i < n / 2; i++) CHECK(hf[i] == 0);
    }
    CHECK(t->get_num_slots() >= nslots);
    for (void *p : {(void *)keys, (void *)absent, (void *)vals, (void *)out, (void *)found, (void *)status,
                    (void *)fails})
        CHECK(hipFree(p) == hipSuccess);
    Table::free_on_device(t);
    std::printf("%s ok (%llu keys)\n", name, (unsigned long long)n);
    return 0;
}

// filter_batch: 48 filters of uneven sizes, every stored key found, fills
// match; then a graph capture replayed twice gives the same answers.
static int batch_checks() {
    std::vector<uint64_t> sizes;
    for (int t = 0; t < 48; t++) sizes.push_back(4096ull + 997ull * t * t);
    filter_batch<fast_tcf> b(sizes);
    CHECK(b.size() == 48);
    std::vector<uint64_t> off(1, 0);
    for (uint64_t s : sizes) off.push_back(off.back() + s * 8 / 10);
    const uint64_t total = off.back();
    uint64_t *keys = nullptr, *d_off = nullptr, *fails = nullptr;
    uint8_t *status = nullptr, *found = nullptr;
    CHECK(hipMalloc(&keys, total * 8) == hipSuccess && hipMalloc(&d_off, off.size() * 8) == hipSuccess);
    CHECK(hipMalloc(&status, total) == hipSuccess && hipMalloc(&found, total) == hipSuccess &&
          hipMalloc(&fails, 8) == hipSuccess);
    CHECK(hipMemset(fails, 0, 8) == hipSuccess);
    CHECK(pg_gen_keys(31, 0, total, keys, nullptr) == PG_OK);
    CHECK(hipMemcpy(d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice) == hipSuccess);
    b.insert(keys, d_off, total, nullptr, status, fails);
    b.query(keys, d_off, total, nullptr, found);
    std::vector<uint8_t> hs(total), hf(total);
    uint64_t nfail = ~0ull;
    CHECK(hipMemcpy(hs.data(), status, total, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(hf.data(), found, total, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(&nfail, fails, 8, hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t stored = 0;
    for (uint64_t i = 0; i < total; i++) {
        stored += hs[i] != 0;
        CHECK(!hs[i] || hf[i]);
    }
    CHECK(nfail == total - stored && nfail <= total / 10000);
    for (uint32_t t = 0; t < 48; t += 7) {
        uint64_t st = 0;
        for (uint64_t i = off[t]; i < off[t + 1]; i++) st += hs[i] != 0;
        CHECK(b.get_fill(t) == st);
    }
    b.capture(keys, d_off, total, keys, d_off, total, nullptr, found, nullptr, true);
    for (int r = 0; r < 2; r++) {
        CHECK(hipMemset(found, 7, total) == hipSuccess);
        b.replay();
        CHECK(hipDeviceSynchronize() == hipSuccess);
        CHECK(hipMemcpy(hf.data(), found, total, hipMemcpyDeviceToHost) == hipSuccess);
        uint64_t hits = 0;
        for (uint64_t i = 0; i < total; i++) hits += hf[i] == 1;
        CHECK(hits + total / 10000 >= total);
    }
    filter_batch<fast_tcf> moved(std::move(b));
    CHECK(moved.size() == 48 && b.handle() == nullptr);
    for (void *p : {(void *)keys, (void *)d_off, (void *)status, (void *)found, (void *)fails})
        CHECK(hipFree(p) == hipSuccess);
    std::printf("filter_batch ok (%llu keys)\n", (unsigned long long)total);
    return 0;
}

// request_server over a filter and a KV table: with per-table ordering, an
// insert and the query of the same keys submitted together give the
// query its one-by-one answer.
static int server_checks() {
    size_in_num_slots<2> s2(1 << 18, 1 << 12);
    size_in_num_slots<1> s1(1 << 18);
    del_tcf *f = del_tcf::generate_on_device(&s2, 42);
    hash_table_64 *kv = hash_table_64::generate_on_device(&s1, 42);
    {
        request_server srv = request_server::over(1 << 24, 4, f, kv);
        srv.set_order(PG_ORDER_TABLE);
        const uint64_t n = 50000;
        uint64_t *keys = nullptr, *vals = nullptr, *out = nullptr;
        uint16_t *fvals = nullptr, *fout = nullptr;
        uint8_t *ff = nullptr, *kf = nullptr;
        CHECK(hipMalloc(&keys, n * 8) == hipSuccess && hipMalloc(&vals, n * 8) == hipSuccess &&
              hipMalloc(&out, n * 8) == hipSuccess && hipMalloc(&fvals, n * 2) == hipSuccess &&
              hipMalloc(&fout, n * 2) == hipSuccess && hipMalloc(&ff, n) == hipSuccess && hipMalloc(&kf, n) == hipSuccess);
        CHECK(pg_gen_keys(41, 0, n, keys, nullptr) == PG_OK);
        std::vector<uint64_t> hk(n), hv(n);
        std::vector<uint16_t> hfv(n);
        CHECK(hipMemcpy(hk.data(), keys, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
        for (uint64_t i = 0; i < n; i++) {
            hv[i] = hk[i] * 7;
            hfv[i] = (uint16_t)(hk[i] >> 20);
        }
        CHECK(hipMemcpy(vals, hv.data(), n * 8, hipMemcpyHostToDevice) == hipSuccess);
        CHECK(hipMemcpy(fvals, hfv.data(), n * 2, hipMemcpyHostToDevice) == hipSuccess);
        std::vector<uint64_t> done;
        srv.submit(0, PG_OP_INSERT, keys, n, fvals);
        srv.submit(1, PG_OP_INSERT, keys, n, vals);
        srv.submit(0, PG_OP_QUERY, keys, n, nullptr, fout, ff);
        srv.submit(1, PG_OP_QUERY, keys, n, nullptr, out, kf);
        srv.drain(done);
        CHECK(done.size() == 4);
        std::vector<uint8_t> hff(n), hkf(n);
        std::vector<uint64_t> ho(n);
        CHECK(hipMemcpy(hff.data(), ff, n, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(hipMemcpy(hkf.data(), kf, n, hipMemcpyDeviceToHost) == hipSuccess);
        CHECK(hipMemcpy(ho.data(), out, n * 8, hipMemcpyDeviceToHost) == hipSuccess);
        for (uint64_t i = 0; i < n; i++) CHECK(hff[i] == 1 && hkf[i] == 1 && ho[i] == hv[i]);
        CHECK(f->get_fill() == n && kv->get_fill() == n);
        srv.run_plain(1, PG_OP_REMOVE, keys, n / 2, nullptr, nullptr, kf);
        CHECK(kv->get_fill() == n - n / 2);
        for (void *p : {(void *)keys, (void *)vals, (void *)out, (void *)fvals, (void *)fout, (void *)ff, (void *)kf})
            CHECK(hipFree(p) == hipSuccess);
    }
    del_tcf::free_on_device(f);
    hash_table_64::free_on_device(kv);
    std::printf("request_server ok\n");
    return 0;
}

// A recursive composition two backing levels deep: the sizing has three
// entries, and keys that the first two levels refuse land in the third.
using tcf3_l2 = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 4, 8, linear_insert_bucket_scheme,
                               64, linearProber, murmurHasher>;
using tcf3_l1 = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 4, 8, linear_insert_bucket_scheme,
                               16, linearProber, murmurHasher, true, tcf3_l2>;
using tcf3 = bucketed_table<uint64_t, uint16_t, bit_grouped_container<16, 16>, 4, 16,
                            power_of_n_insert_shortcut_bucket_scheme, 2, XORPowerOfTwoHasher, murmurHasher, true,
                            tcf3_l1>;
static_assert(tcf3::levels == 3, "three levels");

static int three_level_checks() {
    size_in_num_slots<3> sz(1 << 12, 128, 512);
    tcf3 *t = tcf3::generate_on_device(&sz, 42);
    const uint64_t n = 4400;  // more than the first two levels hold
    uint64_t *keys = nullptr;
    uint8_t *status = nullptr, *found = nullptr;
    CHECK(hipMalloc(&keys, n * 8) == hipSuccess && hipMalloc(&status, n) == hipSuccess &&
          hipMalloc(&found, n) == hipSuccess);
    CHECK(pg_gen_keys(51, 0, n, keys, nullptr) == PG_OK);
    t->set_mode(PG_DETERMINISTIC);
    t->bulk_insert(keys, nullptr, n, status);
    t->bulk_query(keys, n, nullptr, found);
    std::vector<uint8_t> hs(n), hf(n);
    CHECK(hipMemcpy(hs.data(), status, n, hipMemcpyDeviceToHost) == hipSuccess);
    CHECK(hipMemcpy(hf.data(), found, n, hipMemcpyDeviceToHost) == hipSuccess);
    uint64_t per[4] = {0, 0, 0, 0};
    for (uint64_t i = 0; i < n; i++) {
        CHECK(hs[i] <= 3);
        per[hs[i]]++;
        CHECK(!hs[i] || hf[i]);
    }
    CHECK(per[3] > 0 && t->get_fill() == n - per[0]);
    CHECK(t->get_num_slots() == (1 << 12) + 128 + 512);
    for (void *p : {(void *)keys, (void *)status, (void *)found}) CHECK(hipFree(p) == hipSuccess);
    tcf3::free_on_device(t);
    std::printf("tcf3 ok (%llu keys in the third level)\n", (unsigned long long)per[3]);
    return 0;
}

static int gpu_checks() {
    if (roundtrip<del_tcf>("del_tcf", 1 << 20, 1 << 14, 900000)) return 1;
    if (roundtrip<tcf>("flat tcf", 1 << 20, 1 << 14, 900000)) return 1;
    if (roundtrip<fast_flat_tcf>("fast_flat_tcf", 1 << 20, 1 << 14, 900000)) return 1;
    if (roundtrip<p2_table>("p2_table", 1 << 20, 0, 900000)) return 1;
    if (roundtrip<fast_tcf>("fast_tcf", 1 << 20, 1 << 14, 900000)) return 1;
    if (roundtrip<hash_table_64>("hash_table_64", 1 << 20, 0, 900000)) return 1;
    if (roundtrip<small_tcf>("small_tcf", 1 << 20, 1 << 14, 900000)) return 1;
    if (roundtrip<tcqf_twelve_16>("tcqf_twelve_16", 1 << 20, 0, 700000)) return 1;
    if (roundtrip<cms_free_table>("cms_free_table", 1 << 20, 0, 800000)) return 1;
    if (three_level_checks()) return 1;
    if (batch_checks()) return 1;
    if (server_checks()) return 1;
    return 0;
}

int main(int argc, char **argv) {
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    try {
        if (cpu_checks()) return 1;
        if (gpu && gpu_checks()) return 1;
    } catch (const std::exception &e) {
        std::fprintf(stderr, "FAILED: exception %s\n", e.what());
        return 1;
    }
    return 0;
}
