// CPU test of the served-batch planner (poggers_amd/csrc/pg_plan.hpp) on
// random queues, tables, op mixes and limits.  Each plan must respect the
// staging space, key cap and request cap; unordered plans take a prefix of
// the queue; ordered plans take, for every table, a prefix of that table's
// queued requests, and never put two conflicting requests of one table in
// one batch.  Draining the queue plan by plan must take every request, and
// in ordered mode every pair of conflicting requests of one table must land
// in batches in submission order.
#include <cstdio>
#include <deque>
#include <random>
#include <vector>

#include "../../poggers_amd/csrc/pg_plan.hpp"

#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                          \
        }                                                                      \
    } while (0)

struct Req {
    uint32_t table, op;
    uint64_t n, need;
    size_t id;  // submission order
};

static uint64_t meta_bytes(size_t k, uint64_t nk) { return 64 * k + nk / 256 + 256; }

int main() {
    std::mt19937_64 rng(2024);
    for (int trial = 0; trial < 3000; trial++) {
        const uint32_t ntables = 1 + rng() % 8;
        const size_t nreq = 1 + rng() % 300;
        const bool ordered = rng() % 2;
        std::vector<Req> all;
        for (size_t i = 0; i < nreq; i++) {
            const uint64_t u = rng() % 10;
            const uint32_t op = u < 6 ? PG_OP_QUERY : (u < 9 ? PG_OP_INSERT : PG_OP_REMOVE);
            const uint64_t n = rng() % 5000;
            const uint64_t need = (rng() % 3 == 0) ? 0 : (n * 8 + 255) / 256 * 256;  // in place or staged
            all.push_back({(uint32_t)(rng() % ntables), op, n, need, i});
        }
        uint64_t max_need = 0;
        for (const Req &r : all) max_need = r.need > max_need ? r.need : max_need;
        pg::PlanLimits lim;
        lim.max_keys = 1 + rng() % 20000;
        lim.max_reqs = 1 + rng() % 64;
        // always room for any single request, so the queue drains
        lim.free_bytes = max_need + meta_bytes(1, 5000) + rng() % 200000;
        lim.ordered = ordered;

        std::deque<Req> queue(all.begin(), all.end());
        std::vector<uint8_t> acc(ntables, 0), blocked(ntables, 0);
        std::vector<int> batch_of(nreq, -1);
        pg::Plan plan;
        int nbatch = 0;
        while (!queue.empty()) {
            pg::plan_batch(queue, lim, meta_bytes, acc, blocked, plan);
            for (uint32_t t = 0; t < ntables; t++) CHECK(acc[t] == 0 && blocked[t] == 0);
            const size_t take = plan.pick.size();
            CHECK(take >= 1 && take <= lim.max_reqs);
            uint64_t keys = 0, need = 0;
            for (size_t i = 0; i < take; i++) {
                CHECK(plan.pick[i] < queue.size() && (i == 0 || plan.pick[i] > plan.pick[i - 1]));
                keys += queue[plan.pick[i]].n;
                need += queue[plan.pick[i]].need;
            }
            CHECK(keys == plan.keys && need == plan.need);
            CHECK(meta_bytes(take, keys) + need <= lim.free_bytes);
            CHECK(take == 1 || keys <= lim.max_keys);
            if (!ordered) {
                for (size_t i = 0; i < take; i++) CHECK(plan.pick[i] == i);
            } else {
                // per table: the picked requests are the table's first ones
                // in the queue, and no two of them conflict
                std::vector<uint8_t> seen_skip(ntables, 0), a(ntables, 0);
                size_t k = 0;
                for (size_t pos = 0; pos <= plan.pick.back(); pos++) {
                    const Req &r = queue[pos];
                    const bool picked = k < take && plan.pick[k] == pos;
                    if (picked) {
                        k++;
                        CHECK(!seen_skip[r.table]);
                        CHECK(!pg::conflicts(a[r.table], pg::access_of(r.op)));
                        a[r.table] |= pg::access_of(r.op);
                    } else {
                        seen_skip[r.table] = 1;
                    }
                }
                for (const auto &ta : plan.access) CHECK(ta.second == a[ta.first] && ta.second != 0);
                size_t touched = 0;
                for (uint32_t t = 0; t < ntables; t++) touched += a[t] != 0;
                CHECK(touched == plan.access.size());
            }
            std::deque<Req> rest;
            size_t k = 0;
            for (size_t pos = 0; pos < queue.size(); pos++) {
                if (k < take && plan.pick[k] == pos) {
                    batch_of[queue[pos].id] = nbatch;
                    k++;
                } else {
                    rest.push_back(queue[pos]);
                }
            }
            queue.swap(rest);
            nbatch++;
            CHECK(nbatch <= (int)nreq);
        }
        for (size_t i = 0; i < nreq; i++) CHECK(batch_of[i] >= 0);
        if (ordered) {
            for (size_t i = 0; i < nreq; i++)
                for (size_t j = i + 1; j < nreq; j++)
                    if (all[i].table == all[j].table &&
                        pg::conflicts(pg::access_of(all[i].op), pg::access_of(all[j].op)))
                        CHECK(batch_of[i] < batch_of[j]);
        } else {
            for (size_t i = 1; i < nreq; i++) CHECK(batch_of[i] >= batch_of[i - 1]);
        }
    }
    std::printf("plan ok\n");
    return 0;
}
