// CPU test of the serving staging ring (poggers_amd/csrc/pg_ring.hpp): random
// allocation / in-order release sequences, with some allocations undone at
// once (release_back, a launch that failed before its kernel), against a
// byte-occupancy model.  Checks: blocks never overlap and stay inside the
// arena, largest_free() equals what alloc() can actually place, an undo
// restores the exact state before its allocation, and an allocation that
// fits some free run at the head or at the arena start never fails.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../poggers_amd/csrc/pg_ring.hpp"

#define CHECK(c)                                                               \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                          \
        }                                                                      \
    } while (0)

int main() {
    std::mt19937_64 rng(12345);
    for (int trial = 0; trial < 200; trial++) {
        const uint64_t cap = 1 + rng() % 5000;
        pg::Ring r;
        r.cap = cap;
        std::vector<int> owner(cap, -1);  // which live block holds each byte
        int next_id = 0;
        std::vector<std::pair<int, std::pair<uint64_t, uint64_t>>> fifo;
        for (int step = 0; step < 2000; step++) {
            const bool do_alloc = fifo.empty() || (rng() % 100) < 55;
            if (do_alloc) {
                const uint64_t need = 1 + rng() % (cap / 2 + 1);
                const uint64_t lf = r.largest_free();
                uint64_t at = 0;
                const bool ok = r.alloc(need, at);
                CHECK(ok == (need <= lf));
                if (!ok) continue;
                CHECK(at + need <= cap);
                if (rng() % 10 == 0) {  // the launch failed: undo this block
                    r.release_back();
                    CHECK(r.largest_free() == lf);
                    uint64_t at2 = 0;
                    CHECK(r.alloc(need, at2) && at2 == at);  // the same block is found again
                    r.release_back();
                    continue;
                }
                for (uint64_t b = at; b < at + need; b++) {
                    CHECK(owner[b] == -1);
                    owner[b] = next_id;
                }
                fifo.push_back({next_id, {at, at + need}});
                next_id++;
            } else {
                const auto blk = fifo.front();
                fifo.erase(fifo.begin());
                CHECK(r.live.front().first == blk.second.first && r.live.front().second == blk.second.second);
                r.release_front();
                for (uint64_t b = blk.second.first; b < blk.second.second; b++) {
                    CHECK(owner[b] == blk.first);
                    owner[b] = -1;
                }
            }
            // with nothing live the whole arena is free again
            if (fifo.empty()) CHECK(r.largest_free() == cap);
        }
    }
    std::printf("ring ok\n");
    return 0;
}
