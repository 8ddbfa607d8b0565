// Host check of the server's arena allocator (seal-cuda_amd/csrc/range_alloc.h) against
// a granule-by-granule model: blocks never overlap and stay inside the arena, every
// granule is free or taken exactly once, free neighbours coalesce (a fully released
// arena is one block again), first fit takes the lowest offset that fits, and a request
// fails only when no free run is large enough.
#include "range_alloc.h"
#include <cstdio>
#include <random>
#include <vector>

using shc::RangeAllocator;

static int fails = 0;
#define CHECK(c, msg)                                                 \
    do {                                                              \
        if (!(c)) {                                                   \
            std::printf("FAIL %s (line %d)\n", msg, __LINE__);        \
            fails++;                                                  \
        }                                                             \
    } while (0)

int main() {
    const size_t G = RangeAllocator::GRAIN, granules = 4096;
    RangeAllocator a(granules * G + 100);  // the tail below one granule is unusable
    CHECK(a.capacity() == granules * G, "capacity rounds down to granules");
    std::vector<int> owner(granules, -1);  // model: which live block holds each granule
    struct Blk { size_t off, size; };
    std::vector<Blk> live;
    std::mt19937_64 rng(5);
    for (int step = 0; step < 200000 && !fails; step++) {
        if (live.empty() || rng() % 3) {
            const size_t want = 1 + rng() % (64 * G);
            const size_t need = (want + G - 1) / G;
            // the lowest offset of a free run of `need` granules in the model
            size_t first = SIZE_MAX, run = 0, largest = 0;
            for (size_t g = 0; g < granules; g++) {
                run = owner[g] < 0 ? run + 1 : 0;
                largest = std::max(largest, run);
                if (run == need && first == SIZE_MAX) first = g + 1 - need;
            }
            CHECK(a.largest_free() == largest * G, "largest free run");
            const size_t off = a.take(want);
            if (first == SIZE_MAX) {
                CHECK(off == SIZE_MAX, "take fails only when nothing fits");
                continue;
            }
            CHECK(off == first * G, "first fit");
            if (off == SIZE_MAX) continue;
            for (size_t g = off / G; g < off / G + need; g++) {
                CHECK(owner[g] < 0, "blocks never overlap");
                owner[g] = (int)live.size();
            }
            live.push_back({off, need});
        } else {
            const size_t k = rng() % live.size();
            a.give(live[k].off);
            for (size_t g = live[k].off / G; g < live[k].off / G + live[k].size; g++) owner[g] = -1;
            live[k] = live.back();
            live.pop_back();
            for (size_t g = 0; g < granules; g++)
                if (owner[g] >= 0) owner[g] = -2;  // re-tag below
            for (size_t b = 0; b < live.size(); b++)
                for (size_t g = live[b].off / G; g < live[b].off / G + live[b].size; g++) owner[g] = (int)b;
        }
        size_t used = 0;
        for (const Blk &b : live) used += b.size * G;
        CHECK(a.in_use() == used, "bytes in use");
    }
    for (const Blk &b : live) a.give(b.off);
    CHECK(a.in_use() == 0 && a.free_blocks() == 1 && a.largest_free() == granules * G, "full release coalesces");
    bool threw = false;
    try {
        a.give(0);
    } catch (const std::logic_error &) {
        threw = true;
    }
    CHECK(threw, "releasing a block that is not taken throws");
    if (fails) return 1;
    std::printf("ok range allocator\n");
    return 0;
}
