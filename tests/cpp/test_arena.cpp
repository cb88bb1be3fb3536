// Unit test of the serving loop's HBM-budget allocator (host bookkeeping):
// first fit, 256-byte rounding, coalescing, and refusal when nothing fits.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "arena.h"

#define CHECK(c)                                                        \
    do {                                                                \
        if (!(c)) {                                                     \
            std::printf("FAIL line %d: %s\n", __LINE__, #c);            \
            return 1;                                                   \
        }                                                               \
    } while (0)

int main()
{
    tcv_detail::Arena a;
    a.init(10 * 1024);
    CHECK(a.capacity() == 10 * 1024 && a.used() == 0);
    size_t o1, o2, o3, o4;
    CHECK(a.alloc(1000, &o1) && o1 == 0);          // rounds to 1024
    CHECK(a.alloc(2048, &o2) && o2 == 1024);
    CHECK(a.alloc(3000, &o3) && o3 == 3072);       // 3072 bytes
    CHECK(a.used() == 1024 + 2048 + 3072);
    CHECK(!a.alloc(5000, &o4));                    // 4 KiB left: refused, nothing changes
    CHECK(a.used() == 6144);
    a.release(o2, 2048);                           // hole in the middle
    CHECK(a.free_ranges() == 2);
    CHECK(a.alloc(1024, &o4) && o4 == 1024);       // first fit reuses the hole
    a.release(o4, 1024);
    a.release(o1, 1000);                           // coalesces with the hole
    CHECK(a.free_ranges() == 2);
    a.release(o3, 3000);                           // everything coalesces back
    CHECK(a.free_ranges() == 1 && a.used() == 0);
    CHECK(a.fits_ever(10 * 1024) && !a.fits_ever(10 * 1024 + 1));
    // random churn keeps the bookkeeping consistent
    std::vector<std::pair<size_t, size_t>> live;
    srand(7);
    for (int it = 0; it < 20000; ++it) {
        if (live.empty() || rand() % 2) {
            size_t b = 1 + rand() % 1500, off;
            if (a.alloc(b, &off)) {
                for (auto &l : live)   // no overlap with any live block
                    CHECK(off + tcv_detail::Arena::round(b) <= l.first ||
                          l.first + tcv_detail::Arena::round(l.second) <= off);
                live.push_back({off, b});
            }
        } else {
            size_t k = rand() % live.size();
            a.release(live[k].first, live[k].second);
            live.erase(live.begin() + k);
        }
    }
    for (auto &l : live) a.release(l.first, l.second);
    CHECK(a.used() == 0 && a.free_ranges() == 1);
    std::printf("arena ok\n");
    return 0;
}
