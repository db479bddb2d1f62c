// Randomised check of the server arena's range allocator against a byte map
// of which ranges are taken: every allocation lies inside the arena, never
// overlaps another, is found whenever a large enough free run exists (first
// fit), and the free list stays merged.  Built with the address and
// undefined-behaviour sanitizers by tests/test_native.py.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "range_alloc.h"

using trec::RangeAllocator;

int main()
{
    const size_t A = RangeAllocator::ALIGN, N = 512;  // arena of N slots of ALIGN bytes
    RangeAllocator r;
    r.reset(N * A);
    std::vector<int> owner(N, -1);
    std::vector<std::pair<size_t, size_t>> live;  // offset, slots
    std::mt19937_64 rng(1234);
    for (int step = 0; step < 200000; ++step) {
        if (live.empty() || rng() % 100 < 55) {
            const size_t want = 1 + rng() % (rng() % 4 == 0 ? 200 : 24);
            const size_t bytes = want * A - (rng() % A);  // align_up gives `want` slots
            // first fit in the byte map
            size_t first = SIZE_MAX, run = 0;
            for (size_t k = 0; k < N; ++k) {
                run = owner[k] < 0 ? run + 1 : 0;
                if (run == want) {
                    first = k + 1 - want;
                    break;
                }
            }
            const size_t off = r.alloc(bytes);
            if (first == SIZE_MAX) {
                if (off != RangeAllocator::NONE || r.fits(bytes)) {
                    std::printf("step %d: allocated %zu slots with no free run\n", step, want);
                    return 1;
                }
                continue;
            }
            if (off != first * A) {
                std::printf("step %d: offset %zu, first fit %zu\n", step, off, first * A);
                return 1;
            }
            for (size_t k = first; k < first + want; ++k) owner[k] = step;
            live.push_back({off, want});
        } else {
            const size_t i = rng() % live.size();
            const auto [off, want] = live[i];
            r.release(off);
            r.release(off);  // a second release of the same offset is ignored
            for (size_t k = off / A; k < off / A + want; ++k) owner[k] = -1;
            live[i] = live.back();
            live.pop_back();
        }
        if (!r.consistent()) {
            std::printf("step %d: free and used ranges inconsistent\n", step);
            return 1;
        }
    }
    for (auto &l : live) r.release(l.first);
    if (!r.consistent() || !r.fits(N * A) || r.fits(N * A + 1)) {
        std::printf("not whole after releasing everything\n");
        return 1;
    }
    std::printf("range allocator: 200000 random steps consistent\n");
    return 0;
}
