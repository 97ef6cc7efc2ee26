// CPU checks of the serving loop's range placement (csrc/arena.hpp) against
// a brute-force byte map: placed ranges never overlap and stay inside the
// capacity, the lowest fitting offset is chosen, freed neighbours merge, a
// request that fits in the free bytes of an empty arena is always placed, and
// in-use / peak bytes are exact. Prints "ARENA OK" on success.
#include "../../imageprocessingflex_amd/csrc/arena.hpp"

#include <cstdio>
#include <cstdint>
#include <random>
#include <vector>

namespace {

int g_fail = 0;
#define EXPECT(c)                                                    \
    do {                                                             \
        if (!(c)) {                                                  \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++g_fail;                                                \
        }                                                            \
    } while (0)

using ipfg::RangeArena;

// Lowest offset (multiple of kAlign) of `len` free units in the unit map.
size_t model_first_fit(const std::vector<uint8_t>& used, size_t len) {
    size_t run = 0;
    for (size_t u = 0; u < used.size(); ++u) {
        run = used[u] ? 0 : run + 1;
        if (run == len) return u + 1 - len;
    }
    return RangeArena::npos;
}

void basics() {
    RangeArena a(12 * RangeArena::kAlign + 17);  // the tail below one unit is unusable
    EXPECT(a.capacity() == 12 * RangeArena::kAlign);
    EXPECT(RangeArena::rounded(0) == 0 && RangeArena::rounded(1) == 256 && RangeArena::rounded(256) == 256 &&
           RangeArena::rounded(257) == 512);
    EXPECT(a.take(0) == 0 && a.in_use() == 0);  // zero bytes occupy nothing
    const size_t x = a.take(300), y = a.take(256), z = a.take(2000);
    EXPECT(x == 0 && y == 512 && z == 768);
    EXPECT(a.in_use() == 512 + 256 + 2048);
    EXPECT(a.take(2 * 256) == RangeArena::npos);  // 12 - 2 - 1 - 8 = 1 unit left
    a.give(y, 256);
    EXPECT(a.take(100) == 512);  // the hole is reused first
    a.give(512, 100);
    a.give(x, 300);
    EXPECT(a.largest_free() == 768);  // merged [0, 768)
    a.give(z, 2000);
    EXPECT(a.free_ranges() == 1 && a.largest_free() == a.capacity() && a.in_use() == 0);
    EXPECT(a.peak() == 512 + 256 + 2048);
    EXPECT(a.take(a.capacity()) == 0);
    EXPECT(a.take(1) == RangeArena::npos);
    a.give(0, a.capacity());
    EXPECT(a.take(a.capacity() + 1) == RangeArena::npos);
    RangeArena empty(0);
    EXPECT(empty.take(1) == RangeArena::npos && empty.take(0) == 0);
}

void random_against_model(unsigned seed) {
    std::mt19937_64 rng(seed);
    const size_t units = 1 + rng() % 400;
    RangeArena a(units * RangeArena::kAlign);
    std::vector<uint8_t> used(units, 0);
    struct Live {
        size_t off, bytes;
    };
    std::vector<Live> live;
    size_t peak = 0, inuse = 0;
    for (int step = 0; step < 4000; ++step) {
        if (live.empty() || rng() % 3) {
            const size_t bytes = 1 + rng() % (RangeArena::kAlign * (1 + units / 4));
            const size_t len = RangeArena::rounded(bytes) / RangeArena::kAlign;
            const size_t want = model_first_fit(used, len);
            const size_t got = a.take(bytes);
            EXPECT(got == (want == RangeArena::npos ? RangeArena::npos : want * RangeArena::kAlign));
            if (got == RangeArena::npos) continue;
            for (size_t u = got / RangeArena::kAlign; u < got / RangeArena::kAlign + len; ++u) {
                EXPECT(u < units && !used[u]);
                used[u] = 1;
            }
            live.push_back({got, bytes});
            inuse += len * RangeArena::kAlign;
            peak = std::max(peak, inuse);
        } else {
            const size_t k = rng() % live.size();
            const Live l = live[k];
            live[k] = live.back();
            live.pop_back();
            a.give(l.off, l.bytes);
            const size_t len = RangeArena::rounded(l.bytes) / RangeArena::kAlign;
            for (size_t u = l.off / RangeArena::kAlign; u < l.off / RangeArena::kAlign + len; ++u) used[u] = 0;
            inuse -= len * RangeArena::kAlign;
        }
        EXPECT(a.in_use() == inuse && a.peak() == peak);
        if (g_fail) return;
    }
    for (const Live& l : live) a.give(l.off, l.bytes);
    EXPECT(a.in_use() == 0 && a.free_ranges() == 1 && a.largest_free() == a.capacity());
}

}  // namespace

int main() {
    basics();
    for (unsigned s = 1; s <= 40 && !g_fail; ++s) random_against_model(s);
    if (g_fail) {
        std::printf("%d failures\n", g_fail);
        return 1;
    }
    std::printf("ARENA OK\n");
    return 0;
}
