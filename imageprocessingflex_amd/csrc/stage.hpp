// Staging of border-padded source windows into LDS with many loads in flight.
#pragma once

#include "common.hpp"

namespace ipfg {

// Copies a chunk of up to 16 rows of the border-padded source into LDS in
// two steps, so that a caller can overlap the global loads of the next chunk
// with work on the current one:
//   load():  every lane issues 4 * SJ independent global loads;
//   store(): converts and writes them to LDS (this waits for the loads).
// Wave w owns chunk rows w, w + 4, w + 8, w + 12; lane l owns samples
// l + 64 j (j < SJ), so a row of up to 64 * SJ samples is covered and
// consecutive lanes touch consecutive addresses. The chunk's top-left corner
// is padded pixel (gx0, gy0); sample s of chunk row r lands at
// lds[r * pitch + s] (rows stay interleaved like the image).
//
// Every lane issues the same 4 * SJ loads on every call (rows past the chunk
// and samples past the window re-read valid addresses; their values are never
// stored), so no load sits behind a branch and the compiler can wait for each
// chunk's data at its first use while the next chunk's loads are in flight.
// Under BorderMode::Constant the positions outside the image hold arbitrary
// image samples: the caller overwrites them with the border value
// (chunk_outside() says whether a chunk has any).
template <class Ti, int SJ>
struct Stager {
    static constexpr int kRows = 16;
    Ti val[4][SJ];

    __device__ __forceinline__ void load(const Ti* __restrict__ src, size_t sstride, int W, int H, int cn,
                                         int border, int gx0, int gy0, int nr, int xs, int tid) {
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
        if (nr <= 0) return;
        const bool inside = gx0 >= 0 && (gx0 * cn + xs) <= W * cn && gy0 >= 0 && gy0 + nr <= H;
        if (inside) {
            const char* p0 = reinterpret_cast<const char*>(src) + static_cast<size_t>(gy0) * sstride +
                             static_cast<size_t>(gx0) * cn * sizeof(Ti);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int r = min(wave + 4 * i, nr - 1);
                const Ti* rp = reinterpret_cast<const Ti*>(p0 + static_cast<size_t>(r) * sstride);
#pragma unroll
                for (int j = 0; j < SJ; ++j) val[i][j] = rp[min(lane + 64 * j, xs - 1)];
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int gy = gy0 + min(wave + 4 * i, nr - 1);
            const int sy = max(border_index(gy, H, border), 0);  // Constant: any row, overwritten later
            const Ti* rp = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) +
                                                       static_cast<size_t>(sy) * sstride);
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                const int s = min(lane + 64 * j, xs - 1);
                const int px = s / cn, c = s - px * cn;
                const int sx = max(border_index(gx0 + px, W, border), 0);
                val[i][j] = rp[static_cast<size_t>(sx) * cn + c];
            }
        }
    }

    template <class Tl, class Conv>
    __device__ __forceinline__ void store(Tl* lds, int pitch, int nr, int xs, int tid, Conv conv) const {
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = wave + 4 * i;
            if (r >= nr) continue;
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                const int s = lane + 64 * j;
                if (s < xs) lds[r * pitch + s] = conv(val[i][j]);
            }
        }
    }
};

// True when chunk rows [gy0, gy0 + nr) x pixels [gx0, gx0 + npx) reach outside
// the W x H image.
__device__ __forceinline__ bool chunk_outside(int gx0, int gy0, int npx, int nr, int W, int H) {
    return gx0 < 0 || gx0 + npx > W || gy0 < 0 || gy0 + nr > H;
}

}  // namespace ipfg
