// Staging of border-padded source windows into LDS with many loads in flight.
#pragma once

#include "common.hpp"

namespace ipfg {

// Copies a chunk of up to 16 rows of the border-padded source into LDS in
// two steps, so that a caller can overlap the global loads of the next chunk
// with work on the current one:
//   load():  every lane issues up to 4 * SJ independent global loads;
//   store(): converts and writes them to LDS (this waits for the loads).
// Wave w owns chunk rows w, w + 4, w + 8, w + 12; lane l owns samples
// l + 64 j (j < SJ), so a row of up to 64 * SJ samples is covered and
// consecutive lanes touch consecutive addresses. The chunk's top-left corner
// is padded pixel (gx0, gy0); sample s of chunk row r lands at
// lds[r * pitch + s] (rows stay interleaved like the image).
template <class Ti, int SJ>
struct Stager {
    static constexpr int kRows = 16;
    Ti val[4][SJ];

    __device__ __forceinline__ void load(const Ti* __restrict__ src, size_t sstride, int W, int H, int cn,
                                         int border, Ti bval, int gx0, int gy0, int nr, int xs, int tid) {
        const int wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
        const bool inside = gx0 >= 0 && (gx0 * cn + xs) <= W * cn && gy0 >= 0 && gy0 + nr <= H;
        if (inside) {
            const char* p0 = reinterpret_cast<const char*>(src) + static_cast<size_t>(gy0 + wave) * sstride +
                             (static_cast<size_t>(gx0) * cn + lane) * sizeof(Ti);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const Ti* rp = reinterpret_cast<const Ti*>(p0 + static_cast<size_t>(4 * i) * sstride);
                if (wave + 4 * i < nr) {
#pragma unroll
                    for (int j = 0; j < SJ; ++j)
                        if (lane + 64 * j < xs) val[i][j] = rp[64 * j];
                }
            }
            return;
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int r = wave + 4 * i;
            if (r >= nr) continue;
            const int gy = gy0 + r;
            const int sy = (gy >= 0 && gy < H) ? gy : border_index(gy, H, border);
            const Ti* rp = reinterpret_cast<const Ti*>(reinterpret_cast<const char*>(src) +
                                                       static_cast<size_t>(sy < 0 ? 0 : sy) * sstride);
#pragma unroll
            for (int j = 0; j < SJ; ++j) {
                const int s = lane + 64 * j;
                if (s >= xs) continue;
                const int px = s / cn, c = s - px * cn;
                const int gx = gx0 + px;
                const int sx = (gx >= 0 && gx < W) ? gx : border_index(gx, W, border);
                val[i][j] = (sy < 0 || sx < 0) ? bval : rp[static_cast<size_t>(sx) * cn + c];
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

}  // namespace ipfg
