// First-fit placement of byte ranges inside one fixed block of device memory
// (host-side bookkeeping only: no HIP calls, so it builds and is tested on the
// CPU). The serving loop carves request buffers out of a block of exactly its
// device budget, so the memory it holds never exceeds the budget and steady
// state runs without hipMalloc / hipFree.
#pragma once

#include <algorithm>
#include <cstddef>
#include <iterator>
#include <map>

namespace ipfg {

class RangeArena {
public:
    static constexpr size_t kAlign = 256;  // offsets and lengths are multiples of this
    static constexpr size_t npos = static_cast<size_t>(-1);

    explicit RangeArena(size_t capacity = 0) { reset(capacity); }

    void reset(size_t capacity) {
        cap_ = capacity / kAlign * kAlign;
        free_.clear();
        if (cap_) free_.emplace(0, cap_);
        used_ = peak_ = 0;
    }

    // Length a request of `bytes` occupies (0 bytes occupy nothing).
    static size_t rounded(size_t bytes) {
        if (bytes == 0) return 0;
        if (bytes > npos - kAlign) return npos;
        return (bytes + kAlign - 1) / kAlign * kAlign;
    }

    // Offset of a newly placed range of rounded(bytes) (lowest offset that
    // fits), or npos when no free range is long enough. A zero-byte request
    // is placed at offset 0 without occupying anything.
    size_t take(size_t bytes) {
        const size_t len = rounded(bytes);
        if (len == 0) return 0;
        if (len == npos) return npos;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            const size_t off = it->first, rest = it->second - len;
            free_.erase(it);
            if (rest) free_.emplace(off + len, rest);
            used_ += len;
            peak_ = std::max(peak_, used_);
            return off;
        }
        return npos;
    }

    // Returns a range placed by take(bytes) (same bytes); neighbouring free
    // ranges merge.
    void give(size_t off, size_t bytes) {
        size_t len = rounded(bytes);
        if (len == 0 || len == npos) return;
        used_ -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        free_.emplace(off, len);
    }

    size_t capacity() const { return cap_; }
    size_t in_use() const { return used_; }
    size_t peak() const { return peak_; }
    size_t free_ranges() const { return free_.size(); }
    size_t largest_free() const {
        size_t m = 0;
        for (const auto& kv : free_) m = std::max(m, kv.second);
        return m;
    }

private:
    size_t cap_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;  // offset -> length; never adjacent
};

}  // namespace ipfg
