// ipf front end for MI355X: host thread control.
//
// Keeps the reference's include/ipf/parallel.hpp surface (setNumThreads,
// getNumThreads 38-51, parallelFor 55-95, parallelForN 99-107): contiguous
// index ranges, a minimum amount of work per thread, nested regions serial,
// the first worker exception rethrown on the caller. Here it drives only host
// work (Image::convertTo, staging); GPU results never depend on it.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <mutex>
#include <vector>

#ifndef IPF_NO_THREADS
#include <system_error>
#include <thread>
#endif

namespace ipf {

namespace detail {
inline std::atomic<int>& threadSetting() {
    static std::atomic<int> setting{0};
    return setting;
}
inline bool& inParallelRegion() {
    thread_local bool nested = false;
    return nested;
}
}  // namespace detail

/// 0 = all hardware threads, 1 = serial.
inline void setNumThreads(int n) { detail::threadSetting() = std::max(n, 0); }

inline int getNumThreads() {
#ifdef IPF_NO_THREADS
    return 1;
#else
    if (const int n = detail::threadSetting(); n > 0) return n;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? static_cast<int>(hw) : 1;
#endif
}

/// fn(i0, i1) over contiguous sub-ranges of [begin, end); threads only when
/// count * costPerItem exceeds 32768 "pixel operations" per thread.
template <class Fn>
void parallelFor(int begin, int end, Fn&& fn, double costPerItem = 1.0) {
    const int n = end - begin;
    if (n <= 0) return;
    const int byWork = std::max(1, static_cast<int>(n * costPerItem / 32768.0));
    const int parts = std::min({getNumThreads(), n, byWork});
    if (parts <= 1 || detail::inParallelRegion()) {
        fn(begin, end);
        return;
    }
#ifdef IPF_NO_THREADS
    fn(begin, end);
#else
    std::exception_ptr first;
    std::mutex mu;
    auto edge = [&](int k) { return begin + static_cast<int>(static_cast<long long>(n) * k / parts); };
    auto body = [&](int a, int b) {
        detail::inParallelRegion() = true;
        try {
            fn(a, b);
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            if (!first) first = std::current_exception();
        }
        detail::inParallelRegion() = false;
    };
    std::vector<std::thread> pool;
    pool.reserve(static_cast<std::size_t>(parts - 1));
    for (int k = 1; k < parts; ++k) {
        try {
            pool.emplace_back(body, edge(k), edge(k + 1));
        } catch (const std::system_error&) {
            body(edge(k), edge(k + 1));  // no thread available: run the chunk here
        }
    }
    body(edge(0), edge(1));
    for (auto& t : pool) t.join();
    if (first) std::rethrow_exception(first);
#endif
}

/// parallelFor over [0, n) in chunks of 4096 items.
template <class Fn>
void parallelForN(std::size_t n, Fn&& fn, double costPerItem = 1.0) {
    constexpr std::size_t kChunk = 4096;
    const std::size_t chunks = (n + kChunk - 1) / kChunk;
    if (chunks == 0) return;
    parallelFor(0, static_cast<int>(chunks), [&](int a, int b) {
        fn(static_cast<std::size_t>(a) * kChunk, std::min(n, static_cast<std::size_t>(b) * kChunk));
    }, costPerItem * static_cast<double>(kChunk));
}

}  // namespace ipf
