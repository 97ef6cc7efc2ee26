// CPU checks of include/ipf/parallel.hpp (the analogue of the reference's
// tests/test_parallel.cpp:23-50): every index covered once for any thread
// count, nested regions serial on the thread that opened them, the first
// exception rethrown on the caller, regions opened from several user threads
// at once, parallelForN's block-aligned ranges, and serial execution of small
// work. Prints "PARALLEL OK" on success; exits non-zero on the first failure.
#include <ipf/parallel.hpp>

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

int g_fail = 0;
#define EXPECT(c)                                                          \
    do {                                                                   \
        if (!(c)) {                                                        \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);       \
            ++g_fail;                                                      \
        }                                                                  \
    } while (0)

void coverage(int threads) {
    ipf::setNumThreads(threads);
    for (int n : {1, 2, 7, 100, 4097, 100000}) {
        for (int begin : {0, -13, 1000}) {
            std::vector<std::atomic<int>> hits(static_cast<size_t>(n));
            for (auto& h : hits) h = 0;
            std::atomic<int> calls{0};
            ipf::parallelFor(begin, begin + n, [&](int a, int b) {
                ++calls;
                EXPECT(a < b && a >= begin && b <= begin + n);
                for (int i = a; i < b; ++i) ++hits[static_cast<size_t>(i - begin)];
            }, 1e4);
            for (auto& h : hits) EXPECT(h.load() == 1);
            EXPECT(calls.load() >= 1);
        }
    }
    // an empty or reversed range calls nothing
    bool called = false;
    ipf::parallelFor(5, 5, [&](int, int) { called = true; });
    ipf::parallelFor(5, 2, [&](int, int) { called = true; });
    EXPECT(!called);
}

// A result computed over ranges equals the serial one for every thread count.
void independence() {
    const int n = 200003;
    std::vector<double> want(n), got(n);
    for (int i = 0; i < n; ++i) want[i] = static_cast<double>(i) * 0.5 + (i % 7);
    for (int t : {1, 2, 3, 8, 16, 33}) {
        ipf::setNumThreads(t);
        std::fill(got.begin(), got.end(), -1.0);
        ipf::parallelFor(0, n, [&](int a, int b) {
            for (int i = a; i < b; ++i) got[i] = static_cast<double>(i) * 0.5 + (i % 7);
        }, 50.0);
        EXPECT(got == want);
    }
}

void nested() {
    ipf::setNumThreads(8);
    std::atomic<long long> total{0};
    std::atomic<int> nested_threads_differ{0};
    ipf::parallelFor(0, 64, [&](int a, int b) {
        const auto outer = std::this_thread::get_id();
        for (int i = a; i < b; ++i) {
            int inner_calls = 0;
            ipf::parallelFor(0, 1000, [&](int c, int d) {
                ++inner_calls;
                if (std::this_thread::get_id() != outer) ++nested_threads_differ;
                total += d - c;
            }, 1e6);
            EXPECT(inner_calls == 1);  // one call: the whole range, serially
        }
    }, 1e6);
    EXPECT(total.load() == 64000LL);
    EXPECT(nested_threads_differ.load() == 0);
}

void exceptions() {
    ipf::setNumThreads(6);
    bool caught = false;
    try {
        ipf::parallelFor(0, 100000, [](int a, int) {
            if (a > 0) throw std::runtime_error("boom");
        }, 10.0);
    } catch (const std::runtime_error& e) {
        caught = std::string(e.what()) == "boom";
    }
    EXPECT(caught);
    // the first range also throws: still exactly one exception reaches the caller
    int n_caught = 0;
    try {
        ipf::parallelFor(0, 100000, [](int, int) { throw std::logic_error("all"); }, 10.0);
    } catch (const std::logic_error&) {
        ++n_caught;
    }
    EXPECT(n_caught == 1);
    // the crew is usable after a failed region
    std::atomic<long long> s{0};
    ipf::parallelFor(0, 50000, [&](int a, int b) { s += b - a; }, 10.0);
    EXPECT(s.load() == 50000);
    // serial path (one thread) rethrows too
    ipf::setNumThreads(1);
    caught = false;
    try {
        ipf::parallelFor(0, 10, [](int, int) { throw std::runtime_error("serial"); }, 1e9);
    } catch (const std::runtime_error&) {
        caught = true;
    }
    EXPECT(caught);
}

void flat() {
    ipf::setNumThreads(5);
    for (size_t n : {size_t(1), size_t(4095), size_t(4096), size_t(4097), size_t(1000003)}) {
        std::mutex m;
        size_t covered = 0;
        std::set<size_t> starts;
        ipf::parallelForN(n, [&](size_t a, size_t b) {
            std::lock_guard<std::mutex> g(m);
            covered += b - a;
            starts.insert(a);
            EXPECT(a % 4096 == 0);
            EXPECT(b == n || b % 4096 == 0);
        });
        EXPECT(covered == n);
    }
    bool called = false;
    ipf::parallelForN(0, [&](size_t, size_t) { called = true; });
    EXPECT(!called);
}

// Little work stays on the calling thread as one call.
void small_work_serial() {
    ipf::setNumThreads(8);
    int calls = 0;
    std::thread::id who;
    ipf::parallelFor(0, 100, [&](int a, int b) {
        ++calls;
        who = std::this_thread::get_id();
        EXPECT(a == 0 && b == 100);
    }, 1.0);
    EXPECT(calls == 1 && who == std::this_thread::get_id());
}

// Regions opened from several user threads at once all complete correctly
// (one holds the crew, the others run serially on their own threads).
void concurrent_callers() {
    ipf::setNumThreads(4);
    const int users = 6, n = 300000;
    std::vector<long long> sums(users, 0);
    std::vector<std::thread> ts;
    for (int u = 0; u < users; ++u)
        ts.emplace_back([&, u] {
            for (int rep = 0; rep < 20; ++rep) {
                std::atomic<long long> s{0};
                ipf::parallelFor(0, n, [&](int a, int b) {
                    long long t = 0;
                    for (int i = a; i < b; ++i) t += i % 11;
                    s += t;
                }, 20.0);
                sums[u] += s.load();
            }
        });
    for (auto& t : ts) t.join();
    long long one = 0;
    for (int i = 0; i < n; ++i) one += i % 11;
    for (int u = 0; u < users; ++u) EXPECT(sums[u] == 20 * one);
}

void settings() {
    ipf::setNumThreads(3);
    EXPECT(ipf::getNumThreads() == 3);
    ipf::setNumThreads(-4);
    EXPECT(ipf::getNumThreads() >= 1);
    ipf::setNumThreads(0);
    const unsigned hw = std::thread::hardware_concurrency();
    EXPECT(ipf::getNumThreads() == (hw ? static_cast<int>(hw) : 1));
}

}  // namespace

int main() {
    settings();
    for (int t : {1, 2, 4, 7, 16}) coverage(t);
    independence();
    nested();
    exceptions();
    flat();
    small_work_serial();
    concurrent_callers();
    ipf::setNumThreads(0);
    if (g_fail) {
        std::printf("%d failures\n", g_fail);
        return 1;
    }
    std::printf("PARALLEL OK\n");
    return 0;
}
