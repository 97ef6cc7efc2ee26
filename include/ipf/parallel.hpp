// ipf front end for MI355X: host thread control.
//
// API of the reference's include/ipf/parallel.hpp (setNumThreads and
// getNumThreads 38-51, parallelFor 55-95, parallelForN 99-107): 0 selects all
// hardware threads, work is handed to the callback as contiguous index
// ranges, regions entered from inside a region run serially on the calling
// thread, and the first exception a callback throws is rethrown on the
// caller once every range has finished. Here it drives only host work
// (Image::convertTo, staging); no GPU result depends on it.
//
// Design: a process-wide crew of helper threads that sleep between regions
// and grows on demand up to getNumThreads() - 1. A region cuts its range into
// more pieces than participants (load balance for uneven rows) and every
// participant, the caller included, claims pieces from one atomic cursor.
// A second region started while the crew is busy (from another user thread)
// runs serially on its own thread instead of queueing behind the first.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <exception>
#include <mutex>
#include <vector>

#ifndef IPF_NO_THREADS
#include <condition_variable>
#include <memory>
#include <system_error>
#include <thread>
#endif

namespace ipf {

namespace detail {

inline std::atomic<int>& threadSetting() {
    static std::atomic<int> requested{0};
    return requested;
}

// True while the current thread executes a region's callback (crew threads
// are always inside one).
inline bool& inParallelRegion() {
    thread_local bool flag = false;
    return flag;
}

// Pieces per participant: enough to even out rows of unequal cost without
// paying one atomic per row.
constexpr int kPiecesPerLane = 4;
// Estimated "pixel operations" below which an extra participant does not pay.
constexpr double kOpsPerLane = 48000.0;

#ifndef IPF_NO_THREADS
// One region: `pieces` ranges handed out through `cursor`; `run(ctx, piece)`
// processes one of them.
struct Region {
    void (*run)(void*, long long) = nullptr;
    void* ctx = nullptr;
    long long pieces = 0;
    std::atomic<long long> cursor{0};
    std::atomic<bool> failed{false};
    std::exception_ptr error;
    std::mutex error_mu;

    void work() {
        for (;;) {
            if (failed.load(std::memory_order_relaxed)) return;
            const long long p = cursor.fetch_add(1, std::memory_order_relaxed);
            if (p >= pieces) return;
            try {
                run(ctx, p);
            } catch (...) {
                std::lock_guard<std::mutex> g(error_mu);
                if (!error) error = std::current_exception();
                failed.store(true, std::memory_order_relaxed);
            }
        }
    }
};

class Crew {
public:
    static Crew& get() {
        static Crew crew;
        return crew;
    }

    // Runs `r` with up to `helpers` crew threads beside the caller. Returns
    // false (nothing run) when another region holds the crew.
    bool execute(Region& r, int helpers) {
        std::unique_lock<std::mutex> busy(owner_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        helpers = grow(helpers);
        {
            std::lock_guard<std::mutex> g(mu_);
            region_ = &r;
            invited_ = helpers;
            pending_ = helpers;
            ++epoch_;
        }
        wake_.notify_all();
        r.work();
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return pending_ == 0; });
        region_ = nullptr;
        return true;
    }

    ~Crew() {
        {
            std::lock_guard<std::mutex> g(mu_);
            quit_ = true;
        }
        wake_.notify_all();
        for (auto& t : threads_) t->join();
    }

private:
    Crew() = default;

    // Makes sure `want` helpers exist; returns how many do.
    int grow(int want) {
        while (static_cast<int>(threads_.size()) < want) {
            const int slot = static_cast<int>(threads_.size());
            try {
                threads_.push_back(std::make_unique<std::thread>([this, slot] { serve(slot); }));
            } catch (const std::system_error&) {
                break;  // the system refused another thread: make do with the crew we have
            }
        }
        return std::min(want, static_cast<int>(threads_.size()));
    }

    void serve(int slot) {
        inParallelRegion() = true;
        unsigned long long seen = 0;
        std::unique_lock<std::mutex> g(mu_);
        for (;;) {
            wake_.wait(g, [&] { return quit_ || epoch_ != seen; });
            if (quit_) return;
            seen = epoch_;
            if (slot >= invited_) continue;
            Region* r = region_;
            g.unlock();
            r->work();
            g.lock();
            if (--pending_ == 0) done_.notify_all();
        }
    }

    std::mutex owner_;  // held by the caller of the region in progress
    std::mutex mu_;
    std::condition_variable wake_, done_;
    std::vector<std::unique_ptr<std::thread>> threads_;
    Region* region_ = nullptr;
    unsigned long long epoch_ = 0;
    int invited_ = 0;
    int pending_ = 0;
    bool quit_ = false;
};
#endif

// Participants worth using for `count` items of `cost` operations each.
inline int lanesFor(long long count, double cost, int threads) {
    const double ops = static_cast<double>(count) * cost;
    const long long byOps = static_cast<long long>(ops / kOpsPerLane);
    return static_cast<int>(std::max(1LL, std::min({static_cast<long long>(threads), count, byOps})));
}

// Calls span(lo, hi) over `pieces` consecutive ranges of [0, count) using
// `lanes` participants; the ranges are equal to within one item.
template <class Span>
void splitAcross(long long count, int lanes, Span& span) {
#ifdef IPF_NO_THREADS
    (void)lanes;
    span(0LL, count);
#else
    const long long pieces = std::min(count, static_cast<long long>(lanes) * kPiecesPerLane);
    struct Ctx {
        Span* span;
        long long count, pieces;
    } ctx{&span, count, pieces};
    Region r;
    r.pieces = pieces;
    r.ctx = &ctx;
    r.run = [](void* p, long long k) {
        const Ctx& c = *static_cast<const Ctx*>(p);
        (*c.span)(c.count * k / c.pieces, c.count * (k + 1) / c.pieces);
    };
    struct Inside {  // regions the callback opens run serially, also on the fallback below
        Inside() { inParallelRegion() = true; }
        ~Inside() { inParallelRegion() = false; }
    };
    {
        Inside guard;
        if (!Crew::get().execute(r, lanes - 1)) {
            span(0LL, count);  // the crew serves another region: this one runs here
            return;
        }
    }
    if (r.error) std::rethrow_exception(r.error);
#endif
}

}  // namespace detail

/// Sets the number of threads ipf uses (0 = all hardware threads, 1 = serial).
inline void setNumThreads(int n) { detail::threadSetting().store(n > 0 ? n : 0); }

/// Number of threads ipf uses for large operations.
inline int getNumThreads() {
#ifdef IPF_NO_THREADS
    return 1;
#else
    const int n = detail::threadSetting().load();
    if (n > 0) return n;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw > 0 ? static_cast<int>(hw) : 1;
#endif
}

/// Calls fn(i0, i1) on contiguous sub-ranges covering [begin, end), on several
/// threads when count * costPerItem (in pixel operations) is large enough.
/// Exceptions thrown by fn are rethrown on the calling thread.
template <class Fn>
void parallelFor(int begin, int end, Fn&& fn, double costPerItem = 1.0) {
    if (end <= begin) return;
    const long long count = static_cast<long long>(end) - begin;
    const int lanes = detail::inParallelRegion() ? 1 : detail::lanesFor(count, costPerItem, getNumThreads());
    if (lanes <= 1) {
        fn(begin, end);
        return;
    }
    auto span = [&](long long lo, long long hi) { fn(begin + static_cast<int>(lo), begin + static_cast<int>(hi)); };
    detail::splitAcross(count, lanes, span);
}

/// parallelFor over a flat range [0, n) of any size (e.g. all samples of an
/// image): fn(i0, i1) gets std::size_t bounds; ranges are whole multiples of
/// 4096 items except the last.
template <class Fn>
void parallelForN(std::size_t n, Fn&& fn, double costPerItem = 1.0) {
    constexpr std::size_t kBlockItems = 4096;
    if (n == 0) return;
    const long long blocks = static_cast<long long>((n + kBlockItems - 1) / kBlockItems);
    const int lanes = detail::inParallelRegion()
                          ? 1
                          : detail::lanesFor(blocks, costPerItem * static_cast<double>(kBlockItems), getNumThreads());
    if (lanes <= 1) {
        fn(std::size_t{0}, n);
        return;
    }
    auto span = [&](long long lo, long long hi) {
        fn(static_cast<std::size_t>(lo) * kBlockItems, std::min(n, static_cast<std::size_t>(hi) * kBlockItems));
    };
    detail::splitAcross(blocks, lanes, span);
}

}  // namespace ipf
