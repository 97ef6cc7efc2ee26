// C-ABI entry points of libipfg.so (declared in include/ipfg.h): argument
// checks with the reference's messages, host<->device staging, per-call
// parameter upload, precision selection and kernel dispatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <cmath>
#include <cstring>
#include <list>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ipfg.h"
#include "arena.hpp"
#include "batch.hpp"
#include "fft.hpp"
#include "hough.hpp"
#include "pnm.hpp"
#include "kernels.hpp"
#include "sep_kernels.hpp"
#include "tuner.hpp"

namespace ipfg {

// ------------------------------------------------------------------ no-alloc scratch
// Device memory of ipfg_sep_filter_noalloc, which allocates nothing: the
// certification counter, workgroup ticket and fix-up report (bytes 0-23) and
// the parameter block (from kNoallocParamOff: kx and ky as double, then as
// float, each followed by 16 zeros). Each GPU holds its own copy until the
// process ends; every call on a GPU reuses it, so calls that overlap on one
// GPU race.
constexpr int kNoallocMaxTaps = 128;
constexpr size_t kNoallocParamOff = 256;
constexpr size_t kNoallocBytes = 4096;
static_assert(kNoallocParamOff + (sizeof(double) + sizeof(float)) * 2 * (kNoallocMaxTaps + 16) <= kNoallocBytes,
              "parameter block must fit the no-alloc scratch");
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {

// ------------------------------------------------------------------ errors
struct ArgError {
    std::string msg;
};
struct DeviceError {
    hipError_t err;
    std::string where;
};
struct Unsupported {
    std::string msg;
};

thread_local std::string t_error;

void require(bool ok, const char* msg) {
    if (!ok) throw ArgError{std::string("ipf: ") + msg};
}
void abi_check(bool ok, const char* msg) {
    if (!ok) throw ArgError{std::string("ipfg: ") + msg};
}
void hip_check(hipError_t e, const char* where) {
    if (e != hipSuccess) throw DeviceError{e, where};
}
#define HIPCALL(x) hip_check((x), #x)

template <class F>
int guarded(F&& f) {
    try {
        f();
        return IPFG_OK;
    } catch (const ArgError& e) {
        t_error = e.msg;
        return IPFG_ERR_ARG;
    } catch (const Unsupported& e) {
        t_error = e.msg;
        return IPFG_ERR_UNSUPPORTED;
    } catch (const DeviceError& e) {
        t_error = std::string("ipfg: HIP error '") + hipGetErrorString(e.err) + "' in " + e.where;
        return e.err == hipErrorOutOfMemory ? IPFG_ERR_NOMEM : IPFG_ERR_DEVICE;
    } catch (const std::bad_alloc&) {
        t_error = "ipfg: host allocation failed";
        return IPFG_ERR_NOMEM;
    } catch (const std::exception& e) {
        t_error = std::string("ipfg: ") + e.what();
        return IPFG_ERR_DEVICE;
    }
}

// ------------------------------------------------------------------ state
std::mutex g_init_mu;
std::atomic<int> g_device{-1};
std::atomic<int> g_threads{0};
constexpr int kMaxDevices = 64;
void* g_noalloc[kMaxDevices] = {};  // device address of `scratch` per device (under g_init_mu)

void init_device(int device) {
    std::lock_guard<std::mutex> lock(g_init_mu);
    if (g_device.load() == device) return;
    int n = 0;
    HIPCALL(hipGetDeviceCount(&n));
    abi_check(device >= 0 && device < n && device < kMaxDevices, "device index out of range");
    HIPCALL(hipSetDevice(device));
    // The no-alloc entry point's scratch: its address on this device (host
    // code uses this address, never the array's name).
    if (!g_noalloc[device]) {
        void* p = nullptr;
        HIPCALL(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_noalloc[device] = p;
    }
    // Keep freed stream-ordered allocations cached: per-call staging and
    // workspace come from this pool.
    hipMemPool_t pool;
    HIPCALL(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t threshold = UINT64_MAX;
    HIPCALL(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold));
    g_device = device;
}

void ensure_init() {
    if (g_device.load() < 0) init_device(0);
    int cur = -1;
    HIPCALL(hipGetDevice(&cur));
    if (cur != g_device.load()) HIPCALL(hipSetDevice(g_device.load()));
}

// A stream the library owns and never destroys. The runtime keeps a pointer
// to the stream an event was last recorded on and consults that stream when
// the event is synchronised or queried, so a cached event (pinned ring slot,
// parameter or scratch block, batch buffer) must not outlive the stream it
// was recorded on: when a stream is destroyed through the library (its work
// has completed by then) those events are re-recorded here.
hipStream_t home_stream() {
    static hipStream_t h = [] {
        hipStream_t st = nullptr;
        HIPCALL(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        return st;
    }();
    return h;
}

// Per-thread default stream (default priority) and a pinned slot receiving the
// fix-up count of the last certified call.
struct ThreadCtx {
    hipStream_t stream = nullptr;
    unsigned long long* fix_host = nullptr;  // pinned, mapped
    unsigned long long* fix_dev = nullptr;   // the same word's device address
    bool fix_valid = false;
};
thread_local ThreadCtx t_ctx;

hipStream_t pick_stream(ipfg_stream user) {
    if (user) return static_cast<hipStream_t>(user);
    if (!t_ctx.stream) HIPCALL(hipStreamCreateWithFlags(&t_ctx.stream, hipStreamNonBlocking));
    return t_ctx.stream;
}

// Device address of this thread's report word.
unsigned long long* fix_host_slot() {
    if (!t_ctx.fix_host) {
        void* p = nullptr;
        HIPCALL(hipHostMalloc(&p, sizeof(unsigned long long), hipHostMallocMapped));
        t_ctx.fix_host = static_cast<unsigned long long*>(p);
        *t_ctx.fix_host = 0;
        void* d = nullptr;
        HIPCALL(hipHostGetDevicePointer(&d, p, 0));
        t_ctx.fix_dev = static_cast<unsigned long long*>(d);
    }
    return t_ctx.fix_dev;
}

// ------------------------------------------------------------------ pinned ring
// Small per-call parameter blocks (kernel weights, tap lists, lookup tables)
// are copied into a pinned slot and DMA'd to a stream-ordered device
// allocation; a slot is reused once its copy event has completed.
class PinnedRing {
public:
    static constexpr size_t kSlot = 256 << 10;
    static constexpr int kSlots = 64;

    // Copies `bytes` of `data` to `dev` on `s`. Returns false if the block
    // does not fit a slot (the caller then copies from pageable memory).
    bool copy(void* dev, const void* data, size_t bytes, hipStream_t s) {
        if (bytes > kSlot) return false;
        std::lock_guard<std::mutex> lock(mu_);
        if (!base_) {
            void* p = nullptr;
            HIPCALL(hipHostMalloc(&p, kSlot * kSlots, hipHostMallocDefault));
            base_ = static_cast<char*>(p);
            events_.resize(kSlots, nullptr);
            for (auto& e : events_) HIPCALL(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            armed_.assign(kSlots, 0);
            rec_.assign(kSlots, nullptr);
        }
        const int slot = next_;
        next_ = (next_ + 1) % kSlots;
        if (armed_[slot]) HIPCALL(hipEventSynchronize(events_[slot]));
        char* h = base_ + static_cast<size_t>(slot) * kSlot;
        std::memcpy(h, data, bytes);
        HIPCALL(hipMemcpyAsync(dev, h, bytes, hipMemcpyHostToDevice, s));
        HIPCALL(hipEventRecord(events_[slot], s));
        armed_[slot] = 1;
        rec_[slot] = s;
        return true;
    }

    // s has completed its work and is being destroyed: its slots are free.
    void forget(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        for (size_t i = 0; i < rec_.size(); ++i)
            if (rec_[i] == s) {
                armed_[i] = 0;
                rec_[i] = nullptr;
            }
    }

private:
    std::mutex mu_;
    char* base_ = nullptr;
    std::vector<hipEvent_t> events_;
    std::vector<char> armed_;
    std::vector<hipStream_t> rec_;  // stream of each slot's last copy
    int next_ = 0;
};
PinnedRing g_ring;

// ------------------------------------------------------------------ param cache
// Per-call parameter blocks (weights, tap lists, range tables) are cached on
// the device keyed by their bytes, so a repeated call enqueues its kernels
// only. An entry is uploaded once on the calling stream (other streams wait
// on its ready event) and is freed on eviction only after its last user's
// kernels completed.
class ParamCache {
public:
    struct Entry {
        void* dev = nullptr;
        size_t bytes = 0;
        hipEvent_t ready = nullptr;
        hipEvent_t last_use = nullptr;
        hipStream_t ready_on = nullptr, used_on = nullptr;  // streams the two events were recorded on
        int users = 0;
        bool used = false;
        std::list<const std::string*>::iterator pos;
    };

    Entry* acquire(const std::string& key, hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        auto it = map_.find(key);
        if (it != map_.end()) {
            Entry& e = it->second;
            lru_.erase(e.pos);
            lru_.push_front(&it->first);
            e.pos = lru_.begin();
            ++e.users;
            HIPCALL(hipStreamWaitEvent(s, e.ready, 0));
            return &e;
        }
        evict(key.size());
        Entry e;
        e.bytes = key.size();
        HIPCALL(hipMalloc(&e.dev, e.bytes ? e.bytes : 16));
        HIPCALL(hipEventCreateWithFlags(&e.ready, hipEventDisableTiming));
        HIPCALL(hipEventCreateWithFlags(&e.last_use, hipEventDisableTiming));
        if (e.bytes && !g_ring.copy(e.dev, key.data(), e.bytes, s))
            HIPCALL(hipMemcpyAsync(e.dev, key.data(), e.bytes, hipMemcpyHostToDevice, s));
        HIPCALL(hipEventRecord(e.ready, s));
        e.ready_on = s;
        e.users = 1;
        auto ins = map_.emplace(key, e).first;
        lru_.push_front(&ins->first);
        ins->second.pos = lru_.begin();
        total_ += e.bytes;
        return &ins->second;
    }

    // The calling stream's kernels that read the entry have been enqueued.
    void release(Entry* e, hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        (void)hipEventRecord(e->last_use, s);
        e->used_on = s;
        e->used = true;
        --e->users;
    }

    // s has completed its work and is being destroyed.
    void forget(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto& kv : map_) {
            Entry& e = kv.second;
            if (e.ready_on == s) {
                HIPCALL(hipEventRecord(e.ready, home_stream()));
                e.ready_on = home_stream();
            }
            if (e.used_on == s) {
                HIPCALL(hipEventRecord(e.last_use, home_stream()));
                e.used_on = home_stream();
            }
        }
    }

private:
    static constexpr size_t kMaxBytes = 256u << 20;
    static constexpr size_t kMaxEntries = 4096;

    void evict(size_t incoming) {
        while (!lru_.empty() && (total_ + incoming > kMaxBytes || map_.size() >= kMaxEntries)) {
            auto victim = std::prev(lru_.end());
            bool freed = false;
            for (auto it = victim;; --it) {  // least recently used entry nobody holds
                auto mit = map_.find(**it);
                Entry& e = mit->second;
                if (e.users == 0) {
                    if (e.used) HIPCALL(hipEventSynchronize(e.last_use));
                    HIPCALL(hipEventSynchronize(e.ready));
                    HIPCALL(hipFree(e.dev));
                    HIPCALL(hipEventDestroy(e.ready));
                    HIPCALL(hipEventDestroy(e.last_use));
                    total_ -= e.bytes;
                    lru_.erase(it);
                    map_.erase(mit);
                    freed = true;
                    break;
                }
                if (it == lru_.begin()) break;
            }
            if (!freed) return;
        }
    }

    std::mutex mu_;
    std::unordered_map<std::string, Entry> map_;
    std::list<const std::string*> lru_;
    size_t total_ = 0;
};
ParamCache g_params;
Tuner g_tuner;

// ------------------------------------------------------------------ fix-up state
// Each stream owns the counter and workgroup ticket of its certified
// launches. They are zero between launches: zeroed by a kernel
// when the stream is first seen, reset by the last workgroup of each launch.
struct FixState {
    unsigned long long* count = nullptr;
    unsigned int* done = nullptr;
};
std::mutex g_fix_mu;
std::unordered_map<hipStream_t, FixState> g_fix;

FixState fix_state(hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_fix_mu);
    auto it = g_fix.find(s);
    if (it != g_fix.end()) return it->second;
    void* p = nullptr;
    HIPCALL(hipMalloc(&p, 32));
    FixState f;
    f.count = static_cast<unsigned long long*>(p);
    f.done = reinterpret_cast<unsigned int*>(f.count + 1);
    HIPCALL(launch_zero_counter(f.count, f.done, s));
    g_fix.emplace(s, f);
    return f;
}

void forget_stream(hipStream_t s) {
    std::lock_guard<std::mutex> lock(g_fix_mu);
    auto it = g_fix.find(s);
    if (it == g_fix.end()) return;
    (void)hipFree(it->second.count);
    g_fix.erase(it);
}

// ------------------------------------------------------------------ scratch
// Cache of device blocks for per-call scratch (staging copies, intermediate
// images). A block returned by a call records an event on that call's stream;
// the next call to take it makes its own stream wait for that event, so reuse
// is ordered on any stream without the host waiting. Allocation and release
// through the stream-ordered allocator made the host wait for the device on
// every call.
class ScratchPool {
public:
    struct Block {
        void* ptr;
        size_t size;
        hipEvent_t ev;
        hipStream_t on = nullptr;  // stream ev was last recorded on
    };
    static constexpr size_t kGranule = size_t(2) << 20;
    static constexpr size_t kKeepBytes = size_t(8) << 30;  // cached beyond this is freed

    Block take(size_t bytes, hipStream_t s) {
        const size_t need = (std::max<size_t>(bytes, 1) + kGranule - 1) / kGranule * kGranule;
        {
            std::lock_guard<std::mutex> lock(mu_);
            auto it = free_.lower_bound(need);
            if (it != free_.end() && it->first <= 2 * need) {
                Block b = it->second;
                free_.erase(it);
                cached_ -= b.size;
                HIPCALL(hipStreamWaitEvent(s, b.ev, 0));
                return b;
            }
        }
        Block b{nullptr, need, nullptr, nullptr};
        HIPCALL(hipMalloc(&b.ptr, need));
        HIPCALL(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
        return b;
    }

    void give(const Block& b, hipStream_t s) {
        if (hipEventRecord(b.ev, s) != hipSuccess) {  // cannot order a reuse: drop the block
            (void)hipStreamSynchronize(s);
            (void)hipEventDestroy(b.ev);
            (void)hipFree(b.ptr);
            return;
        }
        std::vector<Block> drop;
        {
            std::lock_guard<std::mutex> lock(mu_);
            Block kept = b;
            kept.on = s;
            free_.emplace(kept.size, kept);
            cached_ += b.size;
            while (cached_ > kKeepBytes && !free_.empty()) {
                auto big = std::prev(free_.end());
                drop.push_back(big->second);
                cached_ -= big->first;
                free_.erase(big);
            }
        }
        for (const Block& d : drop) release(d);
    }

    // s has completed its work and is being destroyed.
    void forget(hipStream_t s) {
        std::lock_guard<std::mutex> lock(mu_);
        for (auto& kv : free_)
            if (kv.second.on == s) {
                HIPCALL(hipEventRecord(kv.second.ev, home_stream()));
                kv.second.on = home_stream();
            }
    }

    void clear() {
        std::multimap<size_t, Block> all;
        {
            std::lock_guard<std::mutex> lock(mu_);
            all.swap(free_);
            cached_ = 0;
        }
        for (auto& kv : all) release(kv.second);
    }

private:
    static void release(const Block& b) {
        (void)hipEventSynchronize(b.ev);
        (void)hipEventDestroy(b.ev);
        (void)hipFree(b.ptr);
    }
    std::mutex mu_;
    std::multimap<size_t, Block> free_;
    size_t cached_ = 0;
};
ScratchPool g_scratch;

// ------------------------------------------------------------------ per call
// Device memory for one call (from the scratch cache, ordered on the call's
// stream) plus the cached parameter blocks it reads.
class CallScratch {
public:
    explicit CallScratch(hipStream_t s) : s_(s) {}
    ~CallScratch() {
        for (auto* e : held_) g_params.release(e, s_);
        for (const auto& b : blocks_) g_scratch.give(b, s_);
    }
    void* alloc(size_t bytes) {
        blocks_.push_back(g_scratch.take(bytes, s_));
        return blocks_.back().ptr;
    }
    // Device copy of a parameter block (cached across calls).
    void* params(const std::vector<char>& blob) {
        ParamCache::Entry* e = g_params.acquire(std::string(blob.data(), blob.size()), s_);
        held_.push_back(e);
        return e->dev;
    }
    // The stream's counter and ticket; the count lands in this thread's
    // pinned report word (written by the launch's last workgroup).
    FixList fixlist() {
        const FixState st = fix_state(s_);
        FixList f;
        f.count = st.count;
        f.done = st.done;
        f.report = fix_host_slot();
        return f;
    }
    hipStream_t stream() const { return s_; }

private:
    hipStream_t s_;
    std::vector<ScratchPool::Block> blocks_;
    std::vector<ParamCache::Entry*> held_;
};

size_t packed_stride(const ipfg_image& im) {
    return static_cast<size_t>(im.width) * im.channels * dtype_size(im.dtype);
}

void check_image(const ipfg_image* im, const char* what) {
    abi_check(im != nullptr, what);
    abi_check(im->dtype >= IPFG_U8 && im->dtype <= IPFG_F64, "unsupported pixel type");
    // Image::checkDims (core.hpp:381-388)
    require(im->width >= 0 && im->height >= 0, "Image: negative dimensions");
    require(im->channels >= 1, "Image: channel count must be >= 1");
    const unsigned long long n = static_cast<unsigned long long>(im->width) * im->height * im->channels;
    require(n < (1ULL << 40), "Image: dimensions too large");
    abi_check(n == 0 || im->data != nullptr, "null image data");
    abi_check(im->row_stride_bytes == 0 || im->row_stride_bytes >= packed_stride(*im), "row stride too small");
}

// Whether two strided views share a byte: each spans from its first byte to
// the end of its last row ((h - 1) * stride + row bytes).
bool view_end_overlap(const void* a, size_t as, int ah, size_t arow, const void* b, size_t bs, int bh, size_t brow) {
    if (ah <= 0 || bh <= 0 || arow == 0 || brow == 0) return false;
    const char* a0 = static_cast<const char*>(a);
    const char* a1 = a0 + static_cast<size_t>(ah - 1) * as + arow;
    const char* b0 = static_cast<const char*>(b);
    const char* b1 = b0 + static_cast<size_t>(bh - 1) * bs + brow;
    return a0 < b1 && b0 < a1;
}

// A call's source and destination as device images plus the copy-back step.
class Staged {
public:
    // any_shape: dst's size is checked by the caller (pyramid steps).
    Staged(const ipfg_image* src, const ipfg_image* dst, CallScratch& sc, bool any_shape = false)
        : src_(*src), dst_(*dst), sc_(sc) {
        abi_check(any_shape || (src->width == dst->width && src->height == dst->height && src->channels == dst->channels),
                  "dst must have the shape of src");
        sync_ = !src->on_device || !dst->on_device;
        ds_ = to_device(src_, true);
        dd_ = to_device(dst_, false);
        // In-place on the device: read from a copy of the view's rows (only
        // the bytes each row holds: a view's last row may end before its
        // stride does, at the end of the parent allocation).
        if (src->on_device && dst->on_device && overlaps()) {
            const size_t row = packed_stride(src_);
            const size_t pitch = (row + 255) & ~static_cast<size_t>(255);
            void* copy = sc_.alloc(pitch * static_cast<size_t>(ds_.h));
            HIPCALL(hipMemcpy2DAsync(copy, pitch, ds_.data, ds_.stride, row, static_cast<size_t>(ds_.h),
                                     hipMemcpyDeviceToDevice, sc_.stream()));
            ds_.data = copy;
            ds_.stride = pitch;
        }
    }
    const DevImage& src() const { return ds_; }
    const DevImage& dst() const { return dd_; }
    // Copies the result back for host destinations and blocks when the call
    // is synchronous.
    void finish() {
        const hipStream_t s = sc_.stream();
        if (!dst_.on_device && dd_.h > 0) {
            const size_t row = packed_stride(dst_);
            const size_t hs = dst_.row_stride_bytes ? dst_.row_stride_bytes : row;
            HIPCALL(hipMemcpy2DAsync(dst_.data, hs, dd_.data, dd_.stride, row, dd_.h, hipMemcpyDeviceToHost, s));
        }
        if (sync_) HIPCALL(hipStreamSynchronize(s));
    }
    bool synchronous() const { return sync_; }

private:
    DevImage to_device(const ipfg_image& im, bool copy_in) {
        DevImage d;
        d.w = im.width;
        d.h = im.height;
        d.cn = im.channels;
        d.dt = im.dtype;
        const size_t row = packed_stride(im);
        if (im.on_device) {
            d.data = im.data;
            d.stride = im.row_stride_bytes ? im.row_stride_bytes : row;
            return d;
        }
        d.stride = (row + 255) & ~static_cast<size_t>(255);  // aligned rows for the vector loads and stores
        d.data = sc_.alloc(d.stride * static_cast<size_t>(im.height));
        if (copy_in && im.height > 0 && row > 0) {
            const size_t hs = im.row_stride_bytes ? im.row_stride_bytes : row;
            HIPCALL(hipMemcpy2DAsync(d.data, d.stride, im.data, hs, row, im.height, hipMemcpyHostToDevice,
                                     sc_.stream()));
        }
        return d;
    }
    bool overlaps() const {
        return view_end_overlap(ds_.data, ds_.stride, ds_.h, packed_stride(src_), dd_.data, dd_.stride, dd_.h,
                                packed_stride(dst_));
    }
    ipfg_image src_, dst_;
    CallScratch& sc_;
    DevImage ds_{}, dd_{};
    bool sync_ = false;
};

void record_fixups(const FixList&, hipStream_t) { t_ctx.fix_valid = true; }

constexpr double kU32 = 1.0 / 16777216.0;          // unit roundoff of FP32 (2^-24)
constexpr double kU64 = 1.0 / 9007199254740992.0;  // unit roundoff of FP64 (2^-53)

// Largest magnitude an input sample (or the Constant border value) can have;
// -1 for float images (the kernels measure each tile).
double input_bound(int dt, double bval, int border) {
    if (dt == IPFG_F32 || dt == IPFG_F64) return -1.0;
    double x = dt == IPFG_U8 ? 255.0 : dt == IPFG_U16 ? 65535.0 : dt == IPFG_I16 ? 32768.0 : 2147483648.0;
    if (border == IPFG_BORDER_CONSTANT) x = std::max(x, std::fabs(bval));
    return x;
}

bool is_integral_value(double v) { return std::isfinite(v) && std::floor(v) == v; }

// Weights m_i * 2^-frac with integer m_i (frac <= 64); msum = sum |m_i|.
struct Dyadic {
    bool ok = false;
    int frac = 0;
    double msum = 0.0;
};

Dyadic dyadic_weights(const double* k, int n) {
    Dyadic d;
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(k[i])) return d;
        int f = 0;
        for (double v = k[i]; v != std::floor(v); v *= 2.0)
            if (++f > 64) return d;
        d.frac = std::max(d.frac, f);
    }
    for (int i = 0; i < n; ++i) d.msum += std::fabs(std::ldexp(k[i], d.frac));
    d.ok = true;
    return d;
}

// Exactness of a linear filter on an integer image: with integer samples
// (and an integral Constant value) bounded by X and dyadic weights, every
// product, partial sum and the delta are integer multiples of 2^-(fx + fy).
// When X * sum|m| (+ |delta| in those units) stays below 2^52, FP64 holds
// every intermediate exactly, in any operation order, so the FP64 result is
// the reference's bit for bit (returns 2); below 2^23 in every pass FP32
// does too (returns 3). Otherwise 0.
int exact_sum_level(int src_dt, int border, double bval, double delta, const Dyadic& x, const Dyadic& y) {
    if (!dtype_is_int(src_dt) || !x.ok || !y.ok) return 0;
    if (border == IPFG_BORDER_CONSTANT && !is_integral_value(bval)) return 0;
    const double X = input_bound(src_dt, bval, border);
    const double unit = std::ldexp(1.0, x.frac + y.frac);
    const double dm = delta * unit;
    if (!is_integral_value(dm)) return 0;
    const double total = X * x.msum * y.msum + std::fabs(dm);
    const double hx = X * x.msum;
    if (total <= 8388608.0 && hx <= 8388608.0) return 3;
    if (total <= 4503599627370496.0 && hx <= 8388608.0) return 2;  // the float ring of uint8 outputs stays exact
    return 0;
}

// Kernel family of a linear filter call. DEFAULT accumulates in FP64 and
// certifies every integer or float32 output (bit-identical to the
// reference); FP32 trades that for FP32 arithmetic; EXACT and the cases the
// fast kernels do not read (int32/double images, double outputs) run the
// reference's operation order.
int resolve_mode(int precision, int tin, int tout) {
    if (precision == IPFG_PREC_EXACT) return MODE_EXACT;
    const bool fast_in = tin == IPFG_U8 || tin == IPFG_U16 || tin == IPFG_I16 || tin == IPFG_F32;
    if (!fast_in) return MODE_EXACT;
    if (precision == IPFG_PREC_FP32) return MODE_FAST;
    return tout == IPFG_F64 ? MODE_EXACT : MODE_CERTIFY;
}

// gaussianKernel1D (filter.hpp:129-145), host FP64 with the platform exp.
std::vector<double> gaussian_kernel(double sigma, int ksize) {
    if (ksize <= 0) {
        require(sigma > 0, "gaussianKernel1D: need sigma > 0 or ksize > 0");
        ksize = 2 * std::max(1, static_cast<int>(3.0 * sigma + 0.5)) + 1;
    }
    require(ksize % 2 == 1, "gaussianKernel1D: ksize must be odd");
    if (sigma <= 0) sigma = 0.3 * ((ksize - 1) * 0.5 - 1) + 0.8;
    std::vector<double> k(static_cast<size_t>(ksize));
    const int r = ksize / 2;
    double sum = 0;
    for (int i = 0; i < ksize; ++i) {
        k[i] = std::exp(-0.5 * (i - r) * (i - r) / (sigma * sigma));
        sum += k[i];
    }
    for (double& v : k) v /= sum;
    return k;
}

// Launch description of a separable call, built the same way by every entry
// point: the kernel family, the exact-sum level and the certification bound.
struct SepPlan {
    SepLaunch a{};
    int exact_level = 0;
    std::vector<char> blob;  // kx and ky as double, then as float; each + kWeightPad zeros
    int nx = 0, ny = 0;
};
constexpr int kWeightPad = 16;  // the kernels prefetch weights ahead

SepPlan sep_plan(const DevImage& s, const DevImage& d, int src_dt, int dst_dt, int cn, const double* kx, int kw,
                 const double* ky, int kh, int ax, int ay, int border, double delta, double bval, int precision) {
    SepPlan p;
    SepLaunch& a = p.a;
    a.src = s;
    a.dst = d;
    a.kw = kw;
    a.kh = kh;
    a.ax = ax;
    a.ay = ay;
    a.border = border;
    a.delta = delta;
    a.bval = bval;
    a.mode = resolve_mode(precision, src_dt, dst_dt);
    // Exact sums need no certification; exact FP32 sums take the FP32 kernels.
    p.exact_level = a.mode == MODE_CERTIFY ? exact_sum_level(src_dt, border, bval, delta, dyadic_weights(kx, kw),
                                                             dyadic_weights(ky, kh))
                                           : 0;
    if (p.exact_level == 3 && sep_fast_fits(cn, kw, kh, src_dt, dst_dt, false)) a.mode = MODE_FAST;
    if (a.mode != MODE_EXACT && !sep_fast_fits(cn, kw, kh, src_dt, dst_dt, a.mode == MODE_CERTIFY))
        a.mode = MODE_EXACT;
    p.nx = kw + kWeightPad;
    p.ny = kh + kWeightPad;
    p.blob.assign((sizeof(double) + sizeof(float)) * (p.nx + p.ny), 0);
    double* bd = reinterpret_cast<double*>(p.blob.data());
    float* bf = reinterpret_cast<float*>(bd + p.nx + p.ny);
    double sx = 0, sy = 0;
    for (int i = 0; i < kw; ++i) {
        bd[i] = kx[i];
        bf[i] = static_cast<float>(kx[i]);
        sx += std::fabs(kx[i]);
    }
    for (int j = 0; j < kh; ++j) {
        bd[p.nx + j] = ky[j];
        bf[p.nx + j] = static_cast<float>(ky[j]);
        sy += std::fabs(ky[j]);
    }
    if (a.mode == MODE_CERTIFY) {
        // Both the FP64 FMA chains and the reference's multiply/add chains are
        // within (kw + kh + 1) u * sum|ky| sum|kx| X (+ (kh + 1) u |delta|) of
        // the exact value; a Constant value staged as float adds its own
        // rounding. Factors are padded.
        a.bcoef = 2.02 * (kw + kh + 4) * kU64 * sx * sy;
        a.kabs = sx * sy;
        a.bdelta = 2.02 * (kh + 2) * kU64 * std::fabs(delta) + 1e-300;
        if (border == IPFG_BORDER_CONSTANT && static_cast<double>(static_cast<float>(bval)) != bval)
            a.bdelta += 1.01 * std::fabs(static_cast<double>(static_cast<float>(bval)) - bval) * sx * sy;
        a.xbound = input_bound(src_dt, bval, border);
        if (p.exact_level >= 2) {  // exact sums: only results exactly on a rounding tie are recomputed
            a.bcoef = 0.0;
            a.kabs = 0.0;
            a.bdelta = 1e-300;
            a.xbound = 1.0;
        }
    }
    return p;
}

// Points the launch at a device copy of the plan's parameter block.
void sep_bind(SepPlan& p, const void* dev) {
    p.a.kxd = static_cast<const double*>(dev);
    p.a.kyd = p.a.kxd + p.nx;
    p.a.kxf = reinterpret_cast<const float*>(p.a.kxd + p.nx + p.ny);
    p.a.kyf = p.a.kxf + p.nx;
}

// Enqueues the plan: the register stencil for 3 x 3 / 5 x 5 exact FP32 sums
// with a centred anchor, the separable kernels otherwise.
hipError_t sep_dispatch(const SepPlan& p, const double* kx, const double* ky, hipStream_t s) {
    const SepLaunch& a = p.a;
    const int kw = a.kw, kh = a.kh;
    hipError_t e = hipErrorNotSupported;
    if (p.exact_level == 3 && a.mode == MODE_FAST && kw == kh && a.ax == kw / 2 && a.ay == kh / 2 &&
        (kw == 3 || kw == 5)) {
        float k2[25], fx[5], fy[5];  // kx[i] * ky[j] and the factors: exact in FP32 at this level
        for (int j = 0; j < kh; ++j)
            for (int i = 0; i < kw; ++i) k2[j * kw + i] = static_cast<float>(ky[j] * kx[i]);
        for (int i = 0; i < kw; ++i) fx[i] = static_cast<float>(kx[i]);
        for (int j = 0; j < kh; ++j) fy[j] = static_cast<float>(ky[j]);
        e = launch_stencil_u8(a.src, a.dst, k2, kw, kh, a.ay, a.border, a.delta, a.bval, s, fx, fy);
    }
    if (e == hipErrorNotSupported) {
        (void)hipGetLastError();
        e = launch_sep(a, s);
    }
    return e;
}

void sep_check_args(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky, int kh,
                    int ax, int ay, int border) {
    require(kx != nullptr && ky != nullptr && kw > 0 && kh > 0, "separable filter: empty kernel");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    abi_check(ax >= 0 && ax < kw && ay >= 0 && ay < kh, "anchor outside kernel");
}

void sep_impl(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky, int kh,
              int ax, int ay, int border, double delta, double bval, int precision, ipfg_stream stream) {
    sep_check_args(src, dst, kx, kw, ky, kh, ax, ay, border);
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    SepPlan p = sep_plan(io.src(), io.dst(), src->dtype, dst->dtype, src->channels, kx, kw, ky, kh, ax, ay, border,
                         delta, bval, precision);
    SepLaunch& a = p.a;
    sep_bind(p, sc.params(p.blob));
    if (a.mode == MODE_CERTIFY) a.fix = sc.fixlist();
    if (a.mode == MODE_EXACT) {
        a.tmp = static_cast<double*>(sc.alloc(sep_exact_tmp_bytes(a.src.w, a.src.h, a.src.cn, kh)));
    } else {
        const TuneKey key{"sep", src->dtype, dst->dtype, src->channels, kw, kh, a.mode,
                          ceil_log2(src->width), ceil_log2(src->height)};
        a.sh = g_tuner.pick(key, static_cast<long long>(src->width) * src->height * src->channels,
                            sep_sh_candidates(a), s, [&](int v) {
                                SepLaunch t = a;
                                t.sh = v;
                                return launch_sep(t, s);
                            });
    }
    HIPCALL(sep_dispatch(p, kx, ky, s));
    if (a.mode == MODE_CERTIFY) record_fixups(a.fix, s);
    else t_ctx.fix_valid = false;
    io.finish();
}

void box_impl(const ipfg_image* src, const ipfg_image* dst, int kw, int kh, int normalize, int border, double bval,
              int precision, ipfg_stream stream) {
    require(kw > 0 && kh > 0, "boxFilter: kernel size must be positive");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    BoxLaunch a{};
    a.src = io.src();
    a.dst = io.dst();
    a.kw = kw;
    a.kh = kh;
    a.border = border;
    a.normalize = normalize;
    a.bval = bval;
    // Integer images give exact integer window sums in any order (bit-exact);
    // a fractional Constant value breaks that, as do float images.
    const bool exact_sums = dtype_is_int(src->dtype) && (border != IPFG_BORDER_CONSTANT || is_integral_value(bval));
    a.mode = MODE_FAST;
    if (precision == IPFG_PREC_EXACT) a.mode = MODE_EXACT;
    else if (!exact_sums && precision == IPFG_PREC_DEFAULT && dtype_is_int(dst->dtype)) a.mode = MODE_EXACT;
    const double scale = normalize ? 1.0 / (static_cast<double>(kw) * kh) : 1.0;
    // Window sums on the separable kernels: unit weights, FP64 accumulation
    // (exact integer sums for integer images, so that sum * scale rounds
    // like the reference's), the scale applied to each result.
    const bool sep_ok = a.mode == MODE_FAST && dtype_is_int(dst->dtype) == dtype_is_int(src->dtype) &&
                        (src->dtype == IPFG_U8 || src->dtype == IPFG_U16 || src->dtype == IPFG_I16 ||
                         src->dtype == IPFG_F32) &&
                        sep_fast_fits(src->channels, kw, kh, src->dtype, dst->dtype, precision != IPFG_PREC_FP32);
    if (sep_ok) {
        SepLaunch l{};
        l.src = a.src;
        l.dst = a.dst;
        l.kw = kw;
        l.kh = kh;
        l.ax = kw / 2;
        l.ay = kh / 2;
        l.border = border;
        l.delta = 0.0;
        l.bval = bval;
        l.oscale = scale;
        l.unit = 1;
        l.mode = precision == IPFG_PREC_FP32 ? MODE_FAST : MODE_CERTIFY;
        constexpr int kPad = 16;
        const int nx = kw + kPad, ny = kh + kPad;
        std::vector<char> blob((sizeof(double) + sizeof(float)) * (nx + ny), 0);
        double* bd = reinterpret_cast<double*>(blob.data());
        float* bf = reinterpret_cast<float*>(bd + nx + ny);
        for (int i = 0; i < kw; ++i) bd[i] = bf[i] = 1.0f;
        for (int j = 0; j < kh; ++j) bd[nx + j] = bf[nx + j] = 1.0f;
        char* dev = static_cast<char*>(sc.params(blob));
        l.kxd = reinterpret_cast<const double*>(dev);
        l.kyd = l.kxd + nx;
        l.kxf = reinterpret_cast<const float*>(l.kxd + nx + ny);
        l.kyf = l.kxf + nx;
        if (l.mode == MODE_CERTIFY) {
            // integer sums are exact (also in a float ring: they stay below
            // 2^24); only results exactly on a rounding tie are recomputed
            l.bcoef = 0.0;
            l.kabs = 0.0;
            l.bdelta = 1e-300;
            l.xbound = 1.0;
            l.fix = sc.fixlist();
        }
        const TuneKey key{"sep", src->dtype, dst->dtype, src->channels, kw, kh, l.mode, ceil_log2(src->width),
                          ceil_log2(src->height)};
        l.sh = g_tuner.pick(key, static_cast<long long>(src->width) * src->height * src->channels,
                            sep_sh_candidates(l), s, [&](int v) {
                                SepLaunch t = l;
                                t.sh = v;
                                return launch_sep(t, s);
                            });
        HIPCALL(launch_sep(l, s));
        t_ctx.fix_valid = false;
        io.finish();
        return;
    }
    if (a.mode == MODE_EXACT) a.tmp = static_cast<double*>(sc.alloc(sep_exact_tmp_bytes(a.src.w, a.src.h, a.src.cn, kh)));
    HIPCALL(launch_box(a, s));
    t_ctx.fix_valid = false;
    io.finish();
}

void f2d_impl(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int ax, int ay,
              int border, double delta, double bval, int precision, ipfg_stream stream) {
    require(kernel != nullptr && kw > 0 && kh > 0, "filter2D: kernel must be a non-empty single-channel image");
    if (ax < 0) ax = kw / 2;
    if (ay < 0) ay = kh / 2;
    require(ax < kw && ay < kh, "filter2D: anchor outside kernel");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    const bool mfma = precision == IPFG_PREC_BF16X3 || precision == IPFG_PREC_FP16X2;
    if (mfma && f2d_mfma_lds_bytes(kw, kh, src->dtype != IPFG_U8) > 156 * 1024)
        throw Unsupported{"ipfg: kernel too large for the MFMA filter2D tile"};
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    F2dLaunch a{};
    if (mfma) {
        // split-precision weight fragments in register order (cached like any parameter block)
        std::vector<double> kd(static_cast<size_t>(kw) * kh);
        for (size_t i = 0; i < kd.size(); ++i) kd[i] = static_cast<double>(kernel[i]);
        std::vector<uint16_t> frag;
        double wscale = 1.0;
        f2d_mfma_fragments(kd.data(), kw, kh, precision, frag, wscale);
        std::vector<char> blob(sizeof(uint16_t) * frag.size());
        std::memcpy(blob.data(), frag.data(), blob.size());
        a.src = io.src();
        a.dst = io.dst();
        a.kw = kw;
        a.kh = kh;
        a.ax = ax;
        a.ay = ay;
        a.border = border;
        a.delta = delta;
        a.bval = bval;
        a.precision = precision;
        a.mfma_w = sc.params(blob);
        a.mfma_wscale = wscale;
        HIPCALL(launch_filter2d_mfma(a, s));
        t_ctx.fix_valid = false;
        io.finish();
        return;
    }
    a.src = io.src();
    a.dst = io.dst();
    a.kw = kw;
    a.kh = kh;
    a.ax = ax;
    a.ay = ay;
    a.border = border;
    a.delta = delta;
    a.bval = bval;
    a.precision = precision;
    a.mode = resolve_mode(precision, src->dtype, dst->dtype);
    // Non-zero taps in row-major order (filter.hpp:171-178).
    std::vector<int> taps;
    std::vector<double> tw;
    for (int j = 0; j < kh; ++j)
        for (int i = 0; i < kw; ++i) {
            const float w = kernel[static_cast<size_t>(j) * kw + i];
            if (w != 0.f) {
                taps.push_back(i);
                taps.push_back(j);
                tw.push_back(static_cast<double>(w));
            }
        }
    a.ntaps = static_cast<int>(tw.size());
    a.has_zero = a.ntaps < kw * kh;
    const size_t nk = static_cast<size_t>(kw) * kh;
    // Parameter block: non-zero tap weights (double), dense weights as double
    // and float, then the non-zero tap offsets.
    std::vector<char> blob(sizeof(double) * (tw.size() + nk) + sizeof(float) * nk + sizeof(int) * taps.size());
    char* w = blob.data();
    std::memcpy(w, tw.data(), sizeof(double) * tw.size());
    double* kd = reinterpret_cast<double*>(w + sizeof(double) * tw.size());
    double sk = 0.0;
    for (size_t i = 0; i < nk; ++i) {
        kd[i] = static_cast<double>(kernel[i]);
        sk += std::fabs(kd[i]);
    }
    std::memcpy(w + sizeof(double) * (tw.size() + nk), kernel, sizeof(float) * nk);
    std::memcpy(w + sizeof(double) * (tw.size() + nk) + sizeof(float) * nk, taps.data(), sizeof(int) * taps.size());
    char* dev = static_cast<char*>(sc.params(blob));
    a.tapw = reinterpret_cast<const double*>(dev);
    a.kd = reinterpret_cast<const double*>(dev + sizeof(double) * tw.size());
    a.kf = reinterpret_cast<const float*>(dev + sizeof(double) * (tw.size() + nk));
    a.taps = reinterpret_cast<const int*>(dev + sizeof(double) * (tw.size() + nk) + sizeof(float) * nk);
    // Exact sums need no certification; exact FP32 sums take the FP32 kernels.
    int exact_level = 0;
    if (a.mode == MODE_CERTIFY) {
        Dyadic one;
        one.ok = true;
        one.msum = 1.0;
        exact_level = exact_sum_level(src->dtype, border, bval, delta, dyadic_weights(kd, static_cast<int>(nk)), one);
        if (exact_level == 3) a.mode = MODE_FAST;
    }
    if (a.mode == MODE_CERTIFY) {
        // FP64 FMA chain and the reference's multiply/add chain each within
        // (taps + 1) u (sum|k| X + |delta|) of the exact sum.
        a.bcoef = 2.02 * (a.ntaps + 3) * kU64 * sk;
        a.bdelta = 2.02 * (a.ntaps + 3) * kU64 * std::fabs(delta) + 1e-300;
        if (border == IPFG_BORDER_CONSTANT && static_cast<double>(static_cast<float>(bval)) != bval)
            a.bdelta += 1.01 * std::fabs(static_cast<double>(static_cast<float>(bval)) - bval) * sk;
        a.xbound = input_bound(src->dtype, bval, border);
        if (exact_level >= 2) {  // exact sums: only results exactly on a rounding tie are recomputed
            a.bcoef = 0.0;
            a.bdelta = 1e-300;
            a.xbound = 1.0;
        }
        a.fix = sc.fixlist();
    }
    hipError_t e = hipErrorNotSupported;
    if (exact_level == 3 && ax == kw / 2 && ay == kh / 2)
        e = launch_stencil_u8(a.src, a.dst, kernel, kw, kh, ay, border, delta, bval, s);
    if (e == hipErrorNotSupported) {
        (void)hipGetLastError();
        e = launch_filter2d(a, s);
    }
    HIPCALL(e);
    if (a.mode == MODE_CERTIFY) record_fixups(a.fix, s);
    else t_ctx.fix_valid = false;
    io.finish();
}

void median_impl(const ipfg_image* src, const ipfg_image* dst, int ksize, int border, ipfg_stream stream) {
    require(ksize >= 1 && ksize % 2 == 1, "medianBlur: ksize must be odd and positive");
    require(border != IPFG_BORDER_CONSTANT, "medianBlur: Constant border is not supported");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(src->dtype == dst->dtype, "medianBlur: dst must have the pixel type of src");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    t_ctx.fix_valid = false;
    if (ksize == 1) {  // the reference returns a copy (filter.hpp:307)
        const DevImage& a = io.src();
        const DevImage& b = io.dst();
        const size_t row = static_cast<size_t>(a.w) * a.cn * dtype_size(a.dt);
        HIPCALL(hipMemcpy2DAsync(b.data, b.stride, a.data, a.stride, row, a.h, hipMemcpyDeviceToDevice, s));
        io.finish();
        return;
    }
    if (ksize == 3) {  // uint8, aligned rows: the register-window kernel
        const hipError_t e = launch_median3_u8(io.src(), io.dst(), border, s);
        if (e != hipErrorNotSupported) {
            HIPCALL(e);
            io.finish();
            return;
        }
        (void)hipGetLastError();
    }
    MedianLaunch a{};
    a.src = io.src();
    a.dst = io.dst();
    a.k = ksize;
    a.border = border;
    HIPCALL(launch_median(a, s));
    io.finish();
}

double nominal_max(int dt) {
    switch (dt) {
        case IPFG_U8: return 255.0;
        case IPFG_U16: return 65535.0;
        case IPFG_I16: return 32767.0;
        case IPFG_I32: return 255.0;  // PixelTraits<int32>::max() (core.hpp:58-60)
    }
    return 1.0;
}

void bilateral_impl(const ipfg_image* src, const ipfg_image* dst, int d, double sigma_color, double sigma_space,
                    int border, int precision, ipfg_stream stream) {
    require(sigma_color > 0 && sigma_space > 0, "bilateralFilter: sigmas must be positive");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(src->dtype == dst->dtype, "bilateralFilter: dst must have the pixel type of src");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    const int r = d <= 0 ? static_cast<int>(std::lround(sigma_space * 1.5)) : d / 2;
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    BilateralLaunch a{};
    a.src = io.src();
    a.dst = io.dst();
    a.r = r;
    a.border = border;
    // Taps and FP64 space weights (filter.hpp:349-352).
    std::vector<int> taps;
    std::vector<double> sw;
    for (int j = -r; j <= r; ++j)
        for (int i = -r; i <= r; ++i)
            if (i * i + j * j <= r * r) {
                taps.push_back(i);
                taps.push_back(j);
                sw.push_back(std::exp(-(i * i + j * j) / (2 * sigma_space * sigma_space)));
            }
    a.ntaps = static_cast<int>(sw.size());
    a.cc = -1.0 / (2 * sigma_color * sigma_color);
    // Range table for integer images (filter.hpp:357-364).
    std::vector<double> lut;
    if (dtype_is_int(src->dtype)) {
        const double maxd = nominal_max(src->dtype) * src->channels;
        if (maxd <= (1 << 20)) {
            lut.resize(static_cast<size_t>(maxd) + 1);
            for (size_t k = 0; k < lut.size(); ++k) lut[k] = std::exp(static_cast<double>(k * k) * a.cc);
        }
    }
    a.lut_len = static_cast<int>(lut.size());
    const bool fast_ok = bilateral_fast_ok(src->dtype, src->channels, r, a.lut_len);
    if (precision == IPFG_PREC_EXACT || !fast_ok) a.mode = MODE_EXACT;
    else a.mode = precision == IPFG_PREC_FP32 ? MODE_FAST : MODE_CERTIFY;
    const size_t n = sw.size(), nl = lut.size();
    std::vector<char> blob(sizeof(double) * (n + nl) + sizeof(float) * (n + nl) + sizeof(int) * taps.size());
    char* w = blob.data();
    std::memcpy(w, sw.data(), sizeof(double) * n);
    std::memcpy(w + sizeof(double) * n, lut.data(), sizeof(double) * nl);
    float* f = reinterpret_cast<float*>(w + sizeof(double) * (n + nl));
    for (size_t i = 0; i < n; ++i) f[i] = static_cast<float>(sw[i]);
    for (size_t i = 0; i < nl; ++i) f[n + i] = static_cast<float>(lut[i]);
    std::memcpy(w + sizeof(double) * (n + nl) + sizeof(float) * (n + nl), taps.data(), sizeof(int) * taps.size());
    char* dev = static_cast<char*>(sc.params(blob));
    a.spaced = reinterpret_cast<const double*>(dev);
    a.lutd = nl ? a.spaced + n : nullptr;
    a.spacef = reinterpret_cast<const float*>(dev + sizeof(double) * (n + nl));
    a.lutf = nl ? a.spacef + n : nullptr;
    a.taps = reinterpret_cast<const int*>(dev + sizeof(double) * (n + nl) + sizeof(float) * (n + nl));
    if (a.mode == MODE_CERTIFY) {
        // The FP64 FMA evaluation and the reference differ only in rounding:
        // per side, one product per weight, two sums of ntaps positive terms
        // and a division, i.e. (2 ntaps + 4) u relative to a result <= X.
        const double X = 255.0;
        a.bound = 2.02 * (2.0 * a.ntaps + 6.0) * kU64 * X + 1e-300;
        a.fix = sc.fixlist();
    }
    HIPCALL(launch_bilateral(a, s));
    if (a.mode == MODE_CERTIFY) record_fixups(a.fix, s);
    else t_ctx.fix_valid = false;
    io.finish();
}


// ------------------------------------------------------------------ batch geometry
DevImage to_dev(const ipfg_image& im) {
    DevImage d;
    d.data = im.data;
    d.w = im.width;
    d.h = im.height;
    d.cn = im.channels;
    d.dt = im.dtype;
    d.stride = im.row_stride_bytes ? im.row_stride_bytes : packed_stride(im);
    return d;
}

bool images_overlap(const ipfg_image& a, const ipfg_image& b) {
    const size_t ar = packed_stride(a), br = packed_stride(b);
    return view_end_overlap(a.data, a.row_stride_bytes ? a.row_stride_bytes : ar, a.height, ar, b.data,
                            b.row_stride_bytes ? b.row_stride_bytes : br, b.height, br);
}

// ------------------------------------------------------------------ no-alloc separable filter
// Weights of one axis passed by value and written into the scratch's
// parameter block (with the zero padding the kernels prefetch).
struct NoallocTaps {
    double w[kNoallocMaxTaps];
    int n;
};

__global__ void noalloc_taps_kernel(NoallocTaps t, double* __restrict__ d, float* __restrict__ f) {
    for (int i = threadIdx.x; i < t.n + kWeightPad; i += blockDim.x) {
        const double v = i < t.n ? t.w[i] : 0.0;
        d[i] = v;
        f[i] = static_cast<float>(v);
    }
}

void* noalloc_base() {
    std::lock_guard<std::mutex> lock(g_init_mu);
    void* p = g_noalloc[g_device.load()];
    abi_check(p != nullptr, "no-alloc scratch not set up on this device");
    return p;
}

// ipfg_sep_filter on device-resident images with no allocation: parameters,
// counters and the fix-up report live in `scratch`, the strip height is the
// built-in rule (no measurement launches) and a NULL stream is the device's
// null stream.
void sep_noalloc_impl(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky,
                      int kh, int ax, int ay, int border, double delta, double bval, int precision,
                      ipfg_stream stream) {
    sep_check_args(src, dst, kx, kw, ky, kh, ax, ay, border);
    abi_check(src->on_device && dst->on_device, "the no-alloc separable filter takes device-resident images");
    abi_check(src->width == dst->width && src->height == dst->height && src->channels == dst->channels,
              "dst must have the shape of src");
    if (kw > kNoallocMaxTaps || kh > kNoallocMaxTaps)
        throw Unsupported{"ipfg: the no-alloc separable filter takes at most 128 taps per axis"};
    if (precision == IPFG_PREC_EXACT)
        throw Unsupported{"ipfg: the no-alloc separable filter has no EXACT mode (it needs an FP64 intermediate)"};
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    abi_check(!images_overlap(*src, *dst), "the no-alloc separable filter cannot run in place (dst overlaps src)");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    SepPlan p = sep_plan(to_dev(*src), to_dev(*dst), src->dtype, dst->dtype, src->channels, kx, kw, ky, kh, ax, ay,
                         border, delta, bval, precision);
    if (p.a.mode == MODE_EXACT)
        throw Unsupported{"ipfg: this call needs the exact two-pass path and its FP64 intermediate (types or sizes "
                          "outside the fast kernels); use ipfg_sep_filter"};
    char* base = static_cast<char*>(noalloc_base());
    sep_bind(p, base + kNoallocParamOff);
    NoallocTaps t{};
    t.n = kw;
    std::memcpy(t.w, kx, sizeof(double) * kw);
    hipLaunchKernelGGL(noalloc_taps_kernel, dim3(1), dim3(64), 0, s, t, const_cast<double*>(p.a.kxd),
                       const_cast<float*>(p.a.kxf));
    HIPCALL(hipGetLastError());
    t = NoallocTaps{};
    t.n = kh;
    std::memcpy(t.w, ky, sizeof(double) * kh);
    hipLaunchKernelGGL(noalloc_taps_kernel, dim3(1), dim3(64), 0, s, t, const_cast<double*>(p.a.kyd),
                       const_cast<float*>(p.a.kyf));
    HIPCALL(hipGetLastError());
    if (p.a.mode == MODE_CERTIFY) {
        p.a.fix.count = reinterpret_cast<unsigned long long*>(base);
        p.a.fix.done = reinterpret_cast<unsigned int*>(base + 8);
        p.a.fix.report = reinterpret_cast<unsigned long long*>(base + 16);
        HIPCALL(launch_zero_counter(p.a.fix.count, p.a.fix.done, s));  // a kernel, not a memset
    }
    p.a.sh = 0;
    HIPCALL(sep_dispatch(p, kx, ky, s));
    t_ctx.fix_valid = false;
}

// ------------------------------------------------------------------ derivatives and sharpening
// filter.hpp:388-567: compositions of the separable, 2-D, box and Gaussian
// filters with elementwise steps, which run as small kernels on the device
// (FP64 and float arithmetic in the reference's order; hypotf/atan2f from
// the device math library).

// derivKernel (filter.hpp:93-114): binomial smoothing (1, 1) applied
// ksize - 1 - order times, then the difference (-1, 1) order times; ksize 1
// gives the unsmoothed central differences. Integer weights, exact in FP64.
std::vector<double> deriv_kernel(int order, int ksize) {
    if (ksize == 1) {
        if (order == 0) return {1.0};
        if (order == 1) return {-1.0, 0.0, 1.0};
        if (order == 2) return {1.0, -2.0, 1.0};
        throw ArgError{"ipf: derivKernel: order > 2 requires ksize >= 3"};
    }
    require(ksize % 2 == 1 && ksize >= 3 && ksize <= 31, "derivKernel: ksize must be odd in [3, 31]");
    require(order >= 0 && order < ksize, "derivKernel: invalid derivative order");
    std::vector<double> k{1.0};
    auto step = [&k](double a, double b) {
        std::vector<double> r(k.size() + 1, 0.0);
        for (size_t i = 0; i < k.size(); ++i) {
            r[i] += k[i] * a;
            r[i + 1] += k[i] * b;
        }
        k.swap(r);
    };
    for (int i = 0; i < ksize - 1 - order; ++i) step(1, 1);
    for (int i = 0; i < order; ++i) step(-1, 1);
    return k;
}

ipfg_image dev_view(const DevImage& d) {
    return ipfg_image{d.data, d.w, d.h, d.cn, d.dt, d.stride, 1};
}

// A packed device image of the given shape and type from the call's scratch.
DevImage scratch_image(CallScratch& sc, int w, int h, int cn, int dt) {
    DevImage d{nullptr, w, h, cn, dt, static_cast<size_t>(w) * cn * dtype_size(dt)};
    d.data = sc.alloc(d.stride * static_cast<size_t>(h > 0 ? h : 1));
    return d;
}

// Elementwise kernels: one thread per sample of a W x H x cn image, rows at
// their strides.
__device__ inline bool sample_at(const DevImage& im, size_t i, size_t* off) {
    const size_t row = static_cast<size_t>(im.w) * im.cn;
    if (i >= row * static_cast<size_t>(im.h)) return false;
    const size_t y = i / row;
    *off = y * im.stride + (i - y * row) * dtype_size(im.dt);
    return true;
}

// a += b (float images): laplacian's sum of the two second derivatives.
__global__ __launch_bounds__(kBlock) void add_f32_kernel(DevImage a, DevImage b) {
    const size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    size_t oa, ob;
    if (!sample_at(a, i, &oa) || !sample_at(b, i, &ob)) return;
    float* pa = reinterpret_cast<float*>(static_cast<char*>(a.data) + oa);
    *pa = *pa + *reinterpret_cast<const float*>(static_cast<const char*>(b.data) + ob);
}

// unsharpMask's combine (filter.hpp:476-482): v + amount * (v - blurred) where
// |v - blurred| > threshold, else v; saturated to the image type.
__global__ __launch_bounds__(kBlock) void unsharp_kernel(DevImage src, DevImage blurred, DevImage dst, double amount,
                                                         double threshold) {
    const size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    size_t os, ob, od;
    if (!sample_at(src, i, &os) || !sample_at(blurred, i, &ob) || !sample_at(dst, i, &od)) return;
    const double v = load_sample(src.data, os, src.dt);
    const double diff = v - *reinterpret_cast<const double*>(static_cast<const char*>(blurred.data) + ob);
    store_sample(dst.data, od, dst.dt, fabs(diff) > threshold ? v + amount * diff : v);
}

// Row-wise elementwise launches: blockIdx.y walks rows (grid-stride), each
// thread takes 4 consecutive samples of the row.
constexpr int kEltPer = 4;

dim3 rows_grid(const DevImage& d) {
    const long long L = static_cast<long long>(d.w) * d.cn;
    const unsigned gx = static_cast<unsigned>((L + kBlock * kEltPer - 1) / (kBlock * kEltPer));
    return dim3(gx ? gx : 1, static_cast<unsigned>(std::min(d.h, 65535)));
}

// unsharpMask's combine (filter.hpp:476-482) from an FP32 blur b32 with
// |b32 - blurred| <= bnd, for integer images. Every diff in
// [v - b32 - bnd', v - b32 + bnd'] is possible (bnd' adds the FP64 rounding of
// the difference); when the threshold branch and the rounded result agree
// over that interval, the output is the reference's. Otherwise the wave
// recomputes the sample's blur in the reference's order (sep_exact_wave:
// gaussianBlur<double>, Reflect101), all 64 lanes on one sample at a time.
template <class T>
__global__ __launch_bounds__(kBlock) void unsharp_cert_kernel(DevImage src, DevImage b32, DevImage dst, double amount,
                                                              double threshold, double bnd,
                                                              const double* __restrict__ k, int ks) {
    const int L = src.w * src.cn;
    const int s0 = (blockIdx.x * kBlock + threadIdx.x) * kEltPer;
    const bool active = s0 < L;  // every lane stays for the wave-wide recomputations
    for (int y = blockIdx.y; y < src.h; y += gridDim.y) {
        const T* srow = reinterpret_cast<const T*>(static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride);
        const float* brow = reinterpret_cast<const float*>(static_cast<const char*>(b32.data) + static_cast<size_t>(y) * b32.stride);
        T* drow = reinterpret_cast<T*>(static_cast<char*>(dst.data) + static_cast<size_t>(y) * dst.stride);
        unsigned doubt = 0u;
#pragma unroll
        for (int e = 0; e < kEltPer; ++e) {
            const int sidx = s0 + e;
            if (!active || sidx >= L) break;
            const double v = static_cast<double>(srow[sidx]);
            const double d = v - static_cast<double>(brow[sidx]);
            const double w = bnd + 2.3e-16 * fabs(d);
            const double lo = d - w, hi = d + w;
            const double mina = (lo <= 0.0 && hi >= 0.0) ? 0.0 : fmin(fabs(lo), fabs(hi));
            const double maxa = fmax(fabs(lo), fabs(hi));
            const bool may_keep = !(mina > threshold), may_sharpen = maxa > threshold;
            T out = saturate_from_double<T>(v);
            bool sure = true;
            if (may_sharpen) {
                const double r1 = v + amount * lo, r2 = v + amount * hi;
                const double eps = 4.5e-16 * (fabs(v) + fabs(amount) * maxa) + 1e-300;
                const T o1 = saturate_from_double<T>(fmin(r1, r2) - eps);
                const T o2 = saturate_from_double<T>(fmax(r1, r2) + eps);
                sure = o1 == o2 && (!may_keep || o1 == out);
                out = o1;
            }
            if (sure) drow[sidx] = out;
            else doubt |= 1u << e;
        }
        unsigned long long bal = __ballot(doubt != 0u);
        while (bal) {
            const int l = __builtin_ctzll(bal);
            bal &= bal - 1ull;
            unsigned mm = static_cast<unsigned>(bcast_i(static_cast<int>(doubt), l));
            const int ls0 = bcast_i(s0, l);
            while (mm) {
                const int e = __builtin_ctz(mm);
                mm &= mm - 1u;
                const int sidx = ls0 + e, px = sidx / src.cn, c = sidx - px * src.cn;
                const double b = sepk::sep_exact_wave<T>(src, px, y, c, k, ks, k, ks, ks / 2, ks / 2,
                                                         IPFG_BORDER_REFLECT101, 0.0, 0.0);
                if (lane_id() == l) {
                    const double v = static_cast<double>(srow[sidx]);
                    const double diff = v - b;
                    drow[sidx] = saturate_from_double<T>(fabs(diff) > threshold ? v + amount * diff : v);
                }
            }
        }
    }
}

// adaptiveThreshold's decision (filter.hpp:559-563) on row-wise launches.
template <class T>
__global__ __launch_bounds__(kBlock) void adaptive_rows_kernel(DevImage src, DevImage mean, DevImage dst, T hi, T lo,
                                                               double c) {
    const int L = src.w * src.cn;
    const int s0 = (blockIdx.x * kBlock + threadIdx.x) * kEltPer;
    if (s0 >= L) return;
    for (int y = blockIdx.y; y < src.h; y += gridDim.y) {
        const T* srow = reinterpret_cast<const T*>(static_cast<const char*>(src.data) + static_cast<size_t>(y) * src.stride);
        const T* mrow = reinterpret_cast<const T*>(static_cast<const char*>(mean.data) + static_cast<size_t>(y) * mean.stride);
        T* drow = reinterpret_cast<T*>(static_cast<char*>(dst.data) + static_cast<size_t>(y) * dst.stride);
#pragma unroll
        for (int e = 0; e < kEltPer; ++e) {
            if (s0 + e >= L) break;
            const bool above = static_cast<double>(srow[s0 + e]) - static_cast<double>(mrow[s0 + e]) > -c;
            drow[s0 + e] = above ? hi : lo;
        }
    }
}


// magnitude / phase of two float images (filter.hpp:438-458).
// magnitude: sqrt(x^2 + y^2) evaluated in FP64 (the squares and their sum of
// two floats are exact or rounded once; the square root is correctly rounded)
// and rounded to float; an infinite operand gives +inf, else NaN propagates.
// This is the formulation of glibc's hypotf, so the host and the device agree
// bit for bit. phase: atan2 in FP64 rounded to float stands in for the host's
// atan2f (within an ulp of it); the degree scaling and the wrap to [0, full)
// follow the reference in FP64.
__global__ __launch_bounds__(kBlock) void polar_kernel(DevImage gx, DevImage gy, DevImage dst, int mode) {
    const size_t i = blockIdx.x * static_cast<size_t>(kBlock) + threadIdx.x;
    size_t ox, oy, od;
    if (!sample_at(gx, i, &ox) || !sample_at(gy, i, &oy) || !sample_at(dst, i, &od)) return;
    const float x = *reinterpret_cast<const float*>(static_cast<const char*>(gx.data) + ox);
    const float y = *reinterpret_cast<const float*>(static_cast<const char*>(gy.data) + oy);
    float* o = reinterpret_cast<float*>(static_cast<char*>(dst.data) + od);
    if (mode == 0) {
        const double xd = x, yd = y;
        *o = (__builtin_isinf(x) || __builtin_isinf(y)) ? __builtin_inff()
                                                        : static_cast<float>(__builtin_sqrt(xd * xd + yd * yd));
    } else {
        constexpr double kPi = 3.14159265358979323846;
        const double full = mode == 1 ? 360.0 : 2 * kPi;
        double a = static_cast<float>(atan2(static_cast<double>(y), static_cast<double>(x)));
        if (mode == 1) a *= 180.0 / kPi;
        if (a < 0) a += full;
        *o = static_cast<float>(a);
    }
}

unsigned elementwise_grid(const DevImage& d) {
    const size_t n = static_cast<size_t>(d.w) * d.h * d.cn;
    return static_cast<unsigned>((n + kBlock - 1) / kBlock);
}

void require_float_dst(const ipfg_image* src, const ipfg_image* dst, const char* what) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(dst->dtype == IPFG_F32, what);
}

// sobel (filter.hpp:392-406)
void sobel_impl(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, int ksize, double scale, int border,
                int precision, ipfg_stream stream) {
    require(dx >= 0 && dy >= 0 && dx + dy > 0, "sobel: need dx + dy > 0");
    std::vector<double> kx = deriv_kernel(dx, ksize), ky = deriv_kernel(dy, ksize);
    if (ksize == 1) {  // both kernels odd-length so the anchors stay centred
        if (kx.size() == 1) kx = {0.0, 1.0, 0.0};
        if (ky.size() == 1) ky = {0.0, 1.0, 0.0};
    }
    for (double& v : kx) v *= scale;
    require_float_dst(src, dst, "sobel: dst must be float32");
    const int kw = static_cast<int>(kx.size()), kh = static_cast<int>(ky.size());
    sep_impl(src, dst, kx.data(), kw, ky.data(), kh, kw / 2, kh / 2, border, 0.0, 0.0, precision, stream);
}

// scharr, prewitt (filter.hpp:408-421): a central difference along the
// derivative axis, a smoothing triple across it.
void diff3_impl(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, const std::vector<double>& d,
                const std::vector<double>& s, int border, int precision, ipfg_stream stream, const char* what) {
    require((dx == 1 && dy == 0) || (dx == 0 && dy == 1), what);
    require_float_dst(src, dst, "derivative: dst must be float32");
    const std::vector<double>& kx = dx ? d : s;
    const std::vector<double>& ky = dx ? s : d;
    sep_impl(src, dst, kx.data(), 3, ky.data(), 3, 1, 1, border, 0.0, 0.0, precision, stream);
}

// laplacian (filter.hpp:423-436)
void laplacian_impl(const ipfg_image* src, const ipfg_image* dst, int ksize, double scale, int border, int precision,
                    ipfg_stream stream) {
    require_float_dst(src, dst, "laplacian: dst must be float32");
    if (ksize == 1) {
        const float base[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
        float k[9];
        for (int i = 0; i < 9; ++i) k[i] = static_cast<float>(base[i] * scale);
        f2d_impl(src, dst, k, 3, 3, 1, 1, border, 0.0, 0.0, precision, stream);
        return;
    }
    const std::vector<double> d2 = deriv_kernel(2, ksize);  // sobel's argument errors, before any device work
    const std::vector<double> d0 = deriv_kernel(0, ksize);
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    // When both second derivatives are exact float sums, so is their float
    // sum: one stencil with the combined kernel gives the same bits.
    std::vector<double> d2s = d2, d0s = d0;
    for (double& v : d2s) v *= scale;
    for (double& v : d0s) v *= scale;
    const int n = static_cast<int>(d2.size());
    if (precision != IPFG_PREC_EXACT && n <= 5 &&
        exact_sum_level(src->dtype, border, 0.0, 0.0, dyadic_weights(d2s.data(), n), dyadic_weights(d0.data(), n)) == 3 &&
        exact_sum_level(src->dtype, border, 0.0, 0.0, dyadic_weights(d0s.data(), n), dyadic_weights(d2.data(), n)) == 3) {
        float k2[25];
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) k2[j * n + i] = static_cast<float>(d0[j] * d2s[i] + d2[j] * d0s[i]);
        const hipError_t e = launch_stencil_u8(io.src(), io.dst(), k2, n, n, n / 2, border, 0.0, 0.0, s);
        if (e != hipErrorNotSupported) {
            HIPCALL(e);
            t_ctx.fix_valid = false;
            io.finish();
            return;
        }
        (void)hipGetLastError();
    }
    const ipfg_image ds = dev_view(io.src()), dd = dev_view(io.dst());
    const DevImage b = scratch_image(sc, src->width, src->height, src->channels, IPFG_F32);
    const ipfg_image db = dev_view(b);
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    sobel_impl(&ds, &dd, 2, 0, ksize, scale, border, precision, st);
    sobel_impl(&ds, &db, 0, 2, ksize, scale, border, precision, st);
    hipLaunchKernelGGL(add_f32_kernel, dim3(elementwise_grid(io.dst())), dim3(kBlock), 0, s, io.dst(), b);
    HIPCALL(hipGetLastError());
    t_ctx.fix_valid = false;
    io.finish();
}

// magnitude (mode 0), phase in degrees (1) or radians (2) (filter.hpp:438-458)
void polar_impl(const ipfg_image* gx, const ipfg_image* gy, const ipfg_image* dst, int mode, ipfg_stream stream) {
    check_image(gx, "null src image");
    check_image(gy, "null src image");
    check_image(dst, "null dst image");
    require(gx->width == gy->width && gx->height == gy->height && gx->channels == gy->channels,
            mode == 0 ? "magnitude: shape mismatch" : "phase: shape mismatch");
    abi_check(gx->dtype == IPFG_F32 && gy->dtype == IPFG_F32 && dst->dtype == IPFG_F32, "gradients must be float32");
    ensure_init();
    if (static_cast<long long>(gx->width) * gx->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(gx, dst, sc);
    DevImage y;
    if (gy->on_device) {
        y = to_dev(*gy);
    } else {
        y = scratch_image(sc, gy->width, gy->height, gy->channels, IPFG_F32);
        const size_t row = packed_stride(*gy);
        HIPCALL(hipMemcpy2DAsync(y.data, y.stride, gy->data, gy->row_stride_bytes ? gy->row_stride_bytes : row, row,
                                 gy->height, hipMemcpyHostToDevice, s));
    }
    hipLaunchKernelGGL(polar_kernel, dim3(elementwise_grid(io.dst())), dim3(kBlock), 0, s, io.src(), y, io.dst(), mode);
    HIPCALL(hipGetLastError());
    io.finish();
}

// ssim (metrics.hpp:46-78): the map in one launch (bit-identical to the
// reference's), its sum over the valid region in a second; the call blocks
// until the mean is on the host, as the reference returns it.
void ssim_impl(const ipfg_image* a, const ipfg_image* b, const ipfg_image* map, double* result, ipfg_stream stream) {
    check_image(a, "null image a");
    check_image(b, "null image b");
    abi_check(result != nullptr, "ssim: null result");
    abi_check(a->dtype == b->dtype, "ssim: images must have the same pixel type");
    const bool same = a->width == b->width && a->height == b->height && a->channels == b->channels;
    require(same && static_cast<long long>(a->width) * a->height > 0,
            "ssim: images must be non-empty and have the same shape");
    if (map) {
        check_image(map, "null ssim map image");
        abi_check(map->dtype == IPFG_F32 && map->width == a->width && map->height == a->height &&
                      map->channels == a->channels,
                  "ssim: the map must be a float32 image of the inputs' shape");
    }
    ensure_init();
    // PixelTraits<T>::max() (core.hpp:57-64): 1 for float images, 255 for int32.
    const double L = a->dtype == IPFG_U8 ? 255.0 : a->dtype == IPFG_U16 ? 65535.0 : a->dtype == IPFG_I16 ? 32767.0
                     : a->dtype == IPFG_I32 ? 255.0 : 1.0;
    SsimTaps k;
    const std::vector<double> g = gaussian_kernel(1.5, 11);
    for (int i = 0; i < 11; ++i) k.w[i] = g[static_cast<size_t>(i)];
    k.c1 = (0.01 * L) * (0.01 * L);
    k.c2 = (0.03 * L) * (0.03 * L);
    const int r = 5;
    const bool valid = a->width > 2 * r && a->height > 2 * r;
    const SsimRoi roi = valid ? SsimRoi{r, a->width - r, r, a->height - r} : SsimRoi{0, a->width, 0, a->height};
    const double area = static_cast<double>(roi.x1 - roi.x0) * (roi.y1 - roi.y0) * a->channels;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    ipfg_image none{};
    none.on_device = 1;
    Staged ia(a, map ? map : &none, sc, true), ib(b, &none, sc, true);
    const DevImage& da = ia.src();
    const long long np = ssim_partials(da);
    double* partial = static_cast<double*>(sc.alloc(static_cast<size_t>(np + 1) * sizeof(double)));
    float* dmap = map ? static_cast<float*>(ia.dst().data) : nullptr;
    const size_t mstride = map ? ia.dst().stride / sizeof(float) : 0;
    HIPCALL(launch_ssim(da, ib.src(), k, dmap, mstride, roi, partial, partial + np, s));
    ia.finish();
    ib.finish();
    double sum = 0.0;
    HIPCALL(hipMemcpyAsync(&sum, partial + np, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCALL(hipStreamSynchronize(s));
    *result = sum / area;
}

// pyrDown / pyrUp (geometry.hpp:501-524): one launch, bit-identical for
// every pixel type (pyramid.hip).
void pyr_impl(const ipfg_image* src, const ipfg_image* dst, bool up, ipfg_stream stream) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    const char* name = up ? "pyrUp" : "pyrDown";
    abi_check(src->dtype == dst->dtype, (std::string(name) + ": dst must have the pixel type of src").c_str());
    abi_check(!up || (src->width <= 0x3fffffff && src->height <= 0x3fffffff), "pyrUp: image too large");
    const int w = up ? 2 * src->width : (src->width + 1) / 2;
    const int h = up ? 2 * src->height : (src->height + 1) / 2;
    abi_check(dst->width == w && dst->height == h && dst->channels == src->channels,
              (std::string(name) + (up ? ": dst must be 2w x 2h" : ": dst must be (w + 1) / 2 x (h + 1) / 2")).c_str());
    ensure_init();
    if (static_cast<long long>(w) * h == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc, true);
    HIPCALL(launch_pyr(io.src(), io.dst(), up, s));
    io.finish();
}

// unsharpMask (filter.hpp:472-483). Integer images outside EXACT: uint8 with
// windows up to 13 taps in one launch (unsharp.hip), other sizes and types an
// FP32 blur plus the certified combine; float images and EXACT: the
// reference's FP64 blur (Reflect101) into a double image, then the combine.
void unsharp_impl(const ipfg_image* src, const ipfg_image* dst, double sigma, double amount, double threshold,
                  int precision, ipfg_stream stream) {
    const std::vector<double> k = gaussian_kernel(sigma, 0);
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(src->dtype == dst->dtype, "unsharpMask: dst must have the pixel type of src");
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    const ipfg_image ds = dev_view(io.src());
    const int ks = static_cast<int>(k.size());
    const int dt = src->dtype;
    if (precision != IPFG_PREC_EXACT && (dt == IPFG_U8 || dt == IPFG_U16 || dt == IPFG_I16)) {
        // FP32 blur, then the certified combine (recomputing the rare samples
        // whose result the FP32 error could change). FP32 chains of kw + kh
        // taps with weights rounded to float stay within
        // (2 ks + 4) 2^-24 (sum|k|)^2 X of the exact blur, as does the
        // reference's FP64 blur (with room to spare); X = the type's bound.
        double sk = 0.0;
        for (double v : k) sk += std::fabs(v);
        const double X = input_bound(dt, 0.0, IPFG_BORDER_REFLECT101);
        const double bnd = 1.02 * ((2.0 * ks + 6) * 5.9604644775390625e-08 * sk * sk * X + 1e-12 * X);
        std::vector<char> blob(sizeof(double) * k.size());
        std::memcpy(blob.data(), k.data(), blob.size());
        const double* kd = static_cast<const double*>(sc.params(blob));
        if (dt == IPFG_U8) {  // small windows: blur and combine in one launch
            const hipError_t e = launch_unsharp_u8(io.src(), io.dst(), ks, k.data(), kd, amount, threshold, bnd, s);
            if (e != hipErrorNotSupported) {
                HIPCALL(e);
                t_ctx.fix_valid = false;
                io.finish();
                return;
            }
            (void)hipGetLastError();
        }
        const DevImage b = scratch_image(sc, src->width, src->height, src->channels, IPFG_F32);
        const ipfg_image db = dev_view(b);
        sep_impl(&ds, &db, k.data(), ks, k.data(), ks, ks / 2, ks / 2, IPFG_BORDER_REFLECT101, 0.0, 0.0,
                 IPFG_PREC_FP32, reinterpret_cast<ipfg_stream>(s));
        const dim3 grid = rows_grid(io.dst());
        switch (dt) {
            case IPFG_U8:
                hipLaunchKernelGGL(unsharp_cert_kernel<uint8_t>, grid, dim3(kBlock), 0, s, io.src(), b, io.dst(), amount,
                                   threshold, bnd, kd, ks);
                break;
            case IPFG_U16:
                hipLaunchKernelGGL(unsharp_cert_kernel<uint16_t>, grid, dim3(kBlock), 0, s, io.src(), b, io.dst(),
                                   amount, threshold, bnd, kd, ks);
                break;
            default:
                hipLaunchKernelGGL(unsharp_cert_kernel<int16_t>, grid, dim3(kBlock), 0, s, io.src(), b, io.dst(), amount,
                                   threshold, bnd, kd, ks);
                break;
        }
        HIPCALL(hipGetLastError());
        t_ctx.fix_valid = false;
        io.finish();
        return;
    }
    const DevImage b = scratch_image(sc, src->width, src->height, src->channels, IPFG_F64);
    const ipfg_image db = dev_view(b);
    sep_impl(&ds, &db, k.data(), ks, k.data(), ks, ks / 2, ks / 2, IPFG_BORDER_REFLECT101, 0.0, 0.0, precision,
             reinterpret_cast<ipfg_stream>(s));
    hipLaunchKernelGGL(unsharp_kernel, dim3(elementwise_grid(io.dst())), dim3(kBlock), 0, s, io.src(), b, io.dst(),
                       amount, threshold);
    HIPCALL(hipGetLastError());
    t_ctx.fix_valid = false;
    io.finish();
}

// sharpen, emboss (filter.hpp:485-496): 3 x 3 filter2D to the input type.
void filter3x3_impl(const ipfg_image* src, const ipfg_image* dst, const float (&k)[9], int precision,
                    ipfg_stream stream) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(src->dtype == dst->dtype, "dst must have the pixel type of src");
    f2d_impl(src, dst, k, 3, 3, 1, 1, IPFG_BORDER_REFLECT101, 0.0, 0.0, precision, stream);
}

// adaptiveThreshold (filter.hpp:545-565): local mean (box, or Gaussian of the
// block size) with Replicate borders in the image type, then the decision.
// Float box means take the reference's running sums (EXACT) so that the
// comparison sees the reference's mean.
void adaptive_impl(const ipfg_image* src, const ipfg_image* dst, double maxval, int method, int type, int block,
                   double c, int precision, ipfg_stream stream) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    require(src->channels == 1, "adaptiveThreshold: expected a single-channel image");
    require(block >= 3 && block % 2 == 1, "adaptiveThreshold: blockSize must be odd and >= 3");
    require(type == 0 || type == 1, "adaptiveThreshold: type must be Binary or BinaryInv");
    abi_check(method == 0 || method == 1, "adaptiveThreshold: method must be Mean or Gaussian");
    abi_check(src->dtype == dst->dtype, "adaptiveThreshold: dst must have the pixel type of src");
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    const ipfg_image ds = dev_view(io.src());
    const DevImage m = scratch_image(sc, src->width, src->height, 1, src->dtype);
    const ipfg_image dm = dev_view(m);
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    if (method == 0) {
        const int p = dtype_is_int(src->dtype) ? precision : IPFG_PREC_EXACT;
        box_impl(&ds, &dm, block, block, 1, IPFG_BORDER_REPLICATE, 0.0, p, st);
    } else {
        const std::vector<double> k = gaussian_kernel(0.0, block);
        sep_impl(&ds, &dm, k.data(), block, k.data(), block, block / 2, block / 2, IPFG_BORDER_REPLICATE, 0.0, 0.0,
                 precision, st);
    }
    // the two possible outputs, saturated like the reference's per-sample cast
    const double vhi = type == 0 ? maxval : 0.0, vlo = type == 0 ? 0.0 : maxval;
    const dim3 grid = rows_grid(io.dst());
    switch (src->dtype) {
        case IPFG_U8:
            hipLaunchKernelGGL(adaptive_rows_kernel<uint8_t>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(),
                               saturate_from_double<uint8_t>(vhi), saturate_from_double<uint8_t>(vlo), c);
            break;
        case IPFG_U16:
            hipLaunchKernelGGL(adaptive_rows_kernel<uint16_t>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(),
                               saturate_from_double<uint16_t>(vhi), saturate_from_double<uint16_t>(vlo), c);
            break;
        case IPFG_I16:
            hipLaunchKernelGGL(adaptive_rows_kernel<int16_t>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(),
                               saturate_from_double<int16_t>(vhi), saturate_from_double<int16_t>(vlo), c);
            break;
        case IPFG_I32:
            hipLaunchKernelGGL(adaptive_rows_kernel<int32_t>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(),
                               saturate_from_double<int32_t>(vhi), saturate_from_double<int32_t>(vlo), c);
            break;
        case IPFG_F32:
            hipLaunchKernelGGL(adaptive_rows_kernel<float>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(),
                               static_cast<float>(vhi), static_cast<float>(vlo), c);
            break;
        default:
            hipLaunchKernelGGL(adaptive_rows_kernel<double>, grid, dim3(kBlock), 0, s, io.src(), m, io.dst(), vhi, vlo,
                               c);
            break;
    }
    HIPCALL(hipGetLastError());
    t_ctx.fix_valid = false;
    io.finish();
}

// ------------------------------------------------------------------ features
void check_edge_dst(const ipfg_image* src, const ipfg_image* dst, const char* what) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(dst->dtype == IPFG_U8 && dst->channels == 1 && dst->width == src->width && dst->height == src->height,
              what);
}

// canny after the image is on the device: Sobel gradients with Replicate
// borders (features.hpp:31-32), then suppression and hysteresis.
void canny_device(const DevImage& src, const DevImage& dst, double low, double high, int ksize, bool l2,
                  const unsigned int* sel, CallScratch& sc) {
    const hipStream_t s = sc.stream();
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    const int W = src.w, H = src.h;
    const DevImage gx = scratch_image(sc, W, H, src.cn, IPFG_F32), gy = scratch_image(sc, W, H, src.cn, IPFG_F32);
    const ipfg_image ds = dev_view(src), dgx = dev_view(gx), dgy = dev_view(gy);
    sobel_impl(&ds, &dgx, 1, 0, ksize, 1.0, IPFG_BORDER_REPLICATE, IPFG_PREC_DEFAULT, st);
    sobel_impl(&ds, &dgy, 0, 1, ksize, 1.0, IPFG_BORDER_REPLICATE, IPFG_PREC_DEFAULT, st);
    const size_t n = static_cast<size_t>(W) * H;
    CannyScratch cs;
    cs.parent = static_cast<int*>(sc.alloc(n * sizeof(int)));
    cs.state = static_cast<uint8_t*>(sc.alloc(n));
    cs.flag = static_cast<uint8_t*>(sc.alloc(n));
    const hipError_t e = launch_canny(gx, gy, low, high, l2, sel, cs, dst, s);
    if (e == hipErrorNotSupported) throw Unsupported{"canny: images of 2^31 pixels or more are not supported"};
    HIPCALL(e);
}

// canny (features.hpp:26-90)
void canny_impl(const ipfg_image* src, const ipfg_image* dst, double low, double high, int ksize, int l2,
                ipfg_stream stream) {
    require(low <= high, "canny: lowThreshold must be <= highThreshold");
    check_edge_dst(src, dst, "canny: dst must be a one-channel uint8 image of the source's size");
    deriv_kernel(1, ksize);  // sobel's argument errors, before any device work
    deriv_kernel(0, ksize);
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc, true);
    canny_device(io.src(), io.dst(), low, high, ksize, l2 != 0, nullptr, sc);
    t_ctx.fix_valid = false;
    io.finish();
}

// cannyAuto (features.hpp:92-101): toGray, gaussianBlur<float> (Reflect101),
// the L2 Sobel magnitude (Reflect101), its element of rank 7n/10 as the high
// threshold (selected on the device), canny(smooth, 0.4 high, high, 3, true).
void canny_auto_impl(const ipfg_image* src, const ipfg_image* dst, double sigma, ipfg_stream stream) {
    check_edge_dst(src, dst, "cannyAuto: dst must be a one-channel uint8 image of the source's size");
    if (src->channels > 2)
        require(src->channels == 3 || src->channels == 4, "rgbToGray: expected a 3-channel (RGB) or 4-channel (RGBA) image");
    const std::vector<double> k = gaussian_kernel(sigma, 0);
    const int ks = static_cast<int>(k.size());
    ensure_init();
    const int W = src->width, H = src->height;
    if (static_cast<long long>(W) * H == 0) return;
    const hipStream_t s = pick_stream(stream);
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    CallScratch sc(s);
    Staged io(src, dst, sc, true);
    DevImage gray = io.src();
    if (src->channels > 1) {
        gray = scratch_image(sc, W, H, 1, src->dtype);
        HIPCALL(launch_to_gray(io.src(), gray, s));
    }
    const DevImage smooth = scratch_image(sc, W, H, 1, IPFG_F32);
    const DevImage a = scratch_image(sc, W, H, 1, IPFG_F32), b = scratch_image(sc, W, H, 1, IPFG_F32);
    const DevImage mag = scratch_image(sc, W, H, 1, IPFG_F32);
    const ipfg_image dgray = dev_view(gray), dsm = dev_view(smooth), da = dev_view(a), db = dev_view(b);
    sep_impl(&dgray, &dsm, k.data(), ks, k.data(), ks, ks / 2, ks / 2, IPFG_BORDER_REFLECT101, 0.0, 0.0,
             IPFG_PREC_DEFAULT, st);
    sobel_impl(&dsm, &da, 1, 0, 3, 1.0, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    sobel_impl(&dsm, &db, 0, 1, 3, 1.0, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    hipLaunchKernelGGL(polar_kernel, dim3(elementwise_grid(mag)), dim3(kBlock), 0, s, a, b, mag, 0);
    HIPCALL(hipGetLastError());
    const size_t n = static_cast<size_t>(W) * H;
    unsigned int* sel = static_cast<unsigned int*>(sc.alloc(select_scratch_bytes()));
    HIPCALL(launch_select_rank(static_cast<const float*>(mag.data), n, n * 7 / 10, sel, s));
    canny_device(smooth, io.dst(), 0.0, 0.0, 3, true, sel, sc);
    t_ctx.fix_valid = false;
    io.finish();
}

// cornerHarris / cornerMinEigenVal (features.hpp:107-153): Sobel derivatives
// scaled by 1 / (2^(ksize - 1) block max) (features.hpp:110-112), the
// structure tensor's box sums and the response.
void corner_impl(const ipfg_image* src, const ipfg_image* dst, int block, int ksize, bool harris, double k,
                 ipfg_stream stream) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    require(src->channels == 1, "corner detection: expected a single-channel image");
    const int kk = ksize > 0 ? ksize : 3;
    double scale = std::ldexp(1.0, kk - 1) * block;
    if (dtype_is_int(src->dtype)) scale *= nominal_max(src->dtype);
    scale = 1.0 / scale;
    deriv_kernel(1, ksize);  // sobel's argument errors, then boxFilter's
    deriv_kernel(0, ksize);
    require(block > 0, "boxFilter: kernel size must be positive");
    abi_check(dst->dtype == IPFG_F32 && dst->channels == 1 && dst->width == src->width && dst->height == src->height,
              harris ? "cornerHarris: dst must be a one-channel float32 image of the source's size"
                     : "cornerMinEigenVal: dst must be a one-channel float32 image of the source's size");
    ensure_init();
    const int W = src->width, H = src->height;
    if (static_cast<long long>(W) * H == 0) return;
    const hipStream_t s = pick_stream(stream);
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    const DevImage dx = scratch_image(sc, W, H, 1, IPFG_F32), dy = scratch_image(sc, W, H, 1, IPFG_F32);
    const ipfg_image ds = dev_view(io.src()), ddx = dev_view(dx), ddy = dev_view(dy);
    sobel_impl(&ds, &ddx, 1, 0, ksize, scale, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    sobel_impl(&ds, &ddy, 0, 1, ksize, scale, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    double* rows = static_cast<double*>(sc.alloc(3 * static_cast<size_t>(W) * H * sizeof(double)));
    HIPCALL(launch_structure_response(dx, dy, block, harris, k, rows, io.dst(), s));
    t_ctx.fix_valid = false;
    io.finish();
}

// ------------------------------------------------------------------ houghCircles
// A device copy of a host image (or the device image itself).
DevImage device_source(const ipfg_image* im, CallScratch& sc) {
    if (im->on_device) return to_dev(*im);
    const size_t row = packed_stride(*im);
    DevImage d = scratch_image(sc, im->width, im->height, im->channels, im->dtype);
    HIPCALL(hipMemcpy2DAsync(d.data, d.stride, im->data, im->row_stride_bytes ? im->row_stride_bytes : row, row,
                             im->height, hipMemcpyHostToDevice, sc.stream()));
    return d;
}

// houghCircles (features.hpp:388-504): the filters, votes and per-peak
// refinement on the device; peak ordering, the minimum-distance pass and the
// final de-duplication on the host, in the reference's order. Writes up to
// `capacity` (cx, cy, radius, votes) quadruples; *count receives the number
// of circles. Blocks until the result is on the host.
void hough_impl(const ipfg_image* src, int min_r, int max_r, double canny_high, int center_votes, double min_dist,
                double min_coverage, double* out, int capacity, int* count, ipfg_stream stream) {
    check_image(src, "null src image");
    abi_check(count != nullptr && (capacity <= 0 || out != nullptr), "houghCircles: null result");
    require(src->channels == 1, "houghCircles: expected a single-channel image");
    require(min_r > 0 && max_r >= min_r, "houghCircles: invalid radius range");
    require(canny_high / 2 <= canny_high, "canny: lowThreshold must be <= highThreshold");
    if (max_r > kHoughMaxRadius) throw Unsupported{"ipfg: houghCircles: maxRadius above 15000 is not supported"};
    *count = 0;
    ensure_init();
    const int W = src->width, H = src->height;
    const size_t n = static_cast<size_t>(W) * H;
    if (n == 0) return;
    const hipStream_t s = pick_stream(stream);
    const ipfg_stream st = reinterpret_cast<ipfg_stream>(s);
    CallScratch sc(s);
    const DevImage gray = device_source(src, sc);
    const DevImage edges = scratch_image(sc, W, H, 1, IPFG_U8);
    canny_device(gray, edges, canny_high / 2, canny_high, 3, false, nullptr, sc);
    const std::vector<double> k = gaussian_kernel(1.0, 0);
    const int ks = static_cast<int>(k.size());
    const DevImage smooth = scratch_image(sc, W, H, 1, IPFG_F32);
    const DevImage dx = scratch_image(sc, W, H, 1, IPFG_F32), dy = scratch_image(sc, W, H, 1, IPFG_F32);
    const ipfg_image dg = dev_view(gray), dsm = dev_view(smooth), ddx = dev_view(dx), ddy = dev_view(dy);
    sep_impl(&dg, &dsm, k.data(), ks, k.data(), ks, ks / 2, ks / 2, IPFG_BORDER_REFLECT101, 0.0, 0.0,
             IPFG_PREC_DEFAULT, st);
    sobel_impl(&dsm, &ddx, 1, 0, 3, 1.0, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    sobel_impl(&dsm, &ddy, 0, 1, 3, 1.0, IPFG_BORDER_REFLECT101, IPFG_PREC_DEFAULT, st);
    const size_t chunks = hough_chunks(W, H);
    HoughScratch hs;
    hs.pts = static_cast<HoughPoint*>(sc.alloc(n * sizeof(HoughPoint)));
    hs.grad = static_cast<double*>(sc.alloc(n * 3 * sizeof(double)));
    hs.counts = static_cast<unsigned*>(sc.alloc(2 * chunks * sizeof(unsigned)));
    hs.total = static_cast<unsigned*>(sc.alloc(2 * sizeof(unsigned)));
    hs.acc = static_cast<float*>(sc.alloc(n * sizeof(float)));
    hs.sm = static_cast<float*>(sc.alloc(n * sizeof(float)));
    hs.cand_xy = static_cast<int*>(sc.alloc(n * 2 * sizeof(int)));
    hs.cand_v = static_cast<float*>(sc.alloc(n * sizeof(float)));
    HIPCALL(launch_hough_votes(edges, dx, dy, min_r, max_r, static_cast<float>(center_votes) - 0.5f, hs, s));
    unsigned totals[2];
    HIPCALL(hipMemcpyAsync(totals, hs.total, sizeof(totals), hipMemcpyDeviceToHost, s));
    HIPCALL(hipStreamSynchronize(s));
    const unsigned npts = totals[0], ncand = totals[1];
    std::vector<int> cxy(2 * static_cast<size_t>(ncand));
    std::vector<float> cv(ncand);
    if (ncand) {
        HIPCALL(hipMemcpyAsync(cxy.data(), hs.cand_xy, cxy.size() * sizeof(int), hipMemcpyDeviceToHost, s));
        HIPCALL(hipMemcpyAsync(cv.data(), hs.cand_v, cv.size() * sizeof(float), hipMemcpyDeviceToHost, s));
        HIPCALL(hipStreamSynchronize(s));
    }
    // findPeaks (features.hpp:174-187): strongest first (ties in raster
    // order: the candidates arrive in raster order), each kept unless a kept
    // peak lies closer than minDist. Kept peaks are binned in square cells of
    // side ceil(minDist), so a candidate tests only the kept peaks of its own
    // and the eight neighbouring cells -- the same test, against every peak
    // that can be that close.
    std::vector<unsigned> order(ncand);
    for (unsigned i = 0; i < ncand; ++i) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](unsigned a, unsigned b) { return cv[a] > cv[b] || (cv[a] == cv[b] && a < b); });
    const double md2 = min_dist * min_dist;
    std::vector<int> peaks;
    if (!(md2 > 0.0)) {  // nothing is ever closer than minDist
        for (const unsigned c : order) {
            peaks.push_back(cxy[2 * c]);
            peaks.push_back(cxy[2 * c + 1]);
        }
    } else {
        const long long cell = std::max(1LL, static_cast<long long>(std::ceil(std::min(std::fabs(min_dist), 1e9))));
        const long long gw = (W + cell - 1) / cell + 1, gh = (H + cell - 1) / cell + 1;
        std::vector<int> head(static_cast<size_t>(gw * gh), -1), next;
        for (const unsigned c : order) {
            const int x = cxy[2 * c], y = cxy[2 * c + 1];
            const long long gx = x / cell, gy = y / cell;
            bool ok = true;
            for (long long j = std::max(0LL, gy - 1); ok && j <= std::min(gh - 1, gy + 1); ++j)
                for (long long i = std::max(0LL, gx - 1); ok && i <= std::min(gw - 1, gx + 1); ++i)
                    for (int q = head[static_cast<size_t>(j * gw + i)]; q >= 0; q = next[static_cast<size_t>(q)]) {
                        const double ddx = x - peaks[2 * static_cast<size_t>(q)];
                        const double ddy = y - peaks[2 * static_cast<size_t>(q) + 1];
                        if (ddx * ddx + ddy * ddy < md2) {
                            ok = false;
                            break;
                        }
                    }
            if (ok) {
                const size_t cellid = static_cast<size_t>(gy * gw + gx);
                next.push_back(head[cellid]);
                head[cellid] = static_cast<int>(peaks.size() / 2);
                peaks.push_back(x);
                peaks.push_back(y);
            }
        }
    }
    const int np = static_cast<int>(peaks.size() / 2);
    std::vector<HoughFit> fits(static_cast<size_t>(np));
    if (np) {
        int* dpeaks = static_cast<int*>(sc.alloc(peaks.size() * sizeof(int)));
        HoughFit* dfits = static_cast<HoughFit*>(sc.alloc(fits.size() * sizeof(HoughFit)));
        HIPCALL(hipMemcpyAsync(dpeaks, peaks.data(), peaks.size() * sizeof(int), hipMemcpyHostToDevice, s));
        HIPCALL(launch_hough_refine(hs.pts, npts, hs.acc, W, H, dpeaks, np, min_r, max_r, min_coverage, dfits, s));
        HIPCALL(hipMemcpyAsync(fits.data(), dfits, fits.size() * sizeof(HoughFit), hipMemcpyDeviceToHost, s));
        HIPCALL(hipStreamSynchronize(s));
    }
    std::vector<HoughFit> circles;
    for (const HoughFit& f : fits)
        if (f.accepted) circles.push_back(f);
    std::stable_sort(circles.begin(), circles.end(),
                     [](const HoughFit& a, const HoughFit& b) { return a.support > b.support; });
    std::vector<HoughFit> unique;
    for (const HoughFit& c : circles) {
        bool dup = false;
        for (const HoughFit& u : unique) dup |= std::hypot(c.cx - u.cx, c.cy - u.cy) < min_dist;
        if (!dup) unique.push_back(c);
    }
    *count = static_cast<int>(unique.size());
    for (int i = 0; i < static_cast<int>(unique.size()) && i < capacity; ++i) {
        out[4 * i] = unique[i].cx;
        out[4 * i + 1] = unique[i].cy;
        out[4 * i + 2] = unique[i].r;
        out[4 * i + 3] = unique[i].support;
    }
    t_ctx.fix_valid = false;
}

// ------------------------------------------------------------------ Netpbm / PFM ingest
// The header grammar of codec/pnm.hpp:16-55: tokens separated by whitespace,
// '#' comments to the end of the line; numbers above 2^31 are rejected.
class PnmHeader {
public:
    PnmHeader(const uint8_t* b, size_t n) : b_(b), n_(n) {
        require(n >= 3 && b[0] == 'P' && ((b[1] >= '1' && b[1] <= '6') || b[1] == 'F' || b[1] == 'f'),
                "PNM: not a P1-P6 netpbm file");
        pfm = b[1] == 'F' || b[1] == 'f';
        const char* fmt = pfm ? "PFM" : "PNM";
        kind = pfm ? 0 : b[1] - '0';
        pos_ = 2;
        const long long ww = number(), hh = number();
        if (ww <= 0 || hh <= 0) throw ArgError{std::string("ipf: ") + fmt + ": invalid image dimensions"};
        if (ww > (1 << 24) || hh > (1 << 24) || ww * hh > (1LL << 30))
            throw ArgError{std::string("ipf: ") + fmt + ": image dimensions too large"};
        w = static_cast<int>(ww);
        h = static_cast<int>(hh);
        if (pfm) {
            cn = b[1] == 'F' ? 3 : 1;
            skip();
            const size_t s0 = pos_;
            while (pos_ < n_ && !std::isspace(b_[pos_])) ++pos_;
            const std::string word(reinterpret_cast<const char*>(b_) + s0, pos_ - s0);
            double scale = 0;
            try {
                scale = std::stod(word);
            } catch (...) {
                require(false, "PFM: invalid scale value");
            }
            require(scale != 0, "PFM: invalid scale value");
            little = scale < 0;
            require(pos_ < n_, "PFM: truncated");
            data = pos_ + 1;  // one whitespace byte
            return;
        }
        const bool bitmap = kind == 1 || kind == 4;
        maxval = bitmap ? 1 : static_cast<int>(std::min<long long>(number(), 1LL << 31));
        require(maxval >= 1 && maxval <= 65535, "PNM: invalid maxval");
        cn = (kind == 3 || kind == 6) ? 3 : 1;
        if (kind >= 4) {
            require(pos_ < n_ && std::isspace(b_[pos_]), "PNM: malformed header");
            data = pos_ + 1;
        }
    }
    // ASCII rasters (P1-P3), rescaled to T
    template <class T>
    std::vector<T> ascii_samples() {
        const size_t count = static_cast<size_t>(w) * h * cn;
        std::vector<T> out(count);
        const double tmax = static_cast<double>(std::numeric_limits<T>::max());
        for (size_t i = 0; i < count; ++i) {
            if (kind == 1) {
                skip();
                require(pos_ < n_ && (b_[pos_] == '0' || b_[pos_] == '1'), "PNM: malformed P1 pixel data");
                out[i] = b_[pos_++] == '1' ? T(0) : static_cast<T>(tmax);
            } else {
                long long v = number();
                if (v > maxval) v = maxval;
                out[i] = static_cast<T>((static_cast<double>(v) * tmax) / static_cast<double>(maxval) + 0.5);
            }
        }
        return out;
    }
    size_t raster_bytes() const {
        if (pfm) return static_cast<size_t>(w) * h * cn * 4;
        if (kind == 4) return (static_cast<size_t>(w) + 7) / 8 * h;
        return static_cast<size_t>(w) * h * cn * (maxval < 256 ? 1 : 2);
    }
    int kind = 0, w = 0, h = 0, cn = 1, maxval = 1;
    bool pfm = false, little = false;
    size_t data = 0;

private:
    void skip() {
        while (pos_ < n_) {
            if (b_[pos_] == '#') {
                while (pos_ < n_ && b_[pos_] != '\n' && b_[pos_] != '\r') ++pos_;
            } else if (std::isspace(b_[pos_])) {
                ++pos_;
            } else {
                break;
            }
        }
    }
    long long number() {
        skip();
        require(pos_ < n_ && std::isdigit(b_[pos_]), "PNM: malformed header or pixel data");
        long long v = 0;
        while (pos_ < n_ && std::isdigit(b_[pos_])) {
            v = v * 10 + (b_[pos_++] - '0');
            require(v <= (1LL << 31), "PNM: number too large");
        }
        return v;
    }
    const uint8_t* b_;
    size_t n_;
    size_t pos_ = 0;
};

void pnm_info_impl(const uint8_t* bytes, size_t n, int* w, int* h, int* cn, int* is_pfm) {
    abi_check(bytes != nullptr || n == 0, "null bytes");
    abi_check(w && h && cn && is_pfm, "null result");
    const PnmHeader hd(bytes, n);
    *w = hd.w;
    *h = hd.h;
    *cn = hd.cn;
    *is_pfm = hd.pfm ? 1 : 0;
}

// decodePNM<T> / decodePFM (codec/pnm.hpp:58-107, 143-176) into dst.
void pnm_decode_impl(const uint8_t* bytes, size_t n, const ipfg_image* dst, ipfg_stream stream) {
    abi_check(bytes != nullptr || n == 0, "null bytes");
    check_image(dst, "null dst image");
    PnmHeader hd(bytes, n);
    abi_check(dst->width == hd.w && dst->height == hd.h && dst->channels == hd.cn,
              "decode: dst must have the file's width, height and channels");
    abi_check(hd.pfm ? dst->dtype == IPFG_F32 : (dst->dtype == IPFG_U8 || dst->dtype == IPFG_U16),
              hd.pfm ? "decodePFM: dst must be float32" : "decodePNM: dst must be uint8 or uint16");
    std::vector<uint8_t> ascii;
    if (!hd.pfm && hd.kind <= 3) {
        if (dst->dtype == IPFG_U8) {
            const std::vector<uint8_t> v = hd.ascii_samples<uint8_t>();
            ascii.assign(v.begin(), v.end());
        } else {
            const std::vector<uint16_t> v = hd.ascii_samples<uint16_t>();
            ascii.resize(v.size() * 2);
            std::memcpy(ascii.data(), v.data(), ascii.size());
        }
    } else if (hd.raster_bytes() > n - std::min(n, hd.data)) {
        throw ArgError{std::string("ipf: ") + (hd.pfm ? "PFM" : "PNM") + ": unexpected end of data (truncated file?)"};
    }
    ensure_init();
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    DevImage d;
    const size_t row = packed_stride(*dst);
    if (dst->on_device) {
        d = to_dev(*dst);
    } else {
        d = scratch_image(sc, dst->width, dst->height, dst->channels, dst->dtype);
    }
    if (!ascii.empty()) {
        HIPCALL(hipMemcpy2DAsync(d.data, d.stride, ascii.data(), row, row, d.h, hipMemcpyHostToDevice, s));
        HIPCALL(hipStreamSynchronize(s));  // `ascii` is pageable and leaves scope
    } else {
        const int layout = hd.pfm ? (hd.little ? PFM_LE : PFM_BE)
                                  : hd.kind == 4 ? PNM_BITS : (hd.maxval < 256 ? PNM_BYTES : PNM_WORDS_BE);
        uint8_t* raw = static_cast<uint8_t*>(sc.alloc(hd.raster_bytes()));
        HIPCALL(hipMemcpyAsync(raw, bytes + hd.data, hd.raster_bytes(), hipMemcpyHostToDevice, s));
        HIPCALL(launch_pnm_unpack(raw, layout, hd.maxval, d, s));
    }
    if (!dst->on_device) {
        HIPCALL(hipMemcpy2DAsync(dst->data, dst->row_stride_bytes ? dst->row_stride_bytes : row, d.data, d.stride,
                                 row, d.h, hipMemcpyDeviceToHost, s));
        HIPCALL(hipStreamSynchronize(s));
    }
    t_ctx.fix_valid = false;
}

// ------------------------------------------------------------------ integralImage
// integralImage (filter.hpp:522-538): dst is a float64 (W + 1) x (H + 1) image
// with src's channels.
void integral_impl(const ipfg_image* src, const ipfg_image* dst, int squared, ipfg_stream stream) {
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(dst->dtype == IPFG_F64 && dst->width == src->width + 1 && dst->height == src->height + 1 &&
                  dst->channels == src->channels,
              "integralImage: dst must be a float64 (W + 1) x (H + 1) image with the source's channels");
    abi_check(dst->row_stride_bytes % sizeof(double) == 0, "integralImage: dst row stride must be a multiple of 8");
    ensure_init();
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc, true);
    const DevImage d = io.dst();
    HIPCALL(launch_integral(io.src(), static_cast<double*>(d.data), d.stride / sizeof(double), squared != 0, s));
    t_ctx.fix_valid = false;
    io.finish();
}

// ------------------------------------------------------------------ filter2DFFT
// Twiddles exp(i sign 2 pi k / len) of every radix-2 stage of an n-point
// transform (len = 2 .. n), evaluated as the reference does (fft.hpp:38-42:
// angle sign * 2 * pi * k / len, then cos and sin of the host libm); the
// stage of length len starts at entry len / 2 - 1.
std::vector<double> fft_twiddles(int n, int sign) {
    constexpr double kPi = 3.14159265358979323846;
    std::vector<double> t(2 * static_cast<size_t>(std::max(n - 1, 1)), 0.0);
    for (int len = 2; len <= n; len <<= 1) {
        const int half = len / 2;
        for (int k = 0; k < half; ++k) {
            const double ang = sign * 2.0 * kPi * static_cast<double>(k) / static_cast<double>(len);
            t[2 * static_cast<size_t>(half - 1 + k)] = std::cos(ang);
            t[2 * static_cast<size_t>(half - 1 + k) + 1] = std::sin(ang);
        }
    }
    return t;
}

// filter2DFFT (fft.hpp:196-216)
void fft_impl(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int border,
              double delta, double bval, ipfg_stream stream) {
    require(kernel != nullptr && kw > 0 && kh > 0, "filter2DFFT: kernel must be single-channel");
    check_image(src, "null src image");
    check_image(dst, "null dst image");
    abi_check(border >= IPFG_BORDER_CONSTANT && border <= IPFG_BORDER_WRAP, "invalid border mode");
    require(static_cast<long long>(src->width) * src->height > 0 || border == IPFG_BORDER_CONSTANT,
            "copyMakeBorder: empty source image");
    ensure_init();
    if (static_cast<long long>(src->width) * src->height == 0) return;
    const int lw = ceil_log2(static_cast<long long>(src->width) + kw - 1);
    const int lh = ceil_log2(static_cast<long long>(src->height) + kh - 1);
    if (!fft_shape_ok(lw, lh))
        throw Unsupported{"ipfg: filter2DFFT: the padded image needs more than 2^28 transform points"};
    const hipStream_t s = pick_stream(stream);
    CallScratch sc(s);
    Staged io(src, dst, sc);
    // parameters: four twiddle tables, then the kernel (floats)
    const std::vector<double> t[4] = {fft_twiddles(1 << lw, -1), fft_twiddles(1 << lh, -1), fft_twiddles(1 << lw, 1),
                                      fft_twiddles(1 << lh, 1)};
    size_t off[5] = {0};
    for (int i = 0; i < 4; ++i) off[i + 1] = off[i] + t[i].size() * sizeof(double);
    std::vector<char> blob(off[4] + static_cast<size_t>(kw) * kh * sizeof(float));
    for (int i = 0; i < 4; ++i) std::memcpy(blob.data() + off[i], t[i].data(), t[i].size() * sizeof(double));
    std::memcpy(blob.data() + off[4], kernel, static_cast<size_t>(kw) * kh * sizeof(float));
    char* dev = static_cast<char*>(sc.params(blob));
    FftLaunch a{};
    a.src = io.src();
    a.dst = io.dst();
    a.kw = kw;
    a.kh = kh;
    a.border = border;
    a.bval = bval;
    a.delta = delta;
    a.kernel = reinterpret_cast<const float*>(dev + off[4]);
    a.tw_fwd_w = reinterpret_cast<const double2*>(dev + off[0]);
    a.tw_fwd_h = reinterpret_cast<const double2*>(dev + off[1]);
    a.tw_inv_w = reinterpret_cast<const double2*>(dev + off[2]);
    a.tw_inv_h = reinterpret_cast<const double2*>(dev + off[3]);
    a.lw = lw;
    a.lh = lh;
    const size_t bytes = sizeof(double2) << (lw + lh);
    a.buf0 = static_cast<double2*>(sc.alloc(bytes));
    a.buf1 = static_cast<double2*>(sc.alloc(bytes));
    HIPCALL(launch_filter2d_fft(a, s));
    t_ctx.fix_valid = false;
    io.finish();
}

}  // namespace

size_t batch_lds_bytes(int cls) {
    if (cls < BCLS_MEDIAN) return sep_batch_lds(batch_sep_cn(cls), batch_sep_kmax(cls));
    if (cls == BCLS_MEDIAN) return median_batch_lds();
    return bilateral_batch_lds();
}

int batch_grid(int cls, size_t lds) {
    int dev = 0, cus = 0;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "hipDeviceGetAttribute");
    int per_cu = 1;
    if (cls < BCLS_MEDIAN) per_cu = sep_batch_occupancy(cls, lds);
    else if (cls == BCLS_MEDIAN) per_cu = median_batch_occupancy(lds);
    else per_cu = bilateral_batch_occupancy(lds);
    return std::max(1, cus) * per_cu;
}

void batch_tiles(int cls, int w, int h, int cn, int k, int* tiles_x, int* tiles_y, int* rows) {
    (void)cn;
    (void)k;
    *tiles_x = (w + 63) / 64;
    if (cls < BCLS_MEDIAN) {
        *rows = h <= 128 ? (h + 15) / 16 * 16 : 128;  // strip height
    } else {
        *rows = kBatchTileRows;
    }
    *tiles_y = (h + *rows - 1) / *rows;
}

namespace {

// ------------------------------------------------------------------ batches
// Parameter blocks of one batch job (host side), built like the single-image
// calls build theirs so that results are identical.
struct JobParams {
    std::vector<char> blob;
    int cls = 0;
};

size_t image_bytes(const ipfg_image& im) {
    return packed_stride(im) * static_cast<size_t>(im.height);
}

// A server request may leave dst.data null with on_device set: the server
// then allocates the output in device memory and drops it on completion.
bool sink_output(const ipfg_job& j) { return j.dst.on_device && j.dst.data == nullptr; }

void check_job(const ipfg_job& j, bool device_only) {
    check_image(&j.src, "null src image");
    if (device_only || !sink_output(j)) {
        check_image(&j.dst, "null dst image");
    } else {
        ipfg_image d = j.dst;
        d.data = j.src.data;  // checked for shape only
        check_image(&d, "null dst image");
    }
    abi_check(j.src.dtype == IPFG_U8 && j.dst.dtype == IPFG_U8, "batched jobs take uint8 images");
    abi_check(j.src.channels == 1 || j.src.channels == 3 || j.src.channels == 4, "batched jobs take 1, 3 or 4 channels");
    abi_check(j.src.width == j.dst.width && j.src.height == j.dst.height && j.src.channels == j.dst.channels,
              "src and dst shapes differ");
    abi_check(j.src.width > 0 && j.src.height > 0, "empty image in a batch");
    abi_check(j.border >= IPFG_BORDER_CONSTANT && j.border <= IPFG_BORDER_WRAP, "invalid border mode");
    if (device_only) abi_check(j.src.on_device && j.dst.on_device, "batch images must be device-resident");
    abi_check(j.src.on_device != j.dst.on_device || sink_output(j) || !images_overlap(j.src, j.dst),
              "dst must not alias src");
    switch (j.op) {
        case IPFG_OP_GAUSSIAN: break;
        case IPFG_OP_MEDIAN:
            require(j.ksize > 0 && j.ksize % 2 == 1, "medianBlur: ksize must be odd and positive");
            require(j.border != IPFG_BORDER_CONSTANT, "medianBlur: Constant border is not supported");
            if (j.ksize != 3 && j.ksize != 5) throw Unsupported{"ipfg: batched median takes ksize 3 or 5"};
            break;
        case IPFG_OP_BILATERAL:
            require(j.sigma_color > 0 && j.sigma_space > 0, "bilateralFilter: sigmas must be positive");
            break;
        default: abi_check(false, "unknown batch op");
    }
}

// Builds the class, tiles and device parameters of one job.
void build_job(const ipfg_job& j, BatchJob& out, JobParams& p) {
    out.src = to_dev(j.src);
    out.dst = to_dev(j.dst);
    out.border = j.border;
    const int w = j.src.width, h = j.src.height, cn = j.src.channels;
    if (j.op == IPFG_OP_GAUSSIAN) {
        const std::vector<double> k = gaussian_kernel(j.sigma, j.ksize);
        const int kw = static_cast<int>(k.size());
        if (kw > kBatchSepLargeK) throw Unsupported{"ipfg: batched gaussian takes windows up to 49"};
        p.cls = batch_sep_class(cn, kw);
        constexpr int kPad = 16;
        p.blob.assign(sizeof(double) * 2 * (kw + kPad), 0);
        double* d = reinterpret_cast<double*>(p.blob.data());
        double sx = 0;
        for (int i = 0; i < kw; ++i) {
            d[i] = k[i];
            d[kw + kPad + i] = k[i];
            sx += std::fabs(k[i]);
        }
        out.kw = out.kh = kw;
        // certified bound as for a single call, plus the float ring rounding
        out.bcoef = 2.02 * (2 * kw + 4) * kU64 * sx * sx + 1.01 * 5.9604644775390625e-08 * sx * sx;
        out.bdelta = 1e-300;
    } else if (j.op == IPFG_OP_MEDIAN) {
        p.cls = BCLS_MEDIAN;
        out.k = j.ksize;
    } else {
        const int r = j.d <= 0 ? static_cast<int>(std::lround(j.sigma_space * 1.5)) : j.d / 2;
        if (r > 5) throw Unsupported{"ipfg: batched bilateral takes radii up to 5"};
        p.cls = BCLS_BILATERAL;
        std::vector<int> taps;
        std::vector<double> sw;
        for (int jj = -r; jj <= r; ++jj)
            for (int i = -r; i <= r; ++i)
                if (i * i + jj * jj <= r * r) {
                    taps.push_back(i);
                    taps.push_back(jj);
                    sw.push_back(std::exp(-(i * i + jj * jj) / (2 * j.sigma_space * j.sigma_space)));
                }
        const double cc = -1.0 / (2 * j.sigma_color * j.sigma_color);
        std::vector<double> lut(static_cast<size_t>(255 * cn) + 1);
        for (size_t q = 0; q < lut.size(); ++q) lut[q] = std::exp(static_cast<double>(q * q) * cc);
        p.blob.resize(sizeof(double) * (sw.size() + lut.size()) + sizeof(int) * taps.size());
        std::memcpy(p.blob.data(), sw.data(), sizeof(double) * sw.size());
        std::memcpy(p.blob.data() + sizeof(double) * sw.size(), lut.data(), sizeof(double) * lut.size());
        std::memcpy(p.blob.data() + sizeof(double) * (sw.size() + lut.size()), taps.data(), sizeof(int) * taps.size());
        out.ntaps = static_cast<int>(sw.size());
        out.lut_len = static_cast<int>(lut.size());
        out.r = r;
        out.cc = cc;
        out.bound = 2.02 * (2.0 * out.ntaps + 6.0) * kU64 * 255.0 + 1e-300;
    }
    batch_tiles(p.cls, w, h, cn, out.kw, &out.tiles_x, &out.tiles_y, &out.rows);
}

// Runs batches on one stream: the job table goes from a pinned host buffer
// (two, alternating) to a fixed device arena, then one persistent launch per
// class; with use_graph the copy and the launches are a captured hipGraph.
class BatchExecutor;
// Live executors (their buffer events may be recorded on caller streams).
std::mutex g_exec_mu;
std::set<BatchExecutor*> g_execs;

class BatchExecutor {
public:
    BatchExecutor(int max_jobs, bool use_graph) : max_jobs_(max_jobs), use_graph_(use_graph) {
        abi_check(max_jobs > 0 && max_jobs <= (1 << 20), "max_jobs out of range");
        ensure_init();
        per_class_ = sizeof(BatchClassHeader) + sizeof(BatchJob) * static_cast<size_t>(max_jobs);
        per_class_ = (per_class_ + 255) / 256 * 256;
        arena_bytes_ = per_class_ * BCLS_COUNT;
        HIPCALL(hipMalloc(&dev_arena_, arena_bytes_));
        HIPCALL(hipMalloc(&dev_state_, 256));
        for (int b = 0; b < 2; ++b) {
            HIPCALL(hipHostMalloc(&host_arena_[b], arena_bytes_, hipHostMallocDefault));
            HIPCALL(hipEventCreateWithFlags(&free_[b], hipEventDisableTiming));
        }
        HIPCALL(hipEventCreateWithFlags(&last_done_, hipEventDisableTiming));
        queues_ = static_cast<BatchQueue*>(dev_state_);
        fix_.count = reinterpret_cast<unsigned long long*>(static_cast<char*>(dev_state_) + 64);
        fix_.done = reinterpret_cast<unsigned int*>(static_cast<char*>(dev_state_) + 72);
        fix_.report = reinterpret_cast<unsigned long long*>(static_cast<char*>(dev_state_) + 80);
        HIPCALL(hipMemset(dev_state_, 0, 256));
        HIPCALL(hipDeviceSynchronize());
        for (int c = 0; c < BCLS_COUNT; ++c) {
            lds_[c] = batch_lds_bytes(c);
            grid_[c] = batch_grid(c, lds_[c]);
        }
        std::lock_guard<std::mutex> lock(g_exec_mu);
        g_execs.insert(this);
    }
    ~BatchExecutor() {
        {
            std::lock_guard<std::mutex> lock(g_exec_mu);
            g_execs.erase(this);
        }
        for (int b = 0; b < 2; ++b) {
            if (armed_[b]) (void)hipEventSynchronize(free_[b]);
            if (exec_[b]) (void)hipGraphExecDestroy(exec_[b]);
            release_params(b, nullptr);
            (void)hipEventDestroy(free_[b]);
            (void)hipHostFree(host_arena_[b]);
        }
        if (last_armed_) (void)hipEventSynchronize(last_done_);
        (void)hipEventDestroy(last_done_);
        (void)hipFree(dev_arena_);
        (void)hipFree(dev_state_);
    }

    void run(const ipfg_job* jobs, int n, hipStream_t s) {
        abi_check(n >= 0 && n <= max_jobs_, "too many jobs for this batch executor");
        // Every job is checked and built before anything is acquired, so a
        // bad job leaves the executor as it was.
        std::vector<BatchJob> built(static_cast<size_t>(n));
        std::vector<JobParams> params(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            check_job(jobs[i], true);
            build_job(jobs[i], built[i], params[i]);
        }
        const int b = next_;
        if (armed_[b]) HIPCALL(hipEventSynchronize(free_[b]));  // this buffer's previous batch has been copied
        release_params(b, s);
        // The device arena and tile queues are shared by all runs: a run on
        // another stream starts after the previous run's kernels.
        if (last_armed_ && s != last_stream_) HIPCALL(hipStreamWaitEvent(s, last_done_, 0));
        char* h = static_cast<char*>(host_arena_[b]);
        BatchClassHeader* hdr[BCLS_COUNT];
        BatchJob* tab[BCLS_COUNT];
        for (int c = 0; c < BCLS_COUNT; ++c) {
            hdr[c] = reinterpret_cast<BatchClassHeader*>(h + per_class_ * c);
            tab[c] = reinterpret_cast<BatchJob*>(h + per_class_ * c + sizeof(BatchClassHeader));
            *hdr[c] = BatchClassHeader{};
        }
        struct Held {  // parameter entries of this run; released if the run fails before its launch
            std::vector<ParamCache::Entry*> v;
            hipStream_t s;
            bool armed = true;
            ~Held() {
                if (armed)
                    for (auto* e : v) g_params.release(e, s);
            }
        } held{{}, s};
        for (int i = 0; i < n; ++i) {
            BatchJob& J = built[i];
            const JobParams& p = params[i];
            if (!p.blob.empty()) {
                ParamCache::Entry* e = g_params.acquire(std::string(p.blob.data(), p.blob.size()), s);
                held.v.push_back(e);
                const char* dev = static_cast<const char*>(e->dev);
                if (p.cls == BCLS_BILATERAL) {
                    J.spaced = reinterpret_cast<const double*>(dev);
                    J.lutd = J.spaced + J.ntaps;
                    J.taps = reinterpret_cast<const int*>(dev + sizeof(double) * (J.ntaps + J.lut_len));
                } else {
                    J.kx = reinterpret_cast<const double*>(dev);
                    J.ky = J.kx + J.kw + 16;
                }
            }
            BatchClassHeader& H = *hdr[p.cls];
            J.tile0 = H.ntiles;
            H.ntiles += J.tiles_x * J.tiles_y;
            tab[p.cls][H.njobs++] = J;
        }
        if (use_graph_) {
            if (!exec_[b]) capture(b, s);
            HIPCALL(hipGraphLaunch(exec_[b], s));
        } else {
            enqueue(b, s);
        }
        next_ ^= 1;
        std::lock_guard<std::mutex> lock(ev_mu_);
        HIPCALL(hipEventRecord(free_[b], s));
        armed_[b] = true;
        free_on_[b] = s;
        HIPCALL(hipEventRecord(last_done_, s));
        last_stream_ = s;
        last_armed_ = true;
        held.armed = false;
        for (auto* e : held.v) g_params.release(e, s);
        held_[b] = std::move(held.v);
    }

    // s has completed its work and is being destroyed: events recorded on it
    // move to the library's own stream.
    void forget(hipStream_t s) {
        std::lock_guard<std::mutex> lock(ev_mu_);
        for (int b = 0; b < 2; ++b)
            if (armed_[b] && free_on_[b] == s) {
                HIPCALL(hipEventRecord(free_[b], home_stream()));
                free_on_[b] = home_stream();
            }
        if (last_armed_ && last_stream_ == s) {
            HIPCALL(hipEventRecord(last_done_, home_stream()));
            last_stream_ = home_stream();
        }
    }

private:
    void enqueue(int b, hipStream_t s) {
        HIPCALL(hipMemcpyAsync(dev_arena_, host_arena_[b], arena_bytes_, hipMemcpyHostToDevice, s));
        char* d = static_cast<char*>(dev_arena_);
        for (int c = 0; c < BCLS_COUNT; ++c) {
            const auto* hdr = reinterpret_cast<const BatchClassHeader*>(d + per_class_ * c);
            const auto* jobs = reinterpret_cast<const BatchJob*>(d + per_class_ * c + sizeof(BatchClassHeader));
            BatchQueue* q = queues_ + c;
            if (c < BCLS_MEDIAN)
                HIPCALL(launch_sep_batch(c, hdr, jobs, q, fix_, grid_[c], lds_[c], s));
            else if (c == BCLS_MEDIAN)
                HIPCALL(launch_median_batch(hdr, jobs, q, grid_[c], lds_[c], s));
            else
                HIPCALL(launch_bilateral_batch(hdr, jobs, q, fix_, grid_[c], lds_[c], s));
        }
    }

    void capture(int b, hipStream_t s) {
        hipGraph_t g = nullptr;
        HIPCALL(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            enqueue(b, s);
        } catch (...) {
            (void)hipStreamEndCapture(s, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        HIPCALL(hipStreamEndCapture(s, &g));
        HIPCALL(hipGraphInstantiate(&exec_[b], g, nullptr, nullptr, 0));
        HIPCALL(hipGraphDestroy(g));
    }

    // Parameter blocks of the batch that last used buffer b: their release was
    // recorded after its launches; forget them now.
    void release_params(int b, hipStream_t) { held_[b].clear(); }

    int max_jobs_;
    bool use_graph_;
    size_t per_class_ = 0, arena_bytes_ = 0;
    void* dev_arena_ = nullptr;
    void* dev_state_ = nullptr;
    void* host_arena_[2] = {nullptr, nullptr};
    hipEvent_t free_[2] = {nullptr, nullptr};
    bool armed_[2] = {false, false};
    hipStream_t free_on_[2] = {nullptr, nullptr};  // streams free_ was recorded on
    std::mutex ev_mu_;                              // free_ / last_done_ bookkeeping vs forget()
    hipGraphExec_t exec_[2] = {nullptr, nullptr};
    std::vector<ParamCache::Entry*> held_[2];
    int next_ = 0;
    hipEvent_t last_done_ = nullptr;  // recorded after each run's launches
    hipStream_t last_stream_ = nullptr;
    bool last_armed_ = false;
    BatchQueue* queues_ = nullptr;
    FixList fix_{};
    size_t lds_[BCLS_COUNT] = {};
    int grid_[BCLS_COUNT] = {};
};

// ------------------------------------------------------------------ server
class Server {
public:
    Server(size_t budget, int max_batch, bool use_graph) : budget_(budget), max_batch_(max_batch) {
        abi_check(budget >= RangeArena::kAlign, "device budget must be positive");
        abi_check(max_batch > 0 && max_batch <= 65536, "max_batch out of range");
        ensure_init();
        device_ = g_device.load();
        // Request buffers are ranges of one block of exactly the budget.
        arena_.reset(budget);
        HIPCALL(hipMalloc(&arena_base_, arena_.capacity()));
        HIPCALL(hipStreamCreateWithFlags(&up_, hipStreamNonBlocking));
        for (int i = 0; i < kLanes; ++i) {
            HIPCALL(hipStreamCreateWithFlags(&lane_[i].stream, hipStreamNonBlocking));
            lane_[i].exec = std::make_unique<BatchExecutor>(max_batch, use_graph);
        }
        worker_ = std::thread([this] { loop(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> lock(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        if (worker_.joinable()) worker_.join();
        for (int i = 0; i < kLanes; ++i) {
            (void)hipStreamSynchronize(lane_[i].stream);
            lane_[i].exec.reset();
            (void)hipStreamDestroy(lane_[i].stream);
        }
        (void)hipStreamDestroy(up_);
        (void)hipFree(arena_base_);
    }

    long long submit(const ipfg_job& r) {
        check_job(r, false);
        const size_t need = device_need(r);
        abi_check(need <= arena_.capacity(), "request larger than the device budget");
        std::lock_guard<std::mutex> lock(mu_);
        abi_check(error_.empty(), error_.c_str());
        Req q;
        q.job = r;
        q.id = next_id_++;
        q.t_submit = std::chrono::steady_clock::now();
        q.need = need;
        pending_.push_back(q);
        ++outstanding_;
        cv_.notify_all();
        return q.id;
    }

    int poll(long long* ids, double* lat, int cap) {
        std::lock_guard<std::mutex> lock(mu_);
        abi_check(error_.empty(), error_.c_str());
        int n = 0;
        while (n < cap && !done_.empty()) {
            ids[n] = done_.front().first;
            if (lat) lat[n] = done_.front().second;
            done_.pop_front();
            ++n;
        }
        return n;
    }

    void drain() {
        std::unique_lock<std::mutex> lock(mu_);
        cv_done_.wait(lock, [&] { return outstanding_ == 0 || !error_.empty(); });
        abi_check(error_.empty(), error_.c_str());
    }

    void stats(long long* batches, int* max_jobs) {
        std::lock_guard<std::mutex> lock(mu_);
        if (batches) *batches = batches_;
        if (max_jobs) *max_jobs = max_jobs_seen_;
    }

    void memory(size_t* held, size_t* peak) {
        std::lock_guard<std::mutex> lock(mu_);
        if (held) *held = arena_.capacity();
        if (peak) *peak = peak_in_use_;
    }

private:
    static constexpr int kLanes = 2;  // batches in flight
    struct Req {
        ipfg_job job{};
        long long id = 0;
        std::chrono::steady_clock::time_point t_submit;
        size_t need = 0;
        void* dsrc = nullptr;  // server-owned device copies (ranges of the arena)
        void* ddst = nullptr;
        size_t osrc = RangeArena::npos, odst = RangeArena::npos;
    };
    struct Lane {
        hipStream_t stream = nullptr;
        std::unique_ptr<BatchExecutor> exec;
        std::vector<Req> reqs;
        hipEvent_t done = nullptr;
        bool busy = false;
    };

    // Bytes of server-owned device copies: the source of a host request and
    // the output of a host or server-owned destination.
    static size_t src_bytes(const ipfg_job& r) { return r.src.on_device ? 0 : image_bytes(r.src); }
    static size_t dst_bytes(const ipfg_job& r) { return r.dst.on_device && !sink_output(r) ? 0 : image_bytes(r.dst); }
    static size_t device_need(const ipfg_job& r) {
        return RangeArena::rounded(src_bytes(r)) + RangeArena::rounded(dst_bytes(r));
    }

    // Places the request's buffers in the arena; false (nothing placed) when
    // they do not fit now. Worker thread only.
    bool place(Req& q) {
        const size_t sb = src_bytes(q.job), db = dst_bytes(q.job);
        const size_t os = arena_.take(sb);
        if (os == RangeArena::npos) return false;
        const size_t od = arena_.take(db);
        if (od == RangeArena::npos) {
            arena_.give(os, sb);
            return false;
        }
        q.osrc = sb ? os : RangeArena::npos;
        q.odst = db ? od : RangeArena::npos;
        return true;
    }

    void unplace(Req& q) {
        if (q.osrc != RangeArena::npos) arena_.give(q.osrc, src_bytes(q.job));
        if (q.odst != RangeArena::npos) arena_.give(q.odst, dst_bytes(q.job));
        q.osrc = q.odst = RangeArena::npos;
    }

    void loop() {
        if (hipSetDevice(device_) != hipSuccess) return fail("hipSetDevice failed in the server thread");
        for (int i = 0; i < kLanes; ++i)
            if (hipEventCreateWithFlags(&lane_[i].done, hipEventDisableTiming) != hipSuccess)
                return fail("event creation failed");
        try {
            for (;;) {
                bool progressed = retire();
                std::vector<Req> admitted;
                {
                    std::unique_lock<std::mutex> lock(mu_);
                    if (stop_ && pending_.empty() && !any_busy()) break;
                    // admit in arrival order while the buffers fit in the budget
                    while (!pending_.empty() && static_cast<int>(admitted.size()) < max_batch_ &&
                           place(pending_.front())) {
                        admitted.push_back(pending_.front());
                        pending_.pop_front();
                    }
                    peak_in_use_ = std::max(peak_in_use_, arena_.in_use());
                    if (admitted.empty() && !progressed) {
                        if (any_busy()) {
                            lock.unlock();
                            std::this_thread::sleep_for(std::chrono::microseconds(20));
                        } else {
                            cv_.wait_for(lock, std::chrono::milliseconds(5));
                        }
                        continue;
                    }
                }
                if (!admitted.empty()) launch(std::move(admitted));
            }
        } catch (const ArgError& e) {
            fail(e.msg);
        } catch (const Unsupported& e) {
            fail(e.msg);
        } catch (const DeviceError& e) {
            fail(std::string("ipfg: HIP error '") + hipGetErrorString(e.err) + "' in " + e.where);
        } catch (const std::exception& e) {
            fail(std::string("ipfg: ") + e.what());
        }
        for (int i = 0; i < kLanes; ++i) (void)hipEventDestroy(lane_[i].done);
    }

    bool any_busy() const {
        for (int i = 0; i < kLanes; ++i)
            if (lane_[i].busy) return true;
        return false;
    }

    // Uploads the admitted requests and runs them as one batch on a free lane
    // (waiting for one when both are busy).
    void launch(std::vector<Req> reqs) {
        int li = -1;
        while (li < 0) {
            for (int i = 0; i < kLanes; ++i)
                if (!lane_[i].busy) {
                    li = i;
                    break;
                }
            if (li < 0) {
                HIPCALL(hipEventSynchronize(lane_[next_lane_].done));
                retire();
            }
        }
        Lane& L = lane_[li];
        next_lane_ = li ^ 1;
        std::vector<ipfg_job> jobs;
        jobs.reserve(reqs.size());
        for (Req& q : reqs) {
            ipfg_job j = q.job;
            if (!j.src.on_device) {
                q.dsrc = static_cast<char*>(arena_base_) + q.osrc;
                HIPCALL(hipMemcpy2DAsync(q.dsrc, packed_stride(j.src), j.src.data, host_stride(j.src), packed_stride(j.src),
                                         j.src.height, hipMemcpyHostToDevice, L.stream));
                j.src.data = q.dsrc;
                j.src.row_stride_bytes = packed_stride(q.job.src);
                j.src.on_device = 1;
            }
            if (!j.dst.on_device || sink_output(j)) {
                q.ddst = static_cast<char*>(arena_base_) + q.odst;
                j.dst.data = q.ddst;
                j.dst.row_stride_bytes = packed_stride(q.job.dst);
                j.dst.on_device = 1;
            }
            jobs.push_back(j);
        }
        L.exec->run(jobs.data(), static_cast<int>(jobs.size()), L.stream);
        for (Req& q : reqs) {
            if (q.ddst && !q.job.dst.on_device)
                HIPCALL(hipMemcpy2DAsync(q.job.dst.data, host_stride(q.job.dst), q.ddst, packed_stride(q.job.dst),
                                         packed_stride(q.job.dst), q.job.dst.height, hipMemcpyDeviceToHost, L.stream));
        }
        HIPCALL(hipEventRecord(L.done, L.stream));
        L.reqs = std::move(reqs);
        L.busy = true;
        std::lock_guard<std::mutex> lock(mu_);
        ++batches_;
        max_jobs_seen_ = std::max(max_jobs_seen_, static_cast<int>(L.reqs.size()));
    }

    // Completes the requests of finished lanes. Returns whether any finished.
    bool retire() {
        bool any = false;
        for (int i = 0; i < kLanes; ++i) {
            Lane& L = lane_[i];
            if (!L.busy) continue;
            const hipError_t e = hipEventQuery(L.done);
            if (e == hipErrorNotReady) continue;
            HIPCALL(e);
            const auto now = std::chrono::steady_clock::now();
            std::lock_guard<std::mutex> lock(mu_);
            for (Req& q : L.reqs) {
                unplace(q);  // the lane's work, copies included, has completed
                done_.emplace_back(q.id, std::chrono::duration<double, std::milli>(now - q.t_submit).count());
                --outstanding_;
            }
            L.reqs.clear();
            L.busy = false;
            any = true;
        }
        if (any) cv_done_.notify_all();
        return any;
    }

    void fail(const std::string& msg) {
        std::lock_guard<std::mutex> lock(mu_);
        error_ = msg.empty() ? std::string("ipfg: server failed") : msg;
        cv_done_.notify_all();
    }

    static size_t host_stride(const ipfg_image& im) {
        return im.row_stride_bytes ? im.row_stride_bytes : packed_stride(im);
    }

    size_t budget_;
    int max_batch_;
    int device_ = 0;
    hipStream_t up_ = nullptr;
    Lane lane_[kLanes];
    int next_lane_ = 0;
    std::thread worker_;
    std::mutex mu_;
    std::condition_variable cv_, cv_done_;
    std::deque<Req> pending_;
    std::deque<std::pair<long long, double>> done_;
    RangeArena arena_;          // placement of request buffers in arena_base_ (worker thread)
    void* arena_base_ = nullptr;
    size_t peak_in_use_ = 0;
    long long next_id_ = 0;
    long long outstanding_ = 0;
    long long batches_ = 0;
    int max_jobs_seen_ = 0;
    bool stop_ = false;
    std::string error_;
};

}  // namespace
}  // namespace ipfg

using namespace ipfg;

extern "C" {

int ipfg_init(int device) {
    return guarded([&] { init_device(device); });
}

int ipfg_shutdown(void) {
    return guarded([&] {
        if (g_device.load() >= 0) {
            HIPCALL(hipDeviceSynchronize());
            g_scratch.clear();
        }
    });
}

const char* ipfg_last_error(void) { return t_error.c_str(); }
int ipfg_abi_version(void) { return IPFG_ABI_VERSION; }

int ipfg_last_fixups(long long* count) {
    return guarded([&] {
        abi_check(count != nullptr, "null count");
        *count = (t_ctx.fix_valid && t_ctx.fix_host) ? static_cast<long long>(*t_ctx.fix_host) : 0;
    });
}

int ipfg_autotune(int enable) {
    return guarded([&] { g_tuner.set_enabled(enable != 0); });
}

int ipfg_tune_export(char* buf, size_t capacity, size_t* needed) {
    return guarded([&] {
        abi_check(needed != nullptr, "null length");
        const std::string t = g_tuner.export_text();
        *needed = t.size() + 1;
        if (buf && capacity >= t.size() + 1) std::memcpy(buf, t.c_str(), t.size() + 1);
    });
}

int ipfg_tune_import(const char* text, int* entries) {
    return guarded([&] {
        abi_check(text != nullptr, "null text");
        const int n = g_tuner.import_text(text);
        abi_check(n >= 0, "malformed tuning table");
        if (entries) *entries = n;
    });
}

int ipfg_tune_clear(void) {
    return guarded([&] { g_tuner.clear(); });
}

void ipfg_set_num_threads(int n) { g_threads = n < 0 ? 0 : n; }
int ipfg_get_num_threads(void) {
    const int n = g_threads.load();
    if (n > 0) return n;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? static_cast<int>(hw) : 1;
}

int ipfg_malloc(void** ptr, size_t bytes) {
    return guarded([&] {
        abi_check(ptr != nullptr, "null pointer");
        ensure_init();
        HIPCALL(hipMalloc(ptr, bytes ? bytes : 1));
    });
}
int ipfg_free(void* ptr) {
    return guarded([&] {
        ensure_init();
        if (ptr) HIPCALL(hipFree(ptr));
    });
}
int ipfg_memcpy_h2d(void* dst, const void* src, size_t bytes, ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        const hipStream_t s = pick_stream(stream);
        HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
        HIPCALL(hipStreamSynchronize(s));
    });
}
int ipfg_memcpy_d2h(void* dst, const void* src, size_t bytes, ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        const hipStream_t s = pick_stream(stream);
        HIPCALL(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
        HIPCALL(hipStreamSynchronize(s));
    });
}
int ipfg_memset(void* dst, int value, size_t bytes, ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        HIPCALL(hipMemsetAsync(dst, value, bytes, pick_stream(stream)));
    });
}
int ipfg_stream_create(ipfg_stream* stream) {
    return guarded([&] {
        abi_check(stream != nullptr, "null stream pointer");
        ensure_init();
        hipStream_t s;
        HIPCALL(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}
// Waits for st and moves the cached events last recorded on it to the
// library's own stream (their work is complete once st has drained).
void detach_stream(hipStream_t st) {
    HIPCALL(hipStreamSynchronize(st));
    forget_stream(st);
    g_ring.forget(st);
    g_params.forget(st);
    g_scratch.forget(st);
    std::lock_guard<std::mutex> lock(g_exec_mu);
    for (BatchExecutor* x : g_execs) x->forget(st);
}

int ipfg_stream_destroy(ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        if (!stream) return;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        detach_stream(st);
        HIPCALL(hipStreamDestroy(st));
    });
}
int ipfg_stream_forget(ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        if (stream) detach_stream(static_cast<hipStream_t>(stream));
    });
}
int ipfg_stream_synchronize(ipfg_stream stream) {
    return guarded([&] {
        ensure_init();
        HIPCALL(hipStreamSynchronize(pick_stream(stream)));
    });
}
int ipfg_device_synchronize(void) {
    return guarded([&] {
        ensure_init();
        HIPCALL(hipDeviceSynchronize());
    });
}

int ipfg_gaussian_kernel_1d(double sigma, int ksize, double* out, int capacity) {
    int n = 0;
    const int rc = guarded([&] {
        const std::vector<double> k = gaussian_kernel(sigma, ksize);
        for (int i = 0; i < static_cast<int>(k.size()) && i < capacity; ++i) out[i] = k[i];
        n = static_cast<int>(k.size());
    });
    return rc == IPFG_OK ? n : -rc;
}

int ipfg_sep_filter(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky, int kh,
                    int ax, int ay, int border, double delta, double border_value, int precision,
                    ipfg_stream stream) {
    return guarded([&] { sep_impl(src, dst, kx, kw, ky, kh, ax, ay, border, delta, border_value, precision, stream); });
}

int ipfg_sep_filter_noalloc(const ipfg_image* src, const ipfg_image* dst, const double* kx, int kw, const double* ky,
                            int kh, int ax, int ay, int border, double delta, double border_value, int precision,
                            ipfg_stream stream) {
    return guarded(
        [&] { sep_noalloc_impl(src, dst, kx, kw, ky, kh, ax, ay, border, delta, border_value, precision, stream); });
}

int ipfg_gaussian_blur(const ipfg_image* src, const ipfg_image* dst, double sigma_x, double sigma_y, int ksize,
                       int border, int precision, ipfg_stream stream) {
    return guarded([&] {
        if (sigma_y <= 0) sigma_y = sigma_x;
        const std::vector<double> kx = gaussian_kernel(sigma_x, ksize);
        const std::vector<double> ky = gaussian_kernel(sigma_y, ksize);
        sep_impl(src, dst, kx.data(), static_cast<int>(kx.size()), ky.data(), static_cast<int>(ky.size()),
                 static_cast<int>(kx.size()) / 2, static_cast<int>(ky.size()) / 2, border, 0.0, 0.0, precision,
                 stream);
    });
}

int ipfg_box_filter(const ipfg_image* src, const ipfg_image* dst, int kw, int kh, int normalize, int border,
                    double border_value, int precision, ipfg_stream stream) {
    return guarded([&] { box_impl(src, dst, kw, kh, normalize, border, border_value, precision, stream); });
}

int ipfg_filter2d(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int ax, int ay,
                  int border, double delta, double border_value, int precision, ipfg_stream stream) {
    return guarded(
        [&] { f2d_impl(src, dst, kernel, kw, kh, ax, ay, border, delta, border_value, precision, stream); });
}

int ipfg_median(const ipfg_image* src, const ipfg_image* dst, int ksize, int border, ipfg_stream stream) {
    return guarded([&] { median_impl(src, dst, ksize, border, stream); });
}

int ipfg_bilateral(const ipfg_image* src, const ipfg_image* dst, int d, double sigma_color, double sigma_space,
                   int border, int precision, ipfg_stream stream) {
    return guarded([&] { bilateral_impl(src, dst, d, sigma_color, sigma_space, border, precision, stream); });
}

int ipfg_sobel(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, int ksize, double scale, int border,
               int precision, ipfg_stream stream) {
    return guarded([&] { sobel_impl(src, dst, dx, dy, ksize, scale, border, precision, stream); });
}

int ipfg_scharr(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, double scale, int border, int precision,
                ipfg_stream stream) {
    return guarded([&] {
        diff3_impl(src, dst, dx, dy, {-scale, 0.0, scale}, {3.0, 10.0, 3.0}, border, precision, stream,
                   "scharr: (dx, dy) must be (1,0) or (0,1)");
    });
}

int ipfg_prewitt(const ipfg_image* src, const ipfg_image* dst, int dx, int dy, int border, int precision,
                 ipfg_stream stream) {
    return guarded([&] {
        diff3_impl(src, dst, dx, dy, {-1.0, 0.0, 1.0}, {1.0, 1.0, 1.0}, border, precision, stream,
                   "prewitt: (dx, dy) must be (1,0) or (0,1)");
    });
}

int ipfg_laplacian(const ipfg_image* src, const ipfg_image* dst, int ksize, double scale, int border, int precision,
                   ipfg_stream stream) {
    return guarded([&] { laplacian_impl(src, dst, ksize, scale, border, precision, stream); });
}

int ipfg_magnitude(const ipfg_image* gx, const ipfg_image* gy, const ipfg_image* dst, ipfg_stream stream) {
    return guarded([&] { polar_impl(gx, gy, dst, 0, stream); });
}

int ipfg_phase(const ipfg_image* gx, const ipfg_image* gy, const ipfg_image* dst, int degrees, ipfg_stream stream) {
    return guarded([&] { polar_impl(gx, gy, dst, degrees ? 1 : 2, stream); });
}

int ipfg_ssim(const ipfg_image* a, const ipfg_image* b, const ipfg_image* ssim_map, double* result,
              ipfg_stream stream) {
    return guarded([&] { ssim_impl(a, b, ssim_map, result, stream); });
}

int ipfg_pyr_down(const ipfg_image* src, const ipfg_image* dst, ipfg_stream stream) {
    return guarded([&] { pyr_impl(src, dst, false, stream); });
}

int ipfg_pyr_up(const ipfg_image* src, const ipfg_image* dst, ipfg_stream stream) {
    return guarded([&] { pyr_impl(src, dst, true, stream); });
}

int ipfg_unsharp_mask(const ipfg_image* src, const ipfg_image* dst, double sigma, double amount, double threshold,
                      int precision, ipfg_stream stream) {
    return guarded([&] { unsharp_impl(src, dst, sigma, amount, threshold, precision, stream); });
}

int ipfg_sharpen(const ipfg_image* src, const ipfg_image* dst, double strength, int precision, ipfg_stream stream) {
    return guarded([&] {
        const float s = static_cast<float>(strength);
        const float k[9] = {0, -s, 0, -s, 1 + 4 * s, -s, 0, -s, 0};
        filter3x3_impl(src, dst, k, precision, stream);
    });
}

int ipfg_emboss(const ipfg_image* src, const ipfg_image* dst, int precision, ipfg_stream stream) {
    return guarded([&] {
        const float k[9] = {-2, -1, 0, -1, 1, 1, 0, 1, 2};
        filter3x3_impl(src, dst, k, precision, stream);
    });
}

int ipfg_adaptive_threshold(const ipfg_image* src, const ipfg_image* dst, double maxval, int method, int type,
                            int block_size, double c, int precision, ipfg_stream stream) {
    return guarded([&] { adaptive_impl(src, dst, maxval, method, type, block_size, c, precision, stream); });
}

int ipfg_hough_circles(const ipfg_image* src, int min_radius, int max_radius, double canny_high, int center_votes,
                       double min_dist, double min_coverage, double* circles, int capacity, int* count,
                       ipfg_stream stream) {
    return guarded([&] {
        hough_impl(src, min_radius, max_radius, canny_high, center_votes, min_dist, min_coverage, circles, capacity,
                   count, stream);
    });
}

int ipfg_pnm_info(const uint8_t* bytes, size_t n, int* width, int* height, int* channels, int* is_pfm) {
    return guarded([&] { pnm_info_impl(bytes, n, width, height, channels, is_pfm); });
}

int ipfg_decode_pnm(const uint8_t* bytes, size_t n, const ipfg_image* dst, ipfg_stream stream) {
    return guarded([&] { pnm_decode_impl(bytes, n, dst, stream); });
}

int ipfg_integral_image(const ipfg_image* src, const ipfg_image* dst, int squared, ipfg_stream stream) {
    return guarded([&] { integral_impl(src, dst, squared, stream); });
}

int ipfg_filter2d_fft(const ipfg_image* src, const ipfg_image* dst, const float* kernel, int kw, int kh, int border,
                      double delta, double border_value, ipfg_stream stream) {
    return guarded([&] { fft_impl(src, dst, kernel, kw, kh, border, delta, border_value, stream); });
}

int ipfg_canny(const ipfg_image* src, const ipfg_image* dst, double low_threshold, double high_threshold, int ksize,
               int l2_gradient, ipfg_stream stream) {
    return guarded([&] { canny_impl(src, dst, low_threshold, high_threshold, ksize, l2_gradient, stream); });
}

int ipfg_canny_auto(const ipfg_image* src, const ipfg_image* dst, double blur_sigma, ipfg_stream stream) {
    return guarded([&] { canny_auto_impl(src, dst, blur_sigma, stream); });
}

int ipfg_corner_harris(const ipfg_image* src, const ipfg_image* dst, int block_size, int ksize, double k,
                       ipfg_stream stream) {
    return guarded([&] { corner_impl(src, dst, block_size, ksize, true, k, stream); });
}

int ipfg_corner_min_eigen_val(const ipfg_image* src, const ipfg_image* dst, int block_size, int ksize,
                              ipfg_stream stream) {
    return guarded([&] { corner_impl(src, dst, block_size, ksize, false, 0.0, stream); });
}

int ipfg_batch_create(ipfg_batch* batch, int max_jobs, int use_graph) {
    return guarded([&] {
        abi_check(batch != nullptr, "null batch");
        *batch = new BatchExecutor(max_jobs, use_graph != 0);
    });
}

int ipfg_batch_run(ipfg_batch batch, const ipfg_job* jobs, int njobs, ipfg_stream stream) {
    return guarded([&] {
        abi_check(batch != nullptr, "null batch");
        abi_check(njobs >= 0 && (jobs != nullptr || njobs == 0), "null jobs");
        ensure_init();
        static_cast<BatchExecutor*>(batch)->run(jobs, njobs, pick_stream(stream));
    });
}

int ipfg_batch_destroy(ipfg_batch batch) {
    return guarded([&] { delete static_cast<BatchExecutor*>(batch); });
}

int ipfg_server_create(ipfg_server* server, size_t device_budget_bytes, int max_batch, int use_graph) {
    return guarded([&] {
        abi_check(server != nullptr, "null server");
        *server = new Server(device_budget_bytes, max_batch, use_graph != 0);
    });
}

int ipfg_server_submit(ipfg_server server, const ipfg_job* request, long long* id) {
    return guarded([&] {
        abi_check(server != nullptr && request != nullptr, "null server or request");
        const long long v = static_cast<Server*>(server)->submit(*request);
        if (id) *id = v;
    });
}

int ipfg_server_submit_many(ipfg_server server, const ipfg_job* requests, int n, long long* first_id) {
    return guarded([&] {
        abi_check(server != nullptr && (requests != nullptr || n == 0) && n >= 0, "null server or requests");
        for (int i = 0; i < n; ++i) {
            const long long v = static_cast<Server*>(server)->submit(requests[i]);
            if (i == 0 && first_id) *first_id = v;
        }
    });
}

int ipfg_server_poll(ipfg_server server, long long* ids, double* latency_ms, int capacity, int* count) {
    return guarded([&] {
        abi_check(server != nullptr && count != nullptr && (ids != nullptr || capacity == 0), "null argument");
        *count = static_cast<Server*>(server)->poll(ids, latency_ms, capacity);
    });
}

int ipfg_server_drain(ipfg_server server) {
    return guarded([&] {
        abi_check(server != nullptr, "null server");
        static_cast<Server*>(server)->drain();
    });
}

int ipfg_server_stats(ipfg_server server, long long* batches, int* max_jobs_seen) {
    return guarded([&] {
        abi_check(server != nullptr, "null server");
        static_cast<Server*>(server)->stats(batches, max_jobs_seen);
    });
}

int ipfg_server_memory(ipfg_server server, size_t* held_bytes, size_t* peak_in_use_bytes) {
    return guarded([&] {
        abi_check(server != nullptr, "null server");
        static_cast<Server*>(server)->memory(held_bytes, peak_in_use_bytes);
    });
}

int ipfg_server_destroy(ipfg_server server) {
    return guarded([&] { delete static_cast<Server*>(server); });
}

}  // extern "C"
