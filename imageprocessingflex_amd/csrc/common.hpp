// Shared device/host helpers for the gfx950 filtering kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "../../include/ipfg.h"

namespace ipfg {

// ------------------------------------------------------------------ launch
constexpr int kBlock = 256;  // 4 wave64s per workgroup

// ------------------------------------------------------------------ types
template <int DT>
struct DTypeOf;
template <> struct DTypeOf<IPFG_U8> { using type = uint8_t; };
template <> struct DTypeOf<IPFG_U16> { using type = uint16_t; };
template <> struct DTypeOf<IPFG_I16> { using type = int16_t; };
template <> struct DTypeOf<IPFG_I32> { using type = int32_t; };
template <> struct DTypeOf<IPFG_F32> { using type = float; };
template <> struct DTypeOf<IPFG_F64> { using type = double; };

__host__ __device__ inline size_t dtype_size(int dt) {
    switch (dt) {
        case IPFG_U8: return 1;
        case IPFG_U16: return 2;
        case IPFG_I16: return 2;
        case IPFG_I32: return 4;
        case IPFG_F32: return 4;
        case IPFG_F64: return 8;
    }
    return 0;
}
__host__ __device__ inline bool dtype_is_int(int dt) { return dt <= IPFG_I32; }
// Integer range of a type (only meaningful for integer dtypes).
__host__ __device__ inline double dtype_lo(int dt) {
    switch (dt) {
        case IPFG_U8: return 0.0;
        case IPFG_U16: return 0.0;
        case IPFG_I16: return -32768.0;
        case IPFG_I32: return -2147483648.0;
    }
    return 0.0;
}
__host__ __device__ inline double dtype_hi(int dt) {
    switch (dt) {
        case IPFG_U8: return 255.0;
        case IPFG_U16: return 65535.0;
        case IPFG_I16: return 32767.0;
        case IPFG_I32: return 2147483647.0;
    }
    return 0.0;
}

// ------------------------------------------------------------------ border
// ipf::borderIndex (core.hpp:187-214): maps p into [0, len) or -1 (Constant).
__host__ __device__ inline int border_index(int p, int len, int mode) {
    if (p >= 0 && p < len) return p;
    if (len <= 0 || mode == IPFG_BORDER_CONSTANT) return -1;
    if (mode == IPFG_BORDER_REPLICATE) return p < 0 ? 0 : len - 1;
    if (mode == IPFG_BORDER_WRAP) {
        const int q = p % len;
        return q < 0 ? q + len : q;
    }
    if (mode == IPFG_BORDER_REFLECT) {
        const long long per = 2LL * len;
        long long q = p % per;
        if (q < 0) q += per;
        return static_cast<int>(q < len ? q : per - 1 - q);
    }
    // Reflect101
    if (len == 1) return 0;
    const long long per = 2LL * len - 2;
    long long q = p % per;
    if (q < 0) q += per;
    return static_cast<int>(q < len ? q : per - q);
}

// ------------------------------------------------------------------ saturate
// ipf::saturate_cast<To>(double) (core.hpp:67-103): integers round half away
// from zero and clamp; NaN -> 0; float targets convert with round-to-nearest.
template <class To>
__host__ __device__ inline To saturate_from_double(double v) {
    if constexpr (std::is_integral<To>::value) {
        if (v != v) return To(0);
        constexpr double lo = static_cast<double>(std::numeric_limits<To>::lowest());
        constexpr double hi = static_cast<double>(std::numeric_limits<To>::max());
        const double d = v < 0 ? v - 0.5 : v + 0.5;
        if (d <= lo) return std::numeric_limits<To>::lowest();
        if (d >= hi) return std::numeric_limits<To>::max();
        return static_cast<To>(d);
    } else {
        return static_cast<To>(v);
    }
}

// Stores a double result into an image of runtime dtype `dt`.
__device__ inline void store_sample(void* base, size_t byte_off, int dt, double v) {
    char* p = static_cast<char*>(base) + byte_off;
    switch (dt) {
        case IPFG_U8: *reinterpret_cast<uint8_t*>(p) = saturate_from_double<uint8_t>(v); break;
        case IPFG_U16: *reinterpret_cast<uint16_t*>(p) = saturate_from_double<uint16_t>(v); break;
        case IPFG_I16: *reinterpret_cast<int16_t*>(p) = saturate_from_double<int16_t>(v); break;
        case IPFG_I32: *reinterpret_cast<int32_t*>(p) = saturate_from_double<int32_t>(v); break;
        case IPFG_F32: *reinterpret_cast<float*>(p) = static_cast<float>(v); break;
        case IPFG_F64: *reinterpret_cast<double*>(p) = v; break;
    }
}

// Loads one sample of runtime dtype `dt` as double.
__device__ inline double load_sample(const void* base, size_t byte_off, int dt) {
    const char* p = static_cast<const char*>(base) + byte_off;
    switch (dt) {
        case IPFG_U8: return *reinterpret_cast<const uint8_t*>(p);
        case IPFG_U16: return *reinterpret_cast<const uint16_t*>(p);
        case IPFG_I16: return *reinterpret_cast<const int16_t*>(p);
        case IPFG_I32: return *reinterpret_cast<const int32_t*>(p);
        case IPFG_F32: return *reinterpret_cast<const float*>(p);
        case IPFG_F64: return *reinterpret_cast<const double*>(p);
    }
    return 0.0;
}

// ------------------------------------------------------------------ certify
// Decides whether an approximate result r (|r - v| <= bound, v the value the
// reference computes in FP64) determines saturate_cast<int type>(v) exactly.
// The integer conversion is constant between consecutive half-integers and
// outside [lo + 0.5, hi - 0.5]; anything else must be recomputed exactly.
__device__ inline bool needs_exact(double r, double bound, double lo, double hi) {
    if (!(r == r) || !(bound < 1e300)) return true;  // NaN or no usable bound: recompute
    if (r + bound < lo + 0.5 || r - bound > hi - 0.5) return false;  // saturated either way
    const double t = r + 0.5;
    const double fr = t - floor(t);  // position between the jumps at integer t
    return fr <= bound || fr >= 1.0 - bound;
}

// Same question for a float32 output: static_cast<float>(v) is constant
// between the midpoints of r's float neighbours. The midpoint on r's side of
// f = (float)r is half a float spacing away: ulp(f) / 2 above, and below
// either ulp(f) / 2 or, for a power of two, ulp(f) / 4.
__device__ inline bool needs_exact_f32(double r, double bound) {
    const float f = static_cast<float>(r);
    const unsigned bits = __float_as_uint(f);
    const int e = static_cast<int>((bits >> 23) & 0xffu);
    if (e == 0xff || !(bound < 1e300)) return true;  // inf or NaN
    const double side = r - static_cast<double>(f);
    const bool toward_zero = (side < 0.0) == (f > 0.0f);
    const bool halved = toward_zero && (bits & 0x7fffffu) == 0u && e > 1;
    // 2^(max(e, 1) - 151) is half a spacing; one more halving on the short side
    const int ee = (e > 1 ? e : 1) - (halved ? 1 : 0);
    const double half = __longlong_as_double(static_cast<long long>(ee + 872) << 52);
    return fabs(side) + bound >= half;
}

// Device-side fix-up list of a certified launch. The counter and the
// finished-workgroup ticket belong to the stream (zero between launches); the
// last workgroup of the launch recomputes the listed samples exactly and
// resets both. When more than `capacity` samples are flagged the last
// workgroup recomputes every sample.
struct FixList {
    unsigned long long* count;
    unsigned int* done;
    unsigned long long* report;  // receives the launch's flag count
    unsigned long long* index;
    unsigned long long capacity;
};

// Flags one sample. The index is written with an agent-scope atomic store
// (coherent across the XCDs' L2 caches), so no device-wide fence is needed.
__device__ inline void fix_push(const FixList& f, unsigned long long idx) {
    const unsigned long long slot = atomicAdd(f.count, 1ull);
    if (slot < f.capacity) __hip_atomic_store(f.index + slot, idx, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Called by every thread of every workgroup after its last flag. The last
// workgroup to arrive calls fix(i) for each flagged index (or every index
// below `total` on overflow) and leaves the stream's counters at zero.
// Ordering: each wave waits for its own stores to complete before the
// workgroup takes a ticket, and the last workgroup reads the list with
// agent-scope atomic loads -- no L2 write-back or invalidate is issued.
template <class Fix>
__device__ inline void fix_epilogue(const FixList& f, unsigned long long total, Fix fix) {
    __shared__ int last;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int nblk = gridDim.x * gridDim.y * gridDim.z;
        last = atomicAdd(f.done, 1u) == nblk - 1;
    }
    __syncthreads();
    if (!last) return;
    const unsigned long long n = __hip_atomic_load(f.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) *f.report = n;
    if (n <= f.capacity) {
        for (unsigned long long k = threadIdx.x; k < n; k += blockDim.x)
            fix(__hip_atomic_load(f.index + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    } else {
        for (unsigned long long k = threadIdx.x; k < total; k += blockDim.x) fix(k);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        __hip_atomic_store(f.count, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(f.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Grid used by fix-up kernels (grid-stride loops over the list or the image).
constexpr int kFixBlocks = 1024;

// ------------------------------------------------------------------ image
struct DevImage {
    void* data;
    int w, h, cn, dt;
    size_t stride;  // bytes per row
};

}  // namespace ipfg
