// Shared device/host helpers for the gfx950 filtering kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <limits>
#include <type_traits>

#include "../../include/ipfg_types.h"

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

// Same question for a float32 output. Rounding to float is monotone, so v's
// float is fixed when both ends of [r - b, r + b] round to the same float
// (compared as bits, so -0 and +0 differ). The ends are computed in FP64:
// widening the bound by 4.5e-16 |r| keeps each rounded end outside the exact
// interval whenever b <= 1.5 |r|; a wider interval spans zero, its ends get
// opposite signs and the sample is recomputed. A bound equal to the callers'
// 1e-300 floor means every term and the offset were exact zeros: both the FMA
// chain and the reference's chain then give +0.
__device__ inline bool needs_exact_f32(double r, double bound) {
    if (!(fabs(r) < 1e300) || !(bound < 1e300)) return true;  // inf or NaN
    if (bound <= 1e-300) return false;
    const double b = fma(fabs(r), 4.5e-16, bound);
    return __float_as_uint(static_cast<float>(r - b)) != __float_as_uint(static_cast<float>(r + b));
}

// ------------------------------------------------------------------ wave helpers
__device__ __forceinline__ int lane_id() { return static_cast<int>(__lane_id()); }

// Value of v in lane l (l wave-uniform).
__device__ __forceinline__ int bcast_i(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double bcast_d(double v, int l) {
    const long long b = __double_as_longlong(v);
    const unsigned lo = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b), l));
    const unsigned hi = static_cast<unsigned>(__builtin_amdgcn_readlane(static_cast<int>(b >> 32), l));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Maximum of v over the 64 lanes of the wave, returned wave-uniform. DPP row
// shifts 1, 2, 4, 8 leave each row of 16 lanes with its maximum in its last
// lane; the row broadcasts 15 and 31 carry those into the rows above, so lane
// 63 ends with the wave's maximum. Lanes a shift reads from outside the row
// contribute 0, the identity for unsigned values.
__device__ __forceinline__ unsigned int wave_max_u32(unsigned int v) {
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, false)));
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false)));
    v = max(v, static_cast<unsigned int>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false)));
    return static_cast<unsigned int>(__builtin_amdgcn_readlane(static_cast<int>(v), 63));
}

// Maximum of a non-negative float (or +inf) over the wave, added to *m (an
// LDS word holding float bits) by one lane. Every lane of the wave calls it.
// Non-negative floats order like their bit patterns; a NaN's bits exceed
// +inf's, so a NaN makes the maximum a NaN.
__device__ __forceinline__ void wave_max_into(unsigned int* m, float v) {
    const unsigned int b = wave_max_u32(__float_as_uint(v));
    if (lane_id() == 0) atomicMax(m, b);
}

// ------------------------------------------------------------------ fix-ups
// Certified launches recompute doubtful samples in place: after a wave has
// written a group of outputs, wave_fix() walks the wave's doubtful ones and,
// for each, runs the exact recomputation on all 64 lanes at once (the
// reference's sequential sums are split into lane-parallel products or rows
// and one broadcast chain). Nothing is left for a later pass, so a launch has
// no serial tail however many samples are flagged.
//
// Counting: each wave adds its recomputed samples to a counter that belongs
// to the stream (zero between launches); the last workgroup to finish copies
// it to `report` and resets it and the finished-workgroup ticket.
struct FixList {
    unsigned long long* count;
    unsigned int* done;
    unsigned long long* report;  // receives the launch's count
};

// `mask` bit m set: this lane's output m is doubtful. All 64 lanes of the
// wave must call this together; fix(l, m) is then called by the whole wave
// (l and m uniform) for each doubtful output m of lane l, in lane order.
template <class Fix>
__device__ __forceinline__ void wave_fix(const FixList& f, unsigned int mask, Fix fix) {
    unsigned long long bal = __ballot(mask != 0u);
    if (bal == 0ull) return;
    unsigned int total = 0;
    while (bal) {
        const int l = __builtin_ctzll(bal);
        bal &= bal - 1ull;
        unsigned int mm = static_cast<unsigned int>(bcast_i(static_cast<int>(mask), l));
        total += __builtin_popcount(mm);
        while (mm) {
            const int m = __builtin_ctz(mm);
            mm &= mm - 1u;
            fix(l, m);
        }
    }
    if (lane_id() == 0) atomicAdd(f.count, static_cast<unsigned long long>(total));
}

// Called by every thread of every workgroup of a certified launch after its
// last wave_fix(). Each wave waits for its own memory operations before the
// workgroup takes a ticket; the last workgroup reads the counter with an
// agent-scope atomic load (no L2 write-back or invalidate is issued).
__device__ inline void fix_finish(const FixList& f) {
    __shared__ int last;
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int nblk = gridDim.x * gridDim.y * gridDim.z;
        last = atomicAdd(f.done, 1u) == nblk - 1;
    }
    __syncthreads();
    if (!last || threadIdx.x != 0) return;
    *f.report = __hip_atomic_load(f.count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f.count, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(f.done, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ------------------------------------------------------------------ image
struct DevImage {
    void* data;
    int w, h, cn, dt;
    size_t stride;  // bytes per row
};

}  // namespace ipfg
