// ipf front end for MI355X: the filtering path.
//
// Drop-in for the reference's include/ipf/filter.hpp:118-387 -- same names,
// template parameters, defaults, return-by-value and "ipf: ..." errors:
//   Kernel / makeKernel (118-125), gaussianKernel1D (127-145),
//   gaussianKernel2D (147-155), filter2D (157-194), sepFilter2D (196-203),
//   boxFilter (205-253), blur (255-259), gaussianBlur (261-267),
//   medianBlur (300-332), bilateralFilter (334-387).
// Every call validates its arguments in the reference's order, then runs the
// gfx950 kernels of libipfg.so through the C ABI (ipfg.h) and returns when
// the result is in the output image. Link with -lipfg.
//
// GPU-specific knobs live in namespace ipf::gpu (precision per thread, a
// stream per thread); they do not change the reference's signatures.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "../ipfg.h"
#include "core.hpp"
#include "parallel.hpp"

namespace ipf {

namespace gpu {
/// Arithmetic of the linear and bilateral filters (ipfg_precision).
enum class Precision { Default = IPFG_PREC_DEFAULT, Exact = IPFG_PREC_EXACT, FP32 = IPFG_PREC_FP32,
                       BF16x3 = IPFG_PREC_BF16X3, FP16x2 = IPFG_PREC_FP16X2 };

inline Precision& precisionSetting() {
    thread_local Precision p = Precision::Default;
    return p;
}
/// Precision used by this thread's filter calls (Default: FP64 accumulation
/// certified against the reference's rounding, bit-identical integer and
/// float32 outputs).
inline void setPrecision(Precision p) { precisionSetting() = p; }
inline Precision getPrecision() { return precisionSetting(); }

inline ipfg_stream& streamSetting() {
    thread_local ipfg_stream s = nullptr;
    return s;
}
/// HIP stream for this thread's calls (nullptr: the library's per-thread stream).
inline void setStream(ipfg_stream s) { streamSetting() = s; }

/// Samples recomputed in FP64 by this thread's last certified call.
inline long long lastFixups() {
    long long n = 0;
    ipfg_last_fixups(&n);
    return n;
}
}  // namespace gpu

namespace detail {

template <class TOut, class T>
using OutT = std::conditional_t<std::is_void<TOut>::value, T, TOut>;

template <class T>
constexpr int abiType() {
    if constexpr (std::is_same<T, std::uint8_t>::value) return IPFG_U8;
    else if constexpr (std::is_same<T, std::uint16_t>::value) return IPFG_U16;
    else if constexpr (std::is_same<T, std::int16_t>::value) return IPFG_I16;
    else if constexpr (std::is_same<T, std::int32_t>::value) return IPFG_I32;
    else if constexpr (std::is_same<T, float>::value) return IPFG_F32;
    else if constexpr (std::is_same<T, double>::value) return IPFG_F64;
    else return -1;
}

template <class T>
ipfg_image describe(const Image<T>& img) {
    static_assert(abiType<T>() >= 0,
                  "ipf on MI355X filters uint8, uint16, int16, int32, float and double images");
    ipfg_image d;
    d.data = const_cast<T*>(img.data());
    d.width = img.width();
    d.height = img.height();
    d.channels = img.channels();
    d.dtype = abiType<T>();
    d.row_stride_bytes = 0;
    d.on_device = 0;
    return d;
}

/// Maps a library status to ipf::Error (argument errors carry ipf's text).
inline void check(int status) {
    if (status == IPFG_OK) return;
    const char* msg = ipfg_last_error();
    throw Error(msg && *msg ? msg : "ipf: GPU filter failed");
}

inline int precision() { return static_cast<int>(gpu::getPrecision()); }

}  // namespace detail

/// A correlation kernel is a single-channel float image.
using Kernel = ImageF;

inline Kernel makeKernel(int width, int height, std::initializer_list<float> values) {
    detail::require(static_cast<int>(values.size()) == width * height, "makeKernel: wrong number of values");
    return Kernel(width, height, 1, std::vector<float>(values));
}

/// Normalised 1-D Gaussian. ksize <= 0 takes the radius int(3 sigma + 0.5)
/// (at least 1); sigma <= 0 takes OpenCV's sigma for ksize.
inline std::vector<double> gaussianKernel1D(double sigma, int ksize = 0) {
    int n = ipfg_gaussian_kernel_1d(sigma, ksize, nullptr, 0);
    if (n <= 0) detail::check(-n);
    std::vector<double> k(static_cast<std::size_t>(n));
    n = ipfg_gaussian_kernel_1d(sigma, ksize, k.data(), n);
    if (n <= 0) detail::check(-n);
    return k;
}

/// Outer product of two 1-D Gaussians, as float.
inline Kernel gaussianKernel2D(double sigma, int ksize = 0) {
    const std::vector<double> g = gaussianKernel1D(sigma, ksize);
    const int n = static_cast<int>(g.size());
    Kernel k(n, n, 1);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) k(x, y) = static_cast<float>(g[y] * g[x]);
    return k;
}

/// out(x, y) = delta + sum over non-zero k(i, j) of k(i, j) * src(x + i - ax, y + j - ay).
template <class TOut = void, class T>
Image<detail::OutT<TOut, T>> filter2D(const Image<T>& src, const Kernel& kernel,
                                      BorderMode border = BorderMode::Reflect101, double delta = 0.0,
                                      Point anchor = Point(-1, -1), double borderValue = 0.0) {
    using O = detail::OutT<TOut, T>;
    detail::require(kernel.channels() == 1 && !kernel.empty(),
                    "filter2D: kernel must be a non-empty single-channel image");
    const int kw = kernel.width(), kh = kernel.height();
    const int ax = anchor.x < 0 ? kw / 2 : anchor.x, ay = anchor.y < 0 ? kh / 2 : anchor.y;
    detail::require(ax < kw && ay < kh, "filter2D: anchor outside kernel");
    Image<O> out(src.width(), src.height(), src.channels());
    if (src.empty()) return out;
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_filter2d(&s, &d, kernel.data(), kw, kh, ax, ay, static_cast<int>(border), delta, borderValue,
                                detail::precision(), gpu::streamSetting()));
    return out;
}

/// Horizontal kx then vertical ky correlation, anchors at the centres.
template <class TOut = void, class T>
Image<detail::OutT<TOut, T>> sepFilter2D(const Image<T>& src, const std::vector<double>& kx,
                                         const std::vector<double>& ky, BorderMode border = BorderMode::Reflect101,
                                         double delta = 0.0, double borderValue = 0.0) {
    using O = detail::OutT<TOut, T>;
    detail::require(!kx.empty() && !ky.empty(), "separable filter: empty kernel");
    Image<O> out(src.width(), src.height(), src.channels());
    if (src.empty()) return out;
    const int kw = static_cast<int>(kx.size()), kh = static_cast<int>(ky.size());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_sep_filter(&s, &d, kx.data(), kw, ky.data(), kh, kw / 2, kh / 2, static_cast<int>(border),
                                  delta, borderValue, detail::precision(), gpu::streamSetting()));
    return out;
}

/// Sum (or mean) over a kw x kh window anchored at (kw/2, kh/2).
template <class TOut = void, class T>
Image<detail::OutT<TOut, T>> boxFilter(const Image<T>& src, int kw, int kh, bool normalize = true,
                                       BorderMode border = BorderMode::Reflect101, double borderValue = 0.0) {
    using O = detail::OutT<TOut, T>;
    detail::require(kw > 0 && kh > 0, "boxFilter: kernel size must be positive");
    Image<O> out(src.width(), src.height(), src.channels());
    if (src.empty()) return out;
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_box_filter(&s, &d, kw, kh, normalize ? 1 : 0, static_cast<int>(border), borderValue,
                                  detail::precision(), gpu::streamSetting()));
    return out;
}

/// Mean over a square ksize x ksize window.
template <class TOut = void, class T>
Image<detail::OutT<TOut, T>> blur(const Image<T>& src, int ksize, BorderMode border = BorderMode::Reflect101) {
    return boxFilter<TOut>(src, ksize, ksize, true, border);
}

/// Gaussian blur; sigmaY <= 0 uses sigmaX, ksize <= 0 derives the window.
template <class TOut = void, class T>
Image<detail::OutT<TOut, T>> gaussianBlur(const Image<T>& src, double sigmaX, double sigmaY = 0.0, int ksize = 0,
                                          BorderMode border = BorderMode::Reflect101) {
    if (sigmaY <= 0) sigmaY = sigmaX;
    return sepFilter2D<TOut>(src, gaussianKernel1D(sigmaX, ksize), gaussianKernel1D(sigmaY, ksize), border);
}

/// Exact median of each ksize x ksize window (ksize odd); Constant is rejected.
template <class T>
Image<T> medianBlur(const Image<T>& src, int ksize, BorderMode border = BorderMode::Replicate) {
    detail::require(ksize >= 1 && ksize % 2 == 1, "medianBlur: ksize must be odd and positive");
    detail::require(border != BorderMode::Constant, "medianBlur: Constant border is not supported");
    if (src.empty() || ksize == 1) return src;
    Image<T> out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_median(&s, &d, ksize, static_cast<int>(border), gpu::streamSetting()));
    return out;
}

/// Edge-preserving bilateral filter: circular window of diameter d (d <= 0:
/// radius round(1.5 sigmaSpace)), Gaussian space and L1-colour range weights.
template <class T>
Image<T> bilateralFilter(const Image<T>& src, int d, double sigmaColor, double sigmaSpace,
                         BorderMode border = BorderMode::Reflect101) {
    detail::require(sigmaColor > 0 && sigmaSpace > 0, "bilateralFilter: sigmas must be positive");
    Image<T> out(src.width(), src.height(), src.channels());
    if (src.empty()) return out;
    const ipfg_image s = detail::describe(src), o = detail::describe(out);
    detail::check(ipfg_bilateral(&s, &o, d, sigmaColor, sigmaSpace, static_cast<int>(border), detail::precision(),
                                 gpu::streamSetting()));
    return out;
}

}  // namespace ipf
