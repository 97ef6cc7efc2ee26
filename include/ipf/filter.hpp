// ipf front end for MI355X: the filtering path.
//
// Drop-in for the reference's include/ipf/filter.hpp:118-387 -- same names,
// template parameters, defaults, return-by-value and "ipf: ..." errors:
//   Kernel / makeKernel (118-125), gaussianKernel1D (127-145),
//   gaussianKernel2D (147-155), filter2D (157-194), sepFilter2D (196-203),
//   boxFilter (205-253), blur (255-259), gaussianBlur (261-267),
//   medianBlur (300-332), bilateralFilter (334-387), and the derivative and
//   sharpening filters of 388-565: sobel, scharr, prewitt, laplacian,
//   magnitude, phase, gradientMagnitude, unsharpMask, sharpen, emboss,
//   motionBlurKernel, adaptiveThreshold, integralImage and integralSum.
// Every call validates its arguments in the reference's order, then runs the
// gfx950 kernels of libipfg.so through the C ABI (ipfg.h) and returns when
// the result is in the output image. Link with -lipfg.
//
// GPU-specific knobs live in namespace ipf::gpu (precision per thread, a
// stream per thread); they do not change the reference's signatures.
#pragma once

#include <cmath>
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

/// Sobel derivative of order (dx, dy), odd ksize in [1, 31] (1: central
/// differences without smoothing); scale multiplies the result.
template <class T>
ImageF sobel(const Image<T>& src, int dx, int dy, int ksize = 3, double scale = 1.0,
             BorderMode border = BorderMode::Reflect101) {
    ImageF out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_sobel(&s, &d, dx, dy, ksize, scale, static_cast<int>(border), detail::precision(),
                             gpu::streamSetting()));
    return out;
}

/// 3x3 Scharr derivative; (dx, dy) is (1, 0) or (0, 1).
template <class T>
ImageF scharr(const Image<T>& src, int dx, int dy, double scale = 1.0, BorderMode border = BorderMode::Reflect101) {
    ImageF out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_scharr(&s, &d, dx, dy, scale, static_cast<int>(border), detail::precision(),
                              gpu::streamSetting()));
    return out;
}

/// 3x3 Prewitt derivative (box-smoothed central difference); (dx, dy) is (1, 0) or (0, 1).
template <class T>
ImageF prewitt(const Image<T>& src, int dx, int dy, BorderMode border = BorderMode::Reflect101) {
    ImageF out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_prewitt(&s, &d, dx, dy, static_cast<int>(border), detail::precision(), gpu::streamSetting()));
    return out;
}

/// Laplacian: ksize 1 correlates with [0 1 0; 1 -4 1; 0 1 0] * scale; larger
/// ksize adds the second-order Sobel derivatives in x and y (float sums).
template <class T>
ImageF laplacian(const Image<T>& src, int ksize = 1, double scale = 1.0, BorderMode border = BorderMode::Reflect101) {
    ImageF out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_laplacian(&s, &d, ksize, scale, static_cast<int>(border), detail::precision(),
                                 gpu::streamSetting()));
    return out;
}

/// hypot(gx, gy) per sample.
inline ImageF magnitude(const ImageF& gx, const ImageF& gy) {
    detail::require(gx.sameShape(gy), "magnitude: shape mismatch");
    ImageF out(gx.width(), gx.height(), gx.channels());
    const ipfg_image a = detail::describe(gx), b = detail::describe(gy), d = detail::describe(out);
    detail::check(ipfg_magnitude(&a, &b, &d, gpu::streamSetting()));
    return out;
}

/// atan2(gy, gx) in degrees [0, 360) (radians [0, 2 pi) when degrees is false).
inline ImageF phase(const ImageF& gx, const ImageF& gy, bool degrees = true) {
    detail::require(gx.sameShape(gy), "phase: shape mismatch");
    ImageF out(gx.width(), gx.height(), gx.channels());
    const ipfg_image a = detail::describe(gx), b = detail::describe(gy), d = detail::describe(out);
    detail::check(ipfg_phase(&a, &b, &d, degrees ? 1 : 0, gpu::streamSetting()));
    return out;
}

/// Magnitude of the Sobel gradient (dx = 1 and dy = 1 images of ksize).
template <class T>
ImageF gradientMagnitude(const Image<T>& src, int ksize = 3) {
    return magnitude(sobel(src, 1, 0, ksize), sobel(src, 0, 1, ksize));
}

/// out = src + amount * (src - gaussianBlur(src, sigma)) where
/// |src - blurred| > threshold, src elsewhere; saturated to T.
template <class T>
Image<T> unsharpMask(const Image<T>& src, double sigma = 1.0, double amount = 1.0, double threshold = 0.0) {
    Image<T> out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_unsharp_mask(&s, &d, sigma, amount, threshold, detail::precision(), gpu::streamSetting()));
    return out;
}

/// 3x3 sharpening kernel {0 -s 0; -s 1+4s -s; 0 -s 0}, s = float(strength).
template <class T>
Image<T> sharpen(const Image<T>& src, double strength = 1.0) {
    Image<T> out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_sharpen(&s, &d, strength, detail::precision(), gpu::streamSetting()));
    return out;
}

/// 3x3 emboss kernel {-2 -1 0; -1 1 1; 0 1 2}.
template <class T>
Image<T> emboss(const Image<T>& src) {
    Image<T> out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_emboss(&s, &d, detail::precision(), gpu::streamSetting()));
    return out;
}

/// Normalised line of `length` pixels through the centre of a square kernel
/// (odd side length or length + 1) at angleDegrees, counter-clockwise from
/// horizontal, drawn from 4 * length points. Built on the host.
inline Kernel motionBlurKernel(int length, double angleDegrees) {
    detail::require(length >= 1, "motionBlurKernel: length must be >= 1");
    const int n = length | 1;
    Kernel k(n, n, 1, 0.f);
    const double rad = angleDegrees * detail::kPi / 180.0, mid = 0.5 * (n - 1);
    const double cx = std::cos(rad), sy = std::sin(rad);
    const int points = 4 * length;
    for (int i = 0; i < points; ++i) {
        const double t = (length - 1) * (static_cast<double>(i) / (points - 1) - 0.5);
        const int x = clamp(static_cast<int>(std::lround(mid + t * cx)), 0, n - 1);
        const int y = clamp(static_cast<int>(std::lround(mid - t * sy)), 0, n - 1);
        k(x, y) = 1.f;
    }
    float total = 0.f;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) total += k(x, y);
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) k(x, y) /= total;
    return k;
}

/// Summed-area table with an extra zero row/column: I(x, y) = sum of src over [0,x) x [0,y).
/// Size (W+1) x (H+1), same channel count, double precision.
template <class T>
ImageD integralImage(const Image<T>& src, bool squared = false) {
    ImageD I(src.width() + 1, src.height() + 1, src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(I);
    detail::check(ipfg_integral_image(&s, &d, squared ? 1 : 0, gpu::streamSetting()));
    return I;
}

/// Sum over the rectangle r using an integral image from integralImage().
inline double integralSum(const ImageD& I, Rect r, int channel = 0) {
    return I(r.right(), r.bottom(), channel) - I(r.x, r.bottom(), channel) - I(r.right(), r.y, channel) +
           I(r.x, r.y, channel);
}

/// Threshold types (the reference declares them in pixel_ops.hpp:253-259);
/// adaptiveThreshold takes Binary and BinaryInv.
enum class ThresholdType { Binary, BinaryInv, Trunc, ToZero, ToZeroInv };

/// Local threshold of adaptiveThreshold: the blockSize box mean or the
/// Gaussian-weighted mean (Replicate border), minus C.
enum class AdaptiveMethod { Mean, Gaussian };

/// maxval where src - localMean > -C (Binary; BinaryInv swaps the outputs),
/// 0 elsewhere; single-channel images, odd blockSize >= 3.
template <class T>
Image<T> adaptiveThreshold(const Image<T>& src, double maxval, AdaptiveMethod method, ThresholdType type, int blockSize,
                           double C) {
    detail::require(src.channels() == 1, "adaptiveThreshold: expected a single-channel image");
    detail::require(blockSize >= 3 && blockSize % 2 == 1, "adaptiveThreshold: blockSize must be odd and >= 3");
    detail::require(type == ThresholdType::Binary || type == ThresholdType::BinaryInv,
                    "adaptiveThreshold: type must be Binary or BinaryInv");
    Image<T> out(src.width(), src.height(), 1);
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_adaptive_threshold(&s, &d, maxval, static_cast<int>(method), static_cast<int>(type), blockSize,
                                          C, detail::precision(), gpu::streamSetting()));
    return out;
}

}  // namespace ipf
