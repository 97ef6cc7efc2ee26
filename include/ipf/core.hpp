// ipf front end for MI355X: core types.
//
// Same public surface and semantics as the reference's include/ipf/core.hpp
// (Error 34-44, PixelTraits 54-65, saturate_cast 67-103, geometry 115-171,
// BorderMode/borderIndex 177-214, Image<T> 224-401 with aliases 403-408), so
// code written against ipf compiles unchanged. Filtering itself runs on the
// GPU through libipfg (see ipf/filter.hpp).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <initializer_list>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "parallel.hpp"

#define IPF_VERSION_MAJOR 1
#define IPF_VERSION_MINOR 1
#define IPF_VERSION_PATCH 0

namespace ipf {

/// Every error raised by ipf (argument checks, device failures).
class Error : public std::runtime_error {
public:
    using std::runtime_error::runtime_error;
};

namespace detail {
[[noreturn]] inline void fail(const std::string& msg) { throw Error("ipf: " + msg); }
inline void require(bool ok, const char* msg) {
    if (!ok) fail(msg);
}
constexpr double kPi = 3.14159265358979323846;
}  // namespace detail

/// Nominal pixel range: [0, max()] -- 255 for uint8, 65535 for uint16, 1 for
/// floating point, 255 for 32-bit signed label/accumulator images.
template <class T>
struct PixelTraits {
    static constexpr bool is_float = std::is_floating_point<T>::value;
    static constexpr double max() {
        if constexpr (is_float) {
            return 1.0;
        } else if constexpr (std::is_signed<T>::value && sizeof(T) >= 4) {
            return 255.0;
        } else {
            return static_cast<double>(std::numeric_limits<T>::max());
        }
    }
};

namespace detail {
template <class To, class From>
constexpr To saturate_int_from_int(From v) {
    constexpr bool fs = std::is_signed<From>::value, ts = std::is_signed<To>::value;
    if constexpr (fs && !ts) {
        if (v < 0) return To(0);
        return static_cast<unsigned long long>(v) > static_cast<unsigned long long>(std::numeric_limits<To>::max())
                   ? std::numeric_limits<To>::max()
                   : static_cast<To>(v);
    } else if constexpr (!fs && ts) {
        return static_cast<unsigned long long>(v) > static_cast<unsigned long long>(std::numeric_limits<To>::max())
                   ? std::numeric_limits<To>::max()
                   : static_cast<To>(v);
    } else {
        using Wide = std::conditional_t<fs, long long, unsigned long long>;
        const Wide w = static_cast<Wide>(v);
        if (w < static_cast<Wide>(std::numeric_limits<To>::lowest())) return std::numeric_limits<To>::lowest();
        if (w > static_cast<Wide>(std::numeric_limits<To>::max())) return std::numeric_limits<To>::max();
        return static_cast<To>(v);
    }
}
}  // namespace detail

/// Converts to To: float -> integer rounds half away from zero and clamps
/// (NaN -> 0); integer -> integer clamps; anything else is a plain cast.
template <class To, class From>
constexpr To saturate_cast(From v) {
    if constexpr (std::is_same<To, From>::value) {
        return v;
    } else if constexpr (std::is_integral<To>::value && std::is_floating_point<From>::value) {
        if (v != v) return To(0);
        double d = static_cast<double>(v);
        d += d < 0 ? -0.5 : 0.5;
        if (d <= static_cast<double>(std::numeric_limits<To>::lowest())) return std::numeric_limits<To>::lowest();
        if (d >= static_cast<double>(std::numeric_limits<To>::max())) return std::numeric_limits<To>::max();
        return static_cast<To>(d);
    } else if constexpr (std::is_integral<To>::value && std::is_integral<From>::value) {
        return detail::saturate_int_from_int<To>(v);
    } else {
        return static_cast<To>(v);
    }
}

template <class T>
constexpr T clamp(T v, T lo, T hi) {
    return v < lo ? lo : (hi < v ? hi : v);
}

// ---------------------------------------------------------------- geometry
struct Point {
    int x = 0, y = 0;
    constexpr Point() = default;
    constexpr Point(int px, int py) : x(px), y(py) {}
    friend constexpr bool operator==(Point a, Point b) { return a.x == b.x && a.y == b.y; }
    friend constexpr bool operator!=(Point a, Point b) { return !(a == b); }
    friend constexpr Point operator+(Point a, Point b) { return {a.x + b.x, a.y + b.y}; }
    friend constexpr Point operator-(Point a, Point b) { return {a.x - b.x, a.y - b.y}; }
};

struct PointF {
    double x = 0, y = 0;
    constexpr PointF() = default;
    constexpr PointF(double px, double py) : x(px), y(py) {}
    constexpr PointF(Point p) : x(p.x), y(p.y) {}
    friend constexpr PointF operator+(PointF a, PointF b) { return {a.x + b.x, a.y + b.y}; }
    friend constexpr PointF operator-(PointF a, PointF b) { return {a.x - b.x, a.y - b.y}; }
    friend constexpr PointF operator*(PointF a, double s) { return {a.x * s, a.y * s}; }
    friend constexpr bool operator==(PointF a, PointF b) { return a.x == b.x && a.y == b.y; }
};

struct Size {
    int width = 0, height = 0;
    constexpr Size() = default;
    constexpr Size(int w, int h) : width(w), height(h) {}
    friend constexpr bool operator==(Size a, Size b) { return a.width == b.width && a.height == b.height; }
    friend constexpr bool operator!=(Size a, Size b) { return !(a == b); }
};

struct Rect {
    int x = 0, y = 0, width = 0, height = 0;
    constexpr Rect() = default;
    constexpr Rect(int px, int py, int w, int h) : x(px), y(py), width(w), height(h) {}
    constexpr int right() const { return x + width; }
    constexpr int bottom() const { return y + height; }
    constexpr long long area() const { return static_cast<long long>(width) * height; }
    constexpr bool empty() const { return width <= 0 || height <= 0; }
    constexpr bool contains(Point p) const { return p.x >= x && p.y >= y && p.x < right() && p.y < bottom(); }
    /// Overlap of two rectangles (empty when disjoint).
    constexpr Rect operator&(const Rect& o) const {
        const int l = x > o.x ? x : o.x, t = y > o.y ? y : o.y;
        const int r = right() < o.right() ? right() : o.right(), b = bottom() < o.bottom() ? bottom() : o.bottom();
        return (r > l && b > t) ? Rect(l, t, r - l, b - t) : Rect();
    }
    /// Bounding box of two rectangles (an empty operand is ignored).
    constexpr Rect operator|(const Rect& o) const {
        if (empty()) return o;
        if (o.empty()) return *this;
        const int l = x < o.x ? x : o.x, t = y < o.y ? y : o.y;
        const int r = right() > o.right() ? right() : o.right(), b = bottom() > o.bottom() ? bottom() : o.bottom();
        return Rect(l, t, r - l, b - t);
    }
    friend constexpr bool operator==(const Rect& a, const Rect& b) {
        return a.x == b.x && a.y == b.y && a.width == b.width && a.height == b.height;
    }
    friend constexpr bool operator!=(const Rect& a, const Rect& b) { return !(a == b); }
};

// ---------------------------------------------------------------- borders
/// Out-of-image pixel synthesis (OpenCV names). Reflect101 is the filters'
/// default; Constant uses a border value.
enum class BorderMode { Constant, Replicate, Reflect, Reflect101, Wrap };

/// Index of the image pixel that stands for coordinate p of a line of `len`
/// pixels, or -1 (Constant outside, or an empty line).
inline int borderIndex(int p, int len, BorderMode mode) {
    if (p >= 0 && p < len) return p;
    if (len <= 0) return -1;
    auto fold = [](long long q, long long period) { return ((q % period) + period) % period; };
    switch (mode) {
        case BorderMode::Constant:
            return -1;
        case BorderMode::Replicate:
            return p < 0 ? 0 : len - 1;
        case BorderMode::Reflect: {
            const long long q = fold(p, 2LL * len);
            return static_cast<int>(q < len ? q : 2LL * len - 1 - q);
        }
        case BorderMode::Reflect101: {
            if (len == 1) return 0;
            const long long q = fold(p, 2LL * len - 2);
            return static_cast<int>(q < len ? q : 2LL * len - 2 - q);
        }
        case BorderMode::Wrap:
            return static_cast<int>(fold(p, len));
    }
    return -1;
}

// ---------------------------------------------------------------- image
/// Dense image with interleaved channels, row-major: sample (x, y, c) is
/// data()[(y * width + x) * channels + c] -- a numpy (H, W, C) array.
/// Copies are deep (value semantics).
template <class T>
class Image {
    static_assert(std::is_arithmetic<T>::value, "ipf::Image pixel type must be arithmetic");

public:
    using value_type = T;

    Image() = default;
    Image(int width, int height, int channels = 1, T fill = T{}) { reset(width, height, channels, fill); }
    Image(int width, int height, int channels, std::vector<T> pixels) {
        validate(width, height, channels);
        detail::require(pixels.size() == count(width, height, channels),
                        "Image: pixel vector size does not match dimensions");
        w_ = width;
        h_ = height;
        c_ = channels;
        px_ = std::move(pixels);
    }

    /// Copies width*height*channels interleaved values from a foreign buffer.
    static Image fromData(int width, int height, int channels, const T* src) {
        Image img(width, height, channels);
        if (!img.px_.empty()) {
            detail::require(src != nullptr, "Image::fromData: null source pointer");
            std::memcpy(img.px_.data(), src, img.px_.size() * sizeof(T));
        }
        return img;
    }

    void reset(int width, int height, int channels = 1, T fill = T{}) {
        validate(width, height, channels);
        w_ = width;
        h_ = height;
        c_ = channels;
        px_.assign(count(width, height, channels), fill);
    }

    int width() const { return w_; }
    int height() const { return h_; }
    int channels() const { return c_; }
    Size size() const { return Size(w_, h_); }
    Rect rect() const { return Rect(0, 0, w_, h_); }
    bool empty() const { return px_.empty(); }
    std::size_t pixelCount() const { return static_cast<std::size_t>(w_) * h_; }
    std::size_t sampleCount() const { return px_.size(); }
    std::size_t stride() const { return static_cast<std::size_t>(w_) * c_; }

    T* data() { return px_.data(); }
    const T* data() const { return px_.data(); }
    std::vector<T>& vector() { return px_; }
    const std::vector<T>& vector() const { return px_; }

    T* row(int y) { return px_.data() + stride() * static_cast<std::size_t>(y); }
    const T* row(int y) const { return px_.data() + stride() * static_cast<std::size_t>(y); }
    T* ptr(int x, int y) { return row(y) + static_cast<std::size_t>(x) * c_; }
    const T* ptr(int x, int y) const { return row(y) + static_cast<std::size_t>(x) * c_; }

    T& operator()(int x, int y, int c = 0) { return px_[offset(x, y, c)]; }
    const T& operator()(int x, int y, int c = 0) const { return px_[offset(x, y, c)]; }

    T& at(int x, int y, int c = 0) {
        bounds(x, y, c);
        return px_[offset(x, y, c)];
    }
    const T& at(int x, int y, int c = 0) const {
        bounds(x, y, c);
        return px_[offset(x, y, c)];
    }

    bool inside(int x, int y) const { return x >= 0 && y >= 0 && x < w_ && y < h_; }

    /// Sample with border extrapolation outside the image.
    T atBorder(int x, int y, int c, BorderMode mode, T borderValue = T{}) const {
        const int sx = borderIndex(x, w_, mode), sy = borderIndex(y, h_, mode);
        return (sx < 0 || sy < 0) ? borderValue : px_[offset(sx, sy, c)];
    }

    void fill(T v) { std::fill(px_.begin(), px_.end(), v); }
    void fill(std::initializer_list<T> color) { fill(std::vector<T>(color)); }
    void fill(const std::vector<T>& color) {
        detail::require(static_cast<int>(color.size()) == c_, "Image::fill: color size != channels");
        for (std::size_t i = 0; i < px_.size(); i += static_cast<std::size_t>(c_))
            std::copy(color.begin(), color.end(), px_.begin() + static_cast<std::ptrdiff_t>(i));
    }

    void setPixel(int x, int y, const T* color) { std::copy(color, color + c_, ptr(x, y)); }

    /// Copy of a region, clipped to the image.
    Image crop(Rect r) const {
        r = r & rect();
        if (r.empty()) return Image();
        Image out(r.width, r.height, c_);
        const std::size_t n = static_cast<std::size_t>(r.width) * c_ * sizeof(T);
        for (int y = 0; y < r.height; ++y) std::memcpy(out.row(y), ptr(r.x, r.y + y), n);
        return out;
    }

    /// Writes src with its top-left corner at (x, y), clipped to this image.
    void paste(const Image& src, int x, int y) {
        detail::require(src.channels() == c_, "Image::paste: channel count mismatch");
        const Rect d = Rect(x, y, src.width(), src.height()) & rect();
        const std::size_t n = static_cast<std::size_t>(d.width) * c_ * sizeof(T);
        for (int yy = d.y; yy < d.bottom(); ++yy) std::memcpy(ptr(d.x, yy), src.ptr(d.x - x, yy - y), n);
    }

    bool sameShape(int w, int h, int c) const { return w_ == w && h_ == h && c_ == c; }
    template <class U>
    bool sameShape(const Image<U>& o) const {
        return sameShape(o.width(), o.height(), o.channels());
    }

    /// out = saturate_cast<U>(in * scale + offset).
    template <class U>
    Image<U> convertTo(double scale = 1.0, double offset = 0.0) const {
        Image<U> out(w_, h_, c_);
        const T* s = px_.data();
        U* d = out.data();
        const bool plain = scale == 1.0 && offset == 0.0;
        parallelForN(px_.size(), [&](std::size_t a, std::size_t b) {
            for (std::size_t i = a; i < b; ++i)
                d[i] = plain ? saturate_cast<U>(s[i]) : saturate_cast<U>(static_cast<double>(s[i]) * scale + offset);
        }, 0.5);
        return out;
    }

    auto begin() { return px_.begin(); }
    auto end() { return px_.end(); }
    auto begin() const { return px_.begin(); }
    auto end() const { return px_.end(); }

    friend bool operator==(const Image& a, const Image& b) {
        return a.sameShape(b) && a.px_ == b.px_;
    }
    friend bool operator!=(const Image& a, const Image& b) { return !(a == b); }

private:
    static std::size_t count(int w, int h, int c) {
        return static_cast<std::size_t>(w) * static_cast<std::size_t>(h) * static_cast<std::size_t>(c);
    }
    static void validate(int w, int h, int c) {
        detail::require(w >= 0 && h >= 0, "Image: negative dimensions");
        detail::require(c >= 1, "Image: channel count must be >= 1");
        const unsigned long long n = static_cast<unsigned long long>(w) * static_cast<unsigned long long>(h) *
                                     static_cast<unsigned long long>(c);
        detail::require(n <= std::numeric_limits<std::size_t>::max() / sizeof(T) && n < (1ULL << 40),
                        "Image: dimensions too large");
    }
    std::size_t offset(int x, int y, int c) const {
        return (static_cast<std::size_t>(y) * w_ + static_cast<std::size_t>(x)) * c_ + c;
    }
    void bounds(int x, int y, int c) const {
        if (x >= 0 && y >= 0 && x < w_ && y < h_ && c >= 0 && c < c_) return;
        detail::fail("Image::at: coordinate (" + std::to_string(x) + ", " + std::to_string(y) + ", " +
                     std::to_string(c) + ") out of range for " + std::to_string(w_) + "x" + std::to_string(h_) + "x" +
                     std::to_string(c_) + " image");
    }

    int w_ = 0, h_ = 0, c_ = 1;
    std::vector<T> px_;
};

using ImageU8 = Image<std::uint8_t>;
using ImageU16 = Image<std::uint16_t>;
using ImageI16 = Image<std::int16_t>;
using ImageI32 = Image<std::int32_t>;
using ImageF = Image<float>;
using ImageD = Image<double>;

}  // namespace ipf
