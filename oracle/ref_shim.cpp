// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
//
// C-ABI wrapper around the reference ipf headers, compiled by oracle/Makefile
// from the reference source tree (include/ipf/filter.hpp, geometry.hpp,
// core.hpp, parallel.hpp, features.hpp, color.hpp, fft.hpp, codec/pnm.hpp) into oracle/_ref/libipf_ref.so. It exists to (1) pin the oracle
// restatement bit-for-bit against the real reference, (2) write the golden
// fixtures under tests/golden/, and (3) serve as the "reference" CPU baseline.
// Signatures match oracle/ipf_oracle.cpp's ipfo_* functions one for one.
#include <ipf/features.hpp>
#include <ipf/fft.hpp>
#include <ipf/io.hpp>
#include <ipf/filter.hpp>
#include <ipf/geometry.hpp>
#include <ipf/metrics.hpp>

#include <chrono>
#include <cstring>
#include <string>

namespace {
thread_local std::string g_error;
thread_local double g_seconds = 0.0;

template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_error = e.what();
        return -1;
    }
}

template <class F>
int with_type(int dt, F f) {
    switch (dt) {
        case 0: return f(static_cast<std::uint8_t*>(nullptr));
        case 1: return f(static_cast<std::uint16_t*>(nullptr));
        case 2: return f(static_cast<std::int16_t*>(nullptr));
        case 3: return f(static_cast<std::int32_t*>(nullptr));
        case 4: return f(static_cast<float*>(nullptr));
        case 5: return f(static_cast<double*>(nullptr));
    }
    g_error = "ipf: unsupported pixel type";
    return -1;
}

template <class T>
ipf::Image<T> wrap(const void* p, int w, int h, int cn) {
    return ipf::Image<T>::fromData(w, h, cn, static_cast<const T*>(p));
}

template <class T>
void unwrap(const ipf::Image<T>& img, void* dst) {
    if (!img.empty()) std::memcpy(dst, img.data(), img.sampleCount() * sizeof(T));
}

template <class F>
auto timed(F f) {
    const auto t0 = std::chrono::steady_clock::now();
    auto r = f();
    g_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return r;
}

ipf::BorderMode bm(int b) { return static_cast<ipf::BorderMode>(b); }
}  // namespace

extern "C" {

const char* ipfr_last_error(void) { return g_error.c_str(); }
double ipfr_last_seconds(void) { return g_seconds; }
void ipfr_set_num_threads(int n) { ipf::setNumThreads(n); }
int ipfr_get_num_threads(void) { return ipf::getNumThreads(); }
int ipfr_border_index(int p, int len, int mode) { return ipf::borderIndex(p, len, bm(mode)); }

int ipfr_saturate(double v, int dtype, void* out) {
    return with_type(dtype, [&](auto* tag) {
        using T = std::remove_pointer_t<decltype(tag)>;
        *static_cast<T*>(out) = ipf::saturate_cast<T>(v);
        return 0;
    });
}

int ipfr_gaussian_kernel_1d(double sigma, int ksize, double* out, int cap) {
    return guarded([&] {
        const std::vector<double> k = ipf::gaussianKernel1D(sigma, ksize);
        for (int i = 0; i < static_cast<int>(k.size()) && i < cap; ++i) out[i] = k[i];
        return static_cast<int>(k.size());
    });
}

int ipfr_sep_filter(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const double* kx, int kw,
                    const double* ky, int kh, int ax, int ay, int border, double delta, double bval) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                const auto img = wrap<Ti>(src, w, h, cn);
                const std::vector<double> vx(kx, kx + kw), vy(ky, ky + kh);
                const auto out = timed([&] {
                    return ipf::detail::sepFilterImpl<To>(img, vx, vy, ax, ay, bm(border), delta, bval);
                });
                unwrap(out, dst);
                return 0;
            });
        });
    });
}

int ipfr_box_filter(const void* src, int w, int h, int cn, int tin, void* dst, int tout, int kw, int kh, int normalize,
                    int border, double bval) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                const auto img = wrap<Ti>(src, w, h, cn);
                const auto out =
                    timed([&] { return ipf::boxFilter<To>(img, kw, kh, normalize != 0, bm(border), bval); });
                unwrap(out, dst);
                return 0;
            });
        });
    });
}

int ipfr_filter2d(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const float* K, int kw, int kh,
                  int ax, int ay, int border, double delta, double bval) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                const auto img = wrap<Ti>(src, w, h, cn);
                const ipf::Kernel k = ipf::Kernel::fromData(kw, kh, 1, K);
                const auto out = timed([&] {
                    return ipf::filter2D<To>(img, k, bm(border), delta, ipf::Point(ax, ay), bval);
                });
                unwrap(out, dst);
                return 0;
            });
        });
    });
}

int ipfr_filter2d_fft(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const float* K, int kw,
                      int kh, int border, double delta, double bval) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                const auto img = wrap<Ti>(src, w, h, cn);
                const ipf::ImageF k = ipf::ImageF::fromData(kw, kh, 1, K);
                const auto out = timed([&] { return ipf::filter2DFFT<To>(img, k, bm(border), delta, bval); });
                unwrap(out, dst);
                return 0;
            });
        });
    });
}

int ipfr_integral_image(const void* src, int w, int h, int cn, int dtype, double* dst, int squared) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            const auto out = timed([&] { return ipf::integralImage(img, squared != 0); });
            unwrap(out, dst);
            return 0;
        });
    });
}

int ipfr_hough_circles(const void* src, int w, int h, int cn, int dtype, int min_r, int max_r, double canny_high,
                       int center_votes, double min_dist, double min_coverage, double* out, int capacity, int* count) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            const auto cs = timed([&] {
                return ipf::houghCircles(img, min_r, max_r, canny_high, center_votes, min_dist, min_coverage);
            });
            *count = static_cast<int>(cs.size());
            for (int i = 0; i < static_cast<int>(cs.size()) && i < capacity; ++i) {
                out[4 * i] = cs[i].center.x;
                out[4 * i + 1] = cs[i].center.y;
                out[4 * i + 2] = cs[i].radius;
                out[4 * i + 3] = cs[i].votes;
            }
            return 0;
        });
    });
}

// Netpbm / PFM codecs (codec/pnm.hpp): decode into out (capacity samples),
// shape into w, h, cn; encode into out (capacity bytes), size into *needed.
int ipfr_decode_pnm(const unsigned char* bytes, size_t n, int depth16, void* out, size_t capacity, int* w, int* h,
                    int* cn) {
    return guarded([&] {
        const std::vector<std::uint8_t> b(bytes, bytes + n);
        auto put = [&](const auto& img) {
            *w = img.width();
            *h = img.height();
            *cn = img.channels();
            if (img.sampleCount() <= capacity) unwrap(img, out);
            return 0;
        };
        return depth16 ? put(ipf::decodePNM<std::uint16_t>(b)) : put(ipf::decodePNM<std::uint8_t>(b));
    });
}

int ipfr_decode_pfm(const unsigned char* bytes, size_t n, float* out, size_t capacity, int* w, int* h, int* cn) {
    return guarded([&] {
        const ipf::ImageF img = ipf::decodePFM(std::vector<std::uint8_t>(bytes, bytes + n));
        *w = img.width();
        *h = img.height();
        *cn = img.channels();
        if (img.sampleCount() <= capacity) unwrap(img, out);
        return 0;
    });
}

int ipfr_encode_pnm(const void* src, int w, int h, int cn, int dtype, int binary, unsigned char* out,
                    size_t capacity, size_t* needed) {
    return guarded([&] {
        std::vector<std::uint8_t> b;
        if (dtype == 0) b = ipf::encodePNM(wrap<std::uint8_t>(src, w, h, cn), binary != 0);
        else b = ipf::encodePNM(wrap<std::uint16_t>(src, w, h, cn), binary != 0);
        *needed = b.size();
        if (b.size() <= capacity) std::memcpy(out, b.data(), b.size());
        return 0;
    });
}

int ipfr_encode_pfm(const void* src, int w, int h, int cn, int dtype, unsigned char* out, size_t capacity,
                    size_t* needed) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const std::vector<std::uint8_t> b = ipf::encodePFM(wrap<T>(src, w, h, cn));
            *needed = b.size();
            if (b.size() <= capacity) std::memcpy(out, b.data(), b.size());
            return 0;
        });
    });
}

int ipfr_median(const void* src, int w, int h, int cn, int dtype, void* dst, int ksize, int border) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            const auto out = timed([&] { return ipf::medianBlur(img, ksize, bm(border)); });
            unwrap(out, dst);
            return 0;
        });
    });
}

int ipfr_bilateral(const void* src, int w, int h, int cn, int dtype, void* dst, int d, double sigma_color,
                   double sigma_space, int border) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            const auto out =
                timed([&] { return ipf::bilateralFilter(img, d, sigma_color, sigma_space, bm(border)); });
            unwrap(out, dst);
            return 0;
        });
    });
}

// ---- derivative and sharpening filters (filter.hpp:388-567)
int ipfr_sobel(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, int ksize, double scale,
               int border) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            const auto img = wrap<Ti>(src, w, h, cn);
            unwrap(timed([&] { return ipf::sobel(img, dx, dy, ksize, scale, bm(border)); }), dst);
            return 0;
        });
    });
}

int ipfr_scharr(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, double scale, int border) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            const auto img = wrap<Ti>(src, w, h, cn);
            unwrap(timed([&] { return ipf::scharr(img, dx, dy, scale, bm(border)); }), dst);
            return 0;
        });
    });
}

int ipfr_prewitt(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, int border) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            const auto img = wrap<Ti>(src, w, h, cn);
            unwrap(timed([&] { return ipf::prewitt(img, dx, dy, bm(border)); }), dst);
            return 0;
        });
    });
}

int ipfr_laplacian(const void* src, int w, int h, int cn, int tin, float* dst, int ksize, double scale, int border) {
    return guarded([&] {
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            const auto img = wrap<Ti>(src, w, h, cn);
            unwrap(timed([&] { return ipf::laplacian(img, ksize, scale, bm(border)); }), dst);
            return 0;
        });
    });
}

int ipfr_magnitude(const float* gx, const float* gy, int w, int h, int cn, float* dst) {
    return guarded([&] {
        const auto a = ipf::ImageF::fromData(w, h, cn, gx), b = ipf::ImageF::fromData(w, h, cn, gy);
        unwrap(timed([&] { return ipf::magnitude(a, b); }), dst);
        return 0;
    });
}

int ipfr_phase(const float* gx, const float* gy, int w, int h, int cn, float* dst, int degrees) {
    return guarded([&] {
        const auto a = ipf::ImageF::fromData(w, h, cn, gx), b = ipf::ImageF::fromData(w, h, cn, gy);
        unwrap(timed([&] { return ipf::phase(a, b, degrees != 0); }), dst);
        return 0;
    });
}

int ipfr_unsharp(const void* src, int w, int h, int cn, int dtype, void* dst, double sigma, double amount,
                 double threshold) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::unsharpMask(img, sigma, amount, threshold); }), dst);
            return 0;
        });
    });
}

int ipfr_sharpen(const void* src, int w, int h, int cn, int dtype, void* dst, double strength) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::sharpen(img, strength); }), dst);
            return 0;
        });
    });
}

int ipfr_ssim(const void* pa, const void* pb, int w, int h, int cn, int dtype, float* map, double* result) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto a = wrap<T>(pa, w, h, cn), b = wrap<T>(pb, w, h, cn);
            ipf::ImageF m;
            *result = timed([&] { return ipf::ssim(a, b, map ? &m : nullptr); });
            if (map) unwrap(m, map);
            return 0;
        });
    });
}

int ipfr_pyr_down(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::pyrDown(img); }), dst);
            return 0;
        });
    });
}

int ipfr_pyr_up(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::pyrUp(img); }), dst);
            return 0;
        });
    });
}

int ipfr_emboss(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::emboss(img); }), dst);
            return 0;
        });
    });
}

int ipfr_adaptive_threshold(const void* src, int w, int h, int dtype, void* dst, double maxval, int method, int type,
                            int block, double c) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, 1);
            unwrap(timed([&] {
                       return ipf::adaptiveThreshold(img, maxval, static_cast<ipf::AdaptiveMethod>(method),
                                                     static_cast<ipf::ThresholdType>(type), block, c);
                   }),
                   dst);
            return 0;
        });
    });
}

int ipfr_motion_kernel(int length, double angle, float* out, int cap) {
    return guarded([&] {
        const ipf::Kernel k = ipf::motionBlurKernel(length, angle);
        const int n = static_cast<int>(k.sampleCount());
        for (int i = 0; i < n && i < cap; ++i) out[i] = k.data()[i];
        return k.width();
    });
}

// ---- feature callers of the filters (features.hpp:26-160)
int ipfr_canny(const void* src, int w, int h, int cn, int dtype, std::uint8_t* dst, double low, double high, int ksize,
               int l2) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::canny(img, low, high, ksize, l2 != 0); }), dst);
            return 0;
        });
    });
}

int ipfr_canny_auto(const void* src, int w, int h, int cn, int dtype, std::uint8_t* dst, double sigma) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::cannyAuto(img, sigma); }), dst);
            return 0;
        });
    });
}

int ipfr_corner_harris(const void* src, int w, int h, int cn, int dtype, float* dst, int block, int ksize, double k) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::cornerHarris(img, block, ksize, k); }), dst);
            return 0;
        });
    });
}

int ipfr_corner_min_eigen(const void* src, int w, int h, int cn, int dtype, float* dst, int block, int ksize) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const auto img = wrap<T>(src, w, h, cn);
            unwrap(timed([&] { return ipf::cornerMinEigenVal(img, block, ksize); }), dst);
            return 0;
        });
    });
}

}  // extern "C"
