// TEST INFRASTRUCTURE -- NOT PART OF THE PRODUCT.
//
// CPU restatement ("the oracle") of ipf's filtering path, used only by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker and the
// CPU baseline. The shipped library (libipfg.so) never links or calls it.
//
// Every function restates the reference algorithm with the same FP64 operation
// order, so that results are bit-identical to the reference on the same inputs.
// Build flags: -O2 -ffp-contract=off (no fused multiply-add, no -ffast-math).
//
// Reference citations (paths relative to the ipf source tree):
//   border index ............ include/ipf/core.hpp:187-214
//   saturate_cast ........... include/ipf/core.hpp:67-103
//   parallel row split ...... include/ipf/parallel.hpp:55-95
//   border table ............ include/ipf/filter.hpp:27-32
//   separable filter ........ include/ipf/filter.hpp:34-90, 196-203
//   gaussian kernel ......... include/ipf/filter.hpp:127-145, 261-267
//   filter2D ................ include/ipf/filter.hpp:157-194
//   box filter / blur ....... include/ipf/filter.hpp:205-259
//   median .................. include/ipf/filter.hpp:269-332
//   bilateral ............... include/ipf/filter.hpp:334-387
//   derivative kernels ...... include/ipf/filter.hpp:93-114
//   sobel/scharr/prewitt .... include/ipf/filter.hpp:392-421
//   laplacian ............... include/ipf/filter.hpp:423-436
//   magnitude / phase ....... include/ipf/filter.hpp:438-458
//   unsharp/sharpen/emboss .. include/ipf/filter.hpp:472-496
//   motion-blur kernel ...... include/ipf/filter.hpp:498-516
//   adaptive threshold ...... include/ipf/filter.hpp:545-565
//   pyrDown / pyrUp ......... include/ipf/geometry.hpp:501-524
//   canny / cannyAuto ....... include/ipf/features.hpp:26-101
//   toGray / rgbToGray ...... include/ipf/color.hpp:51-81, core.hpp:465-473
//   corner responses ........ include/ipf/features.hpp:107-160
//   filter2DFFT ............. include/ipf/fft.hpp:21-98, 163-216; core.hpp:526-544
//   integralImage ........... include/ipf/filter.hpp:522-538
//
// Pinning: every function here is pinned by tests/test_oracle.py against the
// reference-built fixtures in tests/golden/ and against closed forms, brute force
// and the reference's own known-answer tests.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <exception>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

namespace {

enum DType { DT_U8 = 0, DT_U16 = 1, DT_I16 = 2, DT_I32 = 3, DT_F32 = 4, DT_F64 = 5 };
enum Border { B_CONSTANT = 0, B_REPLICATE = 1, B_REFLECT = 2, B_REFLECT101 = 3, B_WRAP = 4 };

thread_local std::string g_error;

struct OracleError {
    std::string msg;
};
void check(bool ok, const char* msg) {
    if (!ok) throw OracleError{std::string("ipf: ") + msg};
}

// ---------------------------------------------------------------- threads
// Row-range split with a minimum amount of work per thread
// (include/ipf/parallel.hpp:55-95). Results never depend on the split.
std::atomic<int> g_threads{0};

int hw_threads() {
    const int n = g_threads.load();
    if (n > 0) return n;
    const unsigned hw = std::thread::hardware_concurrency();
    return hw ? static_cast<int>(hw) : 1;
}

template <class Body>
void rows_parallel(int n, double cost_per_row, Body body) {
    if (n <= 0) return;
    int t = std::min(hw_threads(), n);
    t = std::min(t, std::max(1, static_cast<int>(n * cost_per_row / 32768.0)));
    if (t <= 1) {
        body(0, n);
        return;
    }
    std::vector<std::thread> pool;
    std::exception_ptr first;
    std::mutex mu;
    auto run = [&](int a, int b) {
        try {
            body(a, b);
        } catch (...) {
            std::lock_guard<std::mutex> g(mu);
            if (!first) first = std::current_exception();
        }
    };
    for (int k = 1; k < t; ++k) {
        const int a = static_cast<int>(static_cast<long long>(n) * k / t);
        const int b = static_cast<int>(static_cast<long long>(n) * (k + 1) / t);
        pool.emplace_back(run, a, b);
    }
    run(0, static_cast<int>(static_cast<long long>(n) / t));
    for (auto& th : pool) th.join();
    if (first) std::rethrow_exception(first);
}

// ---------------------------------------------------------------- core
// core.hpp:187-214. Far-out coordinates fold by the mode's period.
int border_index(int p, int len, int mode) {
    if (p >= 0 && p < len) return p;
    if (len <= 0) return -1;
    if (mode == B_CONSTANT) return -1;
    if (mode == B_REPLICATE) return p < 0 ? 0 : len - 1;
    if (mode == B_REFLECT) {
        const long long per = 2LL * len;
        long long q = ((p % per) + per) % per;
        return static_cast<int>(q < len ? q : per - 1 - q);
    }
    if (mode == B_REFLECT101) {
        if (len == 1) return 0;
        const long long per = 2LL * len - 2;
        long long q = ((p % per) + per) % per;
        return static_cast<int>(q < len ? q : per - q);
    }
    if (mode == B_WRAP) {
        const int q = p % len;
        return q < 0 ? q + len : q;
    }
    return -1;
}

// filter.hpp:27-32: table[p] = border_index(p - before) for the padded range.
std::vector<int> border_table(int len, int before, int after, int mode) {
    std::vector<int> t(static_cast<size_t>(len + before + after));
    for (size_t p = 0; p < t.size(); ++p) t[p] = border_index(static_cast<int>(p) - before, len, mode);
    return t;
}

// core.hpp:67-103, specialised to a double source (every filter accumulates in
// double): integers round half away from zero and clamp, NaN maps to 0; float
// targets use the plain conversion.
template <class To>
To sat(double v) {
    if constexpr (std::is_integral<To>::value) {
        if (v != v) return To(0);
        const double lo = static_cast<double>(std::numeric_limits<To>::lowest());
        const double hi = static_cast<double>(std::numeric_limits<To>::max());
        const double d = v < 0 ? v - 0.5 : v + 0.5;
        if (d <= lo) return std::numeric_limits<To>::lowest();
        if (d >= hi) return std::numeric_limits<To>::max();
        return static_cast<To>(d);
    } else {
        return static_cast<To>(v);
    }
}

// core.hpp:54-65 (nominal maximum; int32 images use 255).
template <class T>
double nominal_max() {
    if constexpr (std::is_floating_point<T>::value) return 1.0;
    else if constexpr (std::is_signed<T>::value && sizeof(T) >= 4) return 255.0;
    else return static_cast<double>(std::numeric_limits<T>::max());
}

// ---------------------------------------------------------------- separable
// filter.hpp:34-90. Horizontal pass into a double buffer of H+kh-1 padded rows,
// taps accumulated from 0.0 in index order; vertical pass starts at delta and
// accumulates rows in index order; one rounding at the end.
template <class Ti, class To>
void sep_filter(const Ti* src, int W, int H, int cn, To* dst, const double* kx, int kw, const double* ky, int kh,
                int ax, int ay, int border, double delta, double bval) {
    if (static_cast<long long>(W) * H == 0) return;
    const std::vector<int> xt = border_table(W, ax, kw - 1 - ax, border);
    const std::vector<int> yt = border_table(H, ay, kh - 1 - ay, border);
    const int PH = H + kh - 1;
    const size_t row = static_cast<size_t>(W) * cn;
    std::vector<double> mid(static_cast<size_t>(PH) * row);
    rows_parallel(PH, static_cast<double>(row) * kw, [&](int r0, int r1) {
        std::vector<double> line(static_cast<size_t>(W + kw - 1) * cn);
        for (int py = r0; py < r1; ++py) {
            const int sy = yt[py];
            for (int px = 0; px < W + kw - 1; ++px) {
                const int sx = xt[px];
                for (int c = 0; c < cn; ++c)
                    line[static_cast<size_t>(px) * cn + c] =
                        (sy < 0 || sx < 0) ? bval : static_cast<double>(src[(static_cast<size_t>(sy) * W + sx) * cn + c]);
            }
            double* out = &mid[static_cast<size_t>(py) * row];
            for (size_t q = 0; q < row; ++q) out[q] = 0.0;
            for (int i = 0; i < kw; ++i) {
                const double k = kx[i];
                const double* in = &line[static_cast<size_t>(i) * cn];
                for (size_t q = 0; q < row; ++q) out[q] = out[q] + k * in[q];
            }
        }
    });
    rows_parallel(H, static_cast<double>(row) * kh, [&](int y0, int y1) {
        std::vector<double> acc(row);
        for (int y = y0; y < y1; ++y) {
            for (size_t q = 0; q < row; ++q) acc[q] = delta;
            for (int j = 0; j < kh; ++j) {
                const double k = ky[j];
                const double* in = &mid[static_cast<size_t>(y + j) * row];
                for (size_t q = 0; q < row; ++q) acc[q] = acc[q] + k * in[q];
            }
            To* o = dst + static_cast<size_t>(y) * row;
            for (size_t q = 0; q < row; ++q) o[q] = sat<To>(acc[q]);
        }
    });
}

// ---------------------------------------------------------------- box
// filter.hpp:205-253. Horizontal running sums (window seeded by a left-to-right
// sum, then += entering - leaving), vertical running sums per column, then the
// sum times 1/(kw*kh) (or 1) rounded once.
template <class Ti, class To>
void box_filter(const Ti* src, int W, int H, int cn, To* dst, int kw, int kh, bool normalize, int border,
                double bval) {
    if (static_cast<long long>(W) * H == 0) return;
    const int ax = kw / 2, ay = kh / 2;
    const std::vector<int> xt = border_table(W, ax, kw - 1 - ax, border);
    const std::vector<int> yt = border_table(H, ay, kh - 1 - ay, border);
    const int PH = H + kh - 1;
    const size_t row = static_cast<size_t>(W) * cn;
    std::vector<double> hs(static_cast<size_t>(PH) * row);
    rows_parallel(PH, static_cast<double>(row) * 4, [&](int r0, int r1) {
        for (int py = r0; py < r1; ++py) {
            const int sy = yt[py];
            auto at = [&](int px, int c) -> double {
                const int sx = xt[px];
                if (sy < 0 || sx < 0) return bval;
                return static_cast<double>(src[(static_cast<size_t>(sy) * W + sx) * cn + c]);
            };
            double* out = &hs[static_cast<size_t>(py) * row];
            for (int c = 0; c < cn; ++c) {
                double s = 0.0;
                for (int i = 0; i < kw; ++i) s = s + at(i, c);
                out[c] = s;
                for (int x = 1; x < W; ++x) {
                    const double step = at(x + kw - 1, c) - at(x - 1, c);
                    s = s + step;
                    out[static_cast<size_t>(x) * cn + c] = s;
                }
            }
        }
    });
    const double scale = normalize ? 1.0 / (static_cast<double>(kw) * kh) : 1.0;
    rows_parallel(static_cast<int>(row), static_cast<double>(H) * 3, [&](int q0, int q1) {
        std::vector<double> col(static_cast<size_t>(q1 - q0), 0.0);
        for (int j = 0; j < kh; ++j)
            for (int q = q0; q < q1; ++q) col[q - q0] = col[q - q0] + hs[static_cast<size_t>(j) * row + q];
        for (int y = 0; y < H; ++y) {
            To* o = dst + static_cast<size_t>(y) * row;
            for (int q = q0; q < q1; ++q) o[q] = sat<To>(col[q - q0] * scale);
            if (y + 1 < H)
                for (int q = q0; q < q1; ++q) {
                    const double step = hs[static_cast<size_t>(y + kh) * row + q] - hs[static_cast<size_t>(y) * row + q];
                    col[q - q0] = col[q - q0] + step;
                }
        }
    });
}

// ---------------------------------------------------------------- filter2D
// filter.hpp:157-194. Only taps whose float weight compares unequal to zero take
// part, in row-major kernel order; s starts at delta.
template <class Ti, class To>
void filter2d(const Ti* src, int W, int H, int cn, To* dst, const float* K, int kw, int kh, int ax, int ay, int border,
              double delta, double bval) {
    if (static_cast<long long>(W) * H == 0) return;
    std::vector<int> tdx, tdy;
    std::vector<double> tw;
    for (int j = 0; j < kh; ++j)
        for (int i = 0; i < kw; ++i) {
            const float w = K[static_cast<size_t>(j) * kw + i];
            if (w != 0.f) {
                tdx.push_back(i);
                tdy.push_back(j);
                tw.push_back(static_cast<double>(w));
            }
        }
    const std::vector<int> xt = border_table(W, ax, kw - 1 - ax, border);
    const std::vector<int> yt = border_table(H, ay, kh - 1 - ay, border);
    const size_t nt = tw.size();
    rows_parallel(H, static_cast<double>(W) * cn * (nt + 1), [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < cn; ++c) {
                    double s = delta;
                    for (size_t t = 0; t < nt; ++t) {
                        const int sx = xt[x + tdx[t]], sy = yt[y + tdy[t]];
                        const double v = (sx < 0 || sy < 0)
                                             ? bval
                                             : static_cast<double>(src[(static_cast<size_t>(sy) * W + sx) * cn + c]);
                        s = s + tw[t] * v;
                    }
                    dst[(static_cast<size_t>(y) * W + x) * cn + c] = sat<To>(s);
                }
    });
}

// ---------------------------------------------------------------- median
// filter.hpp:300-332. The exact order statistic of rank k*k/2 of each window.
// uint8 uses a sliding 256-bin histogram along x (the order statistic does not
// depend on the method); other types select with nth_element.
template <class T>
void median(const T* src, int W, int H, int cn, T* dst, int k, int border) {
    if (static_cast<long long>(W) * H == 0) return;
    const size_t total = static_cast<size_t>(W) * H * cn;
    if (k == 1) {
        std::memcpy(dst, src, total * sizeof(T));
        return;
    }
    const int r = k / 2;
    const std::vector<int> xt = border_table(W, r, r, border);
    const std::vector<int> yt = border_table(H, r, r, border);
    const int rank = k * k / 2;
    auto S = [&](int px, int py, int c) { return src[(static_cast<size_t>(yt[py]) * W + xt[px]) * cn + c]; };
    if constexpr (std::is_same<T, uint8_t>::value) {
        if (k >= 5) {
            rows_parallel(H, static_cast<double>(W) * cn * k * 2, [&](int y0, int y1) {
                for (int y = y0; y < y1; ++y)
                    for (int c = 0; c < cn; ++c) {
                        int hist[256] = {0};
                        for (int j = 0; j < k; ++j)
                            for (int i = 0; i < k; ++i) ++hist[S(i, y + j, c)];
                        int m = 0, below = 0;  // m: current median, below: #samples < m
                        while (below + hist[m] <= rank) below += hist[m++];
                        dst[(static_cast<size_t>(y) * W) * cn + c] = static_cast<uint8_t>(m);
                        for (int x = 1; x < W; ++x) {
                            for (int j = 0; j < k; ++j) {
                                const int out = S(x - 1, y + j, c), in = S(x + k - 1, y + j, c);
                                --hist[out];
                                if (out < m) --below;
                                ++hist[in];
                                if (in < m) ++below;
                            }
                            while (below > rank) below -= hist[--m];
                            while (below + hist[m] <= rank) below += hist[m++];
                            dst[(static_cast<size_t>(y) * W + x) * cn + c] = static_cast<uint8_t>(m);
                        }
                    }
            });
            return;
        }
    }
    rows_parallel(H, static_cast<double>(W) * cn * k * k * 3, [&](int y0, int y1) {
        std::vector<T> win(static_cast<size_t>(k) * k);
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x)
                for (int c = 0; c < cn; ++c) {
                    size_t n = 0;
                    for (int j = 0; j < k; ++j)
                        for (int i = 0; i < k; ++i) win[n++] = S(x + i, y + j, c);
                    std::nth_element(win.begin(), win.begin() + rank, win.end());
                    dst[(static_cast<size_t>(y) * W + x) * cn + c] = win[rank];
                }
    });
}

// ---------------------------------------------------------------- bilateral
// filter.hpp:334-387. Circular taps (i^2+j^2 <= r^2, j outer, i inner) with
// exp space weights; range weight exp(dist^2 * cc) of the L1 colour distance,
// looked up in a table for integer images; Constant border skips the tap;
// out = acc / wsum rounded once.
//
// Divergence: for int16/int32 images whose L1 distance exceeds the table
// (values outside the nominal range) the reference reads past its table; this
// restatement evaluates exp directly for such distances instead.
template <class T>
void bilateral(const T* src, int W, int H, int cn, T* dst, int d, double sc, double ss, int border) {
    const int r = d <= 0 ? static_cast<int>(std::lround(ss * 1.5)) : d / 2;
    if (static_cast<long long>(W) * H == 0) return;
    std::vector<int> tdx, tdy;
    std::vector<double> tw;
    for (int j = -r; j <= r; ++j)
        for (int i = -r; i <= r; ++i)
            if (i * i + j * j <= r * r) {
                tdx.push_back(i);
                tdy.push_back(j);
                tw.push_back(std::exp(-(i * i + j * j) / (2 * ss * ss)));
            }
    const std::vector<int> xt = border_table(W, r, r, border);
    const std::vector<int> yt = border_table(H, r, r, border);
    const double cc = -1.0 / (2 * sc * sc);
    std::vector<double> lut;
    if constexpr (std::is_integral<T>::value) {
        const double maxd = nominal_max<T>() * cn;
        if (maxd <= (1 << 20)) {
            lut.resize(static_cast<size_t>(maxd) + 1);
            for (size_t k = 0; k < lut.size(); ++k) lut[k] = std::exp(static_cast<double>(k * k) * cc);
        }
    }
    const size_t nt = tw.size();
    rows_parallel(H, static_cast<double>(W) * nt * (cn + 2), [&](int y0, int y1) {
        std::vector<double> acc(static_cast<size_t>(cn));
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x) {
                const T* p0 = src + (static_cast<size_t>(y) * W + x) * cn;
                std::fill(acc.begin(), acc.end(), 0.0);
                double wsum = 0.0;
                for (size_t t = 0; t < nt; ++t) {
                    const int sx = xt[x + r + tdx[t]], sy = yt[y + r + tdy[t]];
                    if (sx < 0 || sy < 0) continue;
                    const T* p = src + (static_cast<size_t>(sy) * W + sx) * cn;
                    double dist = 0.0;
                    for (int c = 0; c < cn; ++c)
                        dist = dist + std::fabs(static_cast<double>(p[c]) - static_cast<double>(p0[c]));
                    double rw;
                    if (!lut.empty() && dist < static_cast<double>(lut.size()))
                        rw = lut[static_cast<size_t>(dist)];
                    else
                        rw = std::exp(dist * dist * cc);
                    const double w = tw[t] * rw;
                    wsum = wsum + w;
                    for (int c = 0; c < cn; ++c) acc[c] = acc[c] + w * static_cast<double>(p[c]);
                }
                T* o = dst + (static_cast<size_t>(y) * W + x) * cn;
                for (int c = 0; c < cn; ++c) o[c] = sat<T>(acc[c] / wsum);
            }
    });
}

// ---------------------------------------------------------------- dispatch
template <class F>
int with_type(int dt, F f) {
    switch (dt) {
        case DT_U8: return f(static_cast<uint8_t*>(nullptr));
        case DT_U16: return f(static_cast<uint16_t*>(nullptr));
        case DT_I16: return f(static_cast<int16_t*>(nullptr));
        case DT_I32: return f(static_cast<int32_t*>(nullptr));
        case DT_F32: return f(static_cast<float*>(nullptr));
        case DT_F64: return f(static_cast<double*>(nullptr));
    }
    throw OracleError{"ipf: unsupported pixel type"};
}

template <class F>
int guarded(F f) {
    try {
        return f();
    } catch (const OracleError& e) {
        g_error = e.msg;
        return -1;
    } catch (const std::exception& e) {
        g_error = e.what();
        return -2;
    }
}

void check_image(int w, int h, int cn) {
    check(w >= 0 && h >= 0, "Image: negative dimensions");
    check(cn >= 1, "Image: channel count must be >= 1");
}

}  // namespace

extern "C" {

const char* ipfo_last_error(void) { return g_error.c_str(); }
void ipfo_set_num_threads(int n) { g_threads = n < 0 ? 0 : n; }
int ipfo_get_num_threads(void) { return hw_threads(); }

int ipfo_border_index(int p, int len, int mode) { return border_index(p, len, mode); }

// Writes saturate_cast<dtype>(v) into *out (core.hpp:67-103, double source).
int ipfo_saturate(double v, int dtype, void* out) {
    return guarded([&] {
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            *static_cast<T*>(out) = sat<T>(v);
            return 0;
        });
    });
}

// gaussianKernel1D (filter.hpp:129-145). Returns the tap count, writes at most
// cap weights; negative on error.
int ipfo_gaussian_kernel_1d(double sigma, int ksize, double* out, int cap) {
    return guarded([&] {
        if (ksize <= 0) {
            check(sigma > 0, "gaussianKernel1D: need sigma > 0 or ksize > 0");
            ksize = 2 * std::max(1, static_cast<int>(3.0 * sigma + 0.5)) + 1;
        }
        check(ksize % 2 == 1, "gaussianKernel1D: ksize must be odd");
        if (sigma <= 0) sigma = 0.3 * ((ksize - 1) * 0.5 - 1) + 0.8;
        std::vector<double> k(static_cast<size_t>(ksize));
        const int r = ksize / 2;
        double sum = 0;
        for (int i = 0; i < ksize; ++i) {
            k[i] = std::exp(-0.5 * (i - r) * (i - r) / (sigma * sigma));
            sum += k[i];
        }
        for (int i = 0; i < ksize; ++i) {
            k[i] /= sum;
            if (i < cap) out[i] = k[i];
        }
        return ksize;
    });
}

int ipfo_sep_filter(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const double* kx, int kw,
                    const double* ky, int kh, int ax, int ay, int border, double delta, double bval) {
    return guarded([&] {
        check(kw > 0 && kh > 0 && kx && ky, "separable filter: empty kernel");
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                sep_filter<Ti, To>(static_cast<const Ti*>(src), w, h, cn, static_cast<To*>(dst), kx, kw, ky, kh, ax,
                                   ay, border, delta, bval);
                return 0;
            });
        });
    });
}

int ipfo_box_filter(const void* src, int w, int h, int cn, int tin, void* dst, int tout, int kw, int kh, int normalize,
                    int border, double bval) {
    return guarded([&] {
        check(kw > 0 && kh > 0, "boxFilter: kernel size must be positive");
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                box_filter<Ti, To>(static_cast<const Ti*>(src), w, h, cn, static_cast<To*>(dst), kw, kh,
                                   normalize != 0, border, bval);
                return 0;
            });
        });
    });
}

int ipfo_filter2d(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const float* K, int kw, int kh,
                  int ax, int ay, int border, double delta, double bval) {
    return guarded([&] {
        check(K != nullptr && kw > 0 && kh > 0, "filter2D: kernel must be a non-empty single-channel image");
        if (ax < 0) ax = kw / 2;
        if (ay < 0) ay = kh / 2;
        check(ax < kw && ay < kh, "filter2D: anchor outside kernel");
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                filter2d<Ti, To>(static_cast<const Ti*>(src), w, h, cn, static_cast<To*>(dst), K, kw, kh, ax, ay,
                                 border, delta, bval);
                return 0;
            });
        });
    });
}

int ipfo_median(const void* src, int w, int h, int cn, int dtype, void* dst, int ksize, int border) {
    return guarded([&] {
        check(ksize >= 1 && ksize % 2 == 1, "medianBlur: ksize must be odd and positive");
        check(border != B_CONSTANT, "medianBlur: Constant border is not supported");
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            median<T>(static_cast<const T*>(src), w, h, cn, static_cast<T*>(dst), ksize, border);
            return 0;
        });
    });
}

int ipfo_bilateral(const void* src, int w, int h, int cn, int dtype, void* dst, int d, double sigma_color,
                   double sigma_space, int border) {
    return guarded([&] {
        check(sigma_color > 0 && sigma_space > 0, "bilateralFilter: sigmas must be positive");
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            bilateral<T>(static_cast<const T*>(src), w, h, cn, static_cast<T*>(dst), d, sigma_color, sigma_space,
                         border);
            return 0;
        });
    });
}

// ---------------------------------------------------------------- derivatives and sharpening
}  // extern "C"

namespace {

constexpr double kPiO = 3.14159265358979323846;

// derivKernel (filter.hpp:93-114): binomial smoothing (1, 1) convolved
// ksize - 1 - order times, then the difference (-1, 1) order times; ksize 1
// gives the unsmoothed central differences.
std::vector<double> deriv_kernel(int order, int ksize) {
    if (ksize == 1) {
        if (order == 0) return {1.0};
        if (order == 1) return {-1.0, 0.0, 1.0};
        if (order == 2) return {1.0, -2.0, 1.0};
        throw OracleError{"ipf: derivKernel: order > 2 requires ksize >= 3"};
    }
    check(ksize % 2 == 1 && ksize >= 3 && ksize <= 31, "derivKernel: ksize must be odd in [3, 31]");
    check(order >= 0 && order < ksize, "derivKernel: invalid derivative order");
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

template <class Ti>
void sep_float(const Ti* src, int w, int h, int cn, float* dst, const std::vector<double>& kx,
               const std::vector<double>& ky, int border) {
    const int kw = static_cast<int>(kx.size()), kh = static_cast<int>(ky.size());
    sep_filter<Ti, float>(src, w, h, cn, dst, kx.data(), kw, ky.data(), kh, kw / 2, kh / 2, border, 0.0, 0.0);
}

// sobel (filter.hpp:392-406)
template <class Ti>
void sobel(const Ti* src, int w, int h, int cn, float* dst, int dx, int dy, int ksize, double scale, int border) {
    check(dx >= 0 && dy >= 0 && dx + dy > 0, "sobel: need dx + dy > 0");
    std::vector<double> kx = deriv_kernel(dx, ksize), ky = deriv_kernel(dy, ksize);
    if (ksize == 1) {  // both kernels odd-length so the anchors stay centred
        if (kx.size() == 1) kx = {0.0, 1.0, 0.0};
        if (ky.size() == 1) ky = {0.0, 1.0, 0.0};
    }
    for (double& v : kx) v *= scale;
    sep_float(src, w, h, cn, dst, kx, ky, border);
}

// filter2D with a 3 x 3 float kernel and the reference's defaults (Reflect101,
// centre anchor, no delta)
template <class Ti, class To>
void filter3x3(const Ti* src, int w, int h, int cn, To* dst, const float (&k)[9], int border) {
    filter2d<Ti, To>(src, w, h, cn, dst, k, 3, 3, 1, 1, border, 0.0, 0.0);
}

}  // namespace

extern "C" {

int ipfo_sobel(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, int ksize, double scale,
               int border) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            sobel<Ti>(static_cast<const Ti*>(src), w, h, cn, dst, dx, dy, ksize, scale, border);
            return 0;
        });
    });
}

// scharr, prewitt (filter.hpp:408-421): a central difference along the
// derivative axis, a smoothing triple across it.
int ipfo_scharr(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, double scale, int border) {
    return guarded([&] {
        check((dx == 1 && dy == 0) || (dx == 0 && dy == 1), "scharr: (dx, dy) must be (1,0) or (0,1)");
        check_image(w, h, cn);
        const std::vector<double> d = {-scale, 0.0, scale}, s = {3.0, 10.0, 3.0};
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            sep_float<Ti>(static_cast<const Ti*>(src), w, h, cn, dst, dx ? d : s, dx ? s : d, border);
            return 0;
        });
    });
}

int ipfo_prewitt(const void* src, int w, int h, int cn, int tin, float* dst, int dx, int dy, int border) {
    return guarded([&] {
        check((dx == 1 && dy == 0) || (dx == 0 && dy == 1), "prewitt: (dx, dy) must be (1,0) or (0,1)");
        check_image(w, h, cn);
        const std::vector<double> d = {-1.0, 0.0, 1.0}, s = {1.0, 1.0, 1.0};
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            sep_float<Ti>(static_cast<const Ti*>(src), w, h, cn, dst, dx ? d : s, dx ? s : d, border);
            return 0;
        });
    });
}

// laplacian (filter.hpp:423-436): ksize 1 is the 4-neighbour 3 x 3 kernel
// (weights scaled in double, rounded to float); otherwise the float sum of the
// second-order Sobel derivatives in x and in y.
int ipfo_laplacian(const void* src, int w, int h, int cn, int tin, float* dst, int ksize, double scale, int border) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            const Ti* s = static_cast<const Ti*>(src);
            if (ksize == 1) {
                const float base[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};
                float k[9];
                for (int i = 0; i < 9; ++i) k[i] = static_cast<float>(base[i] * scale);
                filter3x3<Ti, float>(s, w, h, cn, dst, k, border);
                return 0;
            }
            const size_t n = static_cast<size_t>(w) * h * cn;
            std::vector<float> b(n);
            sobel<Ti>(s, w, h, cn, dst, 2, 0, ksize, scale, border);
            sobel<Ti>(s, w, h, cn, b.data(), 0, 2, ksize, scale, border);
            for (size_t i = 0; i < n; ++i) dst[i] += b[i];
            return 0;
        });
    });
}

// magnitude, phase (filter.hpp:438-458)
int ipfo_magnitude(const float* gx, const float* gy, int w, int h, int cn, float* dst) {
    return guarded([&] {
        check_image(w, h, cn);
        const size_t n = static_cast<size_t>(w) * h * cn;
        for (size_t i = 0; i < n; ++i) dst[i] = std::hypot(gx[i], gy[i]);
        return 0;
    });
}

int ipfo_phase(const float* gx, const float* gy, int w, int h, int cn, float* dst, int degrees) {
    return guarded([&] {
        check_image(w, h, cn);
        const size_t n = static_cast<size_t>(w) * h * cn;
        const double full = degrees ? 360.0 : 2 * kPiO;
        for (size_t i = 0; i < n; ++i) {
            double a = std::atan2(gy[i], gx[i]);
            if (degrees) a *= 180.0 / kPiO;
            if (a < 0) a += full;
            dst[i] = static_cast<float>(a);
        }
        return 0;
    });
}

// unsharpMask (filter.hpp:472-483): FP64 Gaussian blur (Reflect101), then
// v + amount * (v - blurred) where |v - blurred| > threshold.
int ipfo_unsharp(const void* src, int w, int h, int cn, int dtype, void* dst, double sigma, double amount,
                 double threshold) {
    return guarded([&] {
        check_image(w, h, cn);
        const int ks = ipfo_gaussian_kernel_1d(sigma, 0, nullptr, 0);
        check(ks > 0, "gaussianKernel1D: need sigma > 0 or ksize > 0");
        std::vector<double> k(static_cast<size_t>(ks));
        ipfo_gaussian_kernel_1d(sigma, 0, k.data(), ks);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const T* s = static_cast<const T*>(src);
            T* d = static_cast<T*>(dst);
            const size_t n = static_cast<size_t>(w) * h * cn;
            std::vector<double> blurred(n);
            sep_filter<T, double>(s, w, h, cn, blurred.data(), k.data(), ks, k.data(), ks, ks / 2, ks / 2, B_REFLECT101,
                                  0.0, 0.0);
            for (size_t i = 0; i < n; ++i) {
                const double v = s[i], diff = v - blurred[i];
                d[i] = sat<T>(std::fabs(diff) > threshold ? v + amount * diff : v);
            }
            return 0;
        });
    });
}

// ssim (metrics.hpp:46-78): x, y, x*x, y*y, x*y as double images, each
// blurred by the 11-tap sigma 1.5 Gaussian (Reflect101) through sep_filter,
// the map as float of the reference's expression, and the mean of the map over
// the valid region (r = 5) summed in row-major order. L = PixelTraits<T>::max()
// (core.hpp:57-64).
int ipfo_ssim(const void* pa, const void* pb, int w, int h, int cn, int dtype, float* map, double* result) {
    return guarded([&] {
        check_image(w, h, cn);
        check(static_cast<long long>(w) * h > 0, "ssim: images must be non-empty and have the same shape");
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const T* a = static_cast<const T*>(pa);
            const T* b = static_cast<const T*>(pb);
            double L = 1.0;
            if constexpr (std::is_integral<T>::value)
                L = (std::is_signed<T>::value && sizeof(T) >= 4) ? 255.0 : static_cast<double>(std::numeric_limits<T>::max());
            const double C1 = (0.01 * L) * (0.01 * L), C2 = (0.03 * L) * (0.03 * L);
            const size_t n = static_cast<size_t>(w) * h * cn;
            std::vector<double> x(n), y(n), xx(n), yy(n), xy(n);
            for (size_t i = 0; i < n; ++i) {
                x[i] = static_cast<double>(a[i]);
                y[i] = static_cast<double>(b[i]);
                xx[i] = x[i] * x[i];
                yy[i] = y[i] * y[i];
                xy[i] = x[i] * y[i];
            }
            double k[11];
            ipfo_gaussian_kernel_1d(1.5, 11, k, 11);
            std::vector<double> m[5];
            const std::vector<double>* in[5] = {&x, &y, &xx, &yy, &xy};
            for (int t = 0; t < 5; ++t) {
                m[t].resize(n);
                sep_filter<double, double>(in[t]->data(), w, h, cn, m[t].data(), k, 11, k, 11, 5, 5, B_REFLECT101,
                                           0.0, 0.0);
            }
            std::vector<float> mp(n);
            for (size_t i = 0; i < n; ++i) {
                const double ux = m[0][i], uy = m[1][i];
                const double vx = m[2][i] - ux * ux, vy = m[3][i] - uy * uy, cxy = m[4][i] - ux * uy;
                mp[i] = static_cast<float>(((2 * ux * uy + C1) * (2 * cxy + C2)) /
                                           ((ux * ux + uy * uy + C1) * (vx + vy + C2)));
            }
            const int r = 5;
            const bool valid = w > 2 * r && h > 2 * r;
            const int x0 = valid ? r : 0, x1 = valid ? w - r : w, y0 = valid ? r : 0, y1 = valid ? h - r : h;
            double s = 0;
            for (int yy0 = y0; yy0 < y1; ++yy0)
                for (int xx0 = x0; xx0 < x1; ++xx0)
                    for (int c = 0; c < cn; ++c) s += mp[(static_cast<size_t>(yy0) * w + xx0) * cn + c];
            if (map) std::copy(mp.begin(), mp.end(), map);
            *result = s / (static_cast<double>(x1 - x0) * (y1 - y0) * cn);
            return 0;
        });
    });
}

// pyrDown (geometry.hpp:501-512): the separable binomial filter {1, 4, 6, 4, 1}
// / 16 into a double image (Reflect101, delta 0), then saturate_cast of the
// samples at even (x, y) into a (w + 1) / 2 x (h + 1) / 2 image.
int ipfo_pyr_down(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        check_image(w, h, cn);
        const double k[5] = {1 / 16.0, 4 / 16.0, 6 / 16.0, 4 / 16.0, 1 / 16.0};
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            std::vector<double> blurred(static_cast<size_t>(w) * h * cn);
            sep_filter<T, double>(static_cast<const T*>(src), w, h, cn, blurred.data(), k, 5, k, 5, 2, 2,
                                  B_REFLECT101, 0.0, 0.0);
            const int W2 = (w + 1) / 2, H2 = (h + 1) / 2;
            T* out = static_cast<T*>(dst);
            for (int y = 0; y < H2; ++y)
                for (int x = 0; x < W2; ++x)
                    for (int c = 0; c < cn; ++c)
                        out[(static_cast<size_t>(y) * W2 + x) * cn + c] =
                            sat<T>(blurred[(static_cast<size_t>(2 * y) * w + 2 * x) * cn + c]);
            return 0;
        });
    });
}

// pyrUp (geometry.hpp:514-524): a 2w x 2h double image of zeros with sample
// (x, y) at (2x, 2y), filtered with {1, 4, 6, 4, 1} / 8 (Reflect101, delta 0)
// and rounded once into the source type.
int ipfo_pyr_up(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        check_image(w, h, cn);
        const double k[5] = {1 / 8.0, 4 / 8.0, 6 / 8.0, 4 / 8.0, 1 / 8.0};
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const int W2 = 2 * w, H2 = 2 * h;
            std::vector<double> up(static_cast<size_t>(W2) * H2 * cn, 0.0);
            const T* in = static_cast<const T*>(src);
            for (int y = 0; y < h; ++y)
                for (int x = 0; x < w; ++x)
                    for (int c = 0; c < cn; ++c)
                        up[(static_cast<size_t>(2 * y) * W2 + 2 * x) * cn + c] =
                            static_cast<double>(in[(static_cast<size_t>(y) * w + x) * cn + c]);
            sep_filter<double, T>(up.data(), W2, H2, cn, static_cast<T*>(dst), k, 5, k, 5, 2, 2, B_REFLECT101, 0.0,
                                  0.0);
            return 0;
        });
    });
}

// sharpen, emboss (filter.hpp:485-496): 3 x 3 filter2D to the input type.
int ipfo_sharpen(const void* src, int w, int h, int cn, int dtype, void* dst, double strength) {
    return guarded([&] {
        check_image(w, h, cn);
        const float s = static_cast<float>(strength);
        const float k[9] = {0, -s, 0, -s, 1 + 4 * s, -s, 0, -s, 0};
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            filter3x3<T, T>(static_cast<const T*>(src), w, h, cn, static_cast<T*>(dst), k, B_REFLECT101);
            return 0;
        });
    });
}

int ipfo_emboss(const void* src, int w, int h, int cn, int dtype, void* dst) {
    return guarded([&] {
        check_image(w, h, cn);
        const float k[9] = {-2, -1, 0, -1, 1, 1, 0, 1, 2};
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            filter3x3<T, T>(static_cast<const T*>(src), w, h, cn, static_cast<T*>(dst), k, B_REFLECT101);
            return 0;
        });
    });
}

// adaptiveThreshold (filter.hpp:545-565): local mean (box, or Gaussian of
// sigma 0 = from the block size) with Replicate borders, in the input type;
// maxval where (src - mean > -C) matches Binary, else 0. method 0 Mean, 1
// Gaussian; type 0 Binary, 1 BinaryInv.
int ipfo_adaptive_threshold(const void* src, int w, int h, int dtype, void* dst, double maxval, int method, int type,
                            int block, double c) {
    return guarded([&] {
        check_image(w, h, 1);
        check(block >= 3 && block % 2 == 1, "adaptiveThreshold: blockSize must be odd and >= 3");
        check(type == 0 || type == 1, "adaptiveThreshold: type must be Binary or BinaryInv");
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const T* s = static_cast<const T*>(src);
            T* d = static_cast<T*>(dst);
            const size_t n = static_cast<size_t>(w) * h;
            std::vector<T> mean(n);
            if (method == 0) {
                box_filter<T, T>(s, w, h, 1, mean.data(), block, block, true, B_REPLICATE, 0.0);
            } else {
                std::vector<double> k(static_cast<size_t>(block));
                ipfo_gaussian_kernel_1d(0.0, block, k.data(), block);
                sep_filter<T, T>(s, w, h, 1, mean.data(), k.data(), block, k.data(), block, block / 2, block / 2,
                                 B_REPLICATE, 0.0, 0.0);
            }
            for (size_t i = 0; i < n; ++i) {
                const bool above = static_cast<double>(s[i]) - static_cast<double>(mean[i]) > -c;
                d[i] = sat<T>((above == (type == 0)) ? maxval : 0.0);
            }
            return 0;
        });
    });
}

// motionBlurKernel (filter.hpp:498-516): 4 * length samples along the line
// through the centre at the angle, marked 1, then normalised by their float
// sum. Returns the kernel side; writes at most cap weights (row-major).
int ipfo_motion_kernel(int length, double angle, float* out, int cap) {
    return guarded([&] {
        check(length >= 1, "motionBlurKernel: length must be >= 1");
        const int size = length % 2 ? length : length + 1;
        std::vector<float> k(static_cast<size_t>(size) * size, 0.f);
        const double a = angle * kPiO / 180.0, c = (size - 1) / 2.0;
        const int samples = 4 * length;
        for (int i = 0; i < samples; ++i) {
            const double t = (length - 1) * (static_cast<double>(i) / (samples - 1) - 0.5);
            const int x = static_cast<int>(std::lround(c + t * std::cos(a)));
            const int y = static_cast<int>(std::lround(c - t * std::sin(a)));
            k[static_cast<size_t>(std::clamp(y, 0, size - 1)) * size + std::clamp(x, 0, size - 1)] = 1.f;
        }
        float s = 0;
        for (float v : k) s += v;
        for (float& v : k) v /= s;
        for (int i = 0; i < size * size && i < cap; ++i) out[i] = k[i];
        return size;
    });
}

}  // extern "C"

// ---------------------------------------------------------------- features
namespace {

// canny (features.hpp:26-90): Sobel gradients (Replicate border), per pixel
// the channel with the largest float magnitude (L1, or L2 computed in float
// and square-rooted), non-maximum suppression along the quantised gradient
// direction (neighbours outside the image count as 0), then hysteresis: the
// weak candidates 8-connected to a strong one (directly or through other
// candidates) become edges. Output 255 / 0, one channel.
template <class T>
void canny(const T* src, int W, int H, int cn, uint8_t* out, double low, double high, int ksize, bool l2) {
    check(low <= high, "canny: lowThreshold must be <= highThreshold");
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<float> gx(n * cn), gy(n * cn);
    sobel<T>(src, W, H, cn, gx.data(), 1, 0, ksize, 1.0, B_REPLICATE);
    sobel<T>(src, W, H, cn, gy.data(), 0, 1, ksize, 1.0, B_REPLICATE);
    std::vector<float> mag(n), sx(n, 0.f), sy(n, 0.f);
    rows_parallel(H, static_cast<double>(W) * cn * 4, [&](int y0, int y1) {
        for (size_t p = static_cast<size_t>(y0) * W; p < static_cast<size_t>(y1) * W; ++p) {
            float best = -1.f;
            for (int c = 0; c < cn; ++c) {
                const float a = gx[p * cn + c], b = gy[p * cn + c];
                const float m = l2 ? std::sqrt(a * a + b * b) : std::fabs(a) + std::fabs(b);
                if (m > best) {
                    best = m;
                    sx[p] = a;
                    sy[p] = b;
                }
            }
            mag[p] = best;
        }
    });
    std::vector<uint8_t> st(n, 0);  // 0 none, 1 weak candidate, 2 strong
    const float tan22 = 0.41421356f;
    auto at = [&](int x, int y) { return (x < 0 || y < 0 || x >= W || y >= H) ? 0.f : mag[static_cast<size_t>(y) * W + x]; };
    rows_parallel(H, static_cast<double>(W) * 6, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y)
            for (int x = 0; x < W; ++x) {
                const size_t p = static_cast<size_t>(y) * W + x;
                const float m = mag[p];
                if (m <= low) continue;
                const float ax = std::fabs(sx[p]), ay = std::fabs(sy[p]);
                bool peak;
                if (ay <= ax * tan22) {
                    peak = m > at(x - 1, y) && m >= at(x + 1, y);
                } else if (ay * tan22 >= ax) {
                    peak = m > at(x, y - 1) && m >= at(x, y + 1);
                } else {
                    const int s = (sx[p] < 0) != (sy[p] < 0) ? -1 : 1;
                    peak = m > at(x - s, y - 1) && m > at(x + s, y + 1);
                }
                if (peak) st[p] = m > high ? 2 : 1;
            }
    });
    // hysteresis by an explicit stack (the result does not depend on the order)
    std::vector<size_t> stack;
    for (size_t p = 0; p < n; ++p)
        if (st[p] == 2) stack.push_back(p);
    while (!stack.empty()) {
        const size_t p = stack.back();
        stack.pop_back();
        const int x = static_cast<int>(p % W), y = static_cast<int>(p / W);
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                const int qx = x + i, qy = y + j;
                if (qx < 0 || qy < 0 || qx >= W || qy >= H) continue;
                const size_t q = static_cast<size_t>(qy) * W + qx;
                if (st[q] != 1) continue;
                st[q] = 2;
                stack.push_back(q);
            }
    }
    for (size_t p = 0; p < n; ++p) out[p] = st[p] == 2 ? 255 : 0;
}

// toGray (color.hpp:78-81): channel 0 of a 1- or 2-channel image; Luma 601
// of the first three channels (0.299 r + 0.587 g + 0.114 b in FP64, then
// saturate_cast) for 3 or 4 channels (color.hpp:51-67).
template <class T>
std::vector<T> to_gray(const T* src, int W, int H, int cn) {
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<T> g(n);
    if (cn <= 2) {
        for (size_t p = 0; p < n; ++p) g[p] = src[p * cn];
        return g;
    }
    check(cn == 3 || cn == 4, "rgbToGray: expected a 3-channel (RGB) or 4-channel (RGBA) image");
    for (size_t p = 0; p < n; ++p) {
        const double r = src[p * cn], gr = src[p * cn + 1], b = src[p * cn + 2];
        g[p] = sat<T>(0.299 * r + 0.587 * gr + 0.114 * b);
    }
    return g;
}

// cannyAuto (features.hpp:92-101): gray, FP64 Gaussian blur to float
// (Reflect101), Sobel gradient magnitude (ksize 3, Reflect101, hypot), high
// threshold = the element of rank 7n/10 of the magnitudes (at least 1e-6),
// low = 0.4 high, then canny with ksize 3 and the L2 magnitude.
template <class T>
void canny_auto(const T* src, int W, int H, int cn, uint8_t* out, double sigma) {
    const std::vector<T> gray = to_gray<T>(src, W, H, cn);
    const int ks = ipfo_gaussian_kernel_1d(sigma, 0, nullptr, 0);
    check(ks > 0, "gaussianKernel1D: need sigma > 0 or ksize > 0");
    std::vector<double> k(static_cast<size_t>(ks));
    ipfo_gaussian_kernel_1d(sigma, 0, k.data(), ks);
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<float> smooth(n), a(n), b(n);
    sep_filter<T, float>(gray.data(), W, H, 1, smooth.data(), k.data(), ks, k.data(), ks, ks / 2, ks / 2,
                         B_REFLECT101, 0.0, 0.0);
    if (n == 0) return;
    sobel<float>(smooth.data(), W, H, 1, a.data(), 1, 0, 3, 1.0, B_REFLECT101);
    sobel<float>(smooth.data(), W, H, 1, b.data(), 0, 1, 3, 1.0, B_REFLECT101);
    std::vector<float> v(n);
    for (size_t p = 0; p < n; ++p) v[p] = std::hypot(a[p], b[p]);
    const size_t r = n * 7 / 10;
    std::nth_element(v.begin(), v.begin() + static_cast<std::ptrdiff_t>(r), v.end());
    const double high = std::max(1e-6, static_cast<double>(v[r]));
    canny<float>(smooth.data(), W, H, 1, out, 0.4 * high, high, 3, true);
}

// structure tensor (features.hpp:107-126): Sobel derivatives scaled by
// 1 / (2^(ksize - 1) * block * nominal max) (float images: no nominal
// max), float products, unnormalised box sums (Reflect101).
template <class T>
void structure_tensor(const T* src, int W, int H, int cn, int block, int ksize, std::vector<float>& a,
                      std::vector<float>& b, std::vector<float>& c) {
    check(cn == 1, "corner detection: expected a single-channel image");
    double scale = static_cast<double>(1 << ((ksize > 0 ? ksize : 3) - 1)) * block;
    if (!std::is_floating_point<T>::value) scale *= nominal_max<T>();
    scale = 1.0 / scale;
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<float> dx(n), dy(n), xx(n), xy(n), yy(n);
    sobel<T>(src, W, H, 1, dx.data(), 1, 0, ksize, scale, B_REFLECT101);
    sobel<T>(src, W, H, 1, dy.data(), 0, 1, ksize, scale, B_REFLECT101);
    for (size_t i = 0; i < n; ++i) {
        xx[i] = dx[i] * dx[i];
        xy[i] = dx[i] * dy[i];
        yy[i] = dy[i] * dy[i];
    }
    a.resize(n);
    b.resize(n);
    c.resize(n);
    box_filter<float, float>(xx.data(), W, H, 1, a.data(), block, block, false, B_REFLECT101, 0.0);
    box_filter<float, float>(xy.data(), W, H, 1, b.data(), block, block, false, B_REFLECT101, 0.0);
    box_filter<float, float>(yy.data(), W, H, 1, c.data(), block, block, false, B_REFLECT101, 0.0);
}

}  // namespace

extern "C" {

int ipfo_canny(const void* src, int w, int h, int cn, int dtype, uint8_t* dst, double low, double high, int ksize,
               int l2) {
    return guarded([&] {
        check(low <= high, "canny: lowThreshold must be <= highThreshold");
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            canny<T>(static_cast<const T*>(src), w, h, cn, dst, low, high, ksize, l2 != 0);
            return 0;
        });
    });
}

int ipfo_canny_auto(const void* src, int w, int h, int cn, int dtype, uint8_t* dst, double sigma) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            canny_auto<T>(static_cast<const T*>(src), w, h, cn, dst, sigma);
            return 0;
        });
    });
}

// cornerHarris (features.hpp:128-139): A C - B^2 - k (A + C)^2 in FP64, to float.
int ipfo_corner_harris(const void* src, int w, int h, int cn, int dtype, float* dst, int block, int ksize, double k) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            std::vector<float> a, b, c;
            structure_tensor<T>(static_cast<const T*>(src), w, h, cn, block, ksize, a, b, c);
            for (size_t i = 0; i < a.size(); ++i) {
                const double A = a[i], B = b[i], C = c[i];
                dst[i] = static_cast<float>(A * C - B * B - k * (A + C) * (A + C));
            }
            return 0;
        });
    });
}

// cornerMinEigenVal (features.hpp:141-153): (A + C) - sqrt((A - C)^2 + B^2)
// with A, C halved, in FP64, to float.
int ipfo_corner_min_eigen(const void* src, int w, int h, int cn, int dtype, float* dst, int block, int ksize) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            std::vector<float> a, b, c;
            structure_tensor<T>(static_cast<const T*>(src), w, h, cn, block, ksize, a, b, c);
            for (size_t i = 0; i < a.size(); ++i) {
                const double A = a[i] * 0.5, B = b[i], C = c[i] * 0.5;
                dst[i] = static_cast<float>((A + C) - std::sqrt((A - C) * (A - C) + B * B));
            }
            return 0;
        });
    });
}

}  // extern "C"

// ---------------------------------------------------------------- filter2DFFT
namespace {

using cplx = std::complex<double>;
constexpr double kPiFft = 3.14159265358979323846;  // core.hpp:45

size_t pow2_ceil(size_t n) {
    size_t p = 1;
    while (p < n) p <<= 1;
    return p;
}

// Radix-2 decimation in time on a power-of-two length (fft.hpp:29-50): the
// bit-reversal permutation, then stages of length 2, 4, ..., n whose
// twiddles exp(i sign 2 pi k / len) are evaluated directly per stage (angle
// formed as sign * 2 * pi * k / len, left to right).
void fft_radix2(cplx* a, size_t n, int sign) {
    if (n <= 1) return;
    size_t j = 0;
    for (size_t i = 1; i < n; ++i) {
        size_t bit = n >> 1;
        while (j & bit) {
            j ^= bit;
            bit >>= 1;
        }
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    std::vector<cplx> tw;
    for (size_t len = 2; len <= n; len <<= 1) {
        const size_t half = len / 2;
        tw.resize(half);
        for (size_t k = 0; k < half; ++k) {
            const double ang = sign * 2.0 * kPiFft * static_cast<double>(k) / static_cast<double>(len);
            tw[k] = cplx(std::cos(ang), std::sin(ang));
        }
        for (size_t base = 0; base < n; base += len)
            for (size_t k = 0; k < half; ++k) {
                const cplx u = a[base + k];
                const cplx v = a[base + k + half] * tw[k];
                a[base + k] = u + v;
                a[base + k + half] = u - v;
            }
    }
}

// fft2D (fft.hpp:79-92): every row, then every column (power-of-two sides).
void fft_2d(std::vector<cplx>& d, int W, int H, int sign) {
    rows_parallel(H, static_cast<double>(W) * 20, [&](int y0, int y1) {
        for (int y = y0; y < y1; ++y) fft_radix2(&d[static_cast<size_t>(y) * W], static_cast<size_t>(W), sign);
    });
    rows_parallel(W, static_cast<double>(H) * 20, [&](int x0, int x1) {
        std::vector<cplx> col(static_cast<size_t>(H));
        for (int x = x0; x < x1; ++x) {
            for (int y = 0; y < H; ++y) col[y] = d[static_cast<size_t>(y) * W + x];
            fft_radix2(col.data(), static_cast<size_t>(H), sign);
            for (int y = 0; y < H; ++y) d[static_cast<size_t>(y) * W + x] = col[y];
        }
    });
}

// filter2DFFT (fft.hpp:196-216) with correlateFFT (fft.hpp:167-193): the source
// as double, padded by copyMakeBorder (core.hpp:526-544: ay rows above, kh-1-ay
// below, ax columns left, kw-1-ax right, Constant cells = borderValue), placed
// per channel in the real part of a pow2 x pow2 zero buffer with the kernel in
// the imaginary part; one forward 2-D FFT; the two real spectra separated as
// (Z(k) + conj Z(-k)) / 2 and (Z(k) - conj Z(-k)) * (-i / 2); their product
// with the kernel's spectrum conjugated; the inverse 2-D FFT; the real part
// times 1 / (fw fh), plus delta, saturated.
template <class Ti, class To>
void filter2d_fft(const Ti* src, int W, int H, int cn, To* dst, const float* K, int kw, int kh, int border,
                  double delta, double bval) {
    const int ax = kw / 2, ay = kh / 2;
    const int PW = W + kw - 1, PH = H + kh - 1;
    std::vector<double> pad(static_cast<size_t>(PW) * PH * cn, bval);
    for (int y = 0; y < PH; ++y) {
        const int sy = border_index(y - ay, H, border);
        if (sy < 0) continue;
        for (int x = 0; x < PW; ++x) {
            const int sx = border_index(x - ax, W, border);
            if (sx < 0) continue;
            for (int c = 0; c < cn; ++c)
                pad[(static_cast<size_t>(y) * PW + x) * cn + c] =
                    static_cast<double>(src[(static_cast<size_t>(sy) * W + sx) * cn + c]);
        }
    }
    const int fw = static_cast<int>(pow2_ceil(static_cast<size_t>(PW)));
    const int fh = static_cast<int>(pow2_ceil(static_cast<size_t>(PH)));
    std::vector<cplx> A(static_cast<size_t>(fw) * fh), B(A.size());
    const double s = 1.0 / (static_cast<double>(fw) * fh);
    for (int c = 0; c < cn; ++c) {
        std::fill(A.begin(), A.end(), cplx(0, 0));
        for (int y = 0; y < PH; ++y)
            for (int x = 0; x < PW; ++x)
                A[static_cast<size_t>(y) * fw + x].real(pad[(static_cast<size_t>(y) * PW + x) * cn + c]);
        for (int y = 0; y < kh; ++y)
            for (int x = 0; x < kw; ++x)
                A[static_cast<size_t>(y) * fw + x].imag(static_cast<double>(K[static_cast<size_t>(y) * kw + x]));
        fft_2d(A, fw, fh, -1);
        for (int v = 0; v < fh; ++v)
            for (int u = 0; u < fw; ++u) {
                const cplx z = A[static_cast<size_t>(v) * fw + u];
                const cplx zm = std::conj(A[static_cast<size_t>((fh - v) % fh) * fw + (fw - u) % fw]);
                const cplx fi = (z + zm) * 0.5;
                const cplx fk = (z - zm) * cplx(0, -0.5);
                B[static_cast<size_t>(v) * fw + u] = fi * std::conj(fk);
            }
        fft_2d(B, fw, fh, +1);
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x)
                dst[(static_cast<size_t>(y) * W + x) * cn + c] =
                    sat<To>(B[static_cast<size_t>(y) * fw + x].real() * s + delta);
    }
}

}  // namespace

extern "C" {

// filter2DFFT (fft.hpp:196-216); the kernel's anchor is its centre.
int ipfo_filter2d_fft(const void* src, int w, int h, int cn, int tin, void* dst, int tout, const float* K, int kw,
                      int kh, int border, double delta, double bval) {
    return guarded([&] {
        check(K != nullptr && kw > 0 && kh > 0, "filter2DFFT: kernel must be single-channel");
        check(static_cast<long long>(w) * h > 0 || border == B_CONSTANT, "copyMakeBorder: empty source image");
        check_image(w, h, cn);
        return with_type(tin, [&](auto* ti) {
            using Ti = std::remove_pointer_t<decltype(ti)>;
            return with_type(tout, [&](auto* to) {
                using To = std::remove_pointer_t<decltype(to)>;
                filter2d_fft<Ti, To>(static_cast<const Ti*>(src), w, h, cn, static_cast<To*>(dst), K, kw, kh, border,
                                     delta, bval);
                return 0;
            });
        });
    });
}

}  // extern "C"

// ---------------------------------------------------------------- integralImage
namespace {

// integralImage (filter.hpp:522-538): row 0 and column 0 zero; per row a
// running sum (of v or v * v, in FP64) left to right, each entry the one above
// plus that running sum.
template <class T>
void integral_image(const T* src, int W, int H, int cn, double* I, bool squared) {
    const size_t row = static_cast<size_t>(W + 1) * cn;
    std::fill(I, I + row * static_cast<size_t>(H + 1), 0.0);
    std::vector<double> run(static_cast<size_t>(cn));
    for (int y = 0; y < H; ++y) {
        std::fill(run.begin(), run.end(), 0.0);
        const double* above = I + static_cast<size_t>(y) * row;
        double* here = I + static_cast<size_t>(y + 1) * row;
        for (int x = 0; x < W; ++x)
            for (int c = 0; c < cn; ++c) {
                const double v = static_cast<double>(src[(static_cast<size_t>(y) * W + x) * cn + c]);
                run[c] += squared ? v * v : v;
                here[static_cast<size_t>(x + 1) * cn + c] = above[static_cast<size_t>(x + 1) * cn + c] + run[c];
            }
    }
}

}  // namespace

extern "C" {

int ipfo_integral_image(const void* src, int w, int h, int cn, int dtype, double* dst, int squared) {
    return guarded([&] {
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            integral_image<T>(static_cast<const T*>(src), w, h, cn, dst, squared != 0);
            return 0;
        });
    });
}

}  // extern "C"

// ---------------------------------------------------------------- houghCircles
namespace {

// dilate with a 3 x 3 rectangle (morphology.hpp:50-79): the maximum over the
// in-image part of the window, rows then columns.
std::vector<float> dilate3(const std::vector<float>& a, int W, int H) {
    std::vector<float> t(a.size()), o(a.size());
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float v = a[static_cast<size_t>(y) * W + x];
            for (int i = std::max(0, x - 1); i <= std::min(W - 1, x + 1); ++i)
                if (v < a[static_cast<size_t>(y) * W + i]) v = a[static_cast<size_t>(y) * W + i];
            t[static_cast<size_t>(y) * W + x] = v;
        }
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            float v = t[static_cast<size_t>(y) * W + x];
            for (int j = std::max(0, y - 1); j <= std::min(H - 1, y + 1); ++j)
                if (v < t[static_cast<size_t>(j) * W + x]) v = t[static_cast<size_t>(j) * W + x];
            o[static_cast<size_t>(y) * W + x] = v;
        }
    return o;
}

struct Circle {
    double cx, cy, r;
    int votes;
};

// houghCircles (features.hpp:388-504): canny edges vote for centres along
// their gradient; accumulator peaks (3 x 3 box sums, findPeaks
// features.hpp:164-187) are refined: vote-weighted centroid, best radius by
// radially aligned support, two Kasa least-squares fits, angular coverage in
// 36 sectors; then sorted by support and de-duplicated by minDist.
template <class T>
std::vector<Circle> hough_circles(const T* src, int W, int H, int minR, int maxR, double high, int centerVotes,
                                  double minDist, double minCoverage) {
    const size_t n = static_cast<size_t>(W) * H;
    std::vector<uint8_t> edges(n);
    canny<T>(src, W, H, 1, edges.data(), high / 2, high, 3, false);
    const int ks = ipfo_gaussian_kernel_1d(1.0, 0, nullptr, 0);
    std::vector<double> k(static_cast<size_t>(ks));
    ipfo_gaussian_kernel_1d(1.0, 0, k.data(), ks);
    std::vector<float> smooth(n), dx(n), dy(n);
    sep_filter<T, float>(src, W, H, 1, smooth.data(), k.data(), ks, k.data(), ks, ks / 2, ks / 2, B_REFLECT101, 0.0,
                         0.0);
    sobel<float>(smooth.data(), W, H, 1, dx.data(), 1, 0, 3, 1.0, B_REFLECT101);
    sobel<float>(smooth.data(), W, H, 1, dy.data(), 0, 1, 3, 1.0, B_REFLECT101);
    struct Pt {
        double x, y, nx, ny;
    };
    std::vector<Pt> pts;
    std::vector<float> acc(n, 0.f);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t p = static_cast<size_t>(y) * W + x;
            if (!edges[p]) continue;
            const double gx = dx[p], gy = dy[p], g = std::hypot(gx, gy);
            if (g < 1e-9) continue;
            pts.push_back({static_cast<double>(x), static_cast<double>(y), gx / g, gy / g});
            for (int sgn = -1; sgn <= 1; sgn += 2)
                for (int r = minR; r <= maxR; ++r) {
                    const int cx = static_cast<int>(std::lround(x + sgn * r * gx / g));
                    const int cy = static_cast<int>(std::lround(y + sgn * r * gy / g));
                    if (cx >= 0 && cy >= 0 && cx < W && cy < H) acc[static_cast<size_t>(cy) * W + cx] += 1.f;
                }
        }
    const double kAlign = 0.92;
    auto radial = [](const Pt& e, double cx, double cy, double& dist) {
        const double ux = e.x - cx, uy = e.y - cy;
        dist = std::hypot(ux, uy);
        return dist > 1e-9 ? std::fabs(ux * e.nx + uy * e.ny) / dist : 0.0;
    };
    // boxFilter(acc, 3, 3, false, Constant): sums of small integers, exact
    std::vector<float> sm(n);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double s = 0;
            for (int j = -1; j <= 1; ++j)
                for (int i = -1; i <= 1; ++i) {
                    const int sx = x + i, sy = y + j;
                    if (sx >= 0 && sy >= 0 && sx < W && sy < H) s += acc[static_cast<size_t>(sy) * W + sx];
                }
            sm[static_cast<size_t>(y) * W + x] = static_cast<float>(s);
        }
    // findPeaks(sm, centerVotes - 0.5, minDist)
    const std::vector<float> dil = dilate3(sm, W, H);
    const float thr = static_cast<float>(centerVotes) - 0.5f;
    struct Peak {
        int x, y;
        float v;
    };
    std::vector<Peak> cand, peaks;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const float v = sm[static_cast<size_t>(y) * W + x];
            if (v > thr && v >= dil[static_cast<size_t>(y) * W + x]) cand.push_back({x, y, v});
        }
    std::stable_sort(cand.begin(), cand.end(), [](const Peak& a, const Peak& b) { return a.v > b.v; });
    const double md2 = minDist * minDist;
    for (const Peak& p : cand) {
        bool ok = true;
        for (const Peak& o : peaks) {
            const double ddx = p.x - o.x, ddy = p.y - o.y;
            if (ddx * ddx + ddy * ddy < md2) {
                ok = false;
                break;
            }
        }
        if (ok) peaks.push_back(p);
    }
    constexpr double kPi = 3.14159265358979323846;
    std::vector<Circle> circles;
    for (const Peak& p : peaks) {
        double sw = 0, scx = 0, scy = 0;
        for (int j = -2; j <= 2; ++j)
            for (int i = -2; i <= 2; ++i) {
                const int x = p.x + i, y = p.y + j;
                if (x < 0 || y < 0 || x >= W || y >= H) continue;
                const double v = acc[static_cast<size_t>(y) * W + x];
                sw += v;
                scx += v * x;
                scy += v * y;
            }
        double cx = sw > 0 ? scx / sw : p.x, cy = sw > 0 ? scy / sw : p.y;
        std::vector<int> hist(static_cast<size_t>(maxR + 2), 0);
        for (const Pt& e : pts) {
            double d;
            if (radial(e, cx, cy, d) < kAlign) continue;
            const int r = static_cast<int>(std::lround(d));
            if (r >= minR && r <= maxR) ++hist[r];
        }
        int bestR = 0;
        double bestScore = 0;
        for (int r = minR; r <= maxR; ++r) {
            const double score = (hist[r] + 0.5 * (hist[r - 1] + hist[r + 1])) / (2 * kPi * r);
            if (score > bestScore) {
                bestScore = score;
                bestR = r;
            }
        }
        if (!bestR) continue;
        double radius = bestR;
        for (int iter = 0; iter < 2; ++iter) {
            double sx = 0, sy = 0, sxx = 0, syy = 0, sxy = 0, sz = 0, sxz = 0, syz = 0, cnt = 0;
            for (const Pt& e : pts) {
                double d;
                if (radial(e, cx, cy, d) < kAlign || std::fabs(d - radius) > 2.0) continue;
                const double x = e.x - cx, y = e.y - cy, z = x * x + y * y;
                sx += x;
                sy += y;
                sxx += x * x;
                syy += y * y;
                sxy += x * y;
                sz += z;
                sxz += x * z;
                syz += y * z;
                cnt += 1;
            }
            if (cnt < 3) break;
            const double m[9] = {sxx, sxy, sx, sxy, syy, sy, sx, sy, cnt}, rhs[3] = {-sxz, -syz, -sz};
            auto det3 = [](const double* t) {
                return t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                       t[2] * (t[3] * t[7] - t[4] * t[6]);
            };
            const double det = det3(m);
            if (std::fabs(det) < 1e-12) break;
            double sol[3];
            for (int col = 0; col < 3; ++col) {
                double t[9];
                std::copy(m, m + 9, t);
                for (int r = 0; r < 3; ++r) t[r * 3 + col] = rhs[r];
                sol[col] = det3(t) / det;
            }
            const double ox = -sol[0] / 2, oy = -sol[1] / 2, r2 = ox * ox + oy * oy - sol[2];
            if (r2 <= 0) break;
            cx = cx + ox;
            cy = cy + oy;
            radius = std::sqrt(r2);
        }
        if (radius < minR - 1 || radius > maxR + 1) continue;
        bool sector[36] = {};
        int support = 0;
        for (const Pt& e : pts) {
            double d;
            if (radial(e, cx, cy, d) < kAlign || std::fabs(d - radius) > 1.5) continue;
            ++support;
            const double a = std::atan2(e.y - cy, e.x - cx) + kPi;
            sector[std::min(35, static_cast<int>(a / (2 * kPi) * 36))] = true;
        }
        const int covered = static_cast<int>(std::count(std::begin(sector), std::end(sector), true));
        if (covered < minCoverage * 36) continue;
        circles.push_back({cx, cy, radius, support});
    }
    std::stable_sort(circles.begin(), circles.end(), [](const Circle& a, const Circle& b) { return a.votes > b.votes; });
    std::vector<Circle> unique;
    for (const Circle& c : circles) {
        bool dup = false;
        for (const Circle& u : unique) dup |= std::hypot(c.cx - u.cx, c.cy - u.cy) < minDist;
        if (!dup) unique.push_back(c);
    }
    return unique;
}

}  // namespace

extern "C" {

// houghCircles: writes up to `capacity` circles as (cx, cy, radius, votes)
// quadruples of doubles into out and their count into *count.
int ipfo_hough_circles(const void* src, int w, int h, int cn, int dtype, int min_r, int max_r, double canny_high,
                       int center_votes, double min_dist, double min_coverage, double* out, int capacity, int* count) {
    return guarded([&] {
        check(cn == 1, "houghCircles: expected a single-channel image");
        check(min_r > 0 && max_r >= min_r, "houghCircles: invalid radius range");
        check_image(w, h, cn);
        return with_type(dtype, [&](auto* tag) {
            using T = std::remove_pointer_t<decltype(tag)>;
            const std::vector<Circle> cs = hough_circles<T>(static_cast<const T*>(src), w, h, min_r, max_r, canny_high,
                                                            center_votes, min_dist, min_coverage);
            *count = static_cast<int>(cs.size());
            for (int i = 0; i < static_cast<int>(cs.size()) && i < capacity; ++i) {
                out[4 * i] = cs[i].cx;
                out[4 * i + 1] = cs[i].cy;
                out[4 * i + 2] = cs[i].r;
                out[4 * i + 3] = cs[i].votes;
            }
            return 0;
        });
    });
}

}  // extern "C"
