// ipf front end for MI355X: the feature callers of the filters from the
// reference's include/ipf/features.hpp:26-153, 373-504 -- canny, cannyAuto,
// cornerHarris, cornerMinEigenVal, houghCircles. The other functions of that
// header (Hough lines, template matching, local binary patterns, peaks,
// non-maximum suppression of boxes) are not part of this build.
//
// Each call validates its arguments in the reference's order, then runs on the
// GPU through the C ABI (ipfg.h) with every intermediate (gradients, edge
// states, component labels, structure tensor sums) in device memory, and
// returns the reference's result bit for bit. Link with -lipfg.
#pragma once

#include <vector>

#include "core.hpp"
#include "filter.hpp"

namespace ipf {

/// Canny edge detection (OpenCV semantics): Sobel gradients with aperture ksize,
/// non-maximum suppression and hysteresis between `lowThreshold` and `highThreshold`
/// (in gradient-magnitude units; L1 magnitude unless l2Gradient). Color images use the
/// channel with the strongest gradient at each pixel. Returns a 0/255 edge mask.
template <class T>
ImageU8 canny(const Image<T>& src, double lowThreshold, double highThreshold, int ksize = 3,
              bool l2Gradient = false) {
    detail::require(lowThreshold <= highThreshold, "canny: lowThreshold must be <= highThreshold");
    ImageU8 out(src.width(), src.height(), 1);
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_canny(&s, &d, lowThreshold, highThreshold, ksize, l2Gradient ? 1 : 0, gpu::streamSetting()));
    return out;
}

/// Canny with automatic thresholds: Gaussian blur of the gray image, high threshold at
/// the 70th percentile of the gradient magnitude, low threshold = 0.4 * high.
template <class T>
ImageU8 cannyAuto(const Image<T>& src, double blurSigma = 1.4) {
    ImageU8 out(src.width(), src.height(), 1);
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_canny_auto(&s, &d, blurSigma, gpu::streamSetting()));
    return out;
}

/// Harris corner response R = det(M) - k * trace(M)^2 of the structure tensor M.
template <class T>
ImageF cornerHarris(const Image<T>& src, int blockSize = 3, int ksize = 3, double k = 0.04) {
    detail::require(src.channels() == 1, "corner detection: expected a single-channel image");
    ImageF out(src.width(), src.height(), 1);
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_corner_harris(&s, &d, blockSize, ksize, k, gpu::streamSetting()));
    return out;
}

/// Minimum eigenvalue of the structure tensor (Shi-Tomasi corner strength).
template <class T>
ImageF cornerMinEigenVal(const Image<T>& src, int blockSize = 3, int ksize = 3) {
    detail::require(src.channels() == 1, "corner detection: expected a single-channel image");
    ImageF out(src.width(), src.height(), 1);
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_corner_min_eigen_val(&s, &d, blockSize, ksize, gpu::streamSetting()));
    return out;
}

struct HoughCircle {
    PointF center;
    double radius = 0;
    int votes = 0;  ///< number of edge pixels supporting the circle
};

/// Gradient-based Hough circle detection on a grayscale image.
///  1. Canny edges (cannyHigh / 2, cannyHigh) vote for centers along their gradient direction.
///  2. Accumulator peaks with at least `centerVotes` votes (and `minDist` apart) become candidates.
///  3. Each candidate gets the radius in [minRadius, maxRadius] best supported by edge pixels whose
///     gradient points radially, then center and radius are refined by a least-squares circle fit.
///  4. A circle is accepted when its supporting edge pixels cover at least `minCoverage` of the
///     circumference (measured in 10-degree sectors), which rejects straight edges and corners.
/// Results are sorted by support (number of supporting edge pixels).
template <class T>
std::vector<HoughCircle> houghCircles(const Image<T>& gray, int minRadius, int maxRadius, double cannyHigh = 100,
                                      int centerVotes = 20, double minDist = 10, double minCoverage = 0.6) {
    detail::require(gray.channels() == 1, "houghCircles: expected a single-channel image");
    detail::require(minRadius > 0 && maxRadius >= minRadius, "houghCircles: invalid radius range");
    const ipfg_image s = detail::describe(gray);
    std::vector<double> buf(4 * 64);
    int n = 0;
    for (;;) {
        detail::check(ipfg_hough_circles(&s, minRadius, maxRadius, cannyHigh, centerVotes, minDist, minCoverage,
                                         buf.data(), static_cast<int>(buf.size() / 4), &n, gpu::streamSetting()));
        if (4 * static_cast<size_t>(n) <= buf.size()) break;
        buf.resize(4 * static_cast<size_t>(n));
    }
    std::vector<HoughCircle> out(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        out[i] = HoughCircle{PointF(buf[4 * i], buf[4 * i + 1]), buf[4 * i + 2], static_cast<int>(buf[4 * i + 3])};
    return out;
}

}  // namespace ipf
