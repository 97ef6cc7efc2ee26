// ipf front end for MI355X: the feature callers of the filters from the
// reference's include/ipf/features.hpp:26-153 -- canny, cannyAuto,
// cornerHarris, cornerMinEigenVal. The other functions of that header (Hough
// transforms, template matching, local binary patterns, peaks, non-maximum
// suppression of boxes) are not part of this build.
//
// Each call validates its arguments in the reference's order, then runs on the
// GPU through the C ABI (ipfg.h) with every intermediate (gradients, edge
// states, component labels, structure tensor sums) in device memory, and
// returns the reference's result bit for bit. Link with -lipfg.
#pragma once

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

}  // namespace ipf
