// ipf front end for MI355X: ssim of the reference's include/ipf/metrics.hpp:46-78.
// The other metrics of that header (mse, psnr, mae, entropy, sharpness) are not
// part of this build.
//
// ssim runs two gfx950 launches through the C ABI (ipfg.h): the map,
// bit-identical to the reference's, and a fixed-order sum of it over the
// valid region. Link with -lipfg.
#pragma once

#include "core.hpp"
#include "filter.hpp"

namespace ipf {

/// Structural similarity (Wang et al. 2004): 11x11 Gaussian window (sigma 1.5),
/// K1 = 0.01, K2 = 0.03, dynamic range = the type's maximum. Multi-channel images
/// return the mean over channels, taken over the valid region (windows fully
/// inside the image); `ssimMap`, if given, receives the full-size map.
template <class T>
double ssim(const Image<T>& a, const Image<T>& b, ImageF* ssimMap = nullptr) {
    detail::require(a.sameShape(b) && !a.empty(), "ssim: images must be non-empty and have the same shape");
    const ipfg_image da = detail::describe(a), db = detail::describe(b);
    double r = 0.0;
    if (ssimMap) {
        ImageF map(a.width(), a.height(), a.channels());
        const ipfg_image dm = detail::describe(map);
        detail::check(ipfg_ssim(&da, &db, &dm, &r, gpu::streamSetting()));
        *ssimMap = std::move(map);
    } else {
        detail::check(ipfg_ssim(&da, &db, nullptr, &r, gpu::streamSetting()));
    }
    return r;
}

}  // namespace ipf
