// ipf front end for MI355X: the Gaussian pyramid functions of the reference's
// include/ipf/geometry.hpp:501-535 (pyrDown, pyrUp, buildPyramid). The other
// geometry functions of that header (warps, resizing, flips) are not part of
// this build.
//
// pyrDown and pyrUp run one gfx950 launch each through the C ABI (ipfg.h) and
// are bit-identical to the reference for every pixel type; buildPyramid is
// the reference's loop of pyrDown calls. Link with -lipfg.
#pragma once

#include <vector>

#include "core.hpp"
#include "filter.hpp"

namespace ipf {

/// Blurs with the 5-tap binomial kernel and drops every other row/column.
/// Output size ((W + 1) / 2, (H + 1) / 2), as OpenCV's pyrDown.
template <class T>
Image<T> pyrDown(const Image<T>& src) {
    Image<T> out((src.width() + 1) / 2, (src.height() + 1) / 2, src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_pyr_down(&s, &d, gpu::streamSetting()));
    return out;
}

/// Doubles the size: zero-insertion followed by the binomial kernel scaled by 4.
template <class T>
Image<T> pyrUp(const Image<T>& src) {
    Image<T> out(src.width() * 2, src.height() * 2, src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_pyr_up(&s, &d, gpu::streamSetting()));
    return out;
}

/// Builds a Gaussian pyramid with `levels` images (level 0 = the source).
template <class T>
std::vector<Image<T>> buildPyramid(const Image<T>& src, int levels) {
    detail::require(levels >= 1, "buildPyramid: levels must be >= 1");
    std::vector<Image<T>> pyr{src};
    while (static_cast<int>(pyr.size()) < levels && pyr.back().width() > 1 && pyr.back().height() > 1)
        pyr.push_back(pyrDown(pyr.back()));
    return pyr;
}

}  // namespace ipf
