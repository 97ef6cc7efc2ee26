// ipf front end for MI355X: filter2DFFT of the reference's include/ipf/fft.hpp:196-216.
// The other functions of that header (1-D/2-D DFTs, spectra, phase correlation)
// are not part of this build.
//
// filter2DFFT runs on the GPU through the C ABI (ipfg.h): FP64 radix-2 FFTs with
// the reference's butterflies and twiddles, so its results are bit-identical to
// the reference's (finite inputs). Link with -lipfg.
#pragma once

#include <type_traits>

#include "core.hpp"
#include "filter.hpp"

namespace ipf {

/// filter2D computed in the frequency domain: same result as ipf::filter2D (any border mode),
/// much faster for large kernels (roughly beyond 15x15).
template <class TOut = void, class T>
Image<std::conditional_t<std::is_void<TOut>::value, T, TOut>> filter2DFFT(
    const Image<T>& src, const ImageF& kernel, BorderMode border = BorderMode::Reflect101, double delta = 0.0,
    double borderValue = 0.0) {
    using O = std::conditional_t<std::is_void<TOut>::value, T, TOut>;
    detail::require(kernel.channels() == 1 && !kernel.empty(), "filter2DFFT: kernel must be single-channel");
    Image<O> out(src.width(), src.height(), src.channels());
    const ipfg_image s = detail::describe(src), d = detail::describe(out);
    detail::check(ipfg_filter2d_fft(&s, &d, kernel.data(), kernel.width(), kernel.height(), static_cast<int>(border),
                                    delta, borderValue, gpu::streamSetting()));
    return out;
}

}  // namespace ipf
