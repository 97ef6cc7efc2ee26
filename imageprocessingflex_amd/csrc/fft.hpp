// Launch interface of filter2DFFT (fft.hip) for the C-ABI runtime.
#pragma once

#include "common.hpp"

namespace ipfg {

struct FftLaunch {
    DevImage src, dst;       // same size and channels
    int kw, kh, border;      // kernel size (anchor at its centre), border mode
    double bval, delta;
    const float* kernel;     // kw * kh weights, row-major (device)
    double2* buf0;           // 2^lw * 2^lh complex values each (device)
    double2* buf1;
    int lw, lh;              // transform sizes: 2^lw >= W + kw - 1, 2^lh >= H + kh - 1
    // twiddles exp(i sign 2 pi k / len) of every stage, len = 2 .. n, each
    // stage's table at offset len / 2 - 1 (n - 1 values per table)
    const double2* tw_fwd_w;
    const double2* tw_fwd_h;
    const double2* tw_inv_w;
    const double2* tw_inv_h;
};
// Transform sizes this build handles.
bool fft_shape_ok(int lw, int lh);
hipError_t launch_filter2d_fft(const FftLaunch& a, hipStream_t s);

}  // namespace ipfg
