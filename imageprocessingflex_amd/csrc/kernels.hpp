// Launch interfaces between the C-ABI runtime (runtime.hip) and the kernels.
#pragma once

#include <vector>

#include "common.hpp"

namespace ipfg {

enum Mode { MODE_FAST = 0, MODE_CERTIFY = 1, MODE_EXACT = 2 };

// ---------------------------------------------------------------- separable
struct SepLaunch {
    DevImage src, dst;
    const float* kxf;   // kw FP32 weights + 16 zeros (device)
    const float* kyf;   // kh FP32 weights + 16 zeros (device)
    const double* kxd;  // kw FP64 weights + 16 zeros (device)
    const double* kyd;  // kh FP64 weights + 16 zeros (device)
    int kw, kh, ax, ay, border;
    double delta, bval;
    int mode;
    // MODE_CERTIFY: |FP64-FMA result - reference| <= bcoef * X + bdelta where
    // X bounds |input| (xbound, or measured per tile when xbound < 0)
    double bcoef, bdelta, xbound;
    double kabs;  // sum|kx| * sum|ky| (0: the ring holds exact sums)
    double oscale;  // results are multiplied by this (0 means 1)
    int unit;       // every weight is 1 (box filters)
    FixList fix;
    double* tmp;  // MODE_EXACT: (H + kh - 1) * W * cn doubles
    int sh;       // strip height of the fast kernels (0: built-in choice)
};
hipError_t launch_sep(const SepLaunch& a, hipStream_t s);
// Strip heights (launch shapes) worth measuring for this call.
std::vector<int> sep_sh_candidates(const SepLaunch& l);
// The fast separable kernel supports this call (channels, kernel size, types).
bool sep_fast_fits(int cn, int kw, int kh, int src_dt, int dst_dt, bool cert);
size_t sep_exact_tmp_bytes(int w, int h, int cn, int kh);

// ---------------------------------------------------------------- box
struct BoxLaunch {
    DevImage src, dst;
    int kw, kh, border;
    int normalize;
    double bval;
    int mode;      // MODE_FAST (window sums) or MODE_EXACT (reference running sums)
    double* tmp;   // MODE_EXACT: (H + kh - 1) * W * cn doubles
};
hipError_t launch_box(const BoxLaunch& a, hipStream_t s);

// ---------------------------------------------------------------- pyramids
// pyrDown (dst (w + 1) / 2 x (h + 1) / 2) or pyrUp (dst 2w x 2h) of src, same
// pixel type and channels; every output bit-identical to the reference.
hipError_t launch_pyr(const DevImage& src, const DevImage& dst, bool up, hipStream_t s);

// ---------------------------------------------------------------- ssim
// Weights of the 11-tap Gaussian window and the stabilising constants.
struct SsimTaps {
    double w[11];
    double c1, c2;
};
// Region the mean is taken over: columns [x0, x1), rows [y0, y1).
struct SsimRoi {
    int x0, x1, y0, y1;
};
// Number of per-workgroup partial sums launch_ssim writes for a (doubles).
long long ssim_partials(const DevImage& a);
// SSIM map of a and b (same shape and type) into map (float samples, row
// stride map_stride samples; may be null) and the sum of the map over roi
// into sum[0]; partial holds ssim_partials(a) doubles.
hipError_t launch_ssim(const DevImage& a, const DevImage& b, const SsimTaps& k, float* map, size_t map_stride,
                       const SsimRoi& roi, double* partial, double* sum, hipStream_t s);

// ---------------------------------------------------------------- filter2D
struct F2dLaunch {
    DevImage src, dst;
    const float* kf;      // kw*kh FP32 weights, row-major (device)
    const double* kd;     // kw*kh FP64 weights, row-major (device)
    const int* taps;      // non-zero taps: ntaps (dx, dy) pairs (device)
    const double* tapw;   // non-zero tap weights as double (device)
    int ntaps;
    int kw, kh, ax, ay, border;
    int has_zero;         // the kernel has zero weights (they must not take part)
    double delta, bval;
    int mode;
    double bcoef, bdelta, xbound;  // MODE_CERTIFY bound: bcoef * X + bdelta
    FixList fix;
    int precision;        // IPFG_PREC_* for the MFMA variants
    const void* mfma_w;   // split weight fragments for the MFMA path (device)
    double mfma_wscale;   // power of two the fragments' weights were scaled by
};
hipError_t launch_filter2d(const F2dLaunch& a, hipStream_t s);
// 3x3 / 5x5 kernels with a centred anchor on uint8 images (uint8 or float32
// out) whose sums are exact in FP32; k is row-major on the host.
// hipErrorNotSupported when the call does not fit (types, sizes, alignment).
// kx / ky (optional): k is their outer product, applied as two 1-D passes.
hipError_t launch_stencil_u8(const DevImage& src, const DevImage& dst, const float* k, int kw, int kh, int ay,
                             int border, double delta, double bval, hipStream_t s, const float* kx = nullptr,
                             const float* ky = nullptr);
// medianBlur 3 x 3 of uint8 images (1, 3 or 4 channels, 16-byte aligned
// rows, border not Constant) in the stencil layout; hipErrorNotSupported
// otherwise.
hipError_t launch_median3_u8(const DevImage& src, const DevImage& dst, int border, hipStream_t s);
// unsharpMask on uint8 images (ks <= 13, 1/3/4 channels, 16-byte aligned
// rows) in one launch; k_host / k_dev hold the ks FP64 Gaussian weights and
// bnd bounds |FP32 blur - reference blur|. hipErrorNotSupported otherwise.
hipError_t launch_unsharp_u8(const DevImage& src, const DevImage& dst, int ks, const double* k_host,
                             const double* k_dev, double amount, double threshold, double bnd, hipStream_t s);
// MFMA implicit-GEMM filter2D (precision BF16X3 / FP16X2).
hipError_t launch_filter2d_mfma(const F2dLaunch& a, hipStream_t s);
size_t f2d_mfma_lds_bytes(int kw, int kh, bool split_input);
int f2d_mfma_pairs(int kw, int kh);
void f2d_mfma_fragments(const double* kd, int kw, int kh, int precision, std::vector<uint16_t>& out, double& wscale);

// ---------------------------------------------------------------- median
struct MedianLaunch {
    DevImage src, dst;
    int k, border;
};
hipError_t launch_median(const MedianLaunch& a, hipStream_t s);

// ---------------------------------------------------------------- bilateral
struct BilateralLaunch {
    DevImage src, dst;
    int r, ntaps, border;
    const int* taps;        // (dx, dy) pairs (device)
    const double* spaced;   // FP64 space weights (device)
    const float* spacef;    // FP32 space weights (device)
    const double* lutd;     // FP64 range table (integer types) or null
    const float* lutf;      // FP32 range table or null
    int lut_len;
    double cc;              // -1 / (2 sigma_color^2)
    int mode;
    double bound;
    FixList fix;
};
hipError_t launch_bilateral(const BilateralLaunch& a, hipStream_t s);
bool bilateral_fast_ok(int dt, int cn, int r, int lut_len);

// ---------------------------------------------------------------- features
// canny after the gradients: gx, gy float images (cn channels) of the same
// shape; dst a one-channel uint8 image of that shape. Thresholds low / high,
// or (sel non-null) cannyAuto's thresholds from the selected magnitude bits.
struct CannyScratch {
    uint8_t* state;  // W x H
    int* parent;     // W x H
    uint8_t* flag;   // W x H
};
hipError_t launch_canny(const DevImage& gx, const DevImage& gy, double low, double high, bool l2,
                        const unsigned int* sel, const CannyScratch& sc, const DevImage& dst, hipStream_t s);
// toGray (color.hpp:78-81) of src (1-4 channels) into the one-channel dst.
hipError_t launch_to_gray(const DevImage& src, const DevImage& dst, hipStream_t s);
// The element of rank k (ascending) of n non-negative floats; its bits end in
// scratch[0] (scratch: select_scratch_bytes()).
size_t select_scratch_bytes();
hipError_t launch_select_rank(const float* v, size_t n, unsigned long long k, unsigned int* scratch, hipStream_t s);
// cornerHarris (harris) or cornerMinEigenVal response of the structure tensor
// of derivatives dx, dy (one-channel float) over block x block windows into
// the float image dst; rows holds 3 W x H doubles.
hipError_t launch_structure_response(const DevImage& dx, const DevImage& dy, int block, bool harris, double k,
                                     double* rows, const DevImage& dst, hipStream_t s);

// integralImage (filter.hpp:522-538) of src into the (W + 1) x (H + 1) x cn
// double table I (rows istride_samples doubles apart), in the reference's
// summation order.
hipError_t launch_integral(const DevImage& src, double* I, size_t istride_samples, bool squared, hipStream_t s);

// Zeroes a stream's fix-up counter and ticket with a kernel.
hipError_t launch_zero_counter(unsigned long long* count, unsigned int* done, hipStream_t s);

}  // namespace ipfg
