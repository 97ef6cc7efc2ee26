// Host-side checks of the C ABI (libipfg's runtime, include/ipfg.h) that need
// no device: every entry point's argument validation with the reference's
// messages, the tuning table's text format, the Netpbm header parser, the
// Gaussian weights and the thread setting. tests/test_host_cpp.py builds it
// with the runtime's host code under AddressSanitizer and
// UndefinedBehaviorSanitizer and runs it on the CPU.
#include <ipfg.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

int failures = 0;

void expect(int rc, int want_rc, const char* want_msg, const char* what) {
    const std::string msg = ipfg_last_error();
    if (rc != want_rc || (want_msg && msg != want_msg)) {
        std::printf("FAIL %s: rc %d (want %d) msg '%s' (want '%s')\n", what, rc, want_rc, msg.c_str(),
                    want_msg ? want_msg : "");
        ++failures;
    }
}

ipfg_image host_image(void* data, int w, int h, int cn, int dt) {
    return ipfg_image{data, w, h, cn, dt, 0, 0};
}

}  // namespace

int main() {
    std::vector<unsigned char> u8(64 * 64 * 4, 7);
    std::vector<float> f32(64 * 64 * 4, 0.5f);
    const ipfg_image g = host_image(u8.data(), 8, 6, 1, IPFG_U8);
    const ipfg_image g3 = host_image(u8.data(), 8, 6, 3, IPFG_U8);
    ipfg_image out = host_image(u8.data() + 1024, 8, 6, 1, IPFG_U8);
    ipfg_image outf = host_image(f32.data(), 8, 6, 1, IPFG_F32);
    const double k3[3] = {0.25, 0.5, 0.25};
    const float kf[9] = {0, 1, 0, 1, -4, 1, 0, 1, 0};

    // argument checks (reference messages, before any device work)
    expect(ipfg_median(&g, &out, 4, IPFG_BORDER_REPLICATE, nullptr), IPFG_ERR_ARG,
           "ipf: medianBlur: ksize must be odd and positive", "median ksize");
    expect(ipfg_median(&g, &out, 3, IPFG_BORDER_CONSTANT, nullptr), IPFG_ERR_ARG,
           "ipf: medianBlur: Constant border is not supported", "median border");
    expect(ipfg_bilateral(&g, &out, 5, 0.0, 3.0, IPFG_BORDER_REFLECT101, 0, nullptr), IPFG_ERR_ARG,
           "ipf: bilateralFilter: sigmas must be positive", "bilateral sigma");
    expect(ipfg_box_filter(&g, &out, 0, 3, 1, IPFG_BORDER_REFLECT101, 0.0, 0, nullptr), IPFG_ERR_ARG,
           "ipf: boxFilter: kernel size must be positive", "box size");
    expect(ipfg_sep_filter(&g, &out, k3, 0, k3, 3, 0, 1, IPFG_BORDER_REFLECT101, 0.0, 0.0, 0, nullptr), IPFG_ERR_ARG,
           "ipf: separable filter: empty kernel", "sep empty");
    expect(ipfg_filter2d(&g, &out, kf, 3, 3, 3, 1, IPFG_BORDER_REFLECT101, 0.0, 0.0, 0, nullptr), IPFG_ERR_ARG,
           "ipf: filter2D: anchor outside kernel", "filter2D anchor");
    expect(ipfg_gaussian_blur(&g, &out, 0.0, 0.0, 0, IPFG_BORDER_REFLECT101, 0, nullptr), IPFG_ERR_ARG,
           "ipf: gaussianKernel1D: need sigma > 0 or ksize > 0", "gaussian sigma");
    expect(ipfg_sobel(&g, &outf, 0, 0, 3, 1.0, IPFG_BORDER_REFLECT101, 0, nullptr), IPFG_ERR_ARG,
           "ipf: sobel: need dx + dy > 0", "sobel order");
    expect(ipfg_sobel(&g, &outf, 1, 0, 33, 1.0, IPFG_BORDER_REFLECT101, 0, nullptr), IPFG_ERR_ARG,
           "ipf: derivKernel: ksize must be odd in [3, 31]", "sobel ksize");
    expect(ipfg_adaptive_threshold(&g3, &g3, 255.0, 0, 0, 3, 0.0, 0, nullptr), IPFG_ERR_ARG,
           "ipf: adaptiveThreshold: expected a single-channel image", "adaptive channels");
    expect(ipfg_canny(&g, &out, 5.0, 3.0, 3, 0, nullptr), IPFG_ERR_ARG,
           "ipf: canny: lowThreshold must be <= highThreshold", "canny thresholds");
    expect(ipfg_corner_harris(&g3, &outf, 3, 3, 0.04, nullptr), IPFG_ERR_ARG,
           "ipf: corner detection: expected a single-channel image", "harris channels");
    expect(ipfg_filter2d_fft(&g, &out, nullptr, 3, 3, IPFG_BORDER_REFLECT101, 0.0, 0.0, nullptr), IPFG_ERR_ARG,
           "ipf: filter2DFFT: kernel must be single-channel", "fft kernel");
    int count = 0;
    double circles[8];
    expect(ipfg_hough_circles(&g, 5, 4, 100.0, 20, 10.0, 0.6, circles, 2, &count, nullptr), IPFG_ERR_ARG,
           "ipf: houghCircles: invalid radius range", "hough radii");
    // ABI-level checks
    expect(ipfg_median(nullptr, &out, 3, IPFG_BORDER_REPLICATE, nullptr), IPFG_ERR_ARG, "ipfg: null src image",
           "null src");
    ipfg_image bad = g;
    bad.dtype = 17;
    expect(ipfg_median(&bad, &out, 3, IPFG_BORDER_REPLICATE, nullptr), IPFG_ERR_ARG, "ipfg: unsupported pixel type",
           "dtype");
    bad = g;
    bad.width = -1;
    expect(ipfg_median(&bad, &out, 3, IPFG_BORDER_REPLICATE, nullptr), IPFG_ERR_ARG, "ipf: Image: negative dimensions",
           "negative width");
    bad = g;
    bad.row_stride_bytes = 3;
    expect(ipfg_median(&bad, &out, 3, IPFG_BORDER_REPLICATE, nullptr), IPFG_ERR_ARG, "ipfg: row stride too small",
           "stride");

    // Netpbm / PFM headers
    const char* headers[][2] = {
        {"Q5 1 1 255\n", "ipf: PNM: not a P1-P6 netpbm file"},
        {"P5 0 1 255\n", "ipf: PNM: invalid image dimensions"},
        {"P5 2 x", "ipf: PNM: malformed header or pixel data"},
        {"P5 99999999999 1 255\n", "ipf: PNM: number too large"},
        {"P5 2 1 70000\n", "ipf: PNM: invalid maxval"},
        {"Pf 2 1 abc\n", "ipf: PFM: invalid scale value"},
        {"PF 20000000 1 -1.0\n", "ipf: PFM: image dimensions too large"},
        {"P5 #c\n 3 #c\n 2 255\n", nullptr},
    };
    for (const auto& hcase : headers) {
        int w = 0, h = 0, cn = 0, pfm = 0;
        const int rc = ipfg_pnm_info(reinterpret_cast<const uint8_t*>(hcase[0]), std::strlen(hcase[0]), &w, &h, &cn,
                                     &pfm);
        expect(rc, hcase[1] ? IPFG_ERR_ARG : IPFG_OK, hcase[1], hcase[0]);
        if (!hcase[1] && (w != 3 || h != 2 || cn != 1 || pfm != 0)) {
            std::printf("FAIL header fields %d %d %d %d\n", w, h, cn, pfm);
            ++failures;
        }
    }
    const char truncated[] = "P6 4 4 255\n\x01\x02";
    ipfg_image px = host_image(u8.data(), 4, 4, 3, IPFG_U8);
    expect(ipfg_decode_pnm(reinterpret_cast<const uint8_t*>(truncated), sizeof(truncated) - 1, &px, nullptr),
           IPFG_ERR_ARG, "ipf: PNM: unexpected end of data (truncated file?)", "pnm truncated");

    // Gaussian weights (filter.hpp:129-145)
    double kw[64];
    const int n = ipfg_gaussian_kernel_1d(2.0, 0, kw, 64);
    double sum = 0;
    for (int i = 0; i < n; ++i) sum += kw[i];
    if (n != 13 || std::fabs(sum - 1.0) > 1e-15 || kw[0] != kw[12]) {
        std::printf("FAIL gaussian kernel n=%d sum=%.17g\n", n, sum);
        ++failures;
    }
    if (ipfg_gaussian_kernel_1d(1.0, 4, kw, 64) != -IPFG_ERR_ARG) ++failures;

    // tuning table text
    ipfg_tune_clear();
    int entries = 0;
    expect(ipfg_tune_import("sep 0 0 3 13 13 1 11 11 64\n# c\n\nsep 4 4 1 49 49 1 13 13 128\n", &entries), IPFG_OK,
           nullptr, "tune import");
    size_t need = 0;
    char buf[256];
    expect(ipfg_tune_export(buf, sizeof buf, &need), IPFG_OK, nullptr, "tune export");
    if (entries != 2 || need == 0 || std::strstr(buf, "sep 4 4 1 49 49 1 13 13 128") == nullptr) {
        std::printf("FAIL tune table: %d entries, '%s'\n", entries, buf);
        ++failures;
    }
    char tiny[4];
    expect(ipfg_tune_export(tiny, sizeof tiny, &need), IPFG_OK, nullptr, "tune export small");
    expect(ipfg_tune_import("sep 0 0 3 13 13 1 11 11\n", &entries), IPFG_ERR_ARG, nullptr, "tune malformed");
    ipfg_tune_clear();

    // thread setting (parallel.hpp:38-51 contract)
    ipfg_set_num_threads(3);
    if (ipfg_get_num_threads() != 3) ++failures;
    ipfg_set_num_threads(-5);
    if (ipfg_get_num_threads() < 1) ++failures;
    ipfg_set_num_threads(0);

    if (failures) {
        std::printf("%d failures\n", failures);
        return 1;
    }
    std::printf("ABI OK\n");
    return 0;
}
