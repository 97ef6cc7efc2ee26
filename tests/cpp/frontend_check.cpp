// Drop-in check of the C++ front end: code written against the reference's
// ipf API (include/ipf/filter.hpp signatures and defaults) compiled unchanged
// against this repository's headers and libipfg.so.
//
//   frontend_check errors        argument checks only (no GPU needed)
//   frontend_check run <dir>     writes inputs and outputs as raw files
//   frontend_check motion n a..  prints motionBlurKernel(n, a) (no GPU needed)
#include <ipf/ipf.hpp>
#include <ipf/npy.hpp>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <functional>
#include <string>

using namespace ipf;

namespace {
template <class T>
void save(const std::string& path, const Image<T>& img) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(img.data()), static_cast<std::streamsize>(img.sampleCount() * sizeof(T)));
}

int expect_error(const char* what, const std::function<void()>& fn, const std::string& msg) {
    try {
        fn();
    } catch (const Error& e) {
        if (msg == e.what()) return 0;
        std::printf("FAIL %s: got '%s'\n", what, e.what());
        return 1;
    }
    std::printf("FAIL %s: no error\n", what);
    return 1;
}

int errors() {
    const ImageU8 img(8, 6, 1, 3);
    int bad = 0;
    bad += expect_error("median constant", [&] { medianBlur(img, 3, BorderMode::Constant); },
                        "ipf: medianBlur: Constant border is not supported");
    bad += expect_error("median even", [&] { medianBlur(img, 4); }, "ipf: medianBlur: ksize must be odd and positive");
    bad += expect_error("bilateral sigma", [&] { bilateralFilter(img, 5, 0.0, 1.0); },
                        "ipf: bilateralFilter: sigmas must be positive");
    bad += expect_error("box size", [&] { boxFilter(img, 0, 3); }, "ipf: boxFilter: kernel size must be positive");
    bad += expect_error("sep empty", [&] { sepFilter2D(img, {}, {1.0}); }, "ipf: separable filter: empty kernel");
    bad += expect_error("filter2D anchor", [&] { filter2D(img, Kernel(3, 3, 1, 1.f), BorderMode::Reflect101, 0.0, Point(3, 0)); },
                        "ipf: filter2D: anchor outside kernel");
    bad += expect_error("filter2D empty", [&] { filter2D(img, Kernel()); },
                        "ipf: filter2D: kernel must be a non-empty single-channel image");
    bad += expect_error("gauss ksize", [&] { gaussianBlur(img, 1.0, 0.0, 4); }, "ipf: gaussianKernel1D: ksize must be odd");
    bad += expect_error("makeKernel", [&] { makeKernel(2, 2, {1, 2, 3}); }, "ipf: makeKernel: wrong number of values");
    bad += expect_error("sobel order", [&] { sobel(img, 0, 0); }, "ipf: sobel: need dx + dy > 0");
    bad += expect_error("sobel ksize", [&] { sobel(img, 1, 0, 4); },
                        "ipf: derivKernel: ksize must be odd in [3, 31]");
    bad += expect_error("sobel order 3", [&] { sobel(img, 3, 0, 1); },
                        "ipf: derivKernel: order > 2 requires ksize >= 3");
    bad += expect_error("scharr", [&] { scharr(img, 1, 1); }, "ipf: scharr: (dx, dy) must be (1,0) or (0,1)");
    bad += expect_error("prewitt", [&] { prewitt(img, 0, 0); }, "ipf: prewitt: (dx, dy) must be (1,0) or (0,1)");
    bad += expect_error("magnitude", [&] { magnitude(ImageF(3, 2), ImageF(2, 3)); }, "ipf: magnitude: shape mismatch");
    bad += expect_error("phase", [&] { phase(ImageF(3, 2), ImageF(3, 2, 2)); }, "ipf: phase: shape mismatch");
    bad += expect_error("filter2DFFT kernel", [&] { filter2DFFT(img, ImageF()); },
                        "ipf: filter2DFFT: kernel must be single-channel");
    bad += expect_error("filter2DFFT empty", [&] { filter2DFFT(ImageU8(0, 3, 1), ImageF(3, 3, 1, 1.f)); },
                        "ipf: copyMakeBorder: empty source image");
    bad += expect_error("pnm magic", [&] { decodePNM(std::vector<std::uint8_t>{'P', '7', ' '}); },
                        "ipf: PNM: not a P1-P6 netpbm file");
    bad += expect_error("pnm truncated", [&] { decodePNM(std::vector<std::uint8_t>{'P', '5', ' ', '2', ' ', '1', ' ', '9', ' ', 0}); },
                        "ipf: PNM: unexpected end of data (truncated file?)");
    bad += expect_error("pfm channels", [&] { encodePFM(ImageF(2, 2, 2)); }, "ipf: encodePFM: image must have 1 or 3 channels");
    bad += expect_error("hough channels", [&] { houghCircles(ImageU8(8, 8, 3), 2, 4); },
                        "ipf: houghCircles: expected a single-channel image");
    bad += expect_error("hough radii", [&] { houghCircles(img, 5, 4); }, "ipf: houghCircles: invalid radius range");
    bad += expect_error("canny thresholds", [&] { canny(img, 9.0, 3.0); },
                        "ipf: canny: lowThreshold must be <= highThreshold");
    bad += expect_error("canny ksize", [&] { canny(img, 1.0, 3.0, 2); }, "ipf: derivKernel: ksize must be odd in [3, 31]");
    bad += expect_error("cannyAuto gray", [&] { cannyAuto(ImageU8(4, 4, 5)); },
                        "ipf: rgbToGray: expected a 3-channel (RGB) or 4-channel (RGBA) image");
    bad += expect_error("harris channels", [&] { cornerHarris(ImageU8(4, 4, 3)); },
                        "ipf: corner detection: expected a single-channel image");
    bad += expect_error("min eigen block", [&] { cornerMinEigenVal(img, 0); },
                        "ipf: boxFilter: kernel size must be positive");
    bad += expect_error("adaptive channels",
                        [&] { adaptiveThreshold(ImageU8(4, 4, 3), 255, AdaptiveMethod::Mean, ThresholdType::Binary, 3, 0); },
                        "ipf: adaptiveThreshold: expected a single-channel image");
    bad += expect_error("adaptive block",
                        [&] { adaptiveThreshold(img, 255, AdaptiveMethod::Mean, ThresholdType::Binary, 4, 0); },
                        "ipf: adaptiveThreshold: blockSize must be odd and >= 3");
    bad += expect_error("adaptive type",
                        [&] { adaptiveThreshold(img, 255, AdaptiveMethod::Gaussian, ThresholdType::Trunc, 5, 0); },
                        "ipf: adaptiveThreshold: type must be Binary or BinaryInv");
    bad += expect_error("motion", [&] { motionBlurKernel(0, 0.0); }, "ipf: motionBlurKernel: length must be >= 1");
    bad += expect_error("buildPyramid", [&] { buildPyramid(img, 0); }, "ipf: buildPyramid: levels must be >= 1");
    // empty images come back empty without device work; ksize 1 is a copy
    if (!gaussianBlur(ImageU8(0, 5, 3), 2.0).empty()) ++bad;
    if (!(medianBlur(img, 1) == img)) ++bad;
    // core semantics used by the path
    if (saturate_cast<std::uint8_t>(127.5) != 128 || saturate_cast<std::int16_t>(-2.5) != -3) ++bad;
    if (borderIndex(-1, 4, BorderMode::Reflect101) != 1 || borderIndex(4, 4, BorderMode::Wrap) != 0) ++bad;
    std::printf(bad ? "ERRORS FAILED %d\n" : "ERRORS OK\n", bad);
    return bad ? 1 : 0;
}

// motionBlurKernel(length, angle) for each pair of arguments, one kernel a line.
int motion(int argc, char** argv) {
    for (int i = 0; i + 1 < argc; i += 2) {
        const Kernel k = motionBlurKernel(std::atoi(argv[i]), std::atof(argv[i + 1]));
        std::printf("%d", k.width());
        for (float v : k) std::printf(" %.9g", static_cast<double>(v));
        std::printf("\n");
    }
    return 0;
}

int run(const std::string& dir) {
    ImageU8 img(97, 61, 3);
    ImageF imgf(83, 59, 1);
    unsigned s = 12345u;
    for (auto& v : img) {
        s = s * 1664525u + 1013904223u;
        v = static_cast<std::uint8_t>(s >> 24);
    }
    for (auto& v : imgf) {
        s = s * 1664525u + 1013904223u;
        v = static_cast<float>(s >> 8) / 16777216.0f;
    }
    save(dir + "/in_u8.bin", img);
    save(dir + "/in_f32.bin", imgf);
    save(dir + "/o1.bin", gaussianBlur(img, 2.0));
    save(dir + "/o2.bin", gaussianBlur<float>(img, 1.5));
    save(dir + "/o3.bin", sepFilter2D(imgf, {1, 2, 3}, {4, 5, 6, 7, 8}, BorderMode::Constant, 0.5, 0.25));
    save(dir + "/o4.bin", boxFilter(img, 5, 3, true, BorderMode::Reflect));
    save(dir + "/o5.bin", blur(imgf, 15));
    save(dir + "/o6.bin", filter2D(img, makeKernel(3, 3, {0, -1, 0, -1, 5, -1, 0, -1, 0}), BorderMode::Replicate, 1.0));
    save(dir + "/o7.bin", medianBlur(img, 5));
    save(dir + "/o8.bin", medianBlur(img, 3, BorderMode::Reflect101));
    save(dir + "/o9.bin", bilateralFilter(img, 9, 75, 75));
    gpu::setPrecision(gpu::Precision::Exact);
    save(dir + "/o10.bin", gaussianBlur(imgf, 3.0));
    gpu::setPrecision(gpu::Precision::Default);
    ImageU8 gray(97, 61, 1);
    for (int y = 0; y < gray.height(); ++y)
        for (int x = 0; x < gray.width(); ++x) gray(x, y) = img(x, y, 1);
    save(dir + "/o11.bin", sobel(img, 1, 0));
    save(dir + "/o12.bin", scharr(imgf, 0, 1, 2.0, BorderMode::Replicate));
    save(dir + "/o13.bin", laplacian(img, 3));
    const ImageF gx = sobel(imgf, 1, 0), gy = sobel(imgf, 0, 1);
    save(dir + "/o14.bin", magnitude(gx, gy));
    save(dir + "/o15.bin", phase(gx, gy));
    save(dir + "/o16.bin", unsharpMask(img, 1.5, 0.8, 2.0));
    save(dir + "/o17.bin", sharpen(img, 0.5));
    save(dir + "/o18.bin", emboss(img));
    save(dir + "/o19.bin", adaptiveThreshold(gray, 255, AdaptiveMethod::Gaussian, ThresholdType::BinaryInv, 7, 3.0));
    save(dir + "/o20.bin", prewitt(img, 1, 0, BorderMode::Wrap));
    save(dir + "/o21.bin", gradientMagnitude(img));
    // pyramids (geometry.hpp:501-535), with the known answers of the
    // reference's test_geometry.cpp:117-129
    save(dir + "/o22.bin", pyrDown(img));
    save(dir + "/o23.bin", pyrUp(imgf));
    const std::vector<ImageU8> pyr = buildPyramid(img, 9);
    for (size_t i = 0; i < pyr.size(); ++i) save(dir + "/p" + std::to_string(i) + ".bin", pyr[i]);
    std::printf("LEVELS %zu\n", pyr.size());
    const ImageU8 c(13, 8, 3, 90);
    const ImageU8 dn = pyrDown(c), upc = pyrUp(c);
    if (!(dn.width() == 7 && dn.height() == 4 && dn == ImageU8(7, 4, 3, 90))) return 1;
    if (!(upc.width() == 26 && upc == ImageU8(26, 16, 3, 90))) return 1;
    // feature callers (features.hpp:26-153)
    save(dir + "/o24.bin", canny(img, 40.0, 120.0));
    save(dir + "/o25.bin", canny(imgf, 0.5, 1.5, 5, true));
    save(dir + "/o26.bin", cannyAuto(img));
    save(dir + "/o27.bin", cornerHarris(gray, 3, 3, 0.04));
    save(dir + "/o28.bin", cornerMinEigenVal(imgf, 5, 5));
    save(dir + "/o29.bin", filter2DFFT(img, makeKernel(3, 3, {0, -1, 0, -1, 5, -1, 0, -1, 0}), BorderMode::Replicate, 1.0));
    {
        // houghCircles on a drawn disc (radius 14 at (48, 30)) with a little noise
        ImageU8 disc(97, 61, 1, 40);
        for (int y = 0; y < 61; ++y)
            for (int x = 0; x < 97; ++x) {
                const int dx = x - 48, dy = y - 30;
                disc(x, y) = static_cast<std::uint8_t>((dx * dx + dy * dy <= 14 * 14 ? 200 : 40) + (img(x, y, 0) & 7));
            }
        save(dir + "/in_disc.bin", disc);
        const std::vector<HoughCircle> cs = houghCircles(disc, 8, 24);
        std::vector<double> flat;
        for (const HoughCircle& c : cs) flat.insert(flat.end(), {c.center.x, c.center.y, c.radius, double(c.votes)});
        std::ofstream f(dir + "/o33.bin", std::ios::binary);
        f.write(reinterpret_cast<const char*>(flat.data()), static_cast<std::streamsize>(flat.size() * sizeof(double)));
        std::printf("CIRCLES %zu\n", cs.size());
    }
    {
        // decoding on the device: the files the encoders write come back unchanged
        const ImageU8 p6 = decodePNM(encodePNM(img));
        const ImageU16 p5 = decodePNM<std::uint16_t>(encodePNM(gray, false));
        const ImageF pf = decodePFM(encodePFM(imgf));
        if (!(p6 == img) || !(pf == imgf)) return 1;
        for (std::size_t i = 0; i < gray.sampleCount(); ++i)
            if (p5.data()[i] != gray.data()[i] * 257u) return 1;
    }
    save(dir + "/o31.bin", integralImage(img));
    save(dir + "/o32.bin", integralImage(imgf, true));
    {
        const ImageD t = integralImage(img);
        double direct = 0;
        for (int y = 3; y < 3 + 20; ++y)
            for (int x = 5; x < 5 + 30; ++x) direct += img(x, y, 1);
        if (integralSum(t, Rect(5, 3, 30, 20), 1) != direct) return 1;
    }
    save(dir + "/o30.bin", filter2DFFT<double>(imgf, gaussianKernel2D(2.0), BorderMode::Constant, 0.5, 0.25));
    std::printf("RUN OK\n");
    return 0;
}
}  // namespace

// Host-only codec paths: PNM / PFM encoders and .npy files (no device work).
int codec(const std::string& dir) {
    ImageU8 rgb(13, 7, 3);
    ImageU16 g16(9, 4, 1);
    ImageF f(6, 5, 3);
    for (std::size_t i = 0; i < rgb.sampleCount(); ++i) rgb.data()[i] = static_cast<std::uint8_t>(i * 37 + 5);
    for (std::size_t i = 0; i < g16.sampleCount(); ++i) g16.data()[i] = static_cast<std::uint16_t>(i * 2741 + 11);
    for (std::size_t i = 0; i < f.sampleCount(); ++i) f.data()[i] = static_cast<float>(i) * 0.37f - 3.0f;
    writeFile(dir + "/rgb_bin.ppm", encodePNM(rgb));
    writeFile(dir + "/rgb_txt.ppm", encodePNM(rgb, false));
    writeFile(dir + "/g16_bin.pgm", encodePNM(g16));
    writeFile(dir + "/g16_txt.pgm", encodePNM(g16, false));
    writeFile(dir + "/f.pfm", encodePFM(f));
    npy::save(rgb, dir + "/rgb.npy");
    npy::save(g16, dir + "/g16.npy");
    npy::save(f, dir + "/f.npy");
    const ImageF back = npy::load<float>(dir + "/f.npy");
    const ImageU8 rgb2 = npy::load<std::uint8_t>(dir + "/rgb.npy");
    if (!(back == f) || !(rgb2 == rgb)) return 1;
    const ImageD wide = npy::load<double>(dir + "/in.npy");  // written by numpy
    save(dir + "/npy_in.bin", wide);
    std::printf("CODEC OK\n");
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 3 && std::strcmp(argv[1], "codec") == 0) return codec(argv[2]);
    if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) return errors();
    if (argc >= 3 && std::strcmp(argv[1], "run") == 0) return run(argv[2]);
    if (argc >= 2 && std::strcmp(argv[1], "motion") == 0) return motion(argc - 2, argv + 2);
    std::printf("usage: frontend_check errors | run <dir> | motion (length angle)...\n");
    return 2;
}
