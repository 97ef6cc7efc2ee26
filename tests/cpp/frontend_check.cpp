// Drop-in check of the C++ front end: code written against the reference's
// ipf API (include/ipf/filter.hpp signatures and defaults) compiled unchanged
// against this repository's headers and libipfg.so.
//
//   frontend_check errors        argument checks only (no GPU needed)
//   frontend_check run <dir>     writes inputs and outputs as raw files
#include <ipf/ipf.hpp>

#include <cstdio>
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
    // empty images come back empty without device work; ksize 1 is a copy
    if (!gaussianBlur(ImageU8(0, 5, 3), 2.0).empty()) ++bad;
    if (!(medianBlur(img, 1) == img)) ++bad;
    // core semantics used by the path
    if (saturate_cast<std::uint8_t>(127.5) != 128 || saturate_cast<std::int16_t>(-2.5) != -3) ++bad;
    if (borderIndex(-1, 4, BorderMode::Reflect101) != 1 || borderIndex(4, 4, BorderMode::Wrap) != 0) ++bad;
    std::printf(bad ? "ERRORS FAILED %d\n" : "ERRORS OK\n", bad);
    return bad ? 1 : 0;
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
    std::printf("RUN OK\n");
    return 0;
}
}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "errors") == 0) return errors();
    if (argc >= 3 && std::strcmp(argv[1], "run") == 0) return run(argv[2]);
    std::printf("usage: frontend_check errors | run <dir>\n");
    return 2;
}
