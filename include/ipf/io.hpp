// ipf front end for MI355X: the Netpbm (PBM/PGM/PPM, P1-P6) and PFM codecs of
// the reference's include/ipf/codec/pnm.hpp:58-197 and the file helpers of
// codec/common.hpp:196-210. The other codecs of the reference (PNG, JPEG, BMP,
// TGA, ICC handling) are not part of this build.
//
// Decoding parses the header on the host and, for binary rasters, unpacks the
// samples on the GPU through the C ABI (ipfg_decode_pnm); results and error
// messages are the reference's. Encoding formats the bytes on the host, as the
// reference does. Link with -lipfg.
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <type_traits>
#include <vector>

#include "core.hpp"
#include "filter.hpp"

namespace ipf {

/// Reads a whole file.
inline std::vector<std::uint8_t> readFile(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) detail::fail("cannot open file for reading: " + path);
    return std::vector<std::uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

/// Writes a whole file.
inline void writeFile(const std::string& path, const std::vector<std::uint8_t>& bytes) {
    std::ofstream f(path, std::ios::binary);
    if (!f) detail::fail("cannot open file for writing: " + path);
    f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
    if (!f) detail::fail("failed writing file: " + path);
}

namespace detail {
template <class T>
Image<T> decodeNetpbm(const std::vector<std::uint8_t>& bytes) {
    int w = 0, h = 0, cn = 0, pfm = 0;
    check(ipfg_pnm_info(bytes.data(), bytes.size(), &w, &h, &cn, &pfm));
    Image<T> img(w, h, cn);
    const ipfg_image d = describe(img);
    check(ipfg_decode_pnm(bytes.data(), bytes.size(), &d, gpu::streamSetting()));
    return img;
}
}  // namespace detail

/// Decodes PBM/PGM/PPM (P1-P6). Returns 1 channel for P1/P2/P4/P5, 3 for P3/P6.
/// Samples are rescaled from the file's maxval to the full range of T
/// (use std::uint8_t or std::uint16_t).
template <class T = std::uint8_t>
Image<T> decodePNM(const std::vector<std::uint8_t>& bytes) {
    static_assert(std::is_same<T, std::uint8_t>::value || std::is_same<T, std::uint16_t>::value,
                  "decodePNM: T must be std::uint8_t or std::uint16_t in this build");
    if (bytes.size() < 3 || bytes[0] != 'P' || bytes[1] < '1' || bytes[1] > '6')
        detail::fail("PNM: not a P1-P6 netpbm file");
    return detail::decodeNetpbm<T>(bytes);
}

/// Decodes a PFM (Portable Float Map, "PF" = RGB, "Pf" = gray) into a float image.
/// Values are returned as stored (not normalized).
inline ImageF decodePFM(const std::vector<std::uint8_t>& bytes) {
    if (bytes.size() < 3 || bytes[0] != 'P' || (bytes[1] != 'F' && bytes[1] != 'f')) detail::fail("PFM: not a PFM file");
    return detail::decodeNetpbm<float>(bytes);
}

/// Encodes a 1-channel image as PGM (P5/P2) or a 3-channel image as PPM (P6/P3).
/// 2/4-channel images have their alpha dropped. uint16 images are written with
/// maxval 65535, uint8 images with maxval 255.
template <class T>
std::vector<std::uint8_t> encodePNM(const Image<T>& src, bool binary = true) {
    static_assert(std::is_same<T, std::uint8_t>::value || std::is_same<T, std::uint16_t>::value,
                  "encodePNM: image must be uint8 or uint16");
    detail::require(!src.empty(), "encodePNM: empty image");
    const int sc = src.channels(), oc = sc <= 2 ? 1 : 3;  // gray keeps channel 0, colour the first three
    const unsigned maxval = sizeof(T) == 1 ? 255u : 65535u;
    const std::string head = "P" + std::to_string(oc == 1 ? (binary ? 5 : 2) : (binary ? 6 : 3)) + "\n" +
                             std::to_string(src.width()) + " " + std::to_string(src.height()) + "\n" +
                             std::to_string(maxval) + "\n";
    std::vector<std::uint8_t> out(head.begin(), head.end());
    const std::size_t perRow = static_cast<std::size_t>(src.width()) * oc;
    std::size_t k = 0;
    for (std::size_t p = 0; p < src.pixelCount(); ++p)
        for (int c = 0; c < oc; ++c) {
            const unsigned v = src.data()[p * sc + c];
            if (binary) {
                if (sizeof(T) == 2) out.push_back(static_cast<std::uint8_t>(v >> 8));
                out.push_back(static_cast<std::uint8_t>(v & 0xFF));
            } else {
                const std::string s = std::to_string(v);
                out.insert(out.end(), s.begin(), s.end());
                out.push_back(++k % perRow == 0 ? '\n' : ' ');
            }
        }
    return out;
}

/// Encodes a 1- or 3-channel float image as little-endian PFM.
template <class T>
std::vector<std::uint8_t> encodePFM(const Image<T>& src) {
    detail::require(src.channels() == 1 || src.channels() == 3, "encodePFM: image must have 1 or 3 channels");
    detail::require(!src.empty(), "encodePFM: empty image");
    const std::string head = std::string(src.channels() == 3 ? "PF" : "Pf") + "\n" + std::to_string(src.width()) +
                             " " + std::to_string(src.height()) + "\n-1.0\n";
    std::vector<std::uint8_t> out(head.begin(), head.end());
    out.reserve(out.size() + src.sampleCount() * 4);
    const std::size_t row = static_cast<std::size_t>(src.width()) * src.channels();
    for (int y = src.height() - 1; y >= 0; --y)  // bottom row first
        for (std::size_t i = 0; i < row; ++i) {
            const float f = static_cast<float>(src.row(y)[i]);
            std::uint32_t u;
            std::memcpy(&u, &f, 4);
            for (int b = 0; b < 4; ++b) out.push_back(static_cast<std::uint8_t>(u >> (8 * b)));
        }
    return out;
}

}  // namespace ipf
