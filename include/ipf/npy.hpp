// ipf front end for MI355X: .npy arrays for fixtures and test data (the
// reference keeps this in tests/crossval/npy.hpp): an image is saved as a
// (h, w) or (h, w, c) little-endian array of its element type; loading
// accepts 1-D (one row), 2-D and 3-D arrays of u1, u2, i2, i4, f4, f8 (and b1)
// and converts the elements to T. Host code only.
#pragma once

#include <cstdint>
#include <cstring>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "io.hpp"

namespace npy {

template <class T>
const char* descr() {
    if (std::is_same<T, std::uint8_t>::value) return "|u1";
    if (std::is_same<T, std::uint16_t>::value) return "<u2";
    if (std::is_same<T, std::int16_t>::value) return "<i2";
    if (std::is_same<T, std::int32_t>::value) return "<i4";
    if (std::is_same<T, float>::value) return "<f4";
    if (std::is_same<T, double>::value) return "<f8";
    return "??";
}

/// Writes an image as a (h, w, c) array (c axis dropped for single-channel images).
template <class T>
void save(const ipf::Image<T>& img, const std::string& path) {
    std::string dict = std::string("{'descr': '") + descr<T>() + "', 'fortran_order': False, 'shape': (" +
                       std::to_string(img.height()) + ", " + std::to_string(img.width()) +
                       (img.channels() > 1 ? ", " + std::to_string(img.channels()) : std::string()) + "), }";
    while ((10 + dict.size() + 1) % 64) dict += ' ';  // header padded so the data starts 64-aligned
    dict += '\n';
    std::vector<std::uint8_t> out = {0x93, 'N', 'U', 'M', 'P', 'Y', 1, 0, static_cast<std::uint8_t>(dict.size() & 0xFF),
                                     static_cast<std::uint8_t>(dict.size() >> 8)};
    out.insert(out.end(), dict.begin(), dict.end());
    const auto* p = reinterpret_cast<const std::uint8_t*>(img.data());
    out.insert(out.end(), p, p + img.sampleCount() * sizeof(T));
    ipf::writeFile(path, out);
}

struct Array {
    std::string descr;
    std::vector<long long> shape;
    std::vector<std::uint8_t> data;
};

inline Array loadRaw(const std::string& path) {
    const std::vector<std::uint8_t> b = ipf::readFile(path);
    if (b.size() < 10 || b[0] != 0x93 || b[1] != 'N') ipf::detail::fail("npy: not a .npy file: " + path);
    const bool v1 = b[6] == 1;
    const std::size_t off = v1 ? 10 : 12;
    if (b.size() < off) ipf::detail::fail("npy: not a .npy file: " + path);
    std::size_t hlen = b[8] | (b[9] << 8);
    if (!v1) hlen |= (static_cast<std::size_t>(b[10]) << 16) | (static_cast<std::size_t>(b[11]) << 24);
    if (b.size() < off + hlen) ipf::detail::fail("npy: truncated data");
    const std::string head(b.begin() + static_cast<std::ptrdiff_t>(off), b.begin() + static_cast<std::ptrdiff_t>(off + hlen));
    Array a;
    const std::size_t d = head.find("'descr'");
    const std::size_t q1 = head.find('\'', head.find(':', d) + 1);
    a.descr = head.substr(q1 + 1, head.find('\'', q1 + 1) - q1 - 1);
    if (head.find("'fortran_order': True") != std::string::npos) ipf::detail::fail("npy: fortran order unsupported");
    const std::size_t s0 = head.find('(', head.find("'shape'")), s1 = head.find(')', s0);
    std::stringstream dims(head.substr(s0 + 1, s1 - s0 - 1));
    std::string tok;
    while (std::getline(dims, tok, ','))
        if (tok.find_first_not_of(' ') != std::string::npos) a.shape.push_back(std::stoll(tok));
    a.data.assign(b.begin() + static_cast<std::ptrdiff_t>(off + hlen), b.end());
    return a;
}

/// Loads a (h, w) or (h, w, c) array, converting the element type to T.
template <class T>
ipf::Image<T> load(const std::string& path) {
    const Array a = loadRaw(path);
    if (a.shape.empty() || a.shape.size() > 3) ipf::detail::fail("npy: expected a 1D, 2D or 3D array");
    const bool row = a.shape.size() == 1;
    const int h = row ? 1 : static_cast<int>(a.shape[0]), w = static_cast<int>(a.shape[row ? 0 : 1]);
    const int c = a.shape.size() == 3 ? static_cast<int>(a.shape[2]) : 1;
    ipf::Image<T> img(w, h, c);
    const std::size_t n = img.sampleCount();
    auto take = [&](auto tag) {
        using S = decltype(tag);
        if (a.data.size() < n * sizeof(S)) ipf::detail::fail("npy: truncated data");
        for (std::size_t i = 0; i < n; ++i) {
            S v;
            std::memcpy(&v, a.data.data() + i * sizeof(S), sizeof(S));
            img.data()[i] = static_cast<T>(v);
        }
    };
    if (a.descr == "|u1" || a.descr == "|b1") take(std::uint8_t{});
    else if (a.descr == "<u2") take(std::uint16_t{});
    else if (a.descr == "<i2") take(std::int16_t{});
    else if (a.descr == "<i4") take(std::int32_t{});
    else if (a.descr == "<f4") take(float{});
    else if (a.descr == "<f8") take(double{});
    else ipf::detail::fail("npy: unsupported dtype " + a.descr);
    return img;
}

}  // namespace npy
