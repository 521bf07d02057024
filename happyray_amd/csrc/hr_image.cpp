// hr_image.cpp -- framebuffer to 8-bit RGB and PNG files (host code).
//
// Reference behaviour followed:
//   Image::gammaCorrect              src/Utils/ImagePNG.cpp:64-72
//   Image::writePNG (libpng branch)  src/Utils/ImagePNG.cpp:268-330
//   SDLGLApplication::writeScreenShot src/Application/SDLGLApplication.cpp:120-143
//
// The reference writes the PNG with libpng; here the file is assembled
// directly (signature, IHDR, one zlib-compressed IDAT of filter-0 scanlines,
// IEND) with zlib for deflate and CRC-32.  The pixel bytes are the same:
// v^(1/gamma) with powf, clamped to [0, 1], times 255, truncated.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include <zlib.h>

#include "happyray.h"
#include "hr_error.h"

namespace {

using hr::set_error;

inline uint8_t to_byte(float v)
{
    // max(min(v, 1), 0) with the host min/max of the reference
    // (src/Utils/CUDAUtil.h:170-187); NaN, which those pass through and the
    // byte cast leaves undefined, becomes 0
    v = (1.f < v) ? 1.f : v;
    v = (v < 0.f) ? 0.f : v;
    if (!(v == v)) return 0;
    return (uint8_t)(v * 255);
}

void put_be32(std::vector<uint8_t>& o, uint32_t v)
{
    o.push_back((uint8_t)(v >> 24));
    o.push_back((uint8_t)(v >> 16));
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

void put_chunk(std::vector<uint8_t>& o, const char type[4], const uint8_t* data, size_t n)
{
    put_be32(o, (uint32_t)n);
    const size_t at = o.size();
    o.insert(o.end(), type, type + 4);
    if (n) o.insert(o.end(), data, data + n);
    const uLong crc = crc32(crc32(0L, Z_NULL, 0), o.data() + at, (uInt)(n + 4));
    put_be32(o, (uint32_t)crc);
}

}  // namespace

extern "C" {

hr_status hr_image_to_rgb8(const float* rgb, int width, int height, float gamma, uint8_t* out)
{
    if (!rgb || !out) return set_error(HR_ERR_ARG, "null image buffer");
    if (width <= 0 || height <= 0) return set_error(HR_ERR_ARG, "invalid image size");
    const size_t n = 3 * (size_t)width * (size_t)height;
    const float e = gamma > 0.f ? 1.0f / gamma : 1.f;
    for (size_t i = 0; i < n; ++i) {
        const float v = gamma > 0.f ? powf(rgb[i], e) : rgb[i];
        out[i] = to_byte(v);
    }
    return HR_OK;
}

hr_status hr_write_png(const char* path, const uint8_t* rgb8, int width, int height)
{
    if (!path || !rgb8) return set_error(HR_ERR_ARG, "null argument");
    if (width <= 0 || height <= 0 || (int64_t)width * height > (1ll << 30))
        return set_error(HR_ERR_ARG, "invalid image size");
    // scanlines, each led by filter type 0 (none)
    const size_t row = 3 * (size_t)width;
    std::vector<uint8_t> raw((row + 1) * (size_t)height);
    for (int y = 0; y < height; ++y) {
        raw[(row + 1) * (size_t)y] = 0;
        std::memcpy(&raw[(row + 1) * (size_t)y + 1], rgb8 + row * (size_t)y, row);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        return set_error(HR_ERR_IO, "deflate failed");
    std::vector<uint8_t> png = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n' };
    std::vector<uint8_t> ihdr;
    put_be32(ihdr, (uint32_t)width);
    put_be32(ihdr, (uint32_t)height);
    ihdr.push_back(8);   // bit depth
    ihdr.push_back(2);   // colour type RGB (PNG_COLOR_TYPE_RGB)
    ihdr.push_back(0);   // compression (PNG_COMPRESSION_TYPE_BASE)
    ihdr.push_back(0);   // filter method (PNG_FILTER_TYPE_BASE)
    ihdr.push_back(0);   // interlace (PNG_INTERLACE_NONE)
    put_chunk(png, "IHDR", ihdr.data(), ihdr.size());
    put_chunk(png, "IDAT", z.data(), zn);
    put_chunk(png, "IEND", nullptr, 0);
    FILE* f = std::fopen(path, "wb");
    if (!f) return set_error(HR_ERR_IO, std::string("cannot open ") + path + " for writing");
    const size_t w = std::fwrite(png.data(), 1, png.size(), f);
    const int c = std::fclose(f);
    if (w != png.size() || c != 0) return set_error(HR_ERR_IO, std::string("cannot write ") + path);
    return HR_OK;
}

hr_status hr_write_screenshot(const char* path, const float* rgb, int width, int height)
{
    if (!rgb) return set_error(HR_ERR_ARG, "null image buffer");
    if (width <= 0 || height <= 0) return set_error(HR_ERR_ARG, "invalid image size");
    std::vector<uint8_t> b(3 * (size_t)width * (size_t)height);
    if (hr_status st = hr_image_to_rgb8(rgb, width, height, 2.2f, b.data())) return st;
    return hr_write_png(path, b.data(), width, height);
}

}  // extern "C"
