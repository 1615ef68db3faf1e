/*
 * lx_output.cpp -- the output path: saveImage (src/app/save_image.cpp:32-84)
 * and Image::savePNG (src/image.cpp:23-43).
 *
 * Per pixel: radiance sum / spp, host ACES (save_image.cpp:32-39), pow 0.45,
 * clamp to [0, 1] * 255 truncated to a byte, column x written to
 * width - 1 - x (the reference's horizontal mirror).  The PNG is 8-bit RGB,
 * zlib-compressed, no filtering.
 */
#include <zlib.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "lx_internal.h"

namespace {

void put32(std::vector<uint8_t>& v, uint32_t x) {
    v.push_back((uint8_t)(x >> 24));
    v.push_back((uint8_t)(x >> 16));
    v.push_back((uint8_t)(x >> 8));
    v.push_back((uint8_t)x);
}

void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    uint32_t crc = (uint32_t)crc32(0L, out.data() + start, (uInt)(out.size() - start));
    put32(out, crc);
}

float aces(float x) {
    const float a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    float v = (x * (a * x + b)) / (x * (c * x + d) + e);
    return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
}

} // namespace

void lx_write_png_rgb8(const char* path, uint32_t w, uint32_t h, const uint8_t* rgb) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (1 + 3 * (size_t)w));
    for (uint32_t y = 0; y < h; y++) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb + (size_t)y * 3 * w, rgb + (size_t)(y + 1) * 3 * w);
    }
    uLongf zn = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, raw.data(), (uLong)raw.size(), 6) != Z_OK)
        throw LxError{LX_ERR_OOM, "PNG: compression failed"};
    z.resize(zn);
    std::vector<uint8_t> png = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE* f = std::fopen(path, "wb");
    if (!f) throw LxError{LX_ERR_IO, std::string("cannot write ") + path};
    size_t n = std::fwrite(png.data(), 1, png.size(), f);
    std::fclose(f);
    if (n != png.size()) throw LxError{LX_ERR_IO, std::string("short write to ") + path};
}

/* accum: w*h*3 radiance sums, row-major (index x + y*w) */
void lx_tonemap_host(const float* accum, uint32_t w, uint32_t h, uint32_t spp, uint8_t* rgb) {
    float samples = (float)spp;
    for (uint32_t x = 0; x < w; x++)
        for (uint32_t y = 0; y < h; y++) {
            size_t i = (size_t)x + (size_t)y * w;
            size_t o = (size_t)(w - 1 - x) + (size_t)y * w;
            for (int c = 0; c < 3; c++) {
                float p = accum[3 * i + c] / samples;
                p = std::pow(aces(p), 0.45f);
                p = (p < 0.0f ? 0.0f : (p > 1.0f ? 1.0f : p)) * 255.f;
                rgb[3 * o + c] = (uint8_t)p;
            }
        }
}

extern "C" lx_status lx_render_save_png(lx_render* r, const char* path) {
    LX_TRY
    LX_REQUIRE(r && path, LX_ERR_INVALID, "null argument");
    lx_status st = lx_render_sync(r);
    if (st != LX_OK) return st;
    LX_REQUIRE(r->spp > 0, LX_ERR_STATE, "no samples accumulated");
    uint32_t w = (uint32_t)r->cam.width, h = (uint32_t)r->cam.height;
    std::vector<float> acc((size_t)w * h * 3);
    uint32_t spp = 0;
    st = lx_render_read_accum(r, acc.data(), &spp);
    if (st != LX_OK) return st;
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    lx_tonemap_host(acc.data(), w, h, spp, rgb.data());
    lx_write_png_rgb8(path, w, h, rgb.data());
    LX_CATCH
}

extern "C" lx_status lx_debug_write_png(const char* path, const float* accum, uint32_t w, uint32_t h, uint32_t spp) {
    LX_TRY
    LX_REQUIRE(path && accum && w && h && spp, LX_ERR_INVALID, "bad argument");
    std::vector<uint8_t> rgb((size_t)w * h * 3);
    lx_tonemap_host(accum, w, h, spp, rgb.data());
    lx_write_png_rgb8(path, w, h, rgb.data());
    LX_CATCH
}
