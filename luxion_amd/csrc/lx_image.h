/*
 * lx_image.h -- decoders used by the glTF loader (PNG, baseline JPEG, EXR).
 */
#ifndef LX_IMAGE_H
#define LX_IMAGE_H

#include <string>
#include <vector>

#include "lx_internal.h"

/* RGBA texels, 8 or 16 bits per channel (16-bit little-endian in memory). */
struct LxImage {
    uint32_t w = 0, h = 0;
    int bits = 8;
    std::vector<uint8_t> data;
};

/* RGBA32F texels. */
struct LxImageF {
    uint32_t w = 0, h = 0;
    std::vector<float> rgba;
};

LxImage lx_decode_png(const uint8_t* p, size_t n);
LxImage lx_decode_jpeg(const uint8_t* p, size_t n);
LxImage lx_decode_image(const uint8_t* p, size_t n); /* by signature */
LxImageF lx_load_exr(const std::string& path);

#endif
