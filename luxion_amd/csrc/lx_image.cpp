/*
 * lx_image.cpp -- texture and environment-map decoders for the glTF loader.
 *
 * PNG (all colour types, bit depths 1-16, Adam7) and baseline JPEG decode to
 * RGBA8, or RGBA16 for 16-bit PNG, as the reference's image path delivers
 * them (tinygltf asks stb_image for four channels, 16 bits where the PNG has
 * them: include/tinygltf/tiny_gltf.h:2617-2715).  OpenEXR scanline files
 * (uncompressed, RLE, ZIPS, ZIP, PIZ; HALF, FLOAT or UINT channels) decode to
 * RGBA32F with tinyexr's LoadEXR channel rules (include/tinyexr/tinyexr.h:
 * 6263-6460): one channel fills R, G, B and A; otherwise R, G and B are
 * required and A defaults to 1.
 */
#include "lx_image.h"

#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

namespace {

[[noreturn]] void bad(const std::string& what) { throw LxError{LX_ERR_FORMAT, what}; }

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

void inflate_all(const uint8_t* src, size_t n, std::vector<uint8_t>& out, size_t expect, const char* what) {
    out.resize(expect);
    z_stream z;
    std::memset(&z, 0, sizeof(z));
    if (inflateInit(&z) != Z_OK) bad(std::string(what) + ": zlib init failed");
    z.next_in = const_cast<Bytef*>(src);
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)expect;
    int r = inflate(&z, Z_FINISH);
    size_t got = expect - z.avail_out;
    inflateEnd(&z);
    if (r != Z_STREAM_END && !(r == Z_BUF_ERROR && got == expect)) bad(std::string(what) + ": corrupt deflate stream");
    if (got != expect) bad(std::string(what) + ": decompressed size mismatch");
}

/* ------------------------------------------------------------------ PNG */
uint8_t paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

void unfilter(uint8_t* rows, uint32_t h, size_t stride, size_t bpp) {
    /* rows: h * (1 + stride) bytes, filter byte first; in place, the filter
     * bytes are left in place and skipped */
    uint8_t* prev = nullptr;
    for (uint32_t y = 0; y < h; y++) {
        uint8_t* r = rows + y * (stride + 1);
        uint8_t f = r[0];
        uint8_t* c = r + 1;
        for (size_t i = 0; i < stride; i++) {
            int a = i >= bpp ? c[i - bpp] : 0;
            int b = prev ? prev[i] : 0;
            int cc = (prev && i >= bpp) ? prev[i - bpp] : 0;
            switch (f) {
            case 0: break;
            case 1: c[i] = (uint8_t)(c[i] + a); break;
            case 2: c[i] = (uint8_t)(c[i] + b); break;
            case 3: c[i] = (uint8_t)(c[i] + ((a + b) >> 1)); break;
            case 4: c[i] = (uint8_t)(c[i] + paeth(a, b, cc)); break;
            default: bad("PNG: bad filter type");
            }
        }
        prev = c;
    }
}

} // namespace

LxImage lx_decode_png(const uint8_t* p, size_t n) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (n < 8 || std::memcmp(p, sig, 8) != 0) bad("PNG: bad signature");
    size_t i = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool seen_end = false;
    while (i + 12 <= n) {
        uint32_t len = be32(p + i);
        const uint8_t* type = p + i + 4;
        const uint8_t* d = p + i + 8;
        if (i + 12 + (size_t)len > n) bad("PNG: truncated chunk");
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) bad("PNG: short IHDR");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (d[10] != 0 || d[11] != 0) bad("PNG: unknown compression or filter method");
            interlace = d[12];
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            seen_end = true;
            break;
        }
        i += 12 + (size_t)len;
    }
    if (!seen_end && idat.empty()) bad("PNG: no image data");
    if (w == 0 || h == 0 || w > (1u << 16) || h > (1u << 16)) bad("PNG: bad dimensions");
    int chans;
    switch (ctype) {
    case 0: chans = 1; break;
    case 2: chans = 3; break;
    case 3: chans = 1; break;
    case 4: chans = 2; break;
    case 6: chans = 4; break;
    default: bad("PNG: bad colour type");
    }
    bool ok_depth = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                    (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                    ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
    if (!ok_depth) bad("PNG: bad bit depth");
    if (ctype == 3 && plte.empty()) bad("PNG: palette image without PLTE");
    bool wide = depth == 16;
    LxImage img;
    img.w = w;
    img.h = h;
    img.bits = wide ? 16 : 8;
    img.data.assign((size_t)w * h * 4 * (wide ? 2 : 1), 0);
    size_t bpp = std::max<size_t>(1, (size_t)chans * depth / 8);

    /* passes: Adam7 or one full pass */
    struct Pass {
        uint32_t x0, y0, dx, dy;
    };
    const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                           {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const Pass full[1] = {{0, 0, 1, 1}};
    const Pass* passes = interlace ? adam7 : full;
    int npass = interlace ? 7 : 1;
    size_t total = 0;
    for (int k = 0; k < npass; k++) {
        const Pass& ps = passes[k];
        uint32_t pw = w > ps.x0 ? (w - ps.x0 + ps.dx - 1) / ps.dx : 0;
        uint32_t ph = h > ps.y0 ? (h - ps.y0 + ps.dy - 1) / ps.dy : 0;
        if (pw && ph) total += (size_t)ph * (1 + ((size_t)pw * chans * depth + 7) / 8);
    }
    std::vector<uint8_t> raw;
    inflate_all(idat.data(), idat.size(), raw, total, "PNG");

    /* colour key (tRNS for grey / RGB) */
    int key[3] = {-1, -1, -1};
    if (!trns.empty() && ctype == 0 && trns.size() >= 2) key[0] = (trns[0] << 8) | trns[1];
    if (!trns.empty() && ctype == 2 && trns.size() >= 6)
        for (int c = 0; c < 3; c++) key[c] = (trns[2 * c] << 8) | trns[2 * c + 1];

    size_t off = 0;
    for (int k = 0; k < npass; k++) {
        const Pass& ps = passes[k];
        uint32_t pw = w > ps.x0 ? (w - ps.x0 + ps.dx - 1) / ps.dx : 0;
        uint32_t ph = h > ps.y0 ? (h - ps.y0 + ps.dy - 1) / ps.dy : 0;
        if (!pw || !ph) continue;
        size_t stride = ((size_t)pw * chans * depth + 7) / 8;
        uint8_t* rows = raw.data() + off;
        unfilter(rows, ph, stride, bpp);
        off += (size_t)ph * (stride + 1);
        for (uint32_t y = 0; y < ph; y++) {
            const uint8_t* r = rows + (size_t)y * (stride + 1) + 1;
            for (uint32_t x = 0; x < pw; x++) {
                uint32_t v[4] = {0, 0, 0, 0};
                if (depth < 8) {
                    size_t bit = (size_t)x * depth;
                    uint32_t s = (r[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1u);
                    v[0] = s;
                } else if (depth == 8) {
                    for (int c = 0; c < chans; c++) v[c] = r[(size_t)x * chans + c];
                } else {
                    for (int c = 0; c < chans; c++)
                        v[c] = (uint32_t)r[((size_t)x * chans + c) * 2] << 8 | r[((size_t)x * chans + c) * 2 + 1];
                }
                uint32_t rgba[4];
                uint32_t maxv = wide ? 65535u : 255u;
                if (ctype == 3) {
                    uint32_t ix = v[0];
                    if ((size_t)ix * 3 + 2 >= plte.size()) bad("PNG: palette index out of range");
                    rgba[0] = plte[3 * ix];
                    rgba[1] = plte[3 * ix + 1];
                    rgba[2] = plte[3 * ix + 2];
                    rgba[3] = ix < trns.size() ? trns[ix] : 255u;
                } else if (ctype == 0 || ctype == 4) {
                    uint32_t g = v[0];
                    bool keyed = key[0] >= 0 && (int)g == key[0];
                    if (depth < 8) g = g * (255u / ((1u << depth) - 1u));
                    rgba[0] = rgba[1] = rgba[2] = g;
                    rgba[3] = ctype == 4 ? v[1] : (keyed ? 0u : maxv);
                } else {
                    bool keyed = key[0] >= 0 && (int)v[0] == key[0] && (int)v[1] == key[1] && (int)v[2] == key[2];
                    rgba[0] = v[0];
                    rgba[1] = v[1];
                    rgba[2] = v[2];
                    rgba[3] = ctype == 6 ? v[3] : (keyed ? 0u : maxv);
                }
                size_t px = (size_t)(ps.y0 + y * ps.dy) * w + (ps.x0 + x * ps.dx);
                if (wide) {
                    uint16_t* o = (uint16_t*)img.data.data() + 4 * px;
                    for (int c = 0; c < 4; c++) o[c] = (uint16_t)rgba[c];
                } else {
                    uint8_t* o = img.data.data() + 4 * px;
                    for (int c = 0; c < 4; c++) o[c] = (uint8_t)rgba[c];
                }
            }
        }
    }
    return img;
}

/* ----------------------------------------------------------------- JPEG */
namespace {

const uint8_t kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                             41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                             30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huff {
    /* canonical code tables: for code length L, codes [mincode[L], maxcode[L]]
     * map to values starting at valptr[L] */
    int32_t mincode[17], maxcode[18], valptr[17];
    uint8_t vals[256];
    bool present = false;
    void build(const uint8_t* counts, const uint8_t* v, int nv) {
        std::memcpy(vals, v, (size_t)nv);
        int code = 0, k = 0;
        for (int l = 1; l <= 16; l++) {
            valptr[l] = k;
            mincode[l] = code;
            code += counts[l - 1];
            k += counts[l - 1];
            maxcode[l] = counts[l - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        present = true;
    }
};

struct BitReader {
    const uint8_t* p;
    size_t n, i = 0;
    uint32_t acc = 0;
    int bits = 0;
    bool marker = false; /* hit a marker: feed zeros */
    BitReader(const uint8_t* d, size_t len, size_t start) : p(d), n(len), i(start) {}
    void fill() {
        while (bits <= 24) {
            uint32_t b = 0;
            if (!marker && i < n) {
                b = p[i];
                if (b == 0xFF) {
                    uint8_t nx = i + 1 < n ? p[i + 1] : 0;
                    if (nx == 0x00) {
                        i += 2;
                    } else {
                        marker = true;
                        b = 0;
                    }
                } else {
                    i++;
                }
            }
            acc |= b << (24 - bits);
            bits += 8;
        }
    }
    int bit() {
        if (bits < 1) fill();
        int v = (int)(acc >> 31);
        acc <<= 1;
        bits--;
        return v;
    }
    int get(int k) {
        if (k == 0) return 0;
        if (bits < k) fill();
        int v = (int)(acc >> (32 - k));
        acc <<= k;
        bits -= k;
        return v;
    }
    int decode(const Huff& h) {
        int code = 0;
        for (int l = 1; l <= 16; l++) {
            code = (code << 1) | bit();
            if (h.maxcode[l] >= 0 && code <= h.maxcode[l] && code >= h.mincode[l])
                return h.vals[h.valptr[l] + code - h.mincode[l]];
        }
        bad("JPEG: bad Huffman code");
    }
    static int extend(int v, int t) { return (t && v < (1 << (t - 1))) ? v - (1 << t) + 1 : v; }
    void restart() {
        /* drop the partial byte and skip the RSTn marker */
        acc = 0;
        bits = 0;
        marker = false;
        while (i + 1 < n && !(p[i] == 0xFF && p[i + 1] >= 0xD0 && p[i + 1] <= 0xD7)) i++;
        if (i + 1 < n) i += 2;
    }
};

/* Separable 8x8 inverse DCT in 64-bit floating point, rounded to the
 * nearest level and clamped (the exact transform, so the result is the
 * ideal reconstruction of the coefficients). */
void idct8x8(const int32_t* in, uint8_t* out, int stride) {
    static double cosv[8][8];
    static bool init = false;
    if (!init) {
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++)
                cosv[x][u] = (u == 0 ? std::sqrt(0.125) : 0.5) * std::cos((2 * x + 1) * u * M_PI / 16.0);
        init = true;
    }
    double tmp[64];
    for (int y = 0; y < 8; y++)
        for (int u = 0; u < 8; u++) {
            double s = 0;
            for (int v = 0; v < 8; v++) s += cosv[y][v] * in[v * 8 + u];
            tmp[y * 8 + u] = s;
        }
    for (int y = 0; y < 8; y++)
        for (int x = 0; x < 8; x++) {
            double s = 0;
            for (int u = 0; u < 8; u++) s += cosv[x][u] * tmp[y * 8 + u];
            int v = (int)std::lround(s + 128.0);
            out[y * stride + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
        }
}

} // namespace

LxImage lx_decode_jpeg(const uint8_t* p, size_t n) {
    if (n < 4 || p[0] != 0xFF || p[1] != 0xD8) bad("JPEG: bad signature");
    uint16_t qt[4][64] = {};
    Huff hdc[4], hac[4];
    struct Comp {
        int id, h, v, tq, td = 0, ta = 0;
        int bw, bh; /* blocks across / down (padded to the MCU grid) */
        std::vector<uint8_t> pix;
        int dc = 0;
    };
    std::vector<Comp> comps;
    int W = 0, H = 0, restart = 0;
    bool frame = false, done = false;
    size_t i = 2;
    int hmax = 1, vmax = 1;
    while (i + 4 <= n && !done) {
        if (p[i] != 0xFF) bad("JPEG: expected a marker");
        uint8_t m = p[i + 1];
        if (m == 0xFF) {
            i++;
            continue;
        }
        if (m == 0xD8 || (m >= 0xD0 && m <= 0xD7) || m == 0x01) {
            i += 2;
            continue;
        }
        if (m == 0xD9) break;
        size_t len = ((size_t)p[i + 2] << 8) | p[i + 3];
        if (len < 2 || i + 2 + len > n) bad("JPEG: truncated segment");
        const uint8_t* d = p + i + 4;
        size_t dl = len - 2;
        if (m == 0xDB) {
            size_t k = 0;
            while (k < dl) {
                int pq = d[k] >> 4, tq = d[k] & 15;
                if (tq > 3) bad("JPEG: bad quantisation table id");
                k++;
                for (int j = 0; j < 64; j++) {
                    qt[tq][kZigzag[j]] = pq ? (uint16_t)(d[k + 2 * j] << 8 | d[k + 2 * j + 1]) : d[k + j];
                }
                k += pq ? 128 : 64;
            }
        } else if (m == 0xC4) {
            size_t k = 0;
            while (k < dl) {
                int tc = d[k] >> 4, th = d[k] & 15;
                if (th > 3 || tc > 1) bad("JPEG: bad Huffman table id");
                const uint8_t* counts = d + k + 1;
                int nv = 0;
                for (int j = 0; j < 16; j++) nv += counts[j];
                if (nv > 256 || k + 17 + nv > dl) bad("JPEG: bad Huffman table");
                (tc ? hac[th] : hdc[th]).build(counts, d + k + 17, nv);
                k += 17 + nv;
            }
        } else if (m == 0xC0 || m == 0xC1) {
            if (d[0] != 8) bad("JPEG: only 8-bit samples are supported");
            H = d[1] << 8 | d[2];
            W = d[3] << 8 | d[4];
            int nc = d[5];
            if (W <= 0 || H <= 0 || (nc != 1 && nc != 3)) bad("JPEG: unsupported frame");
            comps.resize(nc);
            for (int c = 0; c < nc; c++) {
                comps[c].id = d[6 + 3 * c];
                comps[c].h = d[7 + 3 * c] >> 4;
                comps[c].v = d[7 + 3 * c] & 15;
                comps[c].tq = d[8 + 3 * c] & 3;
                if (comps[c].h < 1 || comps[c].h > 2 || comps[c].v < 1 || comps[c].v > 2)
                    bad("JPEG: unsupported sampling factors");
                hmax = std::max(hmax, comps[c].h);
                vmax = std::max(vmax, comps[c].v);
            }
            frame = true;
        } else if (m >= 0xC2 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) {
            bad("JPEG: progressive, lossless and arithmetic-coded files are not supported");
        } else if (m == 0xDD) {
            restart = d[0] << 8 | d[1];
        } else if (m == 0xDA) {
            if (!frame) bad("JPEG: scan before frame");
            int ns = d[0];
            std::vector<int> sc(ns);
            for (int s = 0; s < ns; s++) {
                int cid = d[1 + 2 * s];
                int k = -1;
                for (size_t c = 0; c < comps.size(); c++)
                    if (comps[c].id == cid) k = (int)c;
                if (k < 0) bad("JPEG: scan names an unknown component");
                comps[k].td = d[2 + 2 * s] >> 4;
                comps[k].ta = d[2 + 2 * s] & 15;
                sc[s] = k;
            }
            int mcux = (W + 8 * hmax - 1) / (8 * hmax), mcuy = (H + 8 * vmax - 1) / (8 * vmax);
            for (auto& c : comps) {
                c.bw = mcux * c.h;
                c.bh = mcuy * c.v;
                c.pix.assign((size_t)c.bw * 8 * c.bh * 8, 0);
                c.dc = 0;
            }
            BitReader br(p, n, i + 2 + len);
            int32_t blk[64];
            auto decode_block = [&](Comp& c, int bx, int by) {
                const Huff& dh = hdc[c.td];
                const Huff& ah = hac[c.ta];
                if (!dh.present || !ah.present) bad("JPEG: missing Huffman table");
                std::memset(blk, 0, sizeof(blk));
                int t = br.decode(dh);
                int diff = BitReader::extend(br.get(t), t);
                c.dc += diff;
                blk[0] = c.dc * qt[c.tq][0];
                for (int k = 1; k < 64;) {
                    int rs = br.decode(ah);
                    int r = rs >> 4, s = rs & 15;
                    if (s == 0) {
                        if (r != 15) break;
                        k += 16;
                        continue;
                    }
                    k += r;
                    if (k > 63) bad("JPEG: coefficient index out of range");
                    int zz = kZigzag[k];
                    blk[zz] = BitReader::extend(br.get(s), s) * qt[c.tq][zz];
                    k++;
                }
                idct8x8(blk, c.pix.data() + (size_t)by * 8 * c.bw * 8 + (size_t)bx * 8, c.bw * 8);
            };
            int nmcu = 0;
            if (ns == 1) {
                /* non-interleaved: blocks covering the component's own size */
                Comp& c = comps[sc[0]];
                int cw = (W * c.h + hmax - 1) / hmax, ch = (H * c.v + vmax - 1) / vmax;
                int bxn = (cw + 7) / 8, byn = (ch + 7) / 8;
                for (int by = 0; by < byn; by++)
                    for (int bx = 0; bx < bxn; bx++) {
                        if (restart && nmcu && nmcu % restart == 0) {
                            br.restart();
                            for (auto& cc : comps) cc.dc = 0;
                        }
                        decode_block(c, bx, by);
                        nmcu++;
                    }
            } else {
                for (int my = 0; my < mcuy; my++)
                    for (int mx = 0; mx < mcux; mx++) {
                        if (restart && nmcu && nmcu % restart == 0) {
                            br.restart();
                            for (auto& cc : comps) cc.dc = 0;
                        }
                        for (int s : sc) {
                            Comp& c = comps[s];
                            for (int v = 0; v < c.v; v++)
                                for (int h = 0; h < c.h; h++) decode_block(c, mx * c.h + h, my * c.v + v);
                        }
                        nmcu++;
                    }
            }
            /* the entropy-coded segment ends at the next marker */
            size_t j = br.i;
            while (j + 1 < n && !(p[j] == 0xFF && p[j + 1] != 0x00 && !(p[j + 1] >= 0xD0 && p[j + 1] <= 0xD7))) j++;
            i = j;
            continue;
        }
        i += 2 + len;
    }
    if (!frame) bad("JPEG: no frame");
    LxImage img;
    img.w = (uint32_t)W;
    img.h = (uint32_t)H;
    img.bits = 8;
    img.data.assign((size_t)W * H * 4, 255);
    /* chroma upsampling: linear interpolation between sample centres */
    auto sample = [&](const Comp& c, int x, int y) -> float {
        int sw = c.bw * 8, sh = c.bh * 8;
        if (c.h == hmax && c.v == vmax) return c.pix[(size_t)y * sw + x];
        float fx = (x + 0.5f) * c.h / hmax - 0.5f, fy = (y + 0.5f) * c.v / vmax - 0.5f;
        int cwid = (W * c.h + hmax - 1) / hmax, chgt = (H * c.v + vmax - 1) / vmax;
        fx = std::min(std::max(fx, 0.0f), (float)(cwid - 1));
        fy = std::min(std::max(fy, 0.0f), (float)(chgt - 1));
        int x0 = (int)fx, y0 = (int)fy;
        int x1 = std::min(x0 + 1, cwid - 1), y1 = std::min(y0 + 1, chgt - 1);
        float ax = fx - x0, ay = fy - y0;
        (void)sh;
        float a = c.pix[(size_t)y0 * sw + x0] * (1 - ax) + c.pix[(size_t)y0 * sw + x1] * ax;
        float b = c.pix[(size_t)y1 * sw + x0] * (1 - ax) + c.pix[(size_t)y1 * sw + x1] * ax;
        return a * (1 - ay) + b * ay;
    };
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            uint8_t* o = img.data.data() + 4 * ((size_t)y * W + x);
            if (comps.size() == 1) {
                uint8_t g = comps[0].pix[(size_t)y * comps[0].bw * 8 + x];
                o[0] = o[1] = o[2] = g;
            } else {
                float Y = sample(comps[0], x, y), cb = sample(comps[1], x, y) - 128.0f,
                      cr = sample(comps[2], x, y) - 128.0f;
                float rgb[3] = {Y + 1.402f * cr, Y - 0.344136f * cb - 0.714136f * cr, Y + 1.772f * cb};
                for (int c = 0; c < 3; c++) {
                    int v = (int)std::lround(rgb[c]);
                    o[c] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
                }
            }
        }
    return img;
}

LxImage lx_decode_image(const uint8_t* p, size_t n) {
    if (n >= 8 && p[0] == 137 && p[1] == 'P' && p[2] == 'N' && p[3] == 'G') return lx_decode_png(p, n);
    if (n >= 3 && p[0] == 0xFF && p[1] == 0xD8) return lx_decode_jpeg(p, n);
    bad("image: not a PNG or JPEG stream");
}

/* ------------------------------------------------------------------ EXR */
namespace {

float half_to_float(uint16_t h) {
    uint32_t s = (uint32_t)(h >> 15) << 31, e = (h >> 10) & 31u, m = h & 1023u, f;
    if (e == 0) {
        if (m == 0) {
            f = s;
        } else {
            /* subnormal: normalise */
            int k = -1;
            do {
                k++;
                m <<= 1;
            } while (!(m & 1024u));
            f = s | ((uint32_t)(127 - 15 - k) << 23) | ((m & 1023u) << 13);
        }
    } else if (e == 31) {
        f = s | 0x7f800000u | (m << 13);
    } else {
        f = s | ((e + 112u) << 23) | (m << 13);
    }
    float r;
    std::memcpy(&r, &f, 4);
    return r;
}

uint32_t le32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }
uint64_t le64(const uint8_t* p) { return (uint64_t)le32(p) | (uint64_t)le32(p + 4) << 32; }

/* ZIP / ZIPS: zlib, then the byte predictor and the two-half interleave. */
void exr_unzip(const uint8_t* src, size_t n, std::vector<uint8_t>& out, size_t expect) {
    std::vector<uint8_t> t;
    inflate_all(src, n, t, expect, "EXR ZIP");
    for (size_t k = 1; k < t.size(); k++) t[k] = (uint8_t)(t[k - 1] + t[k] - 128);
    out.resize(expect);
    size_t half = (expect + 1) / 2;
    for (size_t k = 0; k < expect; k++) out[k] = (k & 1) ? t[half + k / 2] : t[k / 2];
}

void exr_unrle(const uint8_t* src, size_t n, std::vector<uint8_t>& out, size_t expect) {
    std::vector<uint8_t> t;
    t.reserve(expect);
    size_t i = 0;
    while (i < n) {
        int8_t c = (int8_t)src[i++];
        if (c < 0) {
            size_t cnt = (size_t)(-c);
            if (i + cnt > n) bad("EXR RLE: truncated literal run");
            t.insert(t.end(), src + i, src + i + cnt);
            i += cnt;
        } else {
            if (i >= n) bad("EXR RLE: truncated repeat run");
            t.insert(t.end(), (size_t)c + 1, src[i++]);
        }
    }
    if (t.size() != expect) bad("EXR RLE: size mismatch");
    for (size_t k = 1; k < t.size(); k++) t[k] = (uint8_t)(t[k - 1] + t[k] - 128);
    out.resize(expect);
    size_t half = (expect + 1) / 2;
    for (size_t k = 0; k < expect; k++) out[k] = (k & 1) ? t[half + k / 2] : t[k / 2];
}

} // namespace

void lx_exr_piz_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expect, const std::vector<int>& ch_bytes,
                       int width, int lines);

LxImageF lx_load_exr(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f.good()) throw LxError{LX_ERR_IO, "cannot read " + path};
    std::vector<uint8_t> b((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t n = b.size();
    if (n < 8 || le32(b.data()) != 20000630u) bad("EXR: bad magic number in " + path);
    uint32_t ver = le32(b.data() + 4);
    if ((ver & 0xff) != 2) bad("EXR: unsupported version");
    if (ver & 0x200) bad("EXR: tiled files are not supported");
    if (ver & 0x1800) bad("EXR: multi-part and deep files are not supported");
    size_t i = 8;
    struct Ch {
        std::string name;
        int type, xs, ys;
    };
    std::vector<Ch> chans;
    int comp = -1, x0 = 0, y0 = 0, x1 = -1, y1 = -1, order = 0;
    while (true) {
        if (i >= n) bad("EXR: truncated header");
        size_t e = i;
        while (e < n && b[e]) e++;
        if (e >= n) bad("EXR: truncated header");
        std::string name((const char*)&b[i], e - i);
        i = e + 1;
        if (name.empty()) break;
        e = i;
        while (e < n && b[e]) e++;
        if (e + 5 > n) bad("EXR: truncated attribute");
        std::string type((const char*)&b[i], e - i);
        i = e + 1;
        uint32_t sz = le32(&b[i]);
        i += 4;
        if (i + sz > n) bad("EXR: truncated attribute value");
        const uint8_t* v = &b[i];
        if (name == "channels") {
            size_t k = 0;
            while (k < sz && v[k]) {
                size_t q = k;
                while (q < sz && v[q]) q++;
                if (q + 17 > sz) bad("EXR: bad channel list");
                Ch c;
                c.name.assign((const char*)v + k, q - k);
                c.type = (int)le32(v + q + 1);
                c.xs = (int)le32(v + q + 9);
                c.ys = (int)le32(v + q + 13);
                if (c.xs != 1 || c.ys != 1) bad("EXR: subsampled channels are not supported");
                if (c.type < 0 || c.type > 2) bad("EXR: bad pixel type");
                chans.push_back(c);
                k = q + 17;
            }
        } else if (name == "compression") {
            comp = v[0];
        } else if (name == "dataWindow") {
            x0 = (int)le32(v);
            y0 = (int)le32(v + 4);
            x1 = (int)le32(v + 8);
            y1 = (int)le32(v + 12);
        } else if (name == "lineOrder") {
            order = v[0];
        }
        i += sz;
    }
    (void)order;
    if (chans.empty()) bad("EXR: no channels");
    int W = x1 - x0 + 1, H = y1 - y0 + 1;
    if (W <= 0 || H <= 0 || W > 65536 || H > 65536) bad("EXR: bad data window");
    int lpb;
    switch (comp) {
    case 0: case 1: case 2: lpb = 1; break;
    case 3: lpb = 16; break;
    case 4: lpb = 32; break;
    default: bad("EXR: unsupported compression " + std::to_string(comp) + " (NONE, RLE, ZIPS, ZIP, PIZ are)");
    }
    std::vector<int> cb(chans.size());
    size_t line_bytes = 0;
    for (size_t c = 0; c < chans.size(); c++) {
        cb[c] = chans[c].type == 1 ? 2 : 4;
        line_bytes += (size_t)cb[c] * W;
    }
    int nblocks = (H + lpb - 1) / lpb;
    if (i + 8 * (size_t)nblocks > n) bad("EXR: truncated offset table");
    std::vector<std::vector<float>> planes(chans.size(), std::vector<float>((size_t)W * H));
    std::vector<uint8_t> dec;
    for (int k = 0; k < nblocks; k++) {
        uint64_t off = le64(&b[i + 8 * (size_t)k]);
        if (off > n || n - off < 8) bad("EXR: bad block offset");
        int y = (int)le32(&b[off]);
        uint32_t len = le32(&b[off + 4]);
        if (len > n - off - 8) bad("EXR: truncated block");
        int ly = y - y0;
        if (ly < 0 || ly >= H) bad("EXR: block outside the data window");
        int lines = std::min(lpb, H - ly);
        size_t expect = line_bytes * lines;
        const uint8_t* src = &b[off + 8];
        if (len == expect) {
            dec.assign(src, src + expect); /* stored uncompressed */
        } else if (comp == 1) {
            exr_unrle(src, len, dec, expect);
        } else if (comp == 2 || comp == 3) {
            exr_unzip(src, len, dec, expect);
        } else if (comp == 4) {
            dec.resize(expect);
            lx_exr_piz_decode(src, len, dec.data(), expect, cb, W, lines);
        } else {
            bad("EXR: block size mismatch");
        }
        size_t p = 0;
        for (int l = 0; l < lines; l++)
            for (size_t c = 0; c < chans.size(); c++) {
                float* dst = planes[c].data() + (size_t)(ly + l) * W;
                for (int x = 0; x < W; x++) {
                    const uint8_t* s = &dec[p];
                    if (chans[c].type == 1) dst[x] = half_to_float((uint16_t)(s[0] | s[1] << 8));
                    else if (chans[c].type == 2) std::memcpy(&dst[x], s, 4);
                    else dst[x] = (float)le32(s);
                    p += (size_t)cb[c];
                }
            }
    }
    LxImageF out;
    out.w = (uint32_t)W;
    out.h = (uint32_t)H;
    out.rgba.resize((size_t)W * H * 4);
    /* tinyexr's LoadEXR: channels without a layer prefix, the first four */
    std::vector<int> top;
    for (size_t c = 0; c < chans.size(); c++)
        if (chans[c].name.find('.') == std::string::npos) top.push_back((int)c);
    if (top.empty()) bad("EXR: only layered channels (e.g. diffuse.R); no top-level R, G, B");
    int iR = -1, iG = -1, iB = -1, iA = -1;
    for (size_t k = 0; k < top.size() && k < 4; k++) {
        const std::string& nm = chans[top[k]].name;
        if (nm == "R") iR = top[k];
        else if (nm == "G") iG = top[k];
        else if (nm == "B") iB = top[k];
        else if (nm == "A") iA = top[k];
    }
    size_t np = (size_t)W * H;
    if (top.size() == 1) {
        const std::vector<float>& g = planes[top[0]];
        for (size_t q = 0; q < np; q++)
            for (int c = 0; c < 4; c++) out.rgba[4 * q + c] = g[q];
    } else {
        if (iR < 0 || iG < 0 || iB < 0) bad("EXR: R, G or B channel not found");
        for (size_t q = 0; q < np; q++) {
            out.rgba[4 * q + 0] = planes[iR][q];
            out.rgba[4 * q + 1] = planes[iG][q];
            out.rgba[4 * q + 2] = planes[iB][q];
            out.rgba[4 * q + 3] = iA >= 0 ? planes[iA][q] : 1.0f;
        }
    }
    return out;
}
