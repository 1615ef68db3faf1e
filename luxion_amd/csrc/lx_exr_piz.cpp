/*
 * lx_exr_piz.cpp -- OpenEXR PIZ block decoder.
 *
 * PIZ is defined by the OpenEXR file format, so this follows its published
 * steps exactly (any other arithmetic would decode different values):
 *   1. a bitmap of the 16-bit values present and the reverse look-up table
 *      that maps dense indices back to them;
 *   2. Huffman decoding: 6-bit code lengths with zero-run codes (59..63),
 *      canonical codes assigned from the longest length down, and a
 *      run-length symbol (the largest symbol) that repeats the previous value;
 *   3. the 2D Haar-like wavelet inverse per channel plane, in its 14-bit form
 *      when the largest index fits in 14 bits and its 16-bit modular form
 *      otherwise;
 *   4. the look-up table applied, then planes re-interleaved per scanline.
 */
#include "lx_image.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace {

[[noreturn]] void bad(const char* what) { throw LxError{LX_ERR_FORMAT, std::string("EXR PIZ: ") + what}; }

uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

struct Bits {
    const uint8_t* p;
    const uint8_t* end;
    uint64_t buf = 0;
    int n = 0;
    uint64_t take(int k) {
        while (n < k) {
            if (p >= end) bad("truncated code table");
            buf = (buf << 8) | *p++;
            n += 8;
        }
        n -= k;
        return (buf >> n) & ((1ull << k) - 1ull);
    }
};

const int kMaxSym = 65537; /* 65536 values + the run-length symbol */

/* Code lengths for symbols im..iM, then canonical codes. */
void read_table(const uint8_t*& p, const uint8_t* end, uint32_t im, uint32_t iM, std::vector<uint64_t>& len,
                std::vector<uint64_t>& code) {
    len.assign(kMaxSym, 0);
    Bits b{p, end};
    for (uint32_t i = im; i <= iM; i++) {
        uint64_t l = b.take(6);
        if (l == 63) {
            uint32_t run = (uint32_t)b.take(8) + 6;
            if (i + run > iM + 1) bad("code table run overflows");
            for (uint32_t k = 0; k < run; k++) len[i + k] = 0;
            i += run - 1;
        } else if (l >= 59) {
            uint32_t run = (uint32_t)(l - 59 + 2);
            if (i + run > iM + 1) bad("code table run overflows");
            for (uint32_t k = 0; k < run; k++) len[i + k] = 0;
            i += run - 1;
        } else {
            len[i] = l;
        }
    }
    p = b.p;
    /* canonical codes: start codes from the longest length down */
    uint64_t count[59] = {0};
    for (int i = 0; i < kMaxSym; i++) count[len[i]]++;
    uint64_t c = 0;
    for (int l = 58; l > 0; l--) {
        uint64_t nc = (c + count[l]) >> 1;
        count[l] = c;
        c = nc;
    }
    code.assign(kMaxSym, 0);
    for (int i = 0; i < kMaxSym; i++)
        if (len[i] > 0) code[i] = count[len[i]]++;
}

void huf_decode(const uint8_t* src, size_t n, uint16_t* out, size_t nout) {
    if (n < 20) bad("short Huffman block");
    uint32_t im = rd32(src), iM = rd32(src + 4), nbits = rd32(src + 12);
    if (im >= (uint32_t)kMaxSym || iM >= (uint32_t)kMaxSym || im > iM) bad("bad symbol range");
    const uint8_t* p = src + 20;
    const uint8_t* end = src + n;
    std::vector<uint64_t> len, code;
    read_table(p, end, im, iM, len, code);
    if ((size_t)(end - p) * 8 < nbits) bad("truncated bit stream");
    /* fast table for codes of up to FB bits; longer codes by length */
    const int FB = 14;
    std::vector<int32_t> fast((size_t)1 << FB, -1);
    std::vector<std::unordered_map<uint64_t, uint32_t>> slow(59);
    for (uint32_t s = im; s <= iM; s++) {
        int l = (int)len[s];
        if (!l) continue;
        if (l <= FB) {
            uint64_t first = code[s] << (FB - l), cnt = 1ull << (FB - l);
            for (uint64_t k = 0; k < cnt; k++) fast[first + k] = (int32_t)(s << 6 | (uint32_t)l);
        } else {
            slow[l][code[s]] = s;
        }
    }
    uint32_t rlc = iM;
    size_t pos = 0; /* bit position */
    size_t o = 0;
    auto peek = [&](int k) -> uint64_t {
        /* k bits at pos, MSB first, zero past the end of the stream */
        uint64_t v = 0;
        for (int j = 0; j < k; j++) {
            size_t bp = pos + (size_t)j;
            uint64_t bit = bp < nbits ? (p[bp >> 3] >> (7 - (bp & 7))) & 1u : 0u;
            v = (v << 1) | bit;
        }
        return v;
    };
    while (pos < nbits) {
        uint32_t sym;
        int l;
        int32_t f = fast[peek(FB)];
        if (f >= 0) {
            sym = (uint32_t)f >> 6;
            l = f & 63;
        } else {
            sym = 0;
            l = 0;
            for (int L = FB + 1; L <= 58; L++) {
                auto it = slow[L].find(peek(L));
                if (it != slow[L].end()) {
                    sym = it->second;
                    l = L;
                    break;
                }
            }
            if (!l) bad("invalid Huffman code");
        }
        if (pos + (size_t)l > nbits) bad("code runs past the stream");
        pos += (size_t)l;
        if (sym == rlc) {
            if (pos + 8 > nbits) bad("truncated run length");
            uint32_t run = (uint32_t)peek(8);
            pos += 8;
            if (o == 0) bad("run length without a previous value");
            if (o + run > nout) bad("run overflows the output");
            uint16_t v = out[o - 1];
            for (uint32_t k = 0; k < run; k++) out[o++] = v;
        } else {
            if (o >= nout) bad("too many symbols");
            out[o++] = (uint16_t)sym;
        }
    }
    if (o != nout) bad("too few symbols");
}

inline void wdec14(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    int16_t ls = (int16_t)l, hs = (int16_t)h;
    int hi = hs;
    int ai = ls + (hi & 1) + (hi >> 1);
    a = (uint16_t)(int16_t)ai;
    b = (uint16_t)(int16_t)(ai - hi);
}

inline void wdec16(uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
    int m = l, d = h;
    int bb = (m - (d >> 1)) & 0xffff;
    int aa = (d + bb - 0x8000) & 0xffff;
    b = (uint16_t)bb;
    a = (uint16_t)aa;
}

/* in-place inverse wavelet of an nx x ny plane with element stride ox and
 * row stride oy */
void wav2_decode(uint16_t* in, int nx, int ox, int ny, int oy, uint16_t mx) {
    bool w14 = mx < (1 << 14);
    int n = std::min(nx, ny);
    int p = 1;
    while (p <= n) p <<= 1;
    p >>= 1;
    int p2 = p;
    p >>= 1;
    auto dec = [&](uint16_t l, uint16_t h, uint16_t& a, uint16_t& b) {
        if (w14) wdec14(l, h, a, b);
        else wdec16(l, h, a, b);
    };
    while (p >= 1) {
        uint16_t* py = in;
        uint16_t* ey = in + (ptrdiff_t)oy * (ny - p2);
        int oy1 = oy * p, oy2 = oy * p2, ox1 = ox * p, ox2 = ox * p2;
        uint16_t i00, i01, i10, i11;
        for (; py <= ey; py += oy2) {
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                uint16_t* p10 = px + oy1;
                uint16_t* p11 = p10 + ox1;
                dec(*px, *p10, i00, i10);
                dec(*p01, *p11, i01, i11);
                dec(i00, i01, *px, *p01);
                dec(i10, i11, *p10, *p11);
            }
            if (nx & p) {
                uint16_t* p10 = px + oy1;
                dec(*px, *p10, i00, *p10);
                *px = i00;
            }
        }
        if (ny & p) {
            uint16_t* px = py;
            uint16_t* ex = py + (ptrdiff_t)ox * (nx - p2);
            for (; px <= ex; px += ox2) {
                uint16_t* p01 = px + ox1;
                dec(*px, *p01, i00, *p01);
                *px = i00;
            }
        }
        p2 = p;
        p >>= 1;
    }
}

} // namespace

void lx_exr_piz_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expect, const std::vector<int>& ch_bytes,
                       int width, int lines) {
    if (n < 4) bad("short block");
    uint16_t minnz = (uint16_t)(src[0] | src[1] << 8), maxnz = (uint16_t)(src[2] | src[3] << 8);
    size_t i = 4;
    std::vector<uint8_t> bitmap(8192, 0);
    if (maxnz >= 8192) bad("bad bitmap range");
    if (minnz <= maxnz) {
        size_t cnt = (size_t)maxnz - minnz + 1;
        if (i + cnt > n) bad("truncated bitmap");
        std::memcpy(&bitmap[minnz], src + i, cnt);
        i += cnt;
    }
    std::vector<uint16_t> lut(65536, 0);
    int k = 0;
    for (int v = 0; v < 65536; v++)
        if (v == 0 || (bitmap[v >> 3] & (1 << (v & 7)))) lut[k++] = (uint16_t)v;
    uint16_t maxv = (uint16_t)(k - 1);
    if (i + 4 > n) bad("missing Huffman length");
    uint32_t hl = rd32(src + i);
    i += 4;
    if (i + hl > n) bad("truncated Huffman data");
    if (expect % 2) bad("odd block size");
    size_t nv = expect / 2;
    std::vector<uint16_t> tmp(nv);
    huf_decode(src + i, hl, tmp.data(), nv);
    /* planes: channel c occupies lines * width * size values, size = bytes/2 */
    size_t off = 0;
    std::vector<size_t> start(ch_bytes.size());
    for (size_t c = 0; c < ch_bytes.size(); c++) {
        int size = ch_bytes[c] / 2;
        start[c] = off;
        for (int j = 0; j < size; j++) wav2_decode(tmp.data() + off + j, width, size, lines, width * size, maxv);
        off += (size_t)lines * width * size;
    }
    if (off != nv) bad("channel sizes do not match the block");
    for (auto& v : tmp) v = lut[v];
    /* re-interleave: line by line, channel by channel */
    uint8_t* o = out;
    for (int y = 0; y < lines; y++)
        for (size_t c = 0; c < ch_bytes.size(); c++) {
            size_t cnt = (size_t)width * (ch_bytes[c] / 2);
            const uint16_t* s = tmp.data() + start[c] + (size_t)y * cnt;
            for (size_t q = 0; q < cnt; q++) {
                *o++ = (uint8_t)(s[q] & 0xff);
                *o++ = (uint8_t)(s[q] >> 8);
            }
        }
}
