/*
 * lx_exr_piz.cpp -- OpenEXR PIZ block decoder.
 *
 * A PIZ block holds, in order: the byte range and bytes of a 65536-bit
 * "values present" bitmap; the byte length of a Huffman-coded stream; the
 * stream.  Decoding, as the OpenEXR file format defines it:
 *   1. the stream decodes to 16-bit dense indices (symbols; the largest
 *      symbol of the table is the run-length escape: 8 bits follow with a
 *      repeat count of the previous index);
 *   2. each channel's plane (lines x width x words-per-sample) is run through
 *      the inverse of the format's integer Haar wavelet, coarsest level first
 *      (14-bit lifting when the largest dense index fits in 14 bits, 16-bit
 *      modular lifting otherwise);
 *   3. dense indices map back to the values whose bitmap bits are set (the
 *      value 0 is always present); planes are re-interleaved per scanline.
 * The arithmetic of each step is fixed by the format; the code below is
 * organised around explicit (x, y) plane coordinates and per-length code
 * ranges.
 */
#include "lx_image.h"

#include <algorithm>
#include <cstring>
#include <unordered_map>

namespace {

[[noreturn]] void bad(const char* what) { throw LxError{LX_ERR_FORMAT, std::string("EXR PIZ: ") + what}; }

uint32_t rd32(const uint8_t* p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

constexpr uint32_t kSymbols = 65537; /* 65536 dense indices + the run-length escape */
constexpr int kMaxLen = 58;          /* longest code the format allows */

/* MSB-first reader over the packed code-length table. */
class BitReader {
public:
    BitReader(const uint8_t* p, const uint8_t* end) : p_(p), end_(end) {}
    uint32_t get(int k) {
        while (have_ < k) {
            if (p_ >= end_) bad("truncated code table");
            acc_ = (acc_ << 8) | *p_++;
            have_ += 8;
        }
        have_ -= k;
        return (uint32_t)((acc_ >> have_) & ((1u << k) - 1u));
    }
    const uint8_t* next_byte() const { return p_; }

private:
    const uint8_t* p_;
    const uint8_t* end_;
    uint64_t acc_ = 0;
    int have_ = 0;
};

/* Code lengths of symbols lo..hi: a 6-bit length per symbol, except that
 * 59..62 stand for runs of 2..5 zero lengths and 63 is followed by an 8-bit
 * count of a run of 6..261 zero lengths. */
std::vector<uint8_t> unpack_lengths(const uint8_t*& p, const uint8_t* end, uint32_t lo, uint32_t hi) {
    std::vector<uint8_t> len(kSymbols, 0);
    BitReader br(p, end);
    uint32_t s = lo;
    while (s <= hi) {
        uint32_t v = br.get(6);
        uint32_t zeros = 0;
        if (v == 63) zeros = br.get(8) + 6;
        else if (v >= 59) zeros = v - 57;
        if (zeros) {
            if (zeros > hi + 1 - s) bad("code table run overflows");
            s += zeros; /* lengths already zero */
        } else {
            len[s++] = (uint8_t)v;
        }
    }
    p = br.next_byte();
    return len;
}

/* Canonical code values.  For each length L the format gives the codes of
 * that length one contiguous range, in symbol order, starting at the summed
 * weight of every longer code, sum over M > L of count[M] * 2^(58 - M),
 * counted in units of 2^(58 - L) and truncated.  Weights summing past 2^58
 * (Kraft sum above 1) mean an oversubscribed table whose codes would not fit
 * their lengths: rejected. */
std::vector<uint64_t> canonical_codes(const std::vector<uint8_t>& len, uint32_t lo, uint32_t hi) {
    uint64_t per_len[kMaxLen + 1] = {0};
    for (uint32_t s = lo; s <= hi; s++) per_len[len[s]]++;
    uint64_t next[kMaxLen + 1] = {0};
    unsigned __int128 longer = 0; /* weight of the codes longer than the current length */
    for (int L = kMaxLen; L >= 1; --L) {
        next[L] = (uint64_t)(longer >> (kMaxLen - L));
        longer += (unsigned __int128)per_len[L] << (kMaxLen - L);
    }
    if (longer > ((unsigned __int128)1 << kMaxLen)) bad("oversubscribed Huffman code lengths");
    std::vector<uint64_t> code(kSymbols, 0);
    for (uint32_t s = lo; s <= hi; s++) {
        int L = len[s];
        if (!L) continue;
        code[s] = next[L]++;
        if (code[s] >> L) bad("Huffman code does not fit its length");
    }
    return code;
}

/* Decode the Huffman block src[0..n) into exactly nout 16-bit indices. */
void huffman_decode(const uint8_t* src, size_t n, uint16_t* out, size_t nout) {
    if (n < 20) bad("short Huffman block");
    const uint32_t lo = rd32(src), hi = rd32(src + 4), nbits = rd32(src + 12);
    if (lo > hi || hi >= kSymbols) bad("bad symbol range");
    const uint8_t* p = src + 20;
    const uint8_t* end = src + n;
    std::vector<uint8_t> len = unpack_lengths(p, end, lo, hi);
    std::vector<uint64_t> code = canonical_codes(len, lo, hi);
    if ((uint64_t)(end - p) * 8 < nbits) bad("truncated bit stream");
    /* one table look-up for codes of up to kFast bits; longer codes are found
     * per length */
    const int kFast = 14;
    std::vector<int32_t> fast((size_t)1 << kFast, -1);
    std::vector<std::unordered_map<uint64_t, uint32_t>> longer(kMaxLen + 1);
    for (uint32_t s = lo; s <= hi; s++) {
        const int L = len[s];
        if (!L) continue;
        if (L <= kFast) {
            const uint64_t span = 1ull << (kFast - L), base = code[s] << (kFast - L);
            std::fill(fast.begin() + (ptrdiff_t)base, fast.begin() + (ptrdiff_t)(base + span),
                      (int32_t)(s << 6 | (uint32_t)L));
        } else {
            longer[L].emplace(code[s], s);
        }
    }
    const uint32_t escape = hi;
    size_t bit = 0;
    /* k bits at `bit`, MSB first, zeros past the end of the stream */
    auto window = [&](int k) -> uint64_t {
        uint64_t v = 0;
        for (int j = 0; j < k; j++) {
            const size_t b = bit + (size_t)j;
            v = (v << 1) | (b < nbits ? (uint64_t)((p[b >> 3] >> (7 - (b & 7))) & 1u) : 0u);
        }
        return v;
    };
    size_t o = 0;
    while (bit < nbits) {
        uint32_t sym = 0;
        int L = 0;
        const int32_t f = fast[window(kFast)];
        if (f >= 0) {
            sym = (uint32_t)f >> 6;
            L = f & 63;
        } else {
            for (int k = kFast + 1; k <= kMaxLen && !L; k++) {
                auto it = longer[k].find(window(k));
                if (it != longer[k].end()) {
                    sym = it->second;
                    L = k;
                }
            }
            if (!L) bad("invalid Huffman code");
        }
        if (bit + (size_t)L > nbits) bad("code runs past the stream");
        bit += (size_t)L;
        if (sym != escape) {
            if (o >= nout) bad("too many symbols");
            out[o++] = (uint16_t)sym;
            continue;
        }
        if (bit + 8 > nbits) bad("truncated run length");
        const uint32_t repeat = (uint32_t)window(8);
        bit += 8;
        if (o == 0) bad("run length without a previous value");
        if (repeat > nout - o) bad("run overflows the output");
        std::fill(out + o, out + o + repeat, out[o - 1]);
        o += repeat;
    }
    if (o != nout) bad("too few symbols");
}

/* Inverse lifting of one (average, difference) pair, in place: on return
 * (a, b) hold the two original samples. */
struct Unlift14 {
    void operator()(uint16_t& a, uint16_t& b) const {
        const int avg = (int16_t)a, diff = (int16_t)b;
        const int first = avg + (diff & 1) + (diff >> 1);
        a = (uint16_t)first;
        b = (uint16_t)(first - diff);
    }
};
struct Unlift16 {
    void operator()(uint16_t& a, uint16_t& b) const {
        const uint32_t avg = a, diff = b;
        const uint32_t second = (avg - (diff >> 1)) & 0xffffu;
        const uint32_t first = (diff + second + 0x8000u) & 0xffffu; /* - 2^15, mod 2^16 */
        a = (uint16_t)first;
        b = (uint16_t)second;
    }
};

/* Inverse wavelet of one nx x ny plane whose sample (x, y) sits at
 * base[x * sx + y * sy].  Level h (h = largest power of two <= min(nx, ny),
 * halved, down to 1) pairs samples h apart inside 2h x 2h cells: columns
 * first, then rows.  A column of cells cut by the right edge (nx & h) has
 * only its vertical pairs, a row cut by the bottom edge (ny & h) only its
 * horizontal pairs, and the corner cell neither. */
template <class Unlift>
void inverse_wavelet(uint16_t* base, int nx, int ny, ptrdiff_t sx, ptrdiff_t sy, Unlift unlift) {
    auto at = [&](int x, int y) -> uint16_t& { return base[(ptrdiff_t)x * sx + (ptrdiff_t)y * sy]; };
    int top = 1;
    while (top * 2 <= std::min(nx, ny)) top *= 2;
    for (int h = top / 2; h >= 1; h /= 2) {
        const int cell = 2 * h;
        const int cells_x = nx / cell * cell, cells_y = ny / cell * cell;
        const bool cut_x = (nx & h) != 0, cut_y = (ny & h) != 0;
        for (int y = 0; y < cells_y; y += cell) {
            for (int x = 0; x < cells_x; x += cell) {
                unlift(at(x, y), at(x, y + h));
                unlift(at(x + h, y), at(x + h, y + h));
                unlift(at(x, y), at(x + h, y));
                unlift(at(x, y + h), at(x + h, y + h));
            }
            if (cut_x) unlift(at(cells_x, y), at(cells_x, y + h));
        }
        if (cut_y)
            for (int x = 0; x < cells_x; x += cell) unlift(at(x, cells_y), at(x + h, cells_y));
    }
}

} // namespace

void lx_exr_piz_decode(const uint8_t* src, size_t n, uint8_t* out, size_t expect, const std::vector<int>& ch_bytes,
                       int width, int lines) {
    if (n < 4) bad("short block");
    const uint16_t first_byte = (uint16_t)(src[0] | src[1] << 8), last_byte = (uint16_t)(src[2] | src[3] << 8);
    size_t i = 4;
    std::vector<uint8_t> present(8192, 0);
    if (first_byte <= last_byte) {
        if (last_byte >= 8192) bad("bad bitmap range");
        const size_t cnt = (size_t)last_byte - first_byte + 1;
        if (cnt > n - i) bad("truncated bitmap");
        std::memcpy(&present[first_byte], src + i, cnt);
        i += cnt;
    }
    /* dense index -> value: the values whose bit is set, in increasing order */
    std::vector<uint16_t> value_of;
    value_of.reserve(65536);
    for (uint32_t v = 0; v < 65536; v++)
        if (v == 0 || ((present[v >> 3] >> (v & 7)) & 1)) value_of.push_back((uint16_t)v);
    const uint32_t largest_index = (uint32_t)value_of.size() - 1;
    if (n - i < 4) bad("missing Huffman length");
    const uint32_t huf_len = rd32(src + i);
    i += 4;
    if (huf_len > n - i) bad("truncated Huffman data");
    if (expect % 2) bad("odd block size");
    const size_t nwords = expect / 2;
    /* channel planes, one after another: lines x width x words-per-sample */
    std::vector<size_t> plane_at(ch_bytes.size());
    size_t total = 0;
    for (size_t c = 0; c < ch_bytes.size(); c++) {
        plane_at[c] = total;
        total += (size_t)lines * (size_t)width * (size_t)(ch_bytes[c] / 2);
    }
    if (total != nwords) bad("channel sizes do not match the block");
    std::vector<uint16_t> words(nwords);
    huffman_decode(src + i, huf_len, words.data(), nwords);
    for (size_t c = 0; c < ch_bytes.size(); c++) {
        const int per_sample = ch_bytes[c] / 2;
        for (int w = 0; w < per_sample; w++) {
            uint16_t* base = words.data() + plane_at[c] + w;
            const ptrdiff_t sx = per_sample, sy = (ptrdiff_t)width * per_sample;
            if (largest_index < (1u << 14)) inverse_wavelet(base, width, lines, sx, sy, Unlift14{});
            else inverse_wavelet(base, width, lines, sx, sy, Unlift16{});
        }
    }
    for (uint16_t& w : words) {
        if (w > largest_index) bad("index beyond the value table");
        w = value_of[w];
    }
    /* scanline order: line by line, channel by channel, little-endian words */
    uint8_t* o = out;
    for (int y = 0; y < lines; y++)
        for (size_t c = 0; c < ch_bytes.size(); c++) {
            const size_t cnt = (size_t)width * (size_t)(ch_bytes[c] / 2);
            const uint16_t* row = words.data() + plane_at[c] + (size_t)y * cnt;
            for (size_t q = 0; q < cnt; q++) {
                *o++ = (uint8_t)(row[q] & 0xff);
                *o++ = (uint8_t)(row[q] >> 8);
            }
        }
}
