// png.cpp -- PNG decoding to packed RGBA8 for the glTF loader.
//
// The reference loads textures with the image crate and converts them with
// to_rgba8() (src/texture.rs:916-931).  This decoder covers what that
// conversion yields for non-interlaced PNG: colour types 0, 2, 3, 4 and 6 at
// every bit depth PNG allows for them, tRNS transparency (palette alpha or a
// colour key) and the five scanline filters.  zlib streams are inflated here
// (RFC 1950/1951: stored, fixed and dynamic Huffman blocks).

#include "png.h"

#include <cstring>

namespace sr {
namespace {

struct BitReader {
    const uint8_t *p;
    size_t n, pos = 0;
    uint64_t bits = 0;
    int count = 0;
    bool ok = true;
    BitReader(const uint8_t *data, size_t size) : p(data), n(size) {}
    void refill() {
        while (count <= 56) {
            uint64_t b = 0;
            if (pos < n) b = p[pos];
            else if (pos > n + 8) { ok = false; return; }
            pos++;
            bits |= b << count;
            count += 8;
        }
    }
    uint32_t get(int k) {
        if (k == 0) return 0;
        if (count < k) refill();
        uint32_t v = (uint32_t)(bits & ((1ull << k) - 1));
        bits >>= k;
        count -= k;
        return v;
    }
    void align_byte() {
        int drop = count & 7;
        bits >>= drop;
        count -= drop;
    }
    bool overrun() const { return !ok || pos - (size_t)(count / 8) > n; }
};

// Canonical Huffman code (RFC 1951 3.2.2): a 9-bit lookup table for short
// codes (stream bits are least significant first, so the table is indexed by
// bit-reversed codes), and for longer ones the symbols sorted by code length,
// decoded one bit at a time against each length's first code.
struct Huffman {
    static constexpr int FAST = 9;
    uint16_t count[16];
    uint16_t symbol[320];
    uint16_t fast[1 << FAST];   // symbol << 4 | length, 0 = longer code
    bool build(const uint8_t *lengths, int n) {
        memset(count, 0, sizeof(count));
        memset(fast, 0, sizeof(fast));
        for (int i = 0; i < n; i++) count[lengths[i]]++;
        count[0] = 0;
        int left = 1;
        for (int len = 1; len < 16; len++) {
            left <<= 1;
            left -= count[len];
            if (left < 0) return false;   // over-subscribed
        }
        uint16_t offs[16];
        offs[1] = 0;
        for (int len = 1; len < 15; len++) offs[len + 1] = (uint16_t)(offs[len] + count[len]);
        for (int i = 0; i < n; i++)
            if (lengths[i]) symbol[offs[lengths[i]]++] = (uint16_t)i;
        uint32_t next[16];
        uint32_t code = 0;
        for (int len = 1; len < 16; len++) {
            code = (code + count[len - 1]) << 1;
            next[len] = code;
        }
        for (int i = 0; i < n; i++) {
            int len = lengths[i];
            if (!len) continue;
            uint32_t c = next[len]++;
            if (len > FAST) continue;
            uint32_t rev = 0;
            for (int b = 0; b < len; b++) rev |= ((c >> b) & 1u) << (len - 1 - b);
            for (uint32_t k = rev; k < (1u << FAST); k += 1u << len) fast[k] = (uint16_t)((i << 4) | len);
        }
        return true;
    }
    int decode(BitReader &br) const {
        if (br.count < 16) br.refill();
        uint16_t e = fast[br.bits & ((1u << FAST) - 1)];
        if (e) {
            int len = e & 15;
            br.bits >>= len;
            br.count -= len;
            return e >> 4;
        }
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; len++) {
            code |= (int)br.get(1);
            int c = count[len];
            if (code - c < first) return symbol[index + (code - first)];
            index += c;
            first += c;
            first <<= 1;
            code <<= 1;
        }
        return -1;
    }
};

const uint16_t LEN_BASE[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                               35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t LEN_EXTRA[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t DIST_BASE[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193,
                                257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t DIST_EXTRA[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};

bool inflate_codes(BitReader &br, const Huffman &lit, const Huffman &dist, std::vector<uint8_t> &out) {
    for (;;) {
        int sym = lit.decode(br);
        if (sym < 0 || br.overrun()) return false;
        if (sym < 256) {
            out.push_back((uint8_t)sym);
        } else if (sym == 256) {
            return true;
        } else {
            sym -= 257;
            if (sym >= 29) return false;
            size_t len = LEN_BASE[sym] + br.get(LEN_EXTRA[sym]);
            int ds = dist.decode(br);
            if (ds < 0 || ds >= 30) return false;
            size_t d = DIST_BASE[ds] + br.get(DIST_EXTRA[ds]);
            if (d > out.size()) return false;
            size_t from = out.size() - d;
            for (size_t k = 0; k < len; k++) out.push_back(out[from + k]);
        }
    }
}

}  // namespace

bool zlib_inflate(const uint8_t *data, size_t size, std::vector<uint8_t> &out, std::string &err) {
    if (size < 2) { err = "zlib stream too short"; return false; }
    uint8_t cmf = data[0], flg = data[1];
    if ((cmf & 15) != 8 || ((cmf << 8) | flg) % 31 != 0) { err = "not a zlib deflate stream"; return false; }
    if (flg & 0x20) { err = "zlib preset dictionary not supported"; return false; }
    BitReader br(data + 2, size - 2);
    for (;;) {
        uint32_t final_block = br.get(1);
        uint32_t type = br.get(2);
        if (type == 0) {
            br.align_byte();
            uint32_t len = br.get(16), nlen = br.get(16);
            if ((len ^ 0xFFFFu) != nlen) { err = "stored block length check failed"; return false; }
            for (uint32_t k = 0; k < len; k++) out.push_back((uint8_t)br.get(8));
        } else if (type == 1) {
            struct Fixed {
                Huffman lit, dist;
                Fixed() {
                    uint8_t l[288];
                    int i = 0;
                    for (; i < 144; i++) l[i] = 8;
                    for (; i < 256; i++) l[i] = 9;
                    for (; i < 280; i++) l[i] = 7;
                    for (; i < 288; i++) l[i] = 8;
                    lit.build(l, 288);
                    uint8_t d[30];
                    for (int k = 0; k < 30; k++) d[k] = 5;
                    dist.build(d, 30);
                }
            };
            static const Fixed fixed;        // RFC 1951 3.2.6
            const Huffman &lit = fixed.lit, &dist = fixed.dist;
            if (!inflate_codes(br, lit, dist, out)) { err = "corrupt fixed-Huffman block"; return false; }
        } else if (type == 2) {
            int hlit = (int)br.get(5) + 257, hdist = (int)br.get(5) + 1, hclen = (int)br.get(4) + 4;
            static const uint8_t ORDER[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
            uint8_t cl[19] = {0};
            for (int i = 0; i < hclen; i++) cl[ORDER[i]] = (uint8_t)br.get(3);
            Huffman clh;
            if (!clh.build(cl, 19)) { err = "bad code-length code"; return false; }
            uint8_t lens[320] = {0};
            int i = 0;
            while (i < hlit + hdist) {
                int sym = clh.decode(br);
                if (sym < 0 || br.overrun()) { err = "corrupt code lengths"; return false; }
                if (sym < 16) {
                    lens[i++] = (uint8_t)sym;
                } else {
                    int rep = 0;
                    uint8_t val = 0;
                    if (sym == 16) {
                        if (i == 0) { err = "repeat with no previous length"; return false; }
                        val = lens[i - 1];
                        rep = 3 + (int)br.get(2);
                    } else if (sym == 17) {
                        rep = 3 + (int)br.get(3);
                    } else {
                        rep = 11 + (int)br.get(7);
                    }
                    if (i + rep > hlit + hdist) { err = "code lengths overflow"; return false; }
                    while (rep--) lens[i++] = val;
                }
            }
            Huffman lit, dist;
            if (!lit.build(lens, hlit) || !dist.build(lens + hlit, hdist)) { err = "bad Huffman lengths"; return false; }
            if (!inflate_codes(br, lit, dist, out)) { err = "corrupt dynamic-Huffman block"; return false; }
        } else {
            err = "invalid deflate block type";
            return false;
        }
        if (br.overrun()) { err = "truncated deflate stream"; return false; }
        if (final_block) break;
    }
    return true;
}

static uint32_t be32(const uint8_t *p) { return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3]; }

static inline uint8_t paeth(int a, int b, int c) {
    int p = a + b - c;
    int pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return (uint8_t)a;
    if (pb <= pc) return (uint8_t)b;
    return (uint8_t)c;
}

bool png_decode_rgba8(const uint8_t *data, size_t size, uint32_t &width, uint32_t &height, std::vector<uint32_t> &out,
                      std::string &err) {
    static const uint8_t SIG[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (size < 8 || memcmp(data, SIG, 8) != 0) { err = "not a PNG file"; return false; }
    size_t pos = 8;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    bool seen_end = false;
    while (pos + 12 <= size) {
        uint32_t len = be32(data + pos);
        const uint8_t *type = data + pos + 4;
        if (pos + 12 + (size_t)len > size) { err = "truncated PNG chunk"; return false; }
        const uint8_t *body = data + pos + 8;
        if (!memcmp(type, "IHDR", 4)) {
            if (len < 13) { err = "bad IHDR"; return false; }
            w = be32(body);
            h = be32(body + 4);
            depth = body[8];
            ctype = body[9];
            interlace = body[12];
        } else if (!memcmp(type, "PLTE", 4)) {
            plte.assign(body, body + len);
        } else if (!memcmp(type, "tRNS", 4)) {
            trns.assign(body, body + len);
        } else if (!memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), body, body + len);
        } else if (!memcmp(type, "IEND", 4)) {
            seen_end = true;
            break;
        }
        pos += 12 + (size_t)len;
    }
    if (!seen_end || ctype < 0) { err = "PNG has no IHDR or IEND"; return false; }
    if (w == 0 || h == 0 || w > 65536 || h > 65536) { err = "PNG size out of range"; return false; }
    // at most a 512 MiB RGBA8 image (the image crate's default allocation limit)
    if ((uint64_t)w * h > (1ull << 27)) { err = "PNG larger than 2^27 pixels"; return false; }
    if (interlace != 0) { err = "interlaced PNG is not supported"; return false; }
    int channels;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: err = "invalid PNG colour type"; return false;
    }
    bool depth_ok = depth == 8 || (depth == 16 && ctype != 3) || ((depth == 1 || depth == 2 || depth == 4) && (ctype == 0 || ctype == 3));
    if (!depth_ok) { err = "invalid PNG bit depth for its colour type"; return false; }
    if (ctype == 3 && plte.empty()) { err = "palette PNG without PLTE"; return false; }
    std::vector<uint8_t> raw;
    if (!zlib_inflate(idat.data(), idat.size(), raw, err)) return false;
    size_t bits_pp = (size_t)channels * depth;
    size_t stride = ((size_t)w * bits_pp + 7) / 8;
    size_t bpp = bits_pp >= 8 ? bits_pp / 8 : 1;      // filter byte distance
    if (raw.size() < (stride + 1) * h) { err = "PNG image data too short"; return false; }
    std::vector<uint8_t> prev(stride, 0), cur(stride);
    out.resize((size_t)w * h);
    // colour key (tRNS for types 0 and 2), compared at the file's bit depth
    bool key = !trns.empty() && (ctype == 0 || ctype == 2);
    uint16_t kv[3] = {0, 0, 0};
    if (key) {
        if (ctype == 0 && trns.size() >= 2) kv[0] = (uint16_t)((trns[0] << 8) | trns[1]);
        else if (ctype == 2 && trns.size() >= 6)
            for (int c = 0; c < 3; c++) kv[c] = (uint16_t)((trns[2 * c] << 8) | trns[2 * c + 1]);
        else key = false;
    }
    // 16 -> 8 bits rounds to nearest (x * 255 / 65535), 1/2/4-bit grey scales
    // by replication (image crate expansion)
    auto to8 = [&](uint32_t v) -> uint8_t {
        if (depth == 16) return (uint8_t)((v + 128u) / 257u);
        if (depth == 8) return (uint8_t)v;
        return (uint8_t)(v * 255u / ((1u << depth) - 1u));
    };
    for (uint32_t y = 0; y < h; y++) {
        const uint8_t *line = raw.data() + (size_t)y * (stride + 1);
        uint8_t f = line[0];
        const uint8_t *src = line + 1;
        for (size_t i = 0; i < stride; i++) {
            int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int x = src[i];
            switch (f) {
                case 0: break;
                case 1: x += a; break;
                case 2: x += b; break;
                case 3: x += (a + b) >> 1; break;
                case 4: x += paeth(a, b, c); break;
                default: err = "invalid PNG filter type"; return false;
            }
            cur[i] = (uint8_t)x;
        }
        auto sample = [&](uint32_t px, int ch) -> uint32_t {
            size_t bitpos = ((size_t)px * channels + ch) * depth;
            if (depth == 16) return ((uint32_t)cur[bitpos / 8] << 8) | cur[bitpos / 8 + 1];
            if (depth == 8) return cur[bitpos / 8];
            uint32_t byte = cur[bitpos / 8];
            int shift = 8 - depth - (int)(bitpos % 8);
            return (byte >> shift) & ((1u << depth) - 1u);
        };
        for (uint32_t x = 0; x < w; x++) {
            uint8_t r, g, bl, al = 255;
            if (ctype == 3) {
                uint32_t idx = sample(x, 0);
                if ((size_t)idx * 3 + 2 >= plte.size()) { err = "palette index out of range"; return false; }
                r = plte[3 * idx]; g = plte[3 * idx + 1]; bl = plte[3 * idx + 2];
                if (idx < trns.size()) al = trns[idx];
            } else if (ctype == 0 || ctype == 4) {
                uint32_t v = sample(x, 0);
                r = g = bl = to8(v);
                if (ctype == 4) al = to8(sample(x, 1));
                else if (key && v == kv[0]) al = 0;
            } else {
                uint32_t v0 = sample(x, 0), v1 = sample(x, 1), v2 = sample(x, 2);
                r = to8(v0); g = to8(v1); bl = to8(v2);
                if (ctype == 6) al = to8(sample(x, 3));
                else if (key && v0 == kv[0] && v1 == kv[1] && v2 == kv[2]) al = 0;
            }
            out[(size_t)y * w + x] = ((uint32_t)r << 24) | ((uint32_t)g << 16) | ((uint32_t)bl << 8) | al;
        }
        std::swap(prev, cur);
    }
    width = w;
    height = h;
    return true;
}

}  // namespace sr
