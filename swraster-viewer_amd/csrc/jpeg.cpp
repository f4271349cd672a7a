// jpeg.cpp -- JPEG decoding for glTF textures (the reference decodes every
// texture image through the image crate, src/texture.rs:897-1015).
//
// Written from ITU-T T.81: Huffman table construction (Annex C), sequential
// decoding (F.2: SOF0 baseline and SOF1 extended, 8-bit samples),
// progressive decoding (G.1.2: DC and AC scans, first and refinement passes,
// end-of-band runs), restart intervals (RSTn), any sampling factors up to 4.
// Coefficients are dequantized and inverse transformed with the 8x8 DCT-III
// in double precision, rounded to the nearest sample.  Chroma planes sampled
// at half resolution are upsampled with the triangle (3/4, 1/4) filter,
// other ratios by replication; YCbCr becomes RGB with the JFIF equations
// (three components, unless an Adobe APP14 marker says RGB), one component
// is grey.  Four-component (CMYK / YCCK), 12-bit, lossless and arithmetic-
// coded files are rejected with an error.

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "png.h"

namespace sr {
namespace {

const int ZIGZAG[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                        41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                        30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Canonical Huffman table (T.81 C.2): codes of each length are consecutive
// integers; a 9-bit lookup decodes short codes, longer ones go through the
// per-length code ranges.
struct Huff {
    bool present = false;
    uint8_t fast_len[512];
    uint8_t fast_val[512];
    int32_t maxcode[18];
    int32_t valptr[17];
    int32_t mincode[17];
    uint8_t vals[256];
    bool build(const uint8_t bits[16], const uint8_t *v, int nv) {
        memset(fast_len, 0, sizeof(fast_len));
        memcpy(vals, v, (size_t)nv);
        int code = 0, k = 0;
        for (int l = 1; l <= 16; l++) {
            valptr[l] = k;
            mincode[l] = code;
            for (int i = 0; i < bits[l - 1]; i++, k++, code++) {
                if (code >= (1 << l)) return false;   // over-subscribed: no l-bit code left
                if (l <= 9) {
                    int lo = code << (9 - l), n = 1 << (9 - l);
                    for (int j = 0; j < n; j++) {
                        fast_len[lo + j] = (uint8_t)l;
                        fast_val[lo + j] = v[k];
                    }
                }
            }
            maxcode[l] = bits[l - 1] ? code - 1 : -1;
            if (code > (1 << l)) return false;       // over-subscribed lengths
            code <<= 1;
        }
        maxcode[17] = 0x7FFFFFFF;
        present = true;
        return true;
    }
};

// Entropy-coded segment reader: stuffed 0xFF 0x00 bytes are data; a marker
// ends the data (zeros are shifted in after it).
struct Bits {
    const uint8_t *p, *end;
    uint32_t buf = 0;
    int cnt = 0;
    bool marker = false;
    void fill() {
        while (cnt <= 24) {
            uint32_t b = 0;
            if (!marker && p < end) {
                b = *p;
                if (b == 0xFF) {
                    uint8_t n = p + 1 < end ? p[1] : 0;
                    if (n == 0x00) p += 2;
                    else {
                        marker = true;
                        b = 0;
                    }
                } else {
                    p++;
                }
            }
            buf |= b << (24 - cnt);
            cnt += 8;
        }
    }
    uint32_t get(int n) {
        if (n == 0) return 0;
        fill();
        uint32_t v = buf >> (32 - n);
        buf <<= n;
        cnt -= n;
        return v;
    }
    int decode(const Huff &h) {
        fill();
        uint32_t look = buf >> 23;
        int l = h.fast_len[look];
        if (l) {
            buf <<= l;
            cnt -= l;
            return h.fast_val[look];
        }
        int code = (int)(buf >> 22);       // 10 bits
        l = 10;
        while (l <= 16 && code > h.maxcode[l]) {
            l++;
            code = (int)(buf >> (32 - l));
        }
        if (l > 16) return -1;
        buf <<= l;
        cnt -= l;
        int idx = h.valptr[l] + code - h.mincode[l];
        return idx >= 0 && idx < 256 ? h.vals[idx] : -1;
    }
    int extend(int s) {                    // RECEIVE then EXTEND (F.2.2.1)
        if (s == 0) return 0;
        if (s > 16) return 0;
        int v = (int)get(s);
        return v < (1 << (s - 1)) ? v - (1 << s) + 1 : v;
    }
    // At a restart: drop the buffered bits and consume the RSTn marker.
    bool restart() {
        buf = 0;
        cnt = 0;
        while (p + 1 < end && !(p[0] == 0xFF && p[1] >= 0xD0 && p[1] <= 0xD7)) {
            if (p[0] == 0xFF && p[1] != 0x00 && p[1] != 0xFF && !(p[1] >= 0xD0 && p[1] <= 0xD7)) return false;
            p++;
        }
        if (p + 1 >= end) return false;
        p += 2;
        marker = false;
        return true;
    }
};

struct Comp {
    int id = 0, h = 1, v = 1, tq = 0;
    int bw = 0, bh = 0;        // blocks covering the component's samples
    int bstride = 0, brows = 0;  // allocated blocks (MCU-padded)
    int w = 0, hgt = 0;        // component size in samples
    std::vector<int16_t> coef;
    int td = 0, ta = 0, pred = 0;
};

struct Decoder {
    const uint8_t *data;
    size_t size;
    std::string err;
    uint16_t qt[4][64];        // natural order
    Huff dc[4], ac[4];
    std::vector<Comp> comps;
    int W = 0, H = 0, hmax = 1, vmax = 1, mcux = 0, mcuy = 0;
    bool progressive = false, frame = false, adobe = false;
    int adobe_transform = -1;
    int restart_interval = 0;
    int eobrun = 0;

    bool fail(const char *m) {
        if (err.empty()) err = m;
        return false;
    }

    bool sof(const uint8_t *p, int len, bool prog) {
        if (len < 6) return fail("short SOF");
        if (p[0] != 8) return fail("only 8-bit JPEG samples are decoded");
        H = (p[1] << 8) | p[2];
        W = (p[3] << 8) | p[4];
        int n = p[5];
        if (W <= 0 || H <= 0) return fail("JPEG without a defined size (DNL) is not decoded");
        // the header alone sizes the coefficient planes: refuse what a
        // 512 MiB RGBA8 image would not hold (the image crate's default
        // allocation limit) before allocating
        if ((uint64_t)W * (uint64_t)H > (1ull << 27)) return fail("JPEG larger than 2^27 pixels");
        if (n != 1 && n != 3) return fail("only 1- and 3-component JPEGs are decoded (CMYK is not)");
        if (len < 6 + 3 * n) return fail("short SOF");
        comps.assign(n, Comp());
        for (int i = 0; i < n; i++) {
            Comp &c = comps[i];
            c.id = p[6 + 3 * i];
            c.h = p[7 + 3 * i] >> 4;
            c.v = p[7 + 3 * i] & 15;
            c.tq = p[8 + 3 * i] & 3;
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4) return fail("bad sampling factor");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        mcux = (W + 8 * hmax - 1) / (8 * hmax);
        mcuy = (H + 8 * vmax - 1) / (8 * vmax);
        for (Comp &c : comps) {
            c.w = (W * c.h + hmax - 1) / hmax;
            c.hgt = (H * c.v + vmax - 1) / vmax;
            c.bw = (c.w + 7) / 8;
            c.bh = (c.hgt + 7) / 8;
            c.bstride = mcux * c.h;
            c.brows = mcuy * c.v;
            c.coef.assign((size_t)c.bstride * c.brows * 64, 0);
        }
        progressive = prog;
        frame = true;
        return true;
    }

    bool dqt(const uint8_t *p, int len) {
        while (len > 0) {
            int pq = p[0] >> 4, tq = p[0] & 3;
            int need = 1 + 64 * (pq ? 2 : 1);
            if (len < need) return fail("short DQT");
            for (int k = 0; k < 64; k++)
                qt[tq][ZIGZAG[k]] = pq ? (uint16_t)((p[1 + 2 * k] << 8) | p[2 + 2 * k]) : p[1 + k];
            p += need;
            len -= need;
        }
        return true;
    }

    bool dht(const uint8_t *p, int len) {
        while (len > 0) {
            if (len < 17) return fail("short DHT");
            int tc = p[0] >> 4, th = p[0] & 3;
            int nv = 0;
            for (int i = 0; i < 16; i++) nv += p[1 + i];
            if (nv > 256 || len < 17 + nv) return fail("bad DHT");
            Huff &h = tc ? ac[th] : dc[th];
            if (!h.build(p + 1, p + 17, nv)) return fail("bad Huffman code lengths");
            p += 17 + nv;
            len -= 17 + nv;
        }
        return true;
    }

    int16_t *block(Comp &c, int bx, int by) { return &c.coef[((size_t)by * c.bstride + bx) * 64]; }

    // One block of a scan (sequential, or one progressive pass).
    bool decode_block(Bits &b, Comp &c, int16_t *blk, int ss, int se, int ah, int al) {
        if (!progressive) {
            int t = b.decode(dc[c.td]);
            if (t < 0 || t > 11) return fail("bad DC code");
            c.pred += b.extend(t);
            blk[0] = (int16_t)c.pred;
            for (int k = 1; k < 64;) {
                int rs = b.decode(ac[c.ta]);
                if (rs < 0) return fail("bad AC code");
                int r = rs >> 4, s = rs & 15;
                if (s) {
                    k += r;
                    if (k > 63) return fail("AC run past the block");
                    blk[ZIGZAG[k]] = (int16_t)b.extend(s);
                    k++;
                } else {
                    if (r != 15) break;
                    k += 16;
                }
            }
            return true;
        }
        if (ss == 0) {                       // DC scans
            if (ah == 0) {
                int t = b.decode(dc[c.td]);
                if (t < 0 || t > 11) return fail("bad DC code");
                c.pred += b.extend(t);
                blk[0] = (int16_t)(c.pred * (1 << al));
            } else if (b.get(1)) {
                blk[0] = (int16_t)(blk[0] | (1 << al));
            }
            return true;
        }
        if (ah == 0) {                       // AC first pass (G.1.2.2)
            if (eobrun > 0) {
                eobrun--;
                return true;
            }
            for (int k = ss; k <= se;) {
                int rs = b.decode(ac[c.ta]);
                if (rs < 0) return fail("bad AC code");
                int r = rs >> 4, s = rs & 15;
                if (s) {
                    k += r;
                    if (k > 63) return fail("AC run past the band");
                    blk[ZIGZAG[k]] = (int16_t)(b.extend(s) * (1 << al));
                    k++;
                } else {
                    if (r < 15) {
                        eobrun = (1 << r) - 1;
                        if (r) eobrun += (int)b.get(r);
                        break;
                    }
                    k += 16;
                }
            }
            return true;
        }
        // AC refinement pass (G.1.2.3): coefficients already nonzero get a
        // correction bit; newly nonzero ones get +-2^al at the position
        // after `r` still-zero coefficients
        int p1 = 1 << al, m1 = -(1 << al);
        int k = ss;
        auto refine = [&](int16_t *cf) {
            if (b.get(1) && (*cf & p1) == 0) *cf = (int16_t)(*cf >= 0 ? *cf + p1 : *cf + m1);
        };
        if (eobrun == 0) {
            for (; k <= se; k++) {
                int rs = b.decode(ac[c.ta]);
                if (rs < 0) return fail("bad AC code");
                int r = rs >> 4, s = rs & 15;
                int val = 0;
                if (s) {
                    val = b.get(1) ? p1 : m1;
                } else if (r != 15) {
                    eobrun = 1 << r;
                    if (r) eobrun += (int)b.get(r);
                    break;
                }
                while (k <= se) {
                    int16_t *cf = &blk[ZIGZAG[k]];
                    if (*cf != 0) {
                        refine(cf);
                    } else {
                        if (r == 0) break;
                        r--;
                    }
                    k++;
                }
                if (val && k <= se) blk[ZIGZAG[k]] = (int16_t)val;
            }
        }
        if (eobrun > 0) {
            for (; k <= se; k++) {
                int16_t *cf = &blk[ZIGZAG[k]];
                if (*cf != 0) refine(cf);
            }
            eobrun--;
        }
        return true;
    }

    bool sos(const uint8_t *p, int len, const uint8_t *&scan_end) {
        if (!frame) return fail("scan before frame header");
        int n = p[0];
        if (n < 1 || n > 4 || len < 4 + 2 * n) return fail("bad SOS");
        std::vector<Comp *> sc;
        for (int i = 0; i < n; i++) {
            int id = p[1 + 2 * i];
            Comp *c = nullptr;
            for (Comp &x : comps)
                if (x.id == id) c = &x;
            if (!c) return fail("scan names an unknown component");
            c->td = p[2 + 2 * i] >> 4;
            c->ta = p[2 + 2 * i] & 15;
            if (c->td > 3 || c->ta > 3) return fail("bad Huffman table index");
            sc.push_back(c);
        }
        int ss = p[1 + 2 * n], se = p[2 + 2 * n], ah = p[3 + 2 * n] >> 4, al = p[3 + 2 * n] & 15;
        if (!progressive) {
            ss = 0;
            se = 63;
            ah = al = 0;
        } else if (ss > se || se > 63 || (ss == 0 && se != 0) || (ss > 0 && n != 1)) {
            return fail("bad progressive scan parameters");
        }
        for (Comp *c : sc) {
            bool need_dc = !progressive || (ss == 0 && ah == 0);
            bool need_ac = !progressive ? true : ss > 0;
            if (need_dc && !dc[c->td].present) return fail("missing DC Huffman table");
            if (need_ac && !ac[c->ta].present) return fail("missing AC Huffman table");
            c->pred = 0;
        }
        eobrun = 0;
        Bits b;
        b.p = p + len;
        b.end = data + size;
        int units = 0;
        auto unit_done = [&]() -> bool {
            units++;
            if (restart_interval && units % restart_interval == 0) {
                if (!b.restart()) return fail("missing restart marker");
                for (Comp *c : sc) c->pred = 0;
                eobrun = 0;
            }
            return true;
        };
        if (sc.size() == 1) {
            Comp &c = *sc[0];
            int total = c.bw * c.bh;
            for (int i = 0; i < total; i++) {
                int bx = i % c.bw, by = i / c.bw;
                if (!decode_block(b, c, block(c, bx, by), ss, se, ah, al)) return false;
                if (i + 1 < total && !unit_done()) return false;
            }
        } else {
            int total = mcux * mcuy;
            for (int m = 0; m < total; m++) {
                int mx = m % mcux, my = m / mcux;
                for (Comp *cp : sc) {
                    Comp &c = *cp;
                    for (int v = 0; v < c.v; v++)
                        for (int h = 0; h < c.h; h++)
                            if (!decode_block(b, c, block(c, mx * c.h + h, my * c.v + v), ss, se, ah, al)) return false;
                }
                if (m + 1 < total && !unit_done()) return false;
            }
        }
        // the scan's data ends at the next marker that is not a restart
        const uint8_t *q = b.p;
        while (q + 1 < data + size && !(q[0] == 0xFF && q[1] != 0x00 && !(q[1] >= 0xD0 && q[1] <= 0xD7))) q++;
        scan_end = q;
        return true;
    }

    bool parse() {
        if (size < 4 || data[0] != 0xFF || data[1] != 0xD8) return fail("not a JPEG file");
        const uint8_t *p = data + 2, *end = data + size;
        bool scanned = false;
        while (p < end) {
            if (*p != 0xFF) {
                p++;
                continue;
            }
            while (p < end && *p == 0xFF) p++;     // fill bytes
            if (p >= end) break;
            uint8_t m = *p++;
            if (m == 0xD9) break;                 // EOI
            if (m >= 0xD0 && m <= 0xD7) continue;
            if (m == 0x01) continue;
            if (p + 2 > end) return fail("truncated marker segment");
            int len = (p[0] << 8) | p[1];
            if (len < 2 || p + len > end) return fail("truncated marker segment");
            const uint8_t *seg = p + 2;
            int sl = len - 2;
            p += len;
            switch (m) {
                case 0xC0: case 0xC1:
                    if (!sof(seg, sl, false)) return false;
                    break;
                case 0xC2:
                    if (!sof(seg, sl, true)) return false;
                    break;
                case 0xC3: case 0xC5: case 0xC6: case 0xC7: case 0xC9: case 0xCA: case 0xCB: case 0xCD:
                case 0xCE: case 0xCF:
                    return fail("lossless, hierarchical and arithmetic-coded JPEGs are not decoded");
                case 0xC4:
                    if (!dht(seg, sl)) return false;
                    break;
                case 0xDB:
                    if (!dqt(seg, sl)) return false;
                    break;
                case 0xDD:
                    if (sl < 2) return fail("short DRI");
                    restart_interval = (seg[0] << 8) | seg[1];
                    break;
                case 0xEE:
                    if (sl >= 12 && memcmp(seg, "Adobe", 5) == 0) {
                        adobe = true;
                        adobe_transform = seg[11];
                    }
                    break;
                case 0xDA: {
                    const uint8_t *se;
                    if (!sos(seg, sl, se)) return false;
                    p = se;
                    scanned = true;
                    break;
                }
                default:
                    break;                        // APPn, COM, ...
            }
        }
        if (!frame || !scanned) return fail("no image data");
        return true;
    }

    // Dequantize and inverse transform every block into 8-bit planes.
    void planes(std::vector<std::vector<uint8_t>> &out) {
        double cosv[8][8];
        for (int x = 0; x < 8; x++)
            for (int u = 0; u < 8; u++)
                cosv[x][u] = (u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * M_PI / 16.0) * 0.5;
        out.resize(comps.size());
        for (size_t ci = 0; ci < comps.size(); ci++) {
            Comp &c = comps[ci];
            int pw = c.bstride * 8, ph = c.brows * 8;
            out[ci].assign((size_t)pw * ph, 0);
            const uint16_t *q = qt[c.tq];
            for (int by = 0; by < c.brows; by++)
                for (int bx = 0; bx < c.bstride; bx++) {
                    const int16_t *blk = block(c, bx, by);
                    double f[64], t[64];
                    for (int k = 0; k < 64; k++) f[k] = (double)blk[k] * q[k];
                    for (int y = 0; y < 8; y++)            // columns: over v
                        for (int u = 0; u < 8; u++) {
                            double s = 0;
                            for (int v = 0; v < 8; v++) s += cosv[y][v] * f[v * 8 + u];
                            t[y * 8 + u] = s;
                        }
                    for (int y = 0; y < 8; y++)
                        for (int x = 0; x < 8; x++) {
                            double s = 0;
                            for (int u = 0; u < 8; u++) s += cosv[x][u] * t[y * 8 + u];
                            double v = std::floor(s + 128.0 + 0.5);
                            out[ci][(size_t)(by * 8 + y) * pw + bx * 8 + x] = (uint8_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
                        }
                }
        }
    }

    // Component ci at full resolution (W x H).
    void upsample(size_t ci, const std::vector<uint8_t> &pl, std::vector<uint8_t> &full) {
        const Comp &c = comps[ci];
        int pw = c.bstride * 8;
        int fx = hmax / c.h, fy = vmax / c.v;
        full.assign((size_t)W * H, 0);
        bool tri_x = fx == 2 && hmax % c.h == 0, tri_y = fy == 2 && vmax % c.v == 0;
        if ((fx == 1 || tri_x) && (fy == 1 || tri_y) && hmax % c.h == 0 && vmax % c.v == 0) {
            // triangle filter: an output sample lies 1/4 of an input sample
            // from the nearer input, weights 3/4 and 1/4 (edges repeat)
            int cw = c.w, chh = c.hgt;
            std::vector<int> rowbuf((size_t)cw);
            for (int y = 0; y < H; y++) {
                // vertical: 4 x the sample, as 3 x nearer + 1 x farther row
                int sy = fy == 2 ? y / 2 : y;
                for (int x = 0; x < cw; x++) {
                    int v0 = pl[(size_t)std::min(sy, chh - 1) * pw + x];
                    if (fy == 2) {
                        int ny = (y & 1) ? std::min(sy + 1, chh - 1) : std::max(sy - 1, 0);
                        rowbuf[x] = 3 * v0 + pl[(size_t)ny * pw + x];
                    } else {
                        rowbuf[x] = 4 * v0;
                    }
                }
                for (int x = 0; x < W; x++) {
                    int v;
                    if (fx == 2) {
                        int sx = x / 2;
                        int nx = (x & 1) ? std::min(sx + 1, cw - 1) : std::max(sx - 1, 0);
                        v = (3 * rowbuf[std::min(sx, cw - 1)] + rowbuf[nx] + 8) >> 4;
                    } else {
                        v = (rowbuf[std::min(x, cw - 1)] + 2) >> 2;
                    }
                    full[(size_t)y * W + x] = (uint8_t)v;
                }
            }
            return;
        }
        for (int y = 0; y < H; y++)               // replication
            for (int x = 0; x < W; x++) {
                int sx = x * c.h / hmax, sy = y * c.v / vmax;
                full[(size_t)y * W + x] = pl[(size_t)sy * pw + sx];
            }
    }
};

}  // namespace

bool jpeg_decode_rgba8(const uint8_t *data, size_t size, uint32_t &width, uint32_t &height, std::vector<uint32_t> &out,
                       std::string &err) {
    Decoder d;
    d.data = data;
    d.size = size;
    memset(d.qt, 0, sizeof(d.qt));
    if (!d.parse()) {
        err = d.err;
        return false;
    }
    std::vector<std::vector<uint8_t>> pl;
    d.planes(pl);
    std::vector<std::vector<uint8_t>> full(d.comps.size());
    for (size_t i = 0; i < d.comps.size(); i++) d.upsample(i, pl[i], full[i]);
    width = (uint32_t)d.W;
    height = (uint32_t)d.H;
    size_t n = (size_t)d.W * d.H;
    out.resize(n);
    bool rgb = d.comps.size() == 3 &&
               ((d.adobe && d.adobe_transform == 0) ||
                (!d.adobe && d.comps[0].id == 'R' && d.comps[1].id == 'G' && d.comps[2].id == 'B'));
    auto clamp8 = [](double v) -> uint32_t {
        v = std::floor(v + 0.5);
        return (uint32_t)(v < 0 ? 0 : (v > 255 ? 255 : v));
    };
    for (size_t i = 0; i < n; i++) {
        uint32_t r, g, b;
        if (d.comps.size() == 1) {
            r = g = b = full[0][i];
        } else if (rgb) {
            r = full[0][i];
            g = full[1][i];
            b = full[2][i];
        } else {
            double y = full[0][i], cb = full[1][i] - 128.0, cr = full[2][i] - 128.0;
            r = clamp8(y + 1.402 * cr);
            g = clamp8(y - 0.344136 * cb - 0.714136 * cr);
            b = clamp8(y + 1.772 * cb);
        }
        out[i] = (r << 24) | (g << 16) | (b << 8) | 0xFFu;
    }
    return true;
}

}  // namespace sr
