// Texture file decoding to 8-bit RGBA (the reference calls
// stbi_load(path, ..., 4), src/SceneLoader.cpp:199).
//
// PNG (ISO/IEC 15948): every colour type and bit depth, PLTE/tRNS, Adam7
// interlacing; 16-bit samples keep their high byte and low-bit grey is
// scaled to 0..255, as stbi_load does for 8-bit output.
// JPEG (ITU T.81): baseline and extended sequential Huffman, 8-bit,
// 1 or 3 components, any sampling factors, restart intervals.  Chroma is
// upsampled with the triangular ("fancy") filter for 2x2 and 2x1 sampling,
// nearest otherwise; YCbCr -> RGB with the JFIF matrix.  Decoders differ in
// IDCT and upsampling rounding, so JPEG texels can differ from stb_image's by
// a few units; progressive JPEG is refused with an error.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>

#include "scene.h"

namespace rtxh {

namespace {

bool readBytes(const std::string &path, std::vector<uint8_t> &out) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    std::ostringstream ss;
    ss << f.rdbuf();
    std::string s = ss.str();
    out.assign(s.begin(), s.end());
    return true;
}

uint32_t be32(const uint8_t *p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// ------------------------------------------------------------------- PNG

int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

// Undoes the per-row filters of one (sub)image of w x h pixels in place;
// `in` holds (1 + rowBytes) * h bytes.  Returns false on a bad filter type.
bool unfilter(const uint8_t *in, uint8_t *out, size_t rowBytes, uint32_t h, int bpp) {
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t *src = in + y * (rowBytes + 1);
        uint8_t *row = out + y * rowBytes;
        const uint8_t *prev = y ? row - rowBytes : nullptr;
        const int type = src[0];
        ++src;
        for (size_t i = 0; i < rowBytes; ++i) {
            int a = i >= (size_t)bpp ? row[i - bpp] : 0;
            int b = prev ? prev[i] : 0;
            int c = (prev && i >= (size_t)bpp) ? prev[i - bpp] : 0;
            int x = src[i];
            switch (type) {
            case 0: break;
            case 1: x += a; break;
            case 2: x += b; break;
            case 3: x += (a + b) >> 1; break;
            case 4: x += paeth(a, b, c); break;
            default: return false;
            }
            row[i] = (uint8_t)x;
        }
    }
    return true;
}

bool decodePng(const std::vector<uint8_t> &f, Texture &t, std::string &err) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return err = "not a PNG file", false;
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, palette, trns;
    bool haveHeader = false, ended = false;
    size_t p = 8;
    while (p + 12 <= f.size() && !ended) {
        uint32_t len = be32(&f[p]);
        if (len > f.size() - p - 12) return err = "truncated PNG chunk", false;
        const uint8_t *type = &f[p + 4], *data = &f[p + 8];
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len < 13) return err = "bad IHDR", false;
            w = be32(data);
            h = be32(data + 4);
            depth = data[8];
            ctype = data[9];
            interlace = data[12];
            if (data[10] != 0 || data[11] != 0 || interlace > 1) return err = "unsupported PNG compression/filter/interlace", false;
            haveHeader = true;
        } else if (!std::memcmp(type, "PLTE", 4)) {
            palette.assign(data, data + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            trns.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            ended = true;
        } else if (!(type[0] & 0x20)) {
            return err = "unknown critical PNG chunk", false;
        }
        p += 12 + len;
    }
    if (!haveHeader || w == 0 || h == 0 || w > (1u << 24) || h > (1u << 24)) return err = "bad PNG header", false;
    int channels;
    switch (ctype) {
    case 0: channels = 1; break;
    case 2: channels = 3; break;
    case 3: channels = 1; break;
    case 4: channels = 2; break;
    case 6: channels = 4; break;
    default: return err = "bad PNG colour type", false;
    }
    const bool depthOk = depth == 8 || depth == 16 || ((ctype == 0 || ctype == 3) && (depth == 1 || depth == 2 || depth == 4));
    if (!depthOk || (ctype == 3 && depth == 16)) return err = "bad PNG bit depth", false;
    if (ctype == 3 && (palette.empty() || palette.size() % 3)) return err = "missing PNG palette", false;

    const int bitsPerPixel = channels * depth;
    const int bpp = std::max(1, bitsPerPixel / 8);
    // passes: Adam7 or one full image
    struct Pass {
        uint32_t x0, y0, dx, dy;
    };
    static const Pass adam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                  {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    static const Pass whole[1] = {{0, 0, 1, 1}};
    const Pass *passes = interlace ? adam7 : whole;
    const int nPasses = interlace ? 7 : 1;
    size_t expected = 0;
    for (int i = 0; i < nPasses; ++i) {
        uint64_t pw = (w - passes[i].x0 + passes[i].dx - 1) / passes[i].dx;
        uint64_t ph = (h - passes[i].y0 + passes[i].dy - 1) / passes[i].dy;
        if (w <= passes[i].x0 || h <= passes[i].y0) continue;
        expected += (size_t)((pw * bitsPerPixel + 7) / 8 + 1) * ph;
    }
    std::vector<uint8_t> raw(expected);
    uLongf rawLen = (uLongf)expected;
    int zr = uncompress(raw.data(), &rawLen, idat.data(), (uLong)idat.size());
    if (zr != Z_OK || rawLen != expected) return err = "corrupt PNG image data", false;

    // samples of the whole image at their native depth, 16-bit as 0..65535
    std::vector<uint16_t> samples((size_t)w * h * channels);
    size_t off = 0;
    for (int i = 0; i < nPasses; ++i) {
        if (w <= passes[i].x0 || h <= passes[i].y0) continue;
        uint32_t pw = (w - passes[i].x0 + passes[i].dx - 1) / passes[i].dx;
        uint32_t ph = (h - passes[i].y0 + passes[i].dy - 1) / passes[i].dy;
        size_t rowBytes = ((size_t)pw * bitsPerPixel + 7) / 8;
        std::vector<uint8_t> img(rowBytes * ph);
        if (!unfilter(raw.data() + off, img.data(), rowBytes, ph, bpp)) return err = "bad PNG filter type", false;
        off += (rowBytes + 1) * ph;
        for (uint32_t y = 0; y < ph; ++y)
            for (uint32_t x = 0; x < pw; ++x)
                for (int c = 0; c < channels; ++c) {
                    const uint8_t *row = img.data() + y * rowBytes;
                    uint32_t v;
                    if (depth == 16) {
                        size_t b = ((size_t)x * channels + c) * 2;
                        v = (uint32_t)row[b] << 8 | row[b + 1];
                    } else if (depth == 8) {
                        v = row[(size_t)x * channels + c];
                    } else {
                        size_t bit = (size_t)x * depth;
                        v = (row[bit / 8] >> (8 - depth - bit % 8)) & ((1u << depth) - 1);
                    }
                    size_t X = passes[i].x0 + (size_t)x * passes[i].dx, Y = passes[i].y0 + (size_t)y * passes[i].dy;
                    samples[(Y * w + X) * channels + c] = (uint16_t)v;
                }
    }

    t.width = w;
    t.height = h;
    t.rgba.assign((size_t)w * h * 4, 255);
    const int scale = depth == 1 ? 0xFF : depth == 2 ? 0x55 : depth == 4 ? 0x11 : 1;
    auto to8 = [&](uint32_t v) -> uint8_t { return depth == 16 ? (uint8_t)(v >> 8) : (uint8_t)(v * scale); };
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint16_t *s = &samples[i * channels];
        uint8_t *o = &t.rgba[i * 4];
        switch (ctype) {
        case 0:
            o[0] = o[1] = o[2] = to8(s[0]);
            if (trns.size() >= 2 && s[0] == ((uint32_t)trns[0] << 8 | trns[1])) o[3] = 0;
            break;
        case 2:
            for (int c = 0; c < 3; ++c) o[c] = to8(s[c]);
            if (trns.size() >= 6 && s[0] == ((uint32_t)trns[0] << 8 | trns[1]) &&
                s[1] == ((uint32_t)trns[2] << 8 | trns[3]) && s[2] == ((uint32_t)trns[4] << 8 | trns[5]))
                o[3] = 0;
            break;
        case 3: {
            uint32_t k = s[0];
            if ((size_t)k * 3 + 2 >= palette.size()) return err = "PNG palette index out of range", false;
            for (int c = 0; c < 3; ++c) o[c] = palette[k * 3 + c];
            o[3] = k < trns.size() ? trns[k] : 255;
            break;
        }
        case 4:
            o[0] = o[1] = o[2] = to8(s[0]);
            o[3] = to8(s[1]);
            break;
        case 6:
            for (int c = 0; c < 4; ++c) o[c] = to8(s[c]);
            break;
        }
    }
    return true;
}

// ------------------------------------------------------------------ JPEG

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,  12, 19, 26, 33, 40, 48,
                         41, 34, 27, 20, 13, 6,  7,  14, 21, 28, 35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23,
                         30, 37, 44, 51, 58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

struct Huffman {
    // canonical code: for each length, the first code and the index of its first symbol
    int minCode[17], maxCode[17], valPtr[17];
    uint8_t values[256];
    bool present = false;
};

struct Component {
    int id, h, v, tq, td = 0, ta = 0;
    int bw, bh;                 // blocks per line / column (padded to the MCU grid)
    std::vector<uint8_t> plane; // bw*8 x bh*8 samples
    int dcPred = 0;
};

struct JpegDecoder {
    const std::vector<uint8_t> &f;
    size_t p = 0;
    std::string err;
    uint16_t qt[4][64];
    Huffman hdc[4], hac[4];
    std::vector<Component> comps;
    int width = 0, height = 0, hmax = 1, vmax = 1, mcusX = 0, mcusY = 0, restart = 0;
    bool frame = false;
    // entropy-coded segment reader
    uint32_t bitBuf = 0;
    int bitCount = 0;
    bool hitMarker = false;
    float cosTable[8][8];

    explicit JpegDecoder(const std::vector<uint8_t> &file) : f(file) {
        std::memset(qt, 0, sizeof(qt));
        for (int x = 0; x < 8; ++x)
            for (int u = 0; u < 8; ++u)
                cosTable[x][u] = (float)((u == 0 ? std::sqrt(0.5) : 1.0) * std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0));
    }

    bool fail(const std::string &m) {
        if (err.empty()) err = m;
        return false;
    }
    int u8() { return p < f.size() ? f[p++] : 0; }
    int u16() {
        int a = u8();
        return a << 8 | u8();
    }

    bool readDqt(size_t end) {
        while (p < end) {
            int pq = u8(), tq = pq & 15;
            pq >>= 4;
            if (tq > 3 || pq > 1) return fail("bad DQT");
            for (int i = 0; i < 64; ++i) qt[tq][kZigzag[i]] = (uint16_t)(pq ? u16() : u8());
        }
        return true;
    }
    bool readDht(size_t end) {
        while (p < end) {
            int tc = u8(), th = tc & 15;
            tc >>= 4;
            if (tc > 1 || th > 3) return fail("bad DHT");
            Huffman &hf = tc ? hac[th] : hdc[th];
            int counts[17] = {0}, total = 0;
            for (int l = 1; l <= 16; ++l) total += counts[l] = u8();
            if (total > 256) return fail("bad DHT");
            for (int i = 0; i < total; ++i) hf.values[i] = (uint8_t)u8();
            int code = 0, k = 0;
            for (int l = 1; l <= 16; ++l) {
                hf.valPtr[l] = k;
                hf.minCode[l] = code;
                code += counts[l];
                k += counts[l];
                hf.maxCode[l] = counts[l] ? code - 1 : -1;
                code <<= 1;
            }
            hf.present = true;
        }
        return true;
    }
    bool readSof(int marker) {
        if (marker == 0xC2 || marker == 0xC6 || marker == 0xCA || marker == 0xCE)
            return fail("progressive JPEG is not supported");
        if (marker != 0xC0 && marker != 0xC1) return fail("only sequential Huffman JPEG is supported");
        if (u8() != 8) return fail("only 8-bit JPEG is supported");
        height = u16();
        width = u16();
        int n = u8();
        if (width <= 0 || height <= 0 || (n != 1 && n != 3)) return fail("bad JPEG frame header");
        comps.resize(n);
        for (Component &c : comps) {
            c.id = u8();
            int hv = u8();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = u8();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) return fail("bad JPEG component");
            hmax = std::max(hmax, c.h);
            vmax = std::max(vmax, c.v);
        }
        mcusX = (width + 8 * hmax - 1) / (8 * hmax);
        mcusY = (height + 8 * vmax - 1) / (8 * vmax);
        for (Component &c : comps) {
            c.bw = mcusX * c.h;
            c.bh = mcusY * c.v;
            c.plane.assign((size_t)c.bw * 8 * c.bh * 8, 0);
        }
        frame = true;
        return true;
    }

    void resetBits() {
        bitBuf = 0;
        bitCount = 0;
        hitMarker = false;
    }
    // Fills the bit buffer from the entropy-coded data (0xFF00 is a literal
    // 0xFF; any other marker ends the segment and feeds zeros).
    void fill() {
        while (bitCount <= 24) {
            int b = 0;
            if (!hitMarker && p < f.size()) {
                b = f[p];
                if (b == 0xFF) {
                    int nx = p + 1 < f.size() ? f[p + 1] : 0;
                    if (nx == 0) {
                        p += 2;
                    } else {
                        hitMarker = true;
                        b = 0;
                    }
                } else {
                    ++p;
                }
            }
            bitBuf |= (uint32_t)b << (24 - bitCount);
            bitCount += 8;
        }
    }
    int bits(int n) {
        if (n == 0) return 0;
        fill();
        int v = (int)(bitBuf >> (32 - n));
        bitBuf <<= n;
        bitCount -= n;
        return v;
    }
    int decodeSymbol(const Huffman &hf) {
        fill();
        int code = 0;
        for (int l = 1; l <= 16; ++l) {
            code = code << 1 | (int)(bitBuf >> 31);
            bitBuf <<= 1;
            --bitCount;
            if (hf.maxCode[l] >= 0 && code <= hf.maxCode[l] && code >= hf.minCode[l])
                return hf.values[hf.valPtr[l] + code - hf.minCode[l]];
        }
        return -1;
    }
    static int extend(int v, int n) { return n && v < (1 << (n - 1)) ? v - (1 << n) + 1 : v; }

    bool decodeBlock(Component &c, int bx, int by) {
        const Huffman &dc = hdc[c.td], &ac = hac[c.ta];
        if (!dc.present || !ac.present) return fail("missing Huffman table");
        int coef[64] = {0};
        int t = decodeSymbol(dc);
        if (t < 0 || t > 11) return fail("corrupt JPEG data");
        c.dcPred += extend(bits(t), t);
        coef[0] = c.dcPred * qt[c.tq][0];
        for (int k = 1; k < 64;) {
            int rs = decodeSymbol(ac);
            if (rs < 0) return fail("corrupt JPEG data");
            int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (r != 15) break; // end of block
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) return fail("corrupt JPEG data");
            coef[kZigzag[k]] = extend(bits(s), s) * qt[c.tq][kZigzag[k]];
            ++k;
        }
        // separable 8x8 inverse DCT: f(x,y) = 1/4 sum C(u)C(v) F(u,v) cos.. cos..
        float tmp[64];
        for (int v = 0; v < 8; ++v)
            for (int x = 0; x < 8; ++x) {
                float s = 0;
                for (int u = 0; u < 8; ++u) s += cosTable[x][u] * (float)coef[v * 8 + u];
                tmp[v * 8 + x] = s;
            }
        uint8_t *dst = c.plane.data() + (size_t)by * 8 * c.bw * 8 + (size_t)bx * 8;
        for (int y = 0; y < 8; ++y)
            for (int x = 0; x < 8; ++x) {
                float s = 0;
                for (int v = 0; v < 8; ++v) s += cosTable[y][v] * tmp[v * 8 + x];
                int val = (int)std::lround(s * 0.25f + 128.0f);
                dst[(size_t)y * c.bw * 8 + x] = (uint8_t)std::min(255, std::max(0, val));
            }
        return true;
    }

    bool readScan() {
        int n = u8();
        std::vector<Component *> sc;
        for (int i = 0; i < n; ++i) {
            int id = u8(), tt = u8();
            Component *c = nullptr;
            for (Component &k : comps)
                if (k.id == id) c = &k;
            if (!c) return fail("JPEG scan names an unknown component");
            c->td = tt >> 4;
            c->ta = tt & 15;
            if (c->td > 3 || c->ta > 3) return fail("bad JPEG scan header");
            sc.push_back(c);
        }
        int ss = u8(), se = u8(), a = u8();
        if (ss != 0 || se != 63 || a != 0) return fail("progressive JPEG is not supported");
        for (Component *c : sc) c->dcPred = 0;
        resetBits();
        int untilRestart = restart;
        if (sc.size() == 1) { // non-interleaved: the component's own block grid
            Component &c = *sc[0];
            int bw = (width * c.h / hmax + 7) / 8, bh = (height * c.v / vmax + 7) / 8;
            for (int by = 0; by < bh; ++by)
                for (int bx = 0; bx < bw; ++bx) {
                    if (restart && untilRestart-- == 0 && !doRestart(sc, untilRestart)) return false;
                    if (!decodeBlock(c, bx, by)) return false;
                }
        } else {
            for (int my = 0; my < mcusY; ++my)
                for (int mx = 0; mx < mcusX; ++mx) {
                    if (restart && untilRestart-- == 0 && !doRestart(sc, untilRestart)) return false;
                    for (Component *c : sc)
                        for (int v = 0; v < c->v; ++v)
                            for (int h = 0; h < c->h; ++h)
                                if (!decodeBlock(*c, mx * c->h + h, my * c->v + v)) return false;
                }
        }
        // skip to the next marker
        while (p + 1 < f.size() && !(f[p] == 0xFF && f[p + 1] != 0 && !(f[p + 1] >= 0xD0 && f[p + 1] <= 0xD7))) ++p;
        return true;
    }
    bool doRestart(const std::vector<Component *> &sc, int &untilRestart) {
        // find RSTn
        while (p + 1 < f.size() && !(f[p] == 0xFF && f[p + 1] >= 0xD0 && f[p + 1] <= 0xD7)) ++p;
        if (p + 1 >= f.size()) return fail("missing JPEG restart marker");
        p += 2;
        resetBits();
        for (Component *c : sc) c->dcPred = 0;
        untilRestart = restart - 1;
        return true;
    }

    bool decode(Texture &t) {
        if (u16() != 0xFFD8) return fail("not a JPEG file");
        bool scanned = false;
        while (p < f.size()) {
            if (u8() != 0xFF) return fail("bad JPEG marker");
            int m = u8();
            while (m == 0xFF) m = u8();
            if (m == 0xD9) break;
            if (m >= 0xD0 && m <= 0xD7) continue;
            size_t len = (size_t)u16();
            if (len < 2 || p + len - 2 > f.size()) return fail("truncated JPEG segment");
            size_t end = p + len - 2;
            bool ok = true;
            if (m == 0xDB) ok = readDqt(end);
            else if (m == 0xC4) ok = readDht(end);
            else if (m == 0xDD) restart = u16();
            else if (m >= 0xC0 && m <= 0xCF && m != 0xC4 && m != 0xC8 && m != 0xCC) ok = readSof(m);
            else if (m == 0xDA) {
                if (!frame) return fail("JPEG scan before the frame header");
                ok = readScan();
                scanned = true;
                if (!ok) return false;
                continue;
            }
            if (!ok) return false;
            p = end;
        }
        if (!scanned) return fail("JPEG has no image data");
        toRgba(t);
        return true;
    }

    // Chroma plane sample at full resolution: triangular 2x (fancy)
    // upsampling along each axis that is subsampled by 2, nearest otherwise.
    void upsample(const Component &c, std::vector<float> &out) const {
        const int pw = c.bw * 8, ph = c.bh * 8;
        const int fx = hmax / c.h, fy = vmax / c.v;
        out.assign((size_t)width * height, 0.0f);
        auto at = [&](int x, int y) {
            x = std::min(std::max(x, 0), pw - 1);
            y = std::min(std::max(y, 0), ph - 1);
            return (float)c.plane[(size_t)y * pw + x];
        };
        const int cw = (width * c.h + hmax - 1) / hmax, ch = (height * c.v + vmax - 1) / vmax; // real extent
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                float v;
                if ((fx == 1 || fx == 2) && (fy == 1 || fy == 2)) {
                    // position in the plane, and the neighbour towards the pixel centre
                    int sx = x / fx, sy = y / fy;
                    int nx = fx == 2 ? (x % 2 ? sx + 1 : sx - 1) : sx;
                    int ny = fy == 2 ? (y % 2 ? sy + 1 : sy - 1) : sy;
                    nx = std::min(std::max(nx, 0), cw - 1);
                    ny = std::min(std::max(ny, 0), ch - 1);
                    float wx = fx == 2 ? 0.75f : 1.0f, wy = fy == 2 ? 0.75f : 1.0f;
                    v = wx * wy * at(sx, sy) + (1 - wx) * wy * at(nx, sy) + wx * (1 - wy) * at(sx, ny) +
                        (1 - wx) * (1 - wy) * at(nx, ny);
                } else {
                    v = at(x / fx, y / fy);
                }
                out[(size_t)y * width + x] = v;
            }
    }

    void toRgba(Texture &t) const {
        t.width = (uint32_t)width;
        t.height = (uint32_t)height;
        t.rgba.assign((size_t)width * height * 4, 255);
        const Component &Y = comps[0];
        const int yw = Y.bw * 8;
        auto sampleY = [&](int x, int y) -> float {
            if (Y.h == hmax && Y.v == vmax) return (float)Y.plane[(size_t)y * yw + x];
            return (float)Y.plane[(size_t)(y * Y.v / vmax) * yw + x * Y.h / hmax];
        };
        auto clamp8 = [](float v) { return (uint8_t)std::min(255.0f, std::max(0.0f, std::round(v))); };
        if (comps.size() == 1) {
            for (int y = 0; y < height; ++y)
                for (int x = 0; x < width; ++x) {
                    uint8_t g = (uint8_t)sampleY(x, y);
                    uint8_t *o = &t.rgba[((size_t)y * width + x) * 4];
                    o[0] = o[1] = o[2] = g;
                }
            return;
        }
        std::vector<float> cb, cr;
        upsample(comps[1], cb);
        upsample(comps[2], cr);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x) {
                size_t i = (size_t)y * width + x;
                float yy = sampleY(x, y), b = cb[i] - 128.0f, r = cr[i] - 128.0f;
                uint8_t *o = &t.rgba[i * 4];
                o[0] = clamp8(yy + 1.402f * r);
                o[1] = clamp8(yy - 0.344136f * b - 0.714136f * r);
                o[2] = clamp8(yy + 1.772f * b);
            }
    }
};

} // namespace

bool loadImageRgba8(const std::string &path, Texture &out, std::string &err) {
    std::vector<uint8_t> f;
    if (!readBytes(path, f)) {
        err = "cannot open " + path;
        return false;
    }
    out = Texture();
    if (f.size() >= 8 && f[0] == 137 && f[1] == 'P' && f[2] == 'N' && f[3] == 'G') return decodePng(f, out, err);
    if (f.size() >= 3 && f[0] == 0xFF && f[1] == 0xD8) {
        JpegDecoder d(f);
        if (!d.decode(out)) {
            err = d.err;
            return false;
        }
        return true;
    }
    err = "unsupported image format (PNG and JPEG are read)";
    return false;
}

} // namespace rtxh
