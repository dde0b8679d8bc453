// Minimal OpenEXR scanline I/O (the reference links OpenEXR for
// CommonOps::writeEXR / readEXR, src/CommonOps.cpp:12-65; the library is
// absent here, so the file format is written directly from the OpenEXR
// file-layout document).  Writer: RGB FLOAT, no compression.  Reader:
// scanline files with NONE / ZIPS / ZIP compression and HALF / FLOAT /
// UINT channels.
#include <zlib.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

#include "scene.h"

namespace rtxh {

namespace {
void put32(std::vector<char> &b, int32_t v) { b.insert(b.end(), (char *)&v, (char *)&v + 4); }
void putf(std::vector<char> &b, float v) { b.insert(b.end(), (char *)&v, (char *)&v + 4); }
void putStr(std::vector<char> &b, const char *s) { b.insert(b.end(), s, s + std::strlen(s) + 1); }
void attrHeader(std::vector<char> &b, const char *name, const char *type, int32_t size) {
    putStr(b, name);
    putStr(b, type);
    put32(b, size);
}

float halfToFloat(uint16_t h) {
    uint32_t sign = (uint32_t)(h >> 15) << 31;
    int exp = (h >> 10) & 0x1f;
    uint32_t mant = h & 0x3ff;
    uint32_t bits;
    if (exp == 0) {
        if (mant == 0) {
            bits = sign;
        } else { // subnormal
            int e = -1;
            do {
                mant <<= 1;
                e++;
            } while (!(mant & 0x400));
            mant &= 0x3ff;
            bits = sign | (uint32_t)(127 - 15 - e) << 23 | (mant << 13);
        }
    } else if (exp == 31) {
        bits = sign | 0x7f800000u | (mant << 13);
    } else {
        bits = sign | (uint32_t)(exp - 15 + 127) << 23 | (mant << 13);
    }
    float f;
    std::memcpy(&f, &bits, 4);
    return f;
}
} // namespace

bool writeExr(const std::string &path, const float *rgba, uint32_t w, uint32_t h, std::string &err) {
    std::vector<char> hdr;
    put32(hdr, 20000630); // magic
    put32(hdr, 2);        // version 2, scanline
    // channels, alphabetical: B, G, R  (FLOAT = 2)
    std::vector<char> ch;
    for (const char *n : {"B", "G", "R"}) {
        putStr(ch, n);
        put32(ch, 2);
        ch.push_back(0); // pLinear
        ch.push_back(0);
        ch.push_back(0);
        ch.push_back(0);
        put32(ch, 1);
        put32(ch, 1);
    }
    ch.push_back(0);
    attrHeader(hdr, "channels", "chlist", (int32_t)ch.size());
    hdr.insert(hdr.end(), ch.begin(), ch.end());
    attrHeader(hdr, "compression", "compression", 1);
    hdr.push_back(0); // NO_COMPRESSION
    attrHeader(hdr, "dataWindow", "box2i", 16);
    put32(hdr, 0); put32(hdr, 0); put32(hdr, (int32_t)w - 1); put32(hdr, (int32_t)h - 1);
    attrHeader(hdr, "displayWindow", "box2i", 16);
    put32(hdr, 0); put32(hdr, 0); put32(hdr, (int32_t)w - 1); put32(hdr, (int32_t)h - 1);
    attrHeader(hdr, "lineOrder", "lineOrder", 1);
    hdr.push_back(0); // INCREASING_Y
    attrHeader(hdr, "pixelAspectRatio", "float", 4);
    putf(hdr, 1.0f);
    attrHeader(hdr, "screenWindowCenter", "v2f", 8);
    putf(hdr, 0.0f); putf(hdr, 0.0f);
    attrHeader(hdr, "screenWindowWidth", "float", 4);
    putf(hdr, 1.0f);
    hdr.push_back(0); // end of header

    const uint64_t lineBytes = 8 + (uint64_t)w * 3 * 4;
    uint64_t offset = hdr.size() + (uint64_t)h * 8;
    std::vector<char> table;
    for (uint32_t y = 0; y < h; ++y) {
        uint64_t o = offset + (uint64_t)y * lineBytes;
        table.insert(table.end(), (char *)&o, (char *)&o + 8);
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    f.write(hdr.data(), (std::streamsize)hdr.size());
    f.write(table.data(), (std::streamsize)table.size());
    std::vector<float> line((size_t)w * 3);
    for (uint32_t y = 0; y < h; ++y) {
        int32_t yy = (int32_t)y, size = (int32_t)(w * 3 * 4);
        f.write((char *)&yy, 4);
        f.write((char *)&size, 4);
        for (int c = 0; c < 3; ++c) { // B, G, R
            int src = 2 - c;
            for (uint32_t x = 0; x < w; ++x) line[(size_t)c * w + x] = rgba[((size_t)y * w + x) * 4 + src];
        }
        f.write((char *)line.data(), (std::streamsize)(line.size() * 4));
    }
    if (!f) {
        err = "write failed: " + path;
        return false;
    }
    return true;
}

bool readExr(const std::string &path, std::vector<float> &rgba, uint32_t &w, uint32_t &h,
             std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return false;
    }
    std::vector<char> data((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    // Every read is checked against the end of the file: a malformed or
    // truncated file is an error, never a read past the buffer.
    size_t p = 0;
    auto need = [&](size_t n) { return p <= data.size() && n <= data.size() - p; };
    auto rd32 = [&](int32_t &v) {
        if (!need(4)) return false;
        std::memcpy(&v, &data[p], 4);
        p += 4;
        return true;
    };
    auto rdStr = [&](size_t limit, std::string &s) {
        const size_t lim = std::min(limit, data.size());
        if (p >= lim) return false;
        const void *z = std::memchr(&data[p], 0, lim - p);
        if (!z) return false;
        s.assign(&data[p], (size_t)((const char *)z - &data[p]));
        p += s.size() + 1;
        return true;
    };
    auto bad = [&](const char *what) {
        err = what;
        return false;
    };
    int32_t magic = 0, version = 0;
    if (!rd32(magic) || magic != 20000630 || !rd32(version)) return bad("not an OpenEXR file");
    if (version & 0x200) return bad("tiled EXR not supported");
    struct Ch {
        std::string name;
        int type;
    };
    std::vector<Ch> chans;
    int compression = -1;
    int32_t dw[4] = {0, 0, -1, -1};
    while (need(1) && data[p] != 0) {
        std::string name, type;
        int32_t size = 0;
        if (!rdStr(data.size(), name) || !rdStr(data.size(), type) || !rd32(size) || size < 0 ||
            !need((size_t)size))
            return bad("truncated EXR header");
        const size_t end = p + (size_t)size;
        if (name == "channels") {
            while (p < end && data[p] != 0) {
                std::string cn;
                int32_t t = 0;
                if (!rdStr(end, cn) || !rd32(t) || p + 12 > end) return bad("bad EXR channel list");
                if (t < 0 || t > 2) return bad("bad EXR channel type");
                p += 4 + 8;
                chans.push_back({cn, t});
            }
        } else if (name == "compression") {
            if (size < 1) return bad("bad EXR compression attribute");
            compression = (unsigned char)data[p];
        } else if (name == "dataWindow") {
            if (size < 16) return bad("bad EXR dataWindow attribute");
            for (int i = 0; i < 4; ++i) rd32(dw[i]);
        }
        p = end;
    }
    if (!need(1)) return bad("truncated EXR header");
    p++; // end of header
    const int64_t w64 = (int64_t)dw[2] - dw[0] + 1, h64 = (int64_t)dw[3] - dw[1] + 1;
    if (w64 <= 0 || h64 <= 0 || chans.empty()) return bad("bad EXR header");
    // 2^28 pixels (4 GiB of RGBA floats) is far above any environment map.
    if (w64 * h64 > (int64_t(1) << 28)) return bad("EXR image too large");
    w = (uint32_t)w64;
    h = (uint32_t)h64;
    int linesPerBlock;
    if (compression == 0 || compression == 2) linesPerBlock = 1;
    else if (compression == 3) linesPerBlock = 16;
    else {
        err = "unsupported EXR compression " + std::to_string(compression);
        return false;
    }
    uint32_t blocks = (h + linesPerBlock - 1) / linesPerBlock;
    if (!need((size_t)blocks * 8)) return bad("truncated EXR");
    std::vector<uint64_t> offsets(blocks);
    std::memcpy(offsets.data(), &data[p], (size_t)blocks * 8);
    std::vector<int> bpp(chans.size());
    size_t pixelBytes = 0;
    for (size_t c = 0; c < chans.size(); ++c) {
        bpp[c] = chans[c].type == 1 ? 2 : 4;
        pixelBytes += bpp[c];
    }
    std::map<std::string, int> target{{"R", 0}, {"G", 1}, {"B", 2}, {"A", 3}};
    rgba.assign((size_t)w * h * 4, 0.0f);
    for (size_t i = 0; i < (size_t)w * h; ++i) rgba[i * 4 + 3] = 1.0f;
    std::vector<unsigned char> raw, tmp;
    for (uint32_t b = 0; b < blocks; ++b) {
        if (offsets[b] > data.size()) return bad("bad EXR block offset");
        p = (size_t)offsets[b];
        int32_t y = 0, size = 0;
        if (!rd32(y) || !rd32(size) || size < 0 || !need((size_t)size)) return bad("truncated EXR block");
        const int64_t y0 = (int64_t)y - dw[1];
        if (y0 < 0 || y0 >= (int64_t)h) return bad("EXR block outside the data window");
        uint32_t lines = std::min<uint32_t>((uint32_t)linesPerBlock, h - (uint32_t)y0);
        size_t expect = (size_t)lines * w * pixelBytes;
        raw.resize(expect);
        if ((size_t)size == expect) {
            std::memcpy(raw.data(), &data[p], expect);
        } else if (compression == 0) {
            return bad("EXR block of the wrong size");
        } else {
            tmp.resize(expect);
            uLongf outLen = (uLongf)expect;
            if (uncompress(tmp.data(), &outLen, (const Bytef *)&data[p], (uLong)size) != Z_OK ||
                outLen != expect)
                return bad("zlib error in EXR block");
            for (size_t i = 1; i < expect; ++i) tmp[i] = (unsigned char)(tmp[i - 1] + tmp[i] - 128);
            size_t half = (expect + 1) / 2;
            for (size_t i = 0; i < expect; ++i)
                raw[i] = (i % 2 == 0) ? tmp[i / 2] : tmp[half + i / 2];
        }
        size_t q = 0;
        for (uint32_t l = 0; l < lines; ++l) {
            size_t row = (size_t)(y0 + l) * w;
            for (size_t c = 0; c < chans.size(); ++c) {
                auto it = target.find(chans[c].name);
                for (uint32_t x = 0; x < w; ++x) {
                    float v;
                    if (chans[c].type == 1) {
                        uint16_t hv;
                        std::memcpy(&hv, &raw[q], 2);
                        v = halfToFloat(hv);
                    } else if (chans[c].type == 2) {
                        std::memcpy(&v, &raw[q], 4);
                    } else {
                        uint32_t u;
                        std::memcpy(&u, &raw[q], 4);
                        v = (float)u;
                    }
                    q += (size_t)bpp[c];
                    if (it != target.end()) rgba[(row + x) * 4 + it->second] = v;
                }
            }
        }
    }
    return true;
}

} // namespace rtxh
