// PNG reader (all colour types and bit depths 1/2/4/8/16, Adam7 interlace,
// palette and tRNS transparency) and the 8-bit gray PNG writer of the map
// output (src/main.cpp:102-109).  zlib inflates the IDAT stream.
//
// Sample conventions follow the reference's loader (stb_image 2.30, which
// src/main.cpp:37-57 calls with 4 channels forced), so a file decodes to the
// same RGBA bytes:
//   - gray samples of depth < 8 are scaled to 0..255 by 0xff, 0x55, 0x11
//     (depths 1, 2, 4); palette indices are not scaled;
//   - 16-bit samples keep their high byte;
//   - a tRNS colour key makes matching pixels transparent (alpha 0, else
//     255), compared on the scaled 8-bit samples for depths <= 8 (the key is
//     taken modulo 256 and scaled the same way) and on the 16-bit samples for
//     depth 16;
//   - palette tRNS entries give the palette's alpha.
#pragma once

#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>

#include "image_common.hpp"

namespace iqmcli {

inline int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

/* Reverses the per-row filters of one (sub)image of `rows` rows of `stride`
 * bytes, in place; `src` holds rows of 1 filter byte + stride bytes, `bpp`
 * is the filter distance in bytes (>= 1). */
inline void png_unfilter(const uint8_t* src, uint8_t* dst, size_t stride, uint32_t rows, size_t bpp) {
    std::vector<uint8_t> zero(stride, 0);
    for (uint32_t y = 0; y < rows; ++y) {
        const uint8_t ft = src[y * (stride + 1)];
        const uint8_t* s = src + y * (stride + 1) + 1;
        uint8_t* cur = dst + (size_t)y * stride;
        const uint8_t* prev = y ? cur - stride : zero.data();
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = s[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) >> 1; break;
                case 4: v += paeth(a, b, c); break;
                default: fail("Corrupt PNG");
            }
            cur[i] = (uint8_t)v;
        }
    }
}

inline InputImage decode_png(const std::vector<uint8_t>& f) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) fail("Corrupt PNG");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    bool have_ihdr = false;
    std::vector<uint8_t> idat, plte;
    std::vector<uint8_t> trns;
    size_t pos = 8;
    for (;;) {
        if (pos + 12 > f.size()) fail("Corrupt PNG");
        const uint32_t len = be32(&f[pos]);
        if (len > (1u << 30) || pos + 12 + (size_t)len > f.size()) fail("Corrupt PNG");
        const char* type = (const char*)&f[pos + 4];
        const uint8_t* d = &f[pos + 8];
        const bool first = !have_ihdr;
        if (!std::memcmp(type, "IHDR", 4)) {
            if (!first) fail("Corrupt PNG");
            if (len != 13) fail("Corrupt PNG");
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            if (d[10] || d[11]) fail("Corrupt PNG");
            interlace = d[12];
            if (interlace > 1) fail("Corrupt PNG");
            if (w == 0 || h == 0) fail("Corrupt PNG");
            if (w > kMaxSide || h > kMaxSide) fail("Very large image (corrupt?)");
            const bool ok = (ctype == 0 && (depth == 1 || depth == 2 || depth == 4 || depth == 8 || depth == 16)) ||
                            (ctype == 3 && (depth == 1 || depth == 2 || depth == 4 || depth == 8)) ||
                            ((ctype == 2 || ctype == 4 || ctype == 6) && (depth == 8 || depth == 16));
            if (!ok) fail("PNG not supported: unsupported color depth");
            have_ihdr = true;
        } else if (first) {
            fail("Corrupt PNG");
        } else if (!std::memcmp(type, "PLTE", 4)) {
            if (len > 768 || len % 3) fail("Corrupt PNG");
            plte.assign(d, d + len);
        } else if (!std::memcmp(type, "tRNS", 4)) {
            if (!idat.empty()) fail("Corrupt PNG");
            if (ctype == 3) {
                if (plte.empty() || len > plte.size() / 3) fail("Corrupt PNG");
            } else {
                if (ctype == 4 || ctype == 6) fail("Corrupt PNG");
                if (len != (ctype == 0 ? 2u : 6u)) fail("Corrupt PNG");
            }
            trns.assign(d, d + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            if (ctype == 3 && plte.empty()) fail("Corrupt PNG");
            idat.insert(idat.end(), d, d + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            break;
        } else if (!(type[0] & 0x20)) {
            fail(std::string(type, 4) + " PNG chunk not known");
        }
        pos += 12 + (size_t)len;
    }
    if (idat.empty()) fail("Corrupt PNG");
    InputImage img = make_rgba(w, h);
    const int comps = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : 4;
    const size_t bits = (size_t)comps * depth;
    const size_t bpp = std::max<size_t>(1, bits / 8);
    /* Adam7 passes (x0, y0, dx, dy), or one pass covering the image */
    static const int a7[7][4] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4},
                                 {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};
    const int npass = interlace ? 7 : 1;
    size_t raw_len = 0;
    uint32_t pw[7], ph[7];
    for (int p = 0; p < npass; ++p) {
        const int x0 = interlace ? a7[p][0] : 0, y0 = interlace ? a7[p][1] : 0;
        const int dx = interlace ? a7[p][2] : 1, dy = interlace ? a7[p][3] : 1;
        pw[p] = w > (uint32_t)x0 ? (w - x0 + dx - 1) / dx : 0;
        ph[p] = h > (uint32_t)y0 ? (h - y0 + dy - 1) / dy : 0;
        if (pw[p] && ph[p]) raw_len += mul_checked((mul_checked(pw[p], bits) + 7) / 8 + 1, ph[p]);
    }
    std::vector<uint8_t> raw(raw_len);
    {
        z_stream zs{};
        if (inflateInit(&zs) != Z_OK) fail("Corrupt PNG");
        zs.next_in = idat.data();
        zs.avail_in = (uInt)idat.size();
        zs.next_out = raw.data();
        zs.avail_out = (uInt)raw.size();
        const int rc = inflate(&zs, Z_FINISH);
        const size_t got = zs.total_out;
        inflateEnd(&zs);
        /* a stream with more bytes than the image needs is accepted, as the
         * reference does; fewer is corrupt */
        if ((rc != Z_STREAM_END && rc != Z_BUF_ERROR && rc != Z_OK) || got != raw.size()) fail("Corrupt PNG");
    }
    static const uint8_t gray_scale[9] = {0, 0xff, 0x55, 0, 0x11, 0, 0, 0, 0x01};
    const int scale = ctype == 0 ? gray_scale[depth] : 1;
    /* colour key of tRNS for gray/RGB: 8-bit compare key or 16-bit key */
    const bool key = !trns.empty() && ctype != 3;
    uint8_t key8[3] = {0, 0, 0};
    uint16_t key16[3] = {0, 0, 0};
    if (key)
        for (int k = 0; k < comps; ++k) {
            const uint16_t v = (uint16_t)(trns[2 * k] << 8 | trns[2 * k + 1]);
            key16[k] = v;
            key8[k] = (uint8_t)((v & 255) * gray_scale[depth]);
        }
    size_t off = 0;
    std::vector<uint8_t> rows;
    for (int p = 0; p < npass; ++p) {
        if (!pw[p] || !ph[p]) continue;
        const size_t stride = (pw[p] * bits + 7) / 8;
        rows.assign(stride * ph[p], 0);
        png_unfilter(raw.data() + off, rows.data(), stride, ph[p], bpp);
        off += (stride + 1) * ph[p];
        const int x0 = interlace ? a7[p][0] : 0, y0 = interlace ? a7[p][1] : 0;
        const int dx = interlace ? a7[p][2] : 1, dy = interlace ? a7[p][3] : 1;
        for (uint32_t y = 0; y < ph[p]; ++y) {
            const uint8_t* r = &rows[(size_t)y * stride];
            for (uint32_t x = 0; x < pw[p]; ++x) {
                uint16_t s16[4] = {0, 0, 0, 0};
                uint8_t s8[4] = {0, 0, 0, 0};
                for (int k = 0; k < comps; ++k) {
                    if (depth == 16) {
                        const size_t b = ((size_t)x * comps + k) * 2;
                        s16[k] = (uint16_t)(r[b] << 8 | r[b + 1]);
                        s8[k] = r[b];
                    } else if (depth == 8) {
                        s8[k] = r[(size_t)x * comps + k];
                    } else {
                        const size_t bit = (size_t)x * depth;
                        const int v = (r[bit >> 3] >> (8 - depth - (int)(bit & 7))) & ((1 << depth) - 1);
                        s8[k] = (uint8_t)(v * scale);
                    }
                }
                uint8_t* o = &img.data[(((size_t)(y0 + y * dy)) * w + x0 + (size_t)x * dx) * 4];
                if (ctype == 3) {
                    const size_t i = s8[0];
                    if (i * 3 + 2 >= plte.size()) fail("Corrupt PNG");
                    o[0] = plte[i * 3];
                    o[1] = plte[i * 3 + 1];
                    o[2] = plte[i * 3 + 2];
                    o[3] = i < trns.size() ? trns[i] : 255;
                    continue;
                }
                int n = comps;
                if (key) {
                    bool hit = true;
                    for (int k = 0; k < comps; ++k)
                        hit = hit && (depth == 16 ? s16[k] == key16[k] : s8[k] == key8[k]);
                    s8[comps] = hit ? 0 : 255;
                    n = comps + 1;
                }
                to_rgba(s8, n, o);
            }
        }
    }
    return img;
}

inline void put32(std::vector<uint8_t>& v, uint32_t x) {
    for (int s = 24; s >= 0; s -= 8) v.push_back((uint8_t)(x >> s));
}

inline void chunk(std::vector<uint8_t>& out, const char* type, const std::vector<uint8_t>& data) {
    put32(out, (uint32_t)data.size());
    const size_t start = out.size();
    out.insert(out.end(), type, type + 4);
    out.insert(out.end(), data.begin(), data.end());
    uLong crc = crc32(0L, Z_NULL, 0);
    crc = crc32(crc, &out[start], (uInt)(out.size() - start));
    put32(out, (uint32_t)crc);
}

/* 8-bit gray PNG writer; returns false on failure (stbi_write_png's contract) */
inline bool write_png_gray(const std::string& path, int w, int h, const unsigned char* px) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)(w + 1) * h);
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), px + (size_t)y * w, px + (size_t)(y + 1) * w);
    }
    uLongf clen = compressBound((uLong)raw.size());
    std::vector<uint8_t> z(clen);
    if (compress2(z.data(), &clen, raw.data(), (uLong)raw.size(), 6) != Z_OK) return false;
    z.resize(clen);
    std::vector<uint8_t> out = {137, 80, 78, 71, 13, 10, 26, 10};
    std::vector<uint8_t> ihdr;
    put32(ihdr, (uint32_t)w);
    put32(ihdr, (uint32_t)h);
    ihdr.insert(ihdr.end(), {8, 0, 0, 0, 0});
    chunk(out, "IHDR", ihdr);
    chunk(out, "IDAT", z);
    chunk(out, "IEND", {});
    FILE* fp = std::fopen(path.c_str(), "wb");
    if (!fp) return false;
    const bool ok = std::fwrite(out.data(), 1, out.size(), fp) == out.size();
    return std::fclose(fp) == 0 && ok;
}

}  // namespace iqmcli
