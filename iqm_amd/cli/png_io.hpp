// Minimal PNG reader/writer for the CLI (zlib for the deflate streams).
// Reading forces 4 channels like the reference's stbi_load(..., 4)
// (src/main.cpp:37-57): gray -> R=G=B, missing alpha -> 255.  Supported:
// bit depth 8 or 16 (16 keeps the high byte), colour types 0/2/3/4/6,
// non-interlaced.  Writing emits 8-bit gray (the map output of
// src/main.cpp:102-109).  Binary PPM (P6) and PGM (P5) are read as well.
#pragma once

#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "iqm/iqm.hpp"

namespace iqmcli {

inline uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

inline std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("can't fopen");
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

inline int paeth(int a, int b, int c) {
    int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

inline InputImage decode_png(const std::vector<uint8_t>& f) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) throw std::runtime_error("not a PNG");
    uint32_t w = 0, h = 0;
    int depth = 0, ctype = -1, interlace = 0;
    std::vector<uint8_t> idat, plte, trns;
    size_t pos = 8;
    while (pos + 12 <= f.size()) {
        const uint32_t len = be32(&f[pos]);
        const std::string type((const char*)&f[pos + 4], 4);
        if (pos + 12 + len > f.size()) throw std::runtime_error("truncated PNG chunk");
        const uint8_t* d = &f[pos + 8];
        if (type == "IHDR") {
            w = be32(d);
            h = be32(d + 4);
            depth = d[8];
            ctype = d[9];
            interlace = d[12];
        } else if (type == "PLTE") {
            plte.assign(d, d + len);
        } else if (type == "tRNS") {
            trns.assign(d, d + len);
        } else if (type == "IDAT") {
            idat.insert(idat.end(), d, d + len);
        } else if (type == "IEND") {
            break;
        }
        pos += 12 + len;
    }
    if (w == 0 || h == 0) throw std::runtime_error("PNG without IHDR");
    if (interlace) throw std::runtime_error("interlaced PNG not supported");
    if (depth != 8 && depth != 16) throw std::runtime_error("PNG bit depth must be 8 or 16");
    int comps;
    switch (ctype) {
        case 0: comps = 1; break;
        case 2: comps = 3; break;
        case 3: comps = 1; break;
        case 4: comps = 2; break;
        case 6: comps = 4; break;
        default: throw std::runtime_error("unsupported PNG colour type");
    }
    if (ctype == 3 && depth != 8) throw std::runtime_error("palette PNG must be 8-bit");
    const size_t bpp = comps * (depth / 8);
    const size_t stride = bpp * w;
    std::vector<uint8_t> raw((stride + 1) * h);
    uLongf out_len = (uLongf)raw.size();
    if (uncompress(raw.data(), &out_len, idat.data(), (uLong)idat.size()) != Z_OK || out_len != raw.size())
        throw std::runtime_error("corrupt PNG data");
    std::vector<uint8_t> cur(stride), prev(stride, 0);
    InputImage img{(int)w, (int)h, std::vector<unsigned char>((size_t)w * h * 4)};
    for (uint32_t y = 0; y < h; ++y) {
        const uint8_t ft = raw[y * (stride + 1)];
        const uint8_t* s = &raw[y * (stride + 1) + 1];
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
            int v = s[i];
            switch (ft) {
                case 0: break;
                case 1: v += a; break;
                case 2: v += b; break;
                case 3: v += (a + b) / 2; break;
                case 4: v += paeth(a, b, c); break;
                default: throw std::runtime_error("bad PNG filter");
            }
            cur[i] = (uint8_t)v;
        }
        for (uint32_t x = 0; x < w; ++x) {
            uint8_t px[4] = {0, 0, 0, 255};
            auto comp = [&](int k) { return cur[x * bpp + k * (depth / 8)]; };
            if (ctype == 0) { px[0] = px[1] = px[2] = comp(0); }
            else if (ctype == 2) { px[0] = comp(0); px[1] = comp(1); px[2] = comp(2); }
            else if (ctype == 3) {
                const int i = comp(0);
                if ((size_t)i * 3 + 2 >= plte.size()) throw std::runtime_error("bad PNG palette index");
                px[0] = plte[i * 3]; px[1] = plte[i * 3 + 1]; px[2] = plte[i * 3 + 2];
                if ((size_t)i < trns.size()) px[3] = trns[i];
            } else if (ctype == 4) { px[0] = px[1] = px[2] = comp(0); px[3] = comp(1); }
            else { px[0] = comp(0); px[1] = comp(1); px[2] = comp(2); px[3] = comp(3); }
            std::memcpy(&img.data[((size_t)y * w + x) * 4], px, 4);
        }
        std::swap(cur, prev);
    }
    return img;
}

inline InputImage decode_pnm(const std::vector<uint8_t>& f) {
    size_t pos = 2;
    auto next_int = [&]() {
        while (pos < f.size()) {
            if (f[pos] == '#') { while (pos < f.size() && f[pos] != '\n') ++pos; }
            else if (std::isspace(f[pos])) ++pos;
            else break;
        }
        long v = 0;
        while (pos < f.size() && std::isdigit(f[pos])) v = v * 10 + (f[pos++] - '0');
        return v;
    };
    const bool color = f[1] == '6';
    const long w = next_int(), h = next_int(), maxv = next_int();
    ++pos;
    if (w <= 0 || h <= 0 || maxv != 255) throw std::runtime_error("unsupported PNM header");
    const size_t ch = color ? 3 : 1;
    if (pos + (size_t)w * h * ch > f.size()) throw std::runtime_error("truncated PNM");
    InputImage img{(int)w, (int)h, std::vector<unsigned char>((size_t)w * h * 4)};
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        const uint8_t* s = &f[pos + i * ch];
        img.data[i * 4 + 0] = s[0];
        img.data[i * 4 + 1] = color ? s[1] : s[0];
        img.data[i * 4 + 2] = color ? s[2] : s[0];
        img.data[i * 4 + 3] = 255;
    }
    return img;
}

/* load_image (src/main.cpp:37-57) with the reference's failure message */
inline InputImage load_image(const std::string& filename) {
    try {
        std::vector<uint8_t> f = read_file(filename);
        if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return decode_pnm(f);
        return decode_png(f);
    } catch (const std::exception& e) {
        throw std::runtime_error("Failed to load image '" + filename + "', reason: " + e.what());
    }
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
