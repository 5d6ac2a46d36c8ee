// BMP, TGA, GIF (first frame) and binary PNM readers for the CLI, to RGBA8.
// Their conventions are those of the reference's loader (stb_image 2.30, via
// stbi_load(..., 4) in src/main.cpp:37-57), so a file gives the same bytes:
//   BMP  1/4/8-bit palettes, 16/32-bit bit fields (a k-bit field is widened
//        to 8 bits by replicating its bit pattern), 24-bit BGR, bottom-up or
//        top-down rows padded to 4 bytes; 32-bit files without explicit masks
//        read the fourth byte as alpha, and an alpha that is 0 everywhere
//        becomes 255; no RLE;
//   TGA  uncompressed and RLE; colour-mapped (8/16-bit indices; an index past
//        the map reads entry 0), true-colour 15/16/24/32-bit (5-bit fields
//        widened by v * 255 / 31) and 8/16-bit gray; BGR(A) to RGB(A); rows
//        flipped unless the descriptor's top-left bit is set; the colour map
//        starts colormap_first bytes after the image id;
//   GIF  the first image of the stream, LZW-decoded, interlaced or not, over
//        a transparent black canvas; pixels whose palette entry is the
//        transparent index are not drawn; when the background index is
//        non-zero, pixels the image does not cover take the global palette's
//        background entry with its red and blue bytes exchanged (the
//        reference's byte order for that fill);
//   PNM  binary P5 (gray) and P6 (RGB) with '#' comments; samples are copied
//        as they are (no rescaling of maxval); 16-bit samples (maxval > 255)
//        keep the second byte of each big-endian pair (what the reference's
//        loader keeps on a little-endian host).
#pragma once

#include <cctype>

#include "image_common.hpp"

namespace iqmcli {

/* ---- BMP ------------------------------------------------------------- */

inline int bmp_high_bit(uint32_t z) {
    if (!z) return -1;
    int n = 0;
    while (z >>= 1) ++n;
    return n;
}
inline int bmp_bitcount(uint32_t a) {
    int n = 0;
    for (; a; a &= a - 1) ++n;
    return n;
}
/* the masked field, moved so its top bit is bit 7, then its `bits`-bit
 * pattern repeated down to 8 bits */
inline uint32_t bmp_widen(uint32_t v, int shift, int bits) {
    static const uint32_t mul[9] = {0, 0xff, 0x55, 0x49, 0x11, 0x21, 0x41, 0x81, 0x01};
    static const int sh[9] = {0, 0, 0, 1, 0, 2, 4, 6, 0};
    v = shift < 0 ? v << -shift : v >> shift;
    v >>= (8 - bits);
    return (v * mul[bits]) >> sh[bits];
}

/* a negative skip moves to the end of the file, as the reference's reader does */
inline void skip_signed(Reader& r, long n) { r.skip(n < 0 ? r.size() : (size_t)n); }

inline InputImage decode_bmp(const std::vector<uint8_t>& f) {
    Reader r(f.data(), f.size());
    if (r.u8() != 'B' || r.u8() != 'M') fail("Corrupt BMP");
    r.u32le();
    r.u16le();
    r.u16le();
    const int32_t offset = (int32_t)r.u32le();
    const uint32_t hsz = r.u32le();
    if (offset < 0) fail("bad BMP");
    if (hsz != 12 && hsz != 40 && hsz != 56 && hsz != 108 && hsz != 124) fail("BMP type not supported: unknown");
    int32_t w, h;
    if (hsz == 12) {
        w = r.u16le();
        h = r.u16le();
    } else {
        w = (int32_t)r.u32le();
        h = (int32_t)r.u32le();
    }
    if (r.u16le() != 1) fail("bad BMP");
    const int bpp = r.u16le();
    uint32_t mr = 0, mg = 0, mb = 0, ma = 0;
    bool alpha_all_zero_fix = false; /* 32-bit without explicit masks */
    int extra_read = 14;
    if (hsz != 12) {
        const uint32_t compress = r.u32le();
        if (compress == 1 || compress == 2) fail("BMP type not supported: RLE");
        if (compress >= 4) fail("BMP type not supported: unsupported compression");
        if (compress == 3 && bpp != 16 && bpp != 32) fail("bad BMP");
        for (int i = 0; i < 5; ++i) r.u32le();
        auto defaults = [&]() {
            if (bpp == 16) {
                mr = 31u << 10;
                mg = 31u << 5;
                mb = 31u;
            } else if (bpp == 32) {
                mr = 0xffu << 16;
                mg = 0xffu << 8;
                mb = 0xffu;
                ma = 0xffu << 24;
                alpha_all_zero_fix = true;
            } else {
                mr = mg = mb = ma = 0;
            }
        };
        if (hsz == 40 || hsz == 56) {
            if (hsz == 56)
                for (int i = 0; i < 4; ++i) r.u32le();
            if (bpp == 16 || bpp == 32) {
                if (compress == 0) {
                    defaults();
                } else {
                    mr = r.u32le();
                    mg = r.u32le();
                    mb = r.u32le();
                    extra_read += 12;
                    if (mr == mg && mg == mb) fail("bad BMP");
                }
            }
        } else {
            mr = r.u32le();
            mg = r.u32le();
            mb = r.u32le();
            ma = r.u32le();
            if (compress != 3) defaults();
            for (int i = 0; i < 13; ++i) r.u32le();
            if (hsz == 124)
                for (int i = 0; i < 4; ++i) r.u32le();
        }
    }
    const bool flip = h > 0;
    if (h < 0) h = -h;
    if (w <= 0 || h <= 0) fail("Corrupt BMP");
    InputImage img = make_rgba((uint32_t)w, (uint32_t)h);
    int psize = 0;
    if (hsz == 12) {
        if (bpp < 24) psize = (offset - extra_read - 24) / 3;
    } else if (bpp < 16) {
        psize = (offset - extra_read - (int)hsz) >> 2;
    }
    if (psize == 0) {
        const int read = (int)r.pos();
        if (read <= 0 || read > 1024) fail("Corrupt BMP");
        if (offset < read || offset - read > 256 * 4) fail("Corrupt BMP");
        r.skip((size_t)(offset - read));
    }
    uint8_t* out = img.data.data();
    uint32_t alpha_or = alpha_all_zero_fix ? 0 : 255;
    if (bpp < 16) {
        if (psize <= 0 || psize > 256) fail("Corrupt BMP");
        uint8_t pal[256][3] = {};
        for (int i = 0; i < psize; ++i) {
            pal[i][2] = r.u8();
            pal[i][1] = r.u8();
            pal[i][0] = r.u8();
            if (hsz != 12) r.u8();
        }
        skip_signed(r, offset - extra_read - (int)hsz - psize * (hsz == 12 ? 3 : 4));
        int width;
        if (bpp == 1) width = (w + 7) >> 3;
        else if (bpp == 4) width = (w + 1) >> 1;
        else if (bpp == 8) width = w;
        else fail("Corrupt BMP");
        const int pad = (-width) & 3;
        auto put = [&](uint8_t*& o, int v) {
            o[0] = pal[v][0];
            o[1] = pal[v][1];
            o[2] = pal[v][2];
            o[3] = 255;
            o += 4;
        };
        uint8_t* o = out;
        for (int j = 0; j < h; ++j) {
            if (bpp == 1) {
                int bit = 7, v = r.u8();
                for (int i = 0; i < w; ++i) {
                    put(o, (v >> bit) & 1);
                    if (i + 1 == w) break;
                    if (--bit < 0) {
                        bit = 7;
                        v = r.u8();
                    }
                }
            } else {
                for (int i = 0; i < w; i += 2) {
                    int v = r.u8(), v2 = 0;
                    if (bpp == 4) {
                        v2 = v & 15;
                        v >>= 4;
                    }
                    put(o, v);
                    if (i + 1 == w) break;
                    put(o, bpp == 8 ? r.u8() : v2);
                }
            }
            r.skip((size_t)pad);
        }
    } else {
        skip_signed(r, offset - extra_read - (int)hsz);
        const int width = bpp == 24 ? 3 * w : bpp == 16 ? 2 * w : 0;
        const int pad = (-width) & 3;
        int easy = 0;
        if (bpp == 24) easy = 1;
        else if (bpp == 32 && mb == 0xff && mg == 0xff00 && mr == 0x00ff0000 && ma == 0xff000000) easy = 2;
        int rs = 0, gs = 0, bs = 0, as = 0, rc = 0, gc = 0, bc = 0, ac = 0;
        if (!easy) {
            if (!mr || !mg || !mb) fail("Corrupt BMP");
            rs = bmp_high_bit(mr) - 7;
            rc = bmp_bitcount(mr);
            gs = bmp_high_bit(mg) - 7;
            gc = bmp_bitcount(mg);
            bs = bmp_high_bit(mb) - 7;
            bc = bmp_bitcount(mb);
            as = bmp_high_bit(ma) - 7;
            ac = bmp_bitcount(ma);
            if (rc > 8 || gc > 8 || bc > 8 || ac > 8) fail("Corrupt BMP");
        }
        uint8_t* o = out;
        for (int j = 0; j < h; ++j) {
            for (int i = 0; i < w; ++i, o += 4) {
                if (easy) {
                    o[2] = r.u8();
                    o[1] = r.u8();
                    o[0] = r.u8();
                    const uint8_t a = easy == 2 ? r.u8() : 255;
                    alpha_or |= a;
                    o[3] = a;
                } else {
                    const uint32_t v = bpp == 16 ? r.u16le() : r.u32le();
                    o[0] = (uint8_t)bmp_widen(v & mr, rs, rc);
                    o[1] = (uint8_t)bmp_widen(v & mg, gs, gc);
                    o[2] = (uint8_t)bmp_widen(v & mb, bs, bc);
                    const uint32_t a = ma ? bmp_widen(v & ma, as, ac) : 255;
                    alpha_or |= a;
                    o[3] = (uint8_t)a;
                }
            }
            r.skip((size_t)pad);
        }
    }
    if (alpha_or == 0)
        for (size_t i = 3; i < img.data.size(); i += 4) img.data[i] = 255;
    if (flip) {
        const size_t row = (size_t)w * 4;
        for (int j = 0; j < h / 2; ++j)
            std::swap_ranges(out + row * j, out + row * (j + 1), out + row * (h - 1 - j));
    }
    return img;
}

/* ---- TGA ------------------------------------------------------------- */

/* whether the 18-byte header looks like a TGA the reference reads */
inline bool tga_test(const std::vector<uint8_t>& f) {
    if (f.size() < 18) return false;
    const int cmap = f[1], type = f[2];
    if (cmap > 1) return false;
    if (cmap == 1) {
        if (type != 1 && type != 9) return false;
        const int eb = f[7];
        if (eb != 8 && eb != 15 && eb != 16 && eb != 24 && eb != 32) return false;
    } else if (type != 2 && type != 3 && type != 10 && type != 11) {
        return false;
    }
    if ((f[12] | f[13] << 8) < 1 || (f[14] | f[15] << 8) < 1) return false;
    const int bpp = f[16];
    if (cmap == 1 && bpp != 8 && bpp != 16) return false;
    return bpp == 8 || bpp == 15 || bpp == 16 || bpp == 24 || bpp == 32;
}

inline InputImage decode_tga(const std::vector<uint8_t>& f) {
    Reader r(f.data(), f.size());
    const int id_len = r.u8(), indexed = r.u8();
    int type = r.u8();
    const int pal_start = r.u16le(), pal_len = r.u16le(), pal_bits = r.u8();
    r.u16le();
    r.u16le();
    const int w = r.u16le(), h = r.u16le(), bpp = r.u8(), desc = r.u8();
    const bool rle = type >= 8;
    if (rle) type -= 8;
    const bool bottom_up = !((desc >> 5) & 1);
    auto comps_of = [](int bits, bool gray, bool& rgb16) {
        rgb16 = false;
        switch (bits) {
            case 8: return 1;
            case 16:
                if (gray) return 2;
                rgb16 = true;
                return 3;
            case 15: rgb16 = true; return 3;
            case 24: return 3;
            case 32: return 4;
            default: return 0;
        }
    };
    bool rgb16 = false;
    const int comp = indexed ? comps_of(pal_bits, false, rgb16) : comps_of(bpp, type == 3, rgb16);
    if (!comp) fail("Can't find out TGA pixelformat");
    InputImage img = make_rgba((uint32_t)w, (uint32_t)h);
    std::vector<uint8_t> px(mul_checked(mul_checked((size_t)w, (size_t)h), (size_t)comp));
    r.skip((size_t)id_len);
    auto read16 = [&](uint8_t* o) {
        const uint16_t v = r.u16le();
        o[0] = (uint8_t)((((v >> 10) & 31) * 255) / 31);
        o[1] = (uint8_t)((((v >> 5) & 31) * 255) / 31);
        o[2] = (uint8_t)(((v & 31) * 255) / 31);
    };
    if (!indexed && !rle && !rgb16) {
        for (int i = 0; i < h; ++i) {
            const int row = bottom_up ? h - 1 - i : i;
            uint8_t* dst = &px[(size_t)row * w * comp];
            if (!r.read(dst, (size_t)w * comp)) {
                /* short read: what the file has, zeros after */
                const size_t have = r.pos() < r.size() ? r.size() - r.pos() : 0;
                std::memcpy(dst, r.at(r.pos()), std::min(have, (size_t)w * comp));
                r.skip(have);
            }
        }
    } else {
        std::vector<uint8_t> pal;
        if (indexed) {
            if (pal_len == 0) fail("Corrupt TGA");
            r.skip((size_t)pal_start);
            pal.assign((size_t)pal_len * comp, 0);
            if (rgb16) {
                for (int i = 0; i < pal_len; ++i) read16(&pal[(size_t)i * comp]);
            } else if (!r.read(pal.data(), pal.size())) {
                fail("Corrupt TGA");
            }
        }
        uint8_t raw[4] = {0, 0, 0, 0};
        int count = 0;
        bool repeat = false, fresh = true;
        for (size_t i = 0; i < (size_t)w * h; ++i) {
            if (rle) {
                if (count == 0) {
                    const int cmd = r.u8();
                    count = 1 + (cmd & 127);
                    repeat = cmd >> 7;
                    fresh = true;
                } else if (!repeat) {
                    fresh = true;
                }
            } else {
                fresh = true;
            }
            if (fresh) {
                if (indexed) {
                    int idx = bpp == 8 ? r.u8() : r.u16le();
                    if (idx >= pal_len) idx = 0;
                    std::memcpy(raw, &pal[(size_t)idx * comp], (size_t)comp);
                } else if (rgb16) {
                    read16(raw);
                } else {
                    for (int k = 0; k < comp; ++k) raw[k] = r.u8();
                }
                fresh = false;
            }
            std::memcpy(&px[i * comp], raw, (size_t)comp);
            --count;
        }
        if (bottom_up) {
            const size_t row = (size_t)w * comp;
            for (int j = 0; 2 * j < h; ++j) {
                if (j == h - 1 - j) break;
                std::swap_ranges(&px[row * j], &px[row * (j + 1)], &px[row * (h - 1 - j)]);
            }
        }
    }
    if (comp >= 3 && !rgb16)
        for (size_t i = 0; i < (size_t)w * h; ++i) std::swap(px[i * comp], px[i * comp + 2]);
    for (size_t i = 0; i < (size_t)w * h; ++i) to_rgba(&px[i * comp], comp, &img.data[i * 4]);
    return img;
}

/* ---- GIF ------------------------------------------------------------- */

inline InputImage decode_gif(const std::vector<uint8_t>& f) {
    Reader r(f.data(), f.size());
    if (r.u8() != 'G' || r.u8() != 'I' || r.u8() != 'F' || r.u8() != '8') fail("Corrupt GIF");
    const int ver = r.u8();
    if ((ver != '7' && ver != '9') || r.u8() != 'a') fail("Corrupt GIF");
    const int W = r.u16le(), H = r.u16le(), flags = r.u8(), bg = r.u8();
    r.u8();
    InputImage img = make_rgba((uint32_t)W, (uint32_t)H); /* transparent black canvas */
    /* palettes as (B, G, R, A), the order the reference keeps them in */
    uint8_t gpal[256][4] = {}, lpal[256][4] = {};
    auto read_pal = [&](uint8_t pal[256][4], int n, int transp) {
        for (int i = 0; i < n; ++i) {
            pal[i][2] = r.u8();
            pal[i][1] = r.u8();
            pal[i][0] = r.u8();
            pal[i][3] = transp == i ? 0 : 255;
        }
    };
    if (flags & 0x80) read_pal(gpal, 2 << (flags & 7), -1);
    int transparent = -1, eflags = 0;
    std::vector<uint8_t> drawn((size_t)W * H, 0);
    for (;;) {
        const int tag = r.u8();
        if (tag == 0x21) {
            const int ext = r.u8();
            if (ext == 0xf9) {
                const int len = r.u8();
                if (len == 4) {
                    eflags = r.u8();
                    r.u16le();
                    if (transparent >= 0) gpal[transparent][3] = 255;
                    if (eflags & 1) {
                        transparent = r.u8();
                        gpal[transparent][3] = 0;
                    } else {
                        r.skip(1);
                        transparent = -1;
                    }
                } else {
                    r.skip((size_t)len);
                }
            }
            for (int len; (len = r.u8()) != 0;) r.skip((size_t)len);
            if (r.eof()) fail("Corrupt GIF");
            continue;
        }
        if (tag != 0x2c) fail(tag == 0x3b ? "Corrupt GIF" : "Corrupt GIF");
        const int x0 = r.u16le(), y0 = r.u16le(), w = r.u16le(), h = r.u16le();
        if (x0 + w > W || y0 + h > H) fail("Corrupt GIF");
        const int lflags = r.u8();
        uint8_t(*pal)[4] = gpal;
        if (lflags & 0x80) {
            read_pal(lpal, 2 << (lflags & 7), (eflags & 1) ? transparent : -1);
            pal = lpal;
        } else if (!(flags & 0x80)) {
            fail("Corrupt GIF");
        }
        /* raster: rows in order, or the four interlace passes */
        int cx = 0, cy = w == 0 ? h : 0, step = (lflags & 0x40) ? 8 : 1, pass = (lflags & 0x40) ? 3 : 0;
        auto emit = [&](int idx) {
            if (cy >= h) return;
            const size_t p = (size_t)(y0 + cy) * W + x0 + cx;
            drawn[p] = 1;
            const uint8_t* c = pal[idx];
            if (c[3] > 128) {
                img.data[p * 4 + 0] = c[2];
                img.data[p * 4 + 1] = c[1];
                img.data[p * 4 + 2] = c[0];
                img.data[p * 4 + 3] = c[3];
            }
            if (++cx >= w) {
                cx = 0;
                cy += step;
                while (cy >= h && pass > 0) {
                    step = 1 << pass;
                    cy = step >> 1;
                    --pass;
                }
            }
        };
        const int min_size = r.u8();
        if (min_size > 12) fail("Corrupt GIF");
        const int clear = 1 << min_size;
        struct Code {
            int16_t prefix;
            uint8_t first, suffix;
        };
        std::vector<Code> codes(8192);
        for (int i = 0; i < clear; ++i) codes[i] = Code{-1, (uint8_t)i, (uint8_t)i};
        int size = min_size + 1, mask = (1 << size) - 1, avail = clear + 2, old = -1;
        bool first = true;
        uint32_t bits = 0;
        int nbits = 0, len = 0;
        std::vector<uint8_t> stack;
        auto output = [&](int code) {
            stack.clear();
            for (int c = code; c >= 0; c = codes[c].prefix) stack.push_back(codes[c].suffix);
            for (size_t k = stack.size(); k-- > 0;) emit(stack[k]);
        };
        for (;;) {
            if (nbits < size) {
                if (len == 0) {
                    len = r.u8();
                    if (len == 0) break;
                }
                --len;
                bits |= (uint32_t)r.u8() << nbits;
                nbits += 8;
                continue;
            }
            const int code = (int)(bits & (uint32_t)mask);
            bits >>= size;
            nbits -= size;
            if (code == clear) {
                size = min_size + 1;
                mask = (1 << size) - 1;
                avail = clear + 2;
                old = -1;
                first = false;
            } else if (code == clear + 1) {
                r.skip((size_t)len);
                for (int l; (l = r.u8()) > 0;) r.skip((size_t)l);
                break;
            } else if (code <= avail) {
                if (first) fail("Corrupt GIF");
                if (old >= 0) {
                    if (avail >= 8192) fail("Corrupt GIF");
                    Code& p = codes[avail++];
                    p.prefix = (int16_t)old;
                    p.first = codes[old].first;
                    p.suffix = code == avail - 1 ? p.first : codes[code].first;
                } else if (code == avail) {
                    fail("Corrupt GIF");
                }
                output(code);
                if ((avail & mask) == 0 && avail <= 0x0fff) {
                    ++size;
                    mask = (1 << size) - 1;
                }
                old = code;
            } else {
                fail("Corrupt GIF");
            }
            if (r.pos() > r.size() + 16) fail("Corrupt GIF");
        }
        if (bg > 0) {
            uint8_t fill[4] = {gpal[bg][0], gpal[bg][1], gpal[bg][2], 255};
            for (size_t p = 0; p < drawn.size(); ++p)
                if (!drawn[p]) std::memcpy(&img.data[p * 4], fill, 4);
        }
        return img;
    }
}

/* ---- PNM ------------------------------------------------------------- */

inline InputImage decode_pnm(const std::vector<uint8_t>& f) {
    Reader r(f.data(), f.size());
    r.u8();
    const bool color = r.u8() == '6';
    int c = r.u8();
    auto skip_ws = [&]() {
        for (;;) {
            while (!r.eof() && std::isspace(c)) c = r.u8();
            if (r.eof() || c != '#') break;
            while (!r.eof() && c != '\n' && c != '\r') c = r.u8();
        }
    };
    auto integer = [&]() {
        long v = 0;
        while (!r.eof() && std::isdigit(c)) {
            v = v * 10 + (c - '0');
            if (v > 0x7fffffff) fail("PPM image header had zero or overflowing width");
            c = r.u8();
        }
        return v;
    };
    skip_ws();
    const long w = integer();
    if (w == 0) fail("PPM image header had zero or overflowing width");
    skip_ws();
    const long h = integer();
    if (h == 0) fail("PPM image header had zero or overflowing width");
    skip_ws();
    const long maxv = integer();
    if (maxv > 65535) fail("PPM image supports only 8-bit and 16-bit images");
    const int bytes = maxv > 255 ? 2 : 1;
    const int ch = color ? 3 : 1;
    InputImage img = make_rgba((uint32_t)w, (uint32_t)h);
    const size_t need = mul_checked(mul_checked((size_t)w * h, (size_t)ch), (size_t)bytes);
    std::vector<uint8_t> raw(need);
    if (!r.read(raw.data(), need)) fail("PNM file truncated");
    uint8_t px[3];
    for (size_t i = 0; i < (size_t)w * h; ++i) {
        for (int k = 0; k < ch; ++k) px[k] = raw[(i * ch + k) * bytes + (bytes - 1)];
        to_rgba(px, ch, &img.data[i * 4]);
    }
    return img;
}

}  // namespace iqmcli
