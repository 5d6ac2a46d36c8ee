// JPEG reader: baseline and extended sequential (SOF0/SOF1) and progressive
// (SOF2) Huffman JPEG, 8-bit samples, 1/3/4 components, any sampling factors
// 1..4, restart intervals.  Entropy decoding follows ITU-T T.81 (Annex F for
// sequential, Annex G for the spectral-selection and successive-approximation
// scans); any correct decoder gets the same coefficients.
//
// The pixel arithmetic after that is fixed by the reference's loader
// (stb_image 2.30 via stbi_load(..., 4), src/main.cpp:37-57), and it is
// reproduced here so that a JPEG file gives the reference's RGBA bytes:
//   - inverse DCT: the accurate integer ("islow") factorization of Loeffler,
//     Ligtenberg and Moschytz with 12-bit constants, columns first with two
//     extra bits kept (rounding +512, >> 10), then rows with rounding and the
//     +128 level shift folded into one bias before >> 17, clamped to 0..255;
//   - chroma upsampling: the triangular ("fancy") filters for 2x horizontal,
//     2x vertical and 2x2 factors (3:1 weights, +2 >> 2 and +8 >> 4
//     roundings, edges replicated), nearest-neighbour replication otherwise;
//     rows are paired with their nearer and farther source row as the
//     vertical filter walks down;
//   - YCbCr -> RGB in 20-bit fixed point with 12-bit coefficients of 1.402,
//     0.71414, 0.34414 and 1.772 (the Cb term of G truncated to its upper 16
//     bits);
//   - 3 components are RGB (no transform) when their ids are 'R','G','B', or
//     when an Adobe marker says "no transform" and no JFIF marker is present;
//     4 components are CMYK (Adobe transform 0) or YCCK (transform 2),
//     combined with K by the rounded x*y/255 product.
#pragma once

#include "image_common.hpp"

namespace iqmcli {
namespace jpg {

/* natural index of the k-th coefficient in zigzag order */
inline const uint8_t* zigzag() {
    static const uint8_t z[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                                  12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                                  35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                                  58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};
    return z;
}

struct Huffman {
    bool defined = false;
    uint8_t vals[256] = {};
    int32_t maxcode[18] = {};  /* largest code of each length, -1 if none */
    int32_t offset[17] = {};   /* vals index of a length's first code minus that code */
    /* 9-bit lookahead: symbol and code length, length 0 if longer */
    uint8_t fast_sym[512] = {};
    uint8_t fast_len[512] = {};

    /* canonical codes from the 16 length counts (T.81 Annex C) */
    void build(const int counts[16], const uint8_t* symbols, int n) {
        std::memcpy(vals, symbols, (size_t)n);
        int code = 0, k = 0;
        std::memset(fast_len, 0, sizeof(fast_len));
        for (int len = 1; len <= 16; ++len) {
            offset[len] = k - code;
            for (int i = 0; i < counts[len - 1]; ++i, ++k, ++code) {
                if (len <= 9) {
                    const int lo = code << (9 - len), hi = (code + 1) << (9 - len);
                    for (int c = lo; c < hi; ++c) {
                        fast_sym[c] = vals[k];
                        fast_len[c] = (uint8_t)len;
                    }
                }
            }
            if (code > (1 << len)) fail("Corrupt JPEG");
            maxcode[len] = counts[len - 1] ? code - 1 : -1;
            code <<= 1;
        }
        maxcode[17] = 0x7fffffff;
        defined = true;
    }
};

/* Entropy-coded segment reader: bytes 0xFF 0x00 are a stuffed 0xFF; any
 * other 0xFF xx (after repeated 0xFF fill) is a marker, at which point (and
 * at the end of the file) the reader supplies zero bits. */
class Bits {
   public:
    Bits(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    size_t pos = 0;
    int marker = -1; /* marker met inside the segment, -1 none */
    void reset() {
        buf_ = 0;
        cnt_ = 0;
        nomore_ = false;
        marker = -1;
    }
    uint8_t byte() { return pos < n_ ? p_[pos++] : (pos++, 0); }
    bool at_end() const { return pos >= n_; }
    void fill() {
        while (cnt_ <= 24) {
            uint32_t b = 0;
            if (!nomore_) {
                b = byte();
                if (b == 0xff) {
                    int c = byte();
                    while (c == 0xff) c = byte();
                    if (c != 0) {
                        marker = c;
                        nomore_ = true;
                        return;
                    }
                }
            }
            buf_ |= b << (24 - cnt_);
            cnt_ += 8;
        }
    }
    uint32_t peek(int n) {
        if (cnt_ < n) fill();
        return buf_ >> (32 - n);
    }
    void drop(int n) {
        buf_ <<= n;
        cnt_ -= n;
    }
    int get(int n) {
        if (n == 0) return 0;
        if (cnt_ < n) fill();
        const uint32_t v = buf_ >> (32 - n);
        drop(n);
        return (int)v;
    }
    /* n-bit magnitude category value (T.81 F.2.2.1 EXTEND) */
    int receive_extend(int n) {
        if (n == 0) return 0;
        const int v = get(n);
        return v < (1 << (n - 1)) ? v - (1 << n) + 1 : v;
    }
    int decode(const Huffman& h) {
        if (!h.defined) fail("Corrupt JPEG");
        if (cnt_ < 16) fill();
        const uint32_t look = buf_ >> 23;
        if (h.fast_len[look]) {
            const int len = h.fast_len[look];
            if (len > cnt_) fail("Corrupt JPEG");
            drop(len);
            return h.fast_sym[look];
        }
        for (int len = 10; len <= 16; ++len) {
            const int32_t code = (int32_t)(buf_ >> (32 - len));
            if (code <= h.maxcode[len]) {
                if (len > cnt_) fail("Corrupt JPEG");
                drop(len);
                const int idx = h.offset[len] + code;
                if (idx < 0 || idx > 255) fail("Corrupt JPEG");
                return h.vals[idx];
            }
        }
        fail("Corrupt JPEG");
    }

   private:
    const uint8_t* p_;
    size_t n_;
    uint32_t buf_ = 0;
    int cnt_ = 0;
    bool nomore_ = false;
};

struct Component {
    int id = 0, h = 1, v = 1, tq = 0, hd = 0, ha = 0;
    int x = 0, y = 0;   /* samples of this component */
    int w2 = 0, h2 = 0; /* padded to whole MCUs */
    int dc_pred = 0;
    std::vector<uint8_t> data;    /* w2 x h2 samples */
    std::vector<int16_t> coeff;   /* progressive: (w2/8 x h2/8) blocks of 64, natural order */
};

/* One 8x8 inverse DCT (islow factorization, see the header comment) from
 * dequantized coefficients in natural order to samples at out (row pitch). */
inline void idct8x8(const int16_t in[64], uint8_t* out, int pitch) {
    /* 12-bit constants as the reference's loader forms them: float x * 4096,
     * + 0.5, truncated toward zero (so negative constants round up) */
    constexpr auto C = [](float x) { return (int)((double)(x * 4096.0f) + 0.5); };
    auto pass = [&](int s0, int s1, int s2, int s3, int s4, int s5, int s6, int s7, int* even, int* odd) {
        /* even part: rotation of (s2, s6) and butterflies with (s0, s4) */
        const int z1 = (s2 + s6) * C(0.5411961f);
        const int e2 = z1 + s6 * C(-1.847759065f), e3 = z1 + s2 * C(0.765366865f);
        const int e0 = (s0 + s4) * 4096, e1 = (s0 - s4) * 4096;
        even[0] = e0 + e3;
        even[3] = e0 - e3;
        even[1] = e1 + e2;
        even[2] = e1 - e2;
        /* odd part on (s7, s5, s3, s1) */
        const int a = s7, b = s5, c = s3, d = s1;
        const int za = a + c, zb = b + d, zc = a + d, zd = b + c;
        const int z5 = (za + zb) * C(1.175875602f);
        const int pc = z5 + zc * C(-0.899976223f), pd = z5 + zd * C(-2.562915447f);
        const int qa = za * C(-1.961570560f), qb = zb * C(-0.390180644f);
        odd[3] = a * C(0.298631336f) + pc + qa;
        odd[2] = b * C(2.053119869f) + pd + qb;
        odd[1] = c * C(3.072711026f) + pd + qa;
        odd[0] = d * C(1.501321110f) + pc + qb;
    };
    int tmp[64];
    for (int i = 0; i < 8; ++i) {
        const int16_t* d = in + i;
        if (!d[8] && !d[16] && !d[24] && !d[32] && !d[40] && !d[48] && !d[56]) {
            for (int r = 0; r < 8; ++r) tmp[r * 8 + i] = d[0] * 4;
            continue;
        }
        int ev[4], od[4];
        pass(d[0], d[8], d[16], d[24], d[32], d[40], d[48], d[56], ev, od);
        for (int k = 0; k < 4; ++k) ev[k] += 512;
        tmp[0 * 8 + i] = (ev[0] + od[0]) >> 10;
        tmp[7 * 8 + i] = (ev[0] - od[0]) >> 10;
        tmp[1 * 8 + i] = (ev[1] + od[1]) >> 10;
        tmp[6 * 8 + i] = (ev[1] - od[1]) >> 10;
        tmp[2 * 8 + i] = (ev[2] + od[2]) >> 10;
        tmp[5 * 8 + i] = (ev[2] - od[2]) >> 10;
        tmp[3 * 8 + i] = (ev[3] + od[3]) >> 10;
        tmp[4 * 8 + i] = (ev[3] - od[3]) >> 10;
    }
    auto clamp = [](int v) { return (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v); };
    for (int r = 0; r < 8; ++r) {
        const int* v = tmp + r * 8;
        int ev[4], od[4];
        pass(v[0], v[1], v[2], v[3], v[4], v[5], v[6], v[7], ev, od);
        /* 1 << 17 of scale: rounding (1 << 16) and the level shift (128 << 17) */
        for (int k = 0; k < 4; ++k) ev[k] += 65536 + (128 << 17);
        uint8_t* o = out + (size_t)r * pitch;
        o[0] = clamp((ev[0] + od[0]) >> 17);
        o[7] = clamp((ev[0] - od[0]) >> 17);
        o[1] = clamp((ev[1] + od[1]) >> 17);
        o[6] = clamp((ev[1] - od[1]) >> 17);
        o[2] = clamp((ev[2] + od[2]) >> 17);
        o[5] = clamp((ev[2] - od[2]) >> 17);
        o[3] = clamp((ev[3] + od[3]) >> 17);
        o[4] = clamp((ev[3] - od[3]) >> 17);
    }
}

class Decoder {
   public:
    explicit Decoder(const std::vector<uint8_t>& f) : bits_(f.data(), f.size()) {}

    InputImage decode() {
        if (bits_.byte() != 0xff || bits_.byte() != 0xd8) fail("Corrupt JPEG");
        int m = next_marker();
        while (m != 0xc0 && m != 0xc1 && m != 0xc2) {
            marker_segment(m);
            m = next_marker();
            while (m < 0) {
                if (bits_.at_end()) fail("Corrupt JPEG");
                m = next_marker();
            }
        }
        progressive_ = m == 0xc2;
        frame_header();
        m = next_marker();
        while (m != 0xd9) {
            if (m == 0xda) {
                scan_header();
                scan();
                if (bits_.marker < 0) bits_.marker = skip_junk();
                m = next_marker();
                if (m >= 0xd0 && m <= 0xd7) m = next_marker();
            } else if (m == 0xdc) { /* DNL */
                const int ld = u16(), nl = u16();
                if (ld != 4 || nl != ny_) fail("Corrupt JPEG");
                m = next_marker();
            } else {
                /* the reference stops at a marker it cannot use after the
                 * frame and keeps what it has decoded */
                if (!try_marker_segment(m)) break;
                m = next_marker();
            }
        }
        if (progressive_) finish();
        return convert();
    }

   private:
    Bits bits_;
    Huffman dc_[4], ac_[4];
    uint16_t q_[4][64] = {};
    Component comp_[4];
    int n_ = 0, nx_ = 0, ny_ = 0, hmax_ = 1, vmax_ = 1, mcux_ = 0, mcuy_ = 0;
    bool progressive_ = false, jfif_ = false;
    int app14_ = -1, rgb_ids_ = 0, restart_ = 0;
    int scan_n_ = 0, order_[4] = {}, ss_ = 0, se_ = 63, ah_ = 0, al_ = 0;
    int todo_ = 0, eobrun_ = 0;

    int u16() {
        const int a = bits_.byte();
        return (a << 8) | bits_.byte();
    }
    /* the pending marker, else the next 0xFF xx in the stream (-1 if the
     * next byte is not 0xFF) */
    int next_marker() {
        if (bits_.marker >= 0) {
            const int m = bits_.marker;
            bits_.marker = -1;
            return m;
        }
        int x = bits_.byte();
        if (x != 0xff) return -1;
        while (x == 0xff) x = bits_.byte();
        return x;
    }
    /* junk after a scan's data: resume at the next real marker */
    int skip_junk() {
        while (!bits_.at_end()) {
            int x = bits_.byte();
            while (x == 0xff) {
                if (bits_.at_end()) return -1;
                x = bits_.byte();
                if (x != 0x00 && x != 0xff) return x;
            }
        }
        return -1;
    }
    void marker_segment(int m) {
        if (!try_marker_segment(m)) fail("Corrupt JPEG");
    }
    bool try_marker_segment(int m) {
        if (m < 0) return false;
        if (m == 0xdd) { /* DRI */
            if (u16() != 4) return false;
            restart_ = u16();
            return true;
        }
        if (m == 0xdb) { /* DQT */
            int L = u16() - 2;
            while (L > 0) {
                const int pq = bits_.byte();
                const int p = pq >> 4, t = pq & 15;
                if (p > 1 || t > 3) return false;
                for (int i = 0; i < 64; ++i) q_[t][zigzag()[i]] = (uint16_t)(p ? u16() : bits_.byte());
                L -= p ? 129 : 65;
            }
            return L == 0;
        }
        if (m == 0xc4) { /* DHT */
            int L = u16() - 2;
            while (L > 0) {
                const int tcth = bits_.byte();
                const int tc = tcth >> 4, th = tcth & 15;
                if (tc > 1 || th > 3) return false;
                int counts[16], n = 0;
                for (int i = 0; i < 16; ++i) n += counts[i] = bits_.byte();
                if (n > 256) return false;
                uint8_t sym[256];
                for (int i = 0; i < n; ++i) sym[i] = bits_.byte();
                (tc ? ac_[th] : dc_[th]).build(counts, sym, n);
                L -= 17 + n;
            }
            return L == 0;
        }
        if ((m >= 0xe0 && m <= 0xef) || m == 0xfe) { /* APPn, COM */
            int L = u16();
            if (L < 2) return false;
            L -= 2;
            if (m == 0xe0 && L >= 5) {
                static const uint8_t tag[5] = {'J', 'F', 'I', 'F', 0};
                bool ok = true;
                for (int i = 0; i < 5; ++i) ok = (bits_.byte() == tag[i]) && ok;
                L -= 5;
                if (ok) jfif_ = true;
            } else if (m == 0xee && L >= 12) {
                static const uint8_t tag[6] = {'A', 'd', 'o', 'b', 'e', 0};
                bool ok = true;
                for (int i = 0; i < 6; ++i) ok = (bits_.byte() == tag[i]) && ok;
                L -= 6;
                if (ok) {
                    bits_.byte();
                    u16();
                    u16();
                    app14_ = bits_.byte();
                    L -= 6;
                }
            }
            bits_.pos += (size_t)L;
            return true;
        }
        return false;
    }
    void frame_header() {
        const int Lf = u16();
        if (Lf < 11) fail("Corrupt JPEG");
        if (bits_.byte() != 8) fail("JPEG format not supported: 8-bit only");
        ny_ = u16();
        nx_ = u16();
        if (ny_ == 0) fail("JPEG format not supported: delayed height");
        if (nx_ == 0) fail("Corrupt JPEG");
        if ((uint32_t)nx_ > kMaxSide || (uint32_t)ny_ > kMaxSide) fail("Very large image (corrupt?)");
        n_ = bits_.byte();
        if (n_ != 1 && n_ != 3 && n_ != 4) fail("Corrupt JPEG");
        if (Lf != 8 + 3 * n_) fail("Corrupt JPEG");
        static const int rgb[3] = {'R', 'G', 'B'};
        for (int i = 0; i < n_; ++i) {
            Component& c = comp_[i];
            c.id = bits_.byte();
            if (n_ == 3 && c.id == rgb[i]) ++rgb_ids_;
            const int hv = bits_.byte();
            c.h = hv >> 4;
            c.v = hv & 15;
            c.tq = bits_.byte();
            if (c.h < 1 || c.h > 4 || c.v < 1 || c.v > 4 || c.tq > 3) fail("Corrupt JPEG");
            hmax_ = std::max(hmax_, c.h);
            vmax_ = std::max(vmax_, c.v);
        }
        mul_checked(mul_checked((size_t)nx_, (size_t)ny_), (size_t)n_);
        for (int i = 0; i < n_; ++i)
            if (hmax_ % comp_[i].h || vmax_ % comp_[i].v) fail("Corrupt JPEG");
        mcux_ = (nx_ + hmax_ * 8 - 1) / (hmax_ * 8);
        mcuy_ = (ny_ + vmax_ * 8 - 1) / (vmax_ * 8);
        for (int i = 0; i < n_; ++i) {
            Component& c = comp_[i];
            c.x = (nx_ * c.h + hmax_ - 1) / hmax_;
            c.y = (ny_ * c.v + vmax_ - 1) / vmax_;
            c.w2 = mcux_ * c.h * 8;
            c.h2 = mcuy_ * c.v * 8;
            c.data.assign(mul_checked((size_t)c.w2, (size_t)c.h2), 0);
            if (progressive_) c.coeff.assign((size_t)c.w2 * c.h2, 0);
        }
    }
    void scan_header() {
        const int Ls = u16();
        scan_n_ = bits_.byte();
        if (scan_n_ < 1 || scan_n_ > 4 || scan_n_ > n_) fail("Corrupt JPEG");
        if (Ls != 6 + 2 * scan_n_) fail("Corrupt JPEG");
        for (int i = 0; i < scan_n_; ++i) {
            const int id = bits_.byte(), t = bits_.byte();
            int which = 0;
            while (which < n_ && comp_[which].id != id) ++which;
            if (which == n_) fail("Corrupt JPEG");
            comp_[which].hd = t >> 4;
            comp_[which].ha = t & 15;
            if (comp_[which].hd > 3 || comp_[which].ha > 3) fail("Corrupt JPEG");
            order_[i] = which;
        }
        ss_ = bits_.byte();
        se_ = bits_.byte();
        const int a = bits_.byte();
        ah_ = a >> 4;
        al_ = a & 15;
        if (progressive_) {
            if (ss_ > 63 || se_ > 63 || ss_ > se_ || ah_ > 13 || al_ > 13) fail("Corrupt JPEG");
        } else {
            if (ss_ != 0 || ah_ != 0 || al_ != 0) fail("Corrupt JPEG");
            se_ = 63;
        }
    }
    void reset_interval() {
        bits_.reset();
        for (int i = 0; i < n_; ++i) comp_[i].dc_pred = 0;
        todo_ = restart_ ? restart_ : 0x7fffffff;
        eobrun_ = 0;
    }
    /* counts an MCU down; false ends the scan (a restart marker was due and
     * is missing: the reference keeps the data decoded so far) */
    bool mcu_done() {
        if (--todo_ > 0) return true;
        bits_.peek(24);
        if (bits_.marker < 0xd0 || bits_.marker > 0xd7) return false;
        reset_interval();
        return true;
    }
    int next_dc(Component& c) {
        const int t = bits_.decode(dc_[c.hd]);
        if (t > 15) fail("Corrupt JPEG");
        const long dc = (long)c.dc_pred + bits_.receive_extend(t);
        if (dc < -(1L << 30) || dc > (1L << 30)) fail("Corrupt JPEG");
        c.dc_pred = (int)dc;
        return c.dc_pred;
    }
    /* sequential block: dequantized coefficients in natural order */
    void block_seq(Component& c, int16_t out[64]) {
        std::memset(out, 0, 64 * sizeof(int16_t));
        const uint16_t* q = q_[c.tq];
        out[0] = (int16_t)(next_dc(c) * q[0]);
        const Huffman& h = ac_[c.ha];
        for (int k = 1; k < 64;) {
            const int rs = bits_.decode(h);
            const int r = rs >> 4, s = rs & 15;
            if (s == 0) {
                if (rs != 0xf0) break;
                k += 16;
                continue;
            }
            k += r;
            if (k > 63) fail("Corrupt JPEG");
            const int z = zigzag()[k++];
            out[z] = (int16_t)(bits_.receive_extend(s) * q[z]);
        }
    }
    void block_prog_dc(Component& c, int16_t* d) {
        if (se_ != 0) fail("Corrupt JPEG");
        if (ah_ == 0) {
            std::memset(d, 0, 64 * sizeof(int16_t));
            d[0] = (int16_t)(next_dc(c) * (1 << al_));
        } else if (bits_.get(1)) {
            d[0] = (int16_t)(d[0] + (1 << al_));
        }
    }
    void refine(int16_t* p, int bit) {
        if (*p != 0 && bits_.get(1) && (*p & bit) == 0) *p = (int16_t)(*p > 0 ? *p + bit : *p - bit);
    }
    void block_prog_ac(Component& c, int16_t* d) {
        if (ss_ == 0) fail("Corrupt JPEG");
        const Huffman& h = ac_[c.ha];
        if (ah_ == 0) {
            if (eobrun_) {
                --eobrun_;
                return;
            }
            for (int k = ss_; k <= se_;) {
                const int rs = bits_.decode(h);
                const int r = rs >> 4, s = rs & 15;
                if (s == 0) {
                    if (r < 15) {
                        eobrun_ = (1 << r) + bits_.get(r) - 1;
                        break;
                    }
                    k += 16;
                    continue;
                }
                k += r;
                if (k > 63) fail("Corrupt JPEG");
                d[zigzag()[k++]] = (int16_t)(bits_.receive_extend(s) * (1 << al_));
            }
            return;
        }
        const int bit = 1 << al_;
        if (eobrun_) {
            --eobrun_;
            for (int k = ss_; k <= se_; ++k) refine(&d[zigzag()[k]], bit);
            return;
        }
        for (int k = ss_; k <= se_;) {
            const int rs = bits_.decode(h);
            int r = rs >> 4, s = rs & 15, val = 0;
            if (s == 0) {
                if (r < 15) {
                    eobrun_ = (1 << r) - 1 + bits_.get(r);
                    r = 64; /* refine the rest of the band, then stop */
                }
            } else {
                if (s != 1) fail("Corrupt JPEG");
                val = bits_.get(1) ? bit : -bit;
            }
            while (k <= se_) {
                int16_t* p = &d[zigzag()[k++]];
                if (*p != 0) {
                    refine(p, bit);
                } else {
                    if (r == 0) {
                        *p = (int16_t)val;
                        break;
                    }
                    --r;
                }
            }
        }
    }
    void scan() {
        reset_interval();
        int16_t blk[64];
        if (scan_n_ == 1) {
            /* one component: its blocks in raster order, each an MCU */
            Component& c = comp_[order_[0]];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            for (int j = 0; j < bh; ++j)
                for (int i = 0; i < bw; ++i) {
                    if (!progressive_) {
                        block_seq(c, blk);
                        idct8x8(blk, &c.data[(size_t)c.w2 * j * 8 + i * 8], c.w2);
                    } else {
                        int16_t* d = &c.coeff[64 * ((size_t)j * (c.w2 / 8) + i)];
                        if (ss_ == 0) block_prog_dc(c, d);
                        else block_prog_ac(c, d);
                    }
                    if (!mcu_done()) return;
                }
            return;
        }
        for (int j = 0; j < mcuy_; ++j)
            for (int i = 0; i < mcux_; ++i) {
                for (int k = 0; k < scan_n_; ++k) {
                    Component& c = comp_[order_[k]];
                    for (int y = 0; y < c.v; ++y)
                        for (int x = 0; x < c.h; ++x) {
                            const int bx = i * c.h + x, by = j * c.v + y;
                            if (!progressive_) {
                                block_seq(c, blk);
                                idct8x8(blk, &c.data[(size_t)c.w2 * by * 8 + bx * 8], c.w2);
                            } else {
                                block_prog_dc(c, &c.coeff[64 * ((size_t)by * (c.w2 / 8) + bx)]);
                            }
                        }
                }
                if (!mcu_done()) return;
            }
    }
    /* progressive: dequantize and transform the blocks that cover each component */
    void finish() {
        for (int n = 0; n < n_; ++n) {
            Component& c = comp_[n];
            const int bw = (c.x + 7) >> 3, bh = (c.y + 7) >> 3;
            for (int j = 0; j < bh; ++j)
                for (int i = 0; i < bw; ++i) {
                    int16_t* d = &c.coeff[64 * ((size_t)j * (c.w2 / 8) + i)];
                    for (int k = 0; k < 64; ++k) d[k] = (int16_t)(d[k] * q_[c.tq][k]);
                    idct8x8(d, &c.data[(size_t)c.w2 * j * 8 + i * 8], c.w2);
                }
        }
    }
    /* one output row of a component at full resolution (see the header) */
    static const uint8_t* upsample(uint8_t* out, const uint8_t* near, const uint8_t* far, int w, int hs, int vs) {
        if (hs == 1 && vs == 1) return near;
        if (hs == 1 && vs == 2) {
            for (int i = 0; i < w; ++i) out[i] = (uint8_t)((3 * near[i] + far[i] + 2) >> 2);
            return out;
        }
        if (hs == 2 && vs == 1) {
            if (w == 1) {
                out[0] = out[1] = near[0];
                return out;
            }
            out[0] = near[0];
            out[1] = (uint8_t)((near[0] * 3 + near[1] + 2) >> 2);
            int i = 1;
            for (; i < w - 1; ++i) {
                const int c = 3 * near[i] + 2;
                out[2 * i] = (uint8_t)((c + near[i - 1]) >> 2);
                out[2 * i + 1] = (uint8_t)((c + near[i + 1]) >> 2);
            }
            out[2 * i] = (uint8_t)((near[w - 2] * 3 + near[w - 1] + 2) >> 2);
            out[2 * i + 1] = near[w - 1];
            return out;
        }
        if (hs == 2 && vs == 2) {
            if (w == 1) {
                out[0] = out[1] = (uint8_t)((3 * near[0] + far[0] + 2) >> 2);
                return out;
            }
            int t1 = 3 * near[0] + far[0];
            out[0] = (uint8_t)((t1 + 2) >> 2);
            for (int i = 1; i < w; ++i) {
                const int t0 = t1;
                t1 = 3 * near[i] + far[i];
                out[2 * i - 1] = (uint8_t)((3 * t0 + t1 + 8) >> 4);
                out[2 * i] = (uint8_t)((3 * t1 + t0 + 8) >> 4);
            }
            out[2 * w - 1] = (uint8_t)((t1 + 2) >> 2);
            return out;
        }
        for (int i = 0; i < w; ++i)
            for (int j = 0; j < hs; ++j) out[i * hs + j] = near[i];
        return out;
    }
    static uint8_t mul255(int x, int y) { /* rounded x * y / 255 */
        const unsigned t = (unsigned)(x * y + 128);
        return (uint8_t)((t + (t >> 8)) >> 8);
    }
    static void ycc_to_rgb(uint8_t* o, int y, int cb, int cr) {
        constexpr auto F = [](float x) { return (int)(x * 4096.0f + 0.5f) << 8; };
        const int yf = (y << 20) + (1 << 19);
        cr -= 128;
        cb -= 128;
        int r = yf + cr * F(1.40200f);
        int g = yf + cr * -F(0.71414f) + (int)((unsigned)(cb * -F(0.34414f)) & 0xffff0000u);
        int b = yf + cb * F(1.77200f);
        r >>= 20;
        g >>= 20;
        b >>= 20;
        o[0] = (uint8_t)(r < 0 ? 0 : r > 255 ? 255 : r);
        o[1] = (uint8_t)(g < 0 ? 0 : g > 255 ? 255 : g);
        o[2] = (uint8_t)(b < 0 ? 0 : b > 255 ? 255 : b);
    }
    InputImage convert() {
        InputImage img = make_rgba((uint32_t)nx_, (uint32_t)ny_);
        const bool is_rgb = n_ == 3 && (rgb_ids_ == 3 || (app14_ == 0 && !jfif_));
        struct Walk {
            const uint8_t *line0, *line1;
            int hs, vs, ystep, ypos, w_lo;
            std::vector<uint8_t> buf;
        } wk[4];
        for (int k = 0; k < n_; ++k) {
            Walk& r = wk[k];
            r.hs = hmax_ / comp_[k].h;
            r.vs = vmax_ / comp_[k].v;
            r.ystep = r.vs >> 1;
            r.w_lo = (nx_ + r.hs - 1) / r.hs;
            r.ypos = 0;
            r.line0 = r.line1 = comp_[k].data.data();
            r.buf.assign((size_t)nx_ + 3 + (size_t)r.hs, 0);
        }
        const uint8_t* row[4] = {};
        for (int j = 0; j < ny_; ++j) {
            for (int k = 0; k < n_; ++k) {
                Walk& r = wk[k];
                const bool bottom = r.ystep >= (r.vs >> 1);
                row[k] = upsample(r.buf.data(), bottom ? r.line1 : r.line0, bottom ? r.line0 : r.line1, r.w_lo, r.hs,
                                  r.vs);
                if (++r.ystep >= r.vs) {
                    r.ystep = 0;
                    r.line0 = r.line1;
                    if (++r.ypos < comp_[k].y) r.line1 += comp_[k].w2;
                }
            }
            uint8_t* o = &img.data[(size_t)j * nx_ * 4];
            for (int i = 0; i < nx_; ++i, o += 4) {
                o[3] = 255;
                if (n_ == 1) {
                    o[0] = o[1] = o[2] = row[0][i];
                } else if (n_ == 3) {
                    if (is_rgb) {
                        o[0] = row[0][i];
                        o[1] = row[1][i];
                        o[2] = row[2][i];
                    } else {
                        ycc_to_rgb(o, row[0][i], row[1][i], row[2][i]);
                    }
                } else if (app14_ == 0) { /* CMYK */
                    const int m = row[3][i];
                    o[0] = mul255(row[0][i], m);
                    o[1] = mul255(row[1][i], m);
                    o[2] = mul255(row[2][i], m);
                } else if (app14_ == 2) { /* YCCK */
                    ycc_to_rgb(o, row[0][i], row[1][i], row[2][i]);
                    const int m = row[3][i];
                    o[0] = mul255(255 - o[0], m);
                    o[1] = mul255(255 - o[1], m);
                    o[2] = mul255(255 - o[2], m);
                } else { /* YCbCr with a fourth channel the reference ignores */
                    ycc_to_rgb(o, row[0][i], row[1][i], row[2][i]);
                }
            }
        }
        return img;
    }
};

}  // namespace jpg

inline InputImage decode_jpeg(const std::vector<uint8_t>& f) { return jpg::Decoder(f).decode(); }

}  // namespace iqmcli
