// Shared pieces of the CLI's image decoders.  Every decoder returns the
// reference's InputImage (src/input_image.h:11-15): interleaved RGBA8, the
// layout its load_image forces with stbi_load(..., 4) (src/main.cpp:37-57).
// Failures throw DecodeError with a short reason; load_image() prefixes the
// reference's "Failed to load image '<path>', reason: ".
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "iqm/iqm.hpp"

namespace iqmcli {

struct DecodeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

[[noreturn]] inline void fail(const std::string& why) { throw DecodeError(why); }

inline uint32_t be32(const uint8_t* p) {
    return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | (uint32_t)p[3];
}

/* the reference's loader refuses sides above 2^24 (STBI_MAX_DIMENSIONS) */
constexpr uint32_t kMaxSide = 1u << 24;

/* a * b, refusing products beyond 2^31 - 1 bytes (the reference loader
 * indexes its buffers with int) */
inline size_t mul_checked(size_t a, size_t b) {
    if (a != 0 && b > (size_t)0x7fffffff / a) fail("Image too large to decode");
    return a * b;
}

inline void check_dims(uint64_t w, uint64_t h) {
    if (w == 0 || h == 0) fail("Corrupt image: zero width or height");
    if (w > kMaxSide || h > kMaxSide) fail("Very large image (corrupt?)");
    mul_checked(mul_checked((size_t)w, (size_t)h), 4);
}

/* Byte cursor over a file.  Reads past the end give 0, as the reference
 * loader's byte reader does for truncated streams of the formats without a
 * length field (BMP, TGA, GIF); callers that need the bytes check eof(). */
class Reader {
   public:
    Reader(const uint8_t* p, size_t n) : p_(p), n_(n) {}
    uint8_t u8() { return pos_ < n_ ? p_[pos_++] : (pos_++, 0); }
    uint16_t u16le() { const uint16_t a = u8(); return (uint16_t)(a | (uint16_t)u8() << 8); }
    uint16_t u16be() { const uint16_t a = u8(); return (uint16_t)((a << 8) | u8()); }
    uint32_t u32le() { const uint32_t a = u16le(); return a | (uint32_t)u16le() << 16; }
    uint32_t u32be() { const uint32_t a = u16be(); return (a << 16) | u16be(); }
    void skip(size_t k) { pos_ = (k > n_ - std::min(pos_, n_)) ? n_ + 1 : pos_ + k; }
    bool read(uint8_t* dst, size_t k) {
        if (pos_ > n_ || k > n_ - pos_) return false;
        std::memcpy(dst, p_ + pos_, k);
        pos_ += k;
        return true;
    }
    size_t pos() const { return pos_; }
    size_t size() const { return n_; }
    bool eof() const { return pos_ >= n_; }
    const uint8_t* at(size_t i) const { return p_ + i; }

   private:
    const uint8_t* p_;
    size_t n_;
    size_t pos_ = 0;
};

inline InputImage make_rgba(uint32_t w, uint32_t h) {
    check_dims(w, h);
    return InputImage{(int)w, (int)h, std::vector<unsigned char>((size_t)w * h * 4)};
}

/* n-channel 8-bit pixels to RGBA as the reference's format conversion does:
 * gray -> (g, g, g, 255), gray+alpha -> (g, g, g, a), RGB -> (r, g, b, 255) */
inline void to_rgba(const uint8_t* src, int n, uint8_t* dst) {
    switch (n) {
        case 1: dst[0] = dst[1] = dst[2] = src[0]; dst[3] = 255; break;
        case 2: dst[0] = dst[1] = dst[2] = src[0]; dst[3] = src[1]; break;
        case 3: dst[0] = src[0]; dst[1] = src[1]; dst[2] = src[2]; dst[3] = 255; break;
        default: std::memcpy(dst, src, 4); break;
    }
}

}  // namespace iqmcli
