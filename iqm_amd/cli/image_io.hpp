// load_image of the CLI (src/main.cpp:37-57): the file's format is found by
// its header in the reference loader's order (PNG, BMP, GIF, JPEG, PNM, then
// TGA, whose header has no signature), decoded to RGBA8, and failures become
// "Failed to load image '<path>', reason: <reason>".  PSD, Softimage PIC and
// Radiance HDR, which the reference's loader also reads, are refused here
// with a reason that names them.
#pragma once

#include <fstream>
#include <iterator>

#include "image_common.hpp"
#include "jpeg_io.hpp"
#include "png_io.hpp"
#include "raster_io.hpp"

namespace iqmcli {

inline std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("Unable to open file");
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

inline bool starts_with(const std::vector<uint8_t>& f, const char* magic, size_t n) {
    return f.size() >= n && std::memcmp(f.data(), magic, n) == 0;
}

/* decodes an image file's bytes (any supported format) */
inline InputImage decode_image(const std::vector<uint8_t>& f) {
    if (starts_with(f, "\x89PNG\r\n\x1a\n", 8)) return decode_png(f);
    if (f.size() >= 18 && f[0] == 'B' && f[1] == 'M') {
        const uint32_t hsz = f[14] | f[15] << 8 | f[16] << 16 | (uint32_t)f[17] << 24;
        if (hsz == 12 || hsz == 40 || hsz == 56 || hsz == 108 || hsz == 124) return decode_bmp(f);
    }
    if (f.size() >= 6 && (starts_with(f, "GIF87a", 6) || starts_with(f, "GIF89a", 6))) return decode_gif(f);
    if (starts_with(f, "8BPS", 4)) fail("PSD images are not supported by this build");
    if (f.size() >= 92 && f[0] == 0x53 && f[1] == 0x80 && f[2] == 0xf6 && f[3] == 0x34 &&
        std::memcmp(&f[88], "PICT", 4) == 0)
        fail("Softimage PIC images are not supported by this build");
    if (f.size() >= 2 && f[0] == 0xff && f[1] == 0xd8) return decode_jpeg(f);
    if (f.size() >= 2 && f[0] == 'P' && (f[1] == '5' || f[1] == '6')) return decode_pnm(f);
    if (starts_with(f, "#?RADIANCE\n", 11) || starts_with(f, "#?RGBE\n", 7))
        fail("HDR images are not supported by this build");
    if (tga_test(f)) return decode_tga(f);
    fail("Image not of any known type, or corrupt");
}

/* load_image (src/main.cpp:37-57) with the reference's failure message */
inline InputImage load_image(const std::string& filename) {
    try {
        return decode_image(read_file(filename));
    } catch (const DecodeError& e) {
        throw std::runtime_error("Failed to load image '" + filename + "', reason: " + e.what());
    }
}

}  // namespace iqmcli
