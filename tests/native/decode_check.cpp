// Test driver for the CLI's image decoders (iqm_amd/cli/image_io.hpp): for
// each file argument writes <file>.rgba (width, height as two little-endian
// uint32, then RGBA8 rows) or prints "ERR <file>: <reason>".
#include <cstdio>
#include <string>

#include "image_io.hpp"

int main(int argc, char** argv) {
    int bad = 0;
    for (int i = 1; i < argc; ++i) {
        try {
            const InputImage img = iqmcli::decode_image(iqmcli::read_file(argv[i]));
            const std::string out = std::string(argv[i]) + ".rgba";
            FILE* f = std::fopen(out.c_str(), "wb");
            if (!f) return 2;
            const uint32_t wh[2] = {(uint32_t)img.width, (uint32_t)img.height};
            std::fwrite(wh, 4, 2, f);
            std::fwrite(img.data.data(), 1, img.data.size(), f);
            std::fclose(f);
        } catch (const std::exception& e) {
            std::printf("ERR %s: %s\n", argv[i], e.what());
            ++bad;
        }
    }
    return 0;
}
