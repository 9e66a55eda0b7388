// ppm_tool.cpp — the P6 reader and writer of the are:: host API from the
// command line (no GPU needed):
//   ppm_tool load <in.ppm> <out.f32>   are::image_texture(path): prints
//       "ok <w> <h>" and writes the linear texels (float32, rows top-down),
//       or prints "error: <message>" and exits 2 when the loader throws;
//   ppm_tool save <in.f32> <w> <h> <spp> <transfer> <out.ppm>
//       resolves a radiance sum (float32, w*h*3) with aurora_resolve_rgb8
//       and writes it with are::write_ppm.
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "are/aurora.hpp"

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        if (mode == "load" && argc == 4) {
            are::image_texture t{std::string(argv[2])};
            FILE* f = fopen(argv[3], "wb");
            if (!f || fwrite(t.rgb.data(), sizeof(float), t.rgb.size(), f) != t.rgb.size()) return 1;
            fclose(f);
            printf("ok %d %d\n", t.width, t.height);
            return 0;
        }
        if (mode == "save" && argc == 8) {
            const int w = atoi(argv[3]), h = atoi(argv[4]);
            const uint32_t spp = (uint32_t)atoi(argv[5]), transfer = (uint32_t)atoi(argv[6]);
            if (w <= 0 || h <= 0) throw std::invalid_argument("width and height must be positive");
            std::vector<float> sum((size_t)w * h * 3);
            FILE* f = fopen(argv[2], "rb");
            if (!f || fread(sum.data(), sizeof(float), sum.size(), f) != sum.size()) return 1;
            fclose(f);
            std::vector<uint8_t> px(sum.size());
            if (aurora_resolve_rgb8(sum.data(), (uint32_t)w, (uint32_t)h, spp, transfer, px.data()) != AURORA_OK)
                throw std::invalid_argument(aurora_last_error(nullptr));
            are::write_ppm(argv[7], w, h, px);
            printf("ok %d %d\n", w, h);
            return 0;
        }
    } catch (const std::exception& e) {
        printf("error: %s\n", e.what());
        return 2;
    }
    fprintf(stderr, "usage: ppm_tool load <in.ppm> <out.f32> | save <in.f32> <w> <h> <spp> <transfer> <out.ppm>\n");
    return 1;
}
