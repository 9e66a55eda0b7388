// aurora_patch.cpp — are::patch (are/patch.hpp): rt10.cpp's scene types
// flattened into the aurora_patch_* descriptors of include/aurora.h and
// rendered by aurora_patch_render on the GPU.
#include "are/patch.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>

#include "are/aurora.hpp"

namespace are {
namespace patch {

namespace {

aurora_patch_triangle descriptor(const Triangle& t, uint32_t material) {
    aurora_patch_triangle d{};
    for (int i = 0; i < 3; ++i) {
        d.p[i][0] = t.p[i].x;
        d.p[i][1] = t.p[i].y;
        d.p[i][2] = t.p[i].z;
        d.uv[i][0] = t.uv[i].x;
        d.uv[i][1] = t.uv[i].y;
    }
    d.id = t.id;
    d.material = material;
    return d;
}

}  // namespace

Vec3 Image::get(int x, int y) const {
    x = std::max(0, std::min(w - 1, x));
    y = std::max(0, std::min(h - 1, y));
    return pix[size_t(y) * size_t(w) + size_t(x)];
}

void Image::writePPM(const std::string& path, double gamma) const {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("Failed to open output file: " + path);
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    std::vector<unsigned char> row(size_t(w) * 3);
    auto byte = [gamma](double c) {
        c = std::max(0.0, std::min(1.0, c));
        c = std::pow(c, 1.0 / gamma);
        const long r = std::lround(c * 255.0);
        return (unsigned char)std::max(0L, std::min(255L, r));
    };
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const Vec3& c = pix[size_t(y) * size_t(w) + size_t(x)];
            row[3 * size_t(x)] = byte(c.x);
            row[3 * size_t(x) + 1] = byte(c.y);
            row[3 * size_t(x) + 2] = byte(c.z);
        }
        if (std::fwrite(row.data(), 1, row.size(), f) != row.size()) {
            std::fclose(f);
            throw std::runtime_error("Failed to write output file: " + path);
        }
    }
    if (std::fclose(f) != 0) throw std::runtime_error("Failed to write output file: " + path);
}

Image Camera::render(const std::vector<Triangle>& scene, const RenderConfig& cfg) const {
    return render(scene, cfg, nullptr);
}

Image Camera::render(const std::vector<Triangle>& scene, const RenderConfig& cfg, aurora_patch_stats* stats) const {
    // materials by pointer, numbered in first-use order
    std::map<const Material*, uint32_t> ids;
    std::vector<aurora_patch_material> mats;
    std::vector<aurora_patch_triangle> tris;
    tris.reserve(scene.size());
    for (const Triangle& t : scene) {
        uint32_t m = AURORA_PATCH_NO_MATERIAL;
        if (t.mat) {
            auto it = ids.find(t.mat);
            if (it == ids.end()) {
                aurora_patch_material d{};
                d.type = t.mat->type == MaterialType::Reflective ? AURORA_PATCH_REFLECTIVE : AURORA_PATCH_DIFFUSE;
                d.albedo[0] = t.mat->albedo.x;
                d.albedo[1] = t.mat->albedo.y;
                d.albedo[2] = t.mat->albedo.z;
                d.metalness = t.mat->metalness;
                it = ids.emplace(t.mat, uint32_t(mats.size())).first;
                mats.push_back(d);
            }
            m = it->second;
        }
        tris.push_back(descriptor(t, m));
    }
    if (width < 0 || height < 0) throw std::invalid_argument("Camera: negative image size");
    aurora_patch_camera cam{};
    cam.origin[0] = origin.x;
    cam.origin[1] = origin.y;
    cam.origin[2] = origin.z;
    cam.viewport[0] = descriptor(vpA, AURORA_PATCH_NO_MATERIAL);
    cam.viewport[1] = descriptor(vpB, AURORA_PATCH_NO_MATERIAL);
    cam.width = uint32_t(width);
    cam.height = uint32_t(height);
    aurora_patch_config c{};
    c.max_depth = cfg.maxDepth;
    c.max_tex_res = cfg.maxTriTexRes;
    c.min_tex_res = cfg.minTriTexRes;
    c.min_area_px = cfg.minAreaPxToRecurse;
    c.env[0] = cfg.envColor.x;
    c.env[1] = cfg.envColor.y;
    c.env[2] = cfg.envColor.z;
    aurora_ctx* ctx = default_context(device);
    std::vector<double> rgb(size_t(width) * size_t(height) * 3);
    const int rc = aurora_patch_render(ctx, tris.data(), uint32_t(tris.size()), mats.data(), uint32_t(mats.size()),
                                       &cam, &c, rgb.data(), stats);
    if (rc == AURORA_ERR_INVALID) throw std::invalid_argument(aurora_last_error(ctx));
    if (rc != AURORA_OK) throw std::runtime_error(aurora_last_error(ctx));
    Image img(width, height);
    for (size_t i = 0; i < img.pix.size(); ++i) img.pix[i] = Vec3(rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]);
    return img;
}

}  // namespace patch
}  // namespace are
