// are/patch.hpp — host C++ API of the patch-projection renderer
// (namespace are::patch), in the terms of the reference program
// experiments/rt10.cpp: Vec2 / Vec3 in double, MaterialType and Material
// (rt10.cpp:145-155), Triangle with raw Material pointers and integer ids
// (:157-169), RenderConfig (:543-549), Image with writePPM (:79-139) and
// Camera with its two viewport triangles and render() (:670-787).  Code
// written against those types keeps its scene set-up; Camera::render runs
// aurora_patch_render on the GPU and returns the reference's image bit for
// bit.  Errors throw std::invalid_argument (bad descriptors) or
// std::runtime_error (device failures, output files), the exception types
// of the are:: API.
#pragma once

#include <string>
#include <vector>

#include "../../../../include/aurora.h"

namespace are {
namespace patch {

struct Vec2 {
    double x = 0, y = 0;
    Vec2() = default;
    Vec2(double x_, double y_) : x(x_), y(y_) {}
};

struct Vec3 {
    double x = 0, y = 0, z = 0;
    Vec3() = default;
    Vec3(double x_, double y_, double z_) : x(x_), y(y_), z(z_) {}
};

enum class MaterialType { Diffuse, Reflective };

struct Material {
    MaterialType type = MaterialType::Diffuse;
    Vec3 albedo = {1, 1, 1};
    double metalness = 0.0;  // reflective: weight of the tinted reflection, clamped to [0,1]
};

struct Triangle {
    int id = -1;
    Vec3 p[3];
    Vec2 uv[3];
    Material* mat = nullptr;  // nullptr: white diffuse
};

struct RenderConfig {
    int maxDepth = 4;
    double minAreaPxToRecurse = 4.0;
    int maxTriTexRes = 256;
    int minTriTexRes = 16;
    Vec3 envColor = {0.08, 0.08, 0.10};
};

// width x height linear RGB, row 0 at the top
struct Image {
    int w = 0, h = 0;
    std::vector<Vec3> pix;
    Image() = default;
    Image(int w_, int h_, Vec3 fill = {0, 0, 0}) : w(w_), h(h_), pix(size_t(w_) * size_t(h_), fill) {}
    Vec3 get(int x, int y) const;  // coordinates clamped to the image
    // P6 file: clamp to [0,1], c^(1/gamma), 255 c rounded half away from zero
    void writePPM(const std::string& path, double gamma = 2.2) const;
};

struct Camera {
    Vec3 origin;
    Triangle vpA;  // viewport triangles: their uv place them in the image
    Triangle vpB;
    int width = 800;
    int height = 600;
    int device = 0;  // GPU the render runs on
    Image render(const std::vector<Triangle>& scene, const RenderConfig& cfg) const;
    // the same, with the plan and raster statistics of the call
    Image render(const std::vector<Triangle>& scene, const RenderConfig& cfg, aurora_patch_stats* stats) const;
};

}  // namespace patch
}  // namespace are
