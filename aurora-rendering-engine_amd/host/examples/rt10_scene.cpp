// rt10_scene — the reference program experiments/rt10.cpp's own scene
// (scene and camera constants of rt10.cpp:837-922: an open-front Cornell
// room with two reflective boxes, 900x650, depth 4) built with the
// are::patch types and rendered on the GPU; writes the P6 file the
// reference program writes (default output_rt10.ppm).
//
//   rt10_scene [out.ppm] [--stats]
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "are/patch.hpp"

using are::patch::Camera;
using are::patch::Material;
using are::patch::MaterialType;
using are::patch::RenderConfig;
using are::patch::Triangle;
using are::patch::Vec2;
using are::patch::Vec3;

namespace {

// a quad as two triangles (p00, p10, p01) and (p10, p11, p01), uv (0,0)-(1,1)
void quad(std::vector<Triangle>& s, int& id, Vec3 p00, Vec3 p10, Vec3 p11, Vec3 p01, Material* m) {
    Triangle a, b;
    a.id = id++;
    a.p[0] = p00, a.p[1] = p10, a.p[2] = p01;
    a.uv[0] = Vec2(0, 0), a.uv[1] = Vec2(1, 0), a.uv[2] = Vec2(0, 1);
    a.mat = m;
    b.id = id++;
    b.p[0] = p10, b.p[1] = p11, b.p[2] = p01;
    b.uv[0] = Vec2(1, 0), b.uv[1] = Vec2(1, 1), b.uv[2] = Vec2(0, 1);
    b.mat = m;
    s.push_back(a);
    s.push_back(b);
}

// an axis-aligned box: faces -x, +x, -y, +y, -z, +z
void box(std::vector<Triangle>& s, int& id, Vec3 lo, Vec3 hi, Material* m) {
    auto c = [&](int i, int j, int k) { return Vec3(i ? hi.x : lo.x, j ? hi.y : lo.y, k ? hi.z : lo.z); };
    quad(s, id, c(0, 0, 0), c(0, 0, 1), c(0, 1, 1), c(0, 1, 0), m);
    quad(s, id, c(1, 0, 0), c(1, 1, 0), c(1, 1, 1), c(1, 0, 1), m);
    quad(s, id, c(0, 0, 0), c(1, 0, 0), c(1, 0, 1), c(0, 0, 1), m);
    quad(s, id, c(0, 1, 0), c(0, 1, 1), c(1, 1, 1), c(1, 1, 0), m);
    quad(s, id, c(0, 0, 0), c(0, 1, 0), c(1, 1, 0), c(1, 0, 0), m);
    quad(s, id, c(0, 0, 1), c(1, 0, 1), c(1, 1, 1), c(0, 1, 1), m);
}

}  // namespace

int main(int argc, char** argv) {
    std::string out = "output_rt10.ppm";
    bool stats = false;
    for (int i = 1; i < argc; ++i) {
        if (!std::strcmp(argv[i], "--stats"))
            stats = true;
        else
            out = argv[i];
    }
    try {
        Material white{MaterialType::Diffuse, {0.85, 0.85, 0.85}, 0.0};
        Material red{MaterialType::Diffuse, {0.85, 0.25, 0.25}, 0.0};
        Material green{MaterialType::Diffuse, {0.25, 0.85, 0.25}, 0.0};
        Material blue{MaterialType::Reflective, {0.25, 0.45, 1.0}, 0.90};
        Material yellow{MaterialType::Reflective, {1.0, 0.92, 0.20}, 0.92};
        std::vector<Triangle> scene;
        int id = 1;
        quad(scene, id, {-1, 0, 0}, {1, 0, 0}, {1, 0, 2}, {-1, 0, 2}, &white);  // floor
        quad(scene, id, {-1, 2, 0}, {-1, 2, 2}, {1, 2, 2}, {1, 2, 0}, &white);  // ceiling
        quad(scene, id, {-1, 0, 2}, {1, 0, 2}, {1, 2, 2}, {-1, 2, 2}, &white);  // back wall
        quad(scene, id, {-1, 0, 0}, {-1, 0, 2}, {-1, 2, 2}, {-1, 2, 0}, &red);  // left wall
        quad(scene, id, {1, 0, 0}, {1, 2, 0}, {1, 2, 2}, {1, 0, 2}, &green);    // right wall
        box(scene, id, {-0.70, 0.0, 0.80}, {-0.15, 0.60, 1.30}, &blue);
        box(scene, id, {0.15, 0.0, 1.00}, {0.70, 1.10, 1.65}, &yellow);

        Camera cam;
        cam.width = 900;
        cam.height = 650;
        cam.origin = Vec3(0.0, 1.0, -3.0);
        const double z = -2.0, vh = 1.6, vw = vh * (double(cam.width) / double(cam.height));
        const Vec3 tl(0.0 - vw * 0.5, 1.0 - vh * 0.5, z), tr(0.0 + vw * 0.5, 1.0 - vh * 0.5, z);
        const Vec3 bl(0.0 - vw * 0.5, 1.0 + vh * 0.5, z), br(0.0 + vw * 0.5, 1.0 + vh * 0.5, z);
        cam.vpA.id = -100;
        cam.vpA.p[0] = tl, cam.vpA.p[1] = tr, cam.vpA.p[2] = bl;
        cam.vpA.uv[0] = Vec2(0, 0), cam.vpA.uv[1] = Vec2(1, 0), cam.vpA.uv[2] = Vec2(0, 1);
        cam.vpB.id = -101;
        cam.vpB.p[0] = tr, cam.vpB.p[1] = br, cam.vpB.p[2] = bl;
        cam.vpB.uv[0] = Vec2(1, 0), cam.vpB.uv[1] = Vec2(1, 1), cam.vpB.uv[2] = Vec2(0, 1);

        RenderConfig cfg;
        cfg.maxDepth = 4;
        cfg.minAreaPxToRecurse = 6.0;
        cfg.maxTriTexRes = 256;
        cfg.minTriTexRes = 16;
        cfg.envColor = Vec3(0.06, 0.07, 0.09);

        aurora_patch_stats st{};
        const are::patch::Image img = cam.render(scene, cfg, &st);
        img.writePPM(out);
        if (stats)
            std::printf("textures %u rasterised %u levels %u warps %u plan %.3f ms raster %.3f ms\n", st.textures,
                        st.raster_textures, st.levels, st.warps, st.plan_ms, st.kernel_ms);
        std::fprintf(stderr, "Done. Wrote %s\n", out.c_str());
        return 0;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "Fatal: %s\n", e.what());
        return 1;
    }
}
