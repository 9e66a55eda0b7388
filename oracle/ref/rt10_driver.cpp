// rt10_driver.cpp — a C entry point over the reference's patch-projection
// renderer, experiments/rt10.cpp, compiled from its source in this
// translation unit (the Makefile renames the program's main so the file
// links as a library).  TEST INFRASTRUCTURE ONLY: the oracle's restatement
// (oracle/orc_patch.cpp) is checked against it image for image.
//
// Wrapped: Camera::render (rt10.cpp:755-787) over a scene of rt10
// Triangles (rt10.cpp:157-169) with Materials (:145-155) and a
// RenderConfig (:543-549), filled from the aurora_patch_* descriptors.
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "../../include/aurora.h"
#include "rt10.cpp"

namespace {
Triangle from_desc(const aurora_patch_triangle& d, std::vector<Material>& mats) {
    Triangle t;
    t.id = d.id;
    for (int i = 0; i < 3; ++i) {
        t.p[i] = Vec3(d.p[i][0], d.p[i][1], d.p[i][2]);
        t.uv[i] = Vec2(d.uv[i][0], d.uv[i][1]);
    }
    t.mat = d.material == AURORA_PATCH_NO_MATERIAL ? nullptr : &mats[d.material];
    return t;
}
}  // namespace

extern "C" int ref_rt10_render(const aurora_patch_triangle* tris, uint32_t n_tris,
                               const aurora_patch_material* mats, uint32_t n_mats,
                               const aurora_patch_camera* cam, const aurora_patch_config* cfg, double* rgb) {
    try {
        std::vector<Material> M(n_mats);
        for (uint32_t i = 0; i < n_mats; ++i) {
            M[i].type = mats[i].type == AURORA_PATCH_REFLECTIVE ? MaterialType::Reflective : MaterialType::Diffuse;
            M[i].albedo = Vec3(mats[i].albedo[0], mats[i].albedo[1], mats[i].albedo[2]);
            M[i].metalness = mats[i].metalness;
        }
        std::vector<Triangle> scene;
        for (uint32_t i = 0; i < n_tris; ++i) scene.push_back(from_desc(tris[i], M));
        Camera c;
        c.origin = Vec3(cam->origin[0], cam->origin[1], cam->origin[2]);
        std::vector<Material> none;
        c.vpA = from_desc(cam->viewport[0], none);
        c.vpB = from_desc(cam->viewport[1], none);
        c.vpA.mat = c.vpB.mat = nullptr;
        c.vpA.id = cam->viewport[0].id;
        c.vpB.id = cam->viewport[1].id;
        c.width = int(cam->width);
        c.height = int(cam->height);
        RenderConfig rc;
        rc.maxDepth = cfg->max_depth;
        rc.minAreaPxToRecurse = cfg->min_area_px;
        rc.maxTriTexRes = cfg->max_tex_res;
        rc.minTriTexRes = cfg->min_tex_res;
        rc.envColor = Vec3(cfg->env[0], cfg->env[1], cfg->env[2]);
        const Image img = c.render(scene, rc);
        for (size_t i = 0; i < img.pix.size(); ++i) {
            rgb[3 * i] = img.pix[i].x;
            rgb[3 * i + 1] = img.pix[i].y;
            rgb[3 * i + 2] = img.pix[i].z;
        }
        return 0;
    } catch (const std::exception&) {
        return 2;
    }
}
