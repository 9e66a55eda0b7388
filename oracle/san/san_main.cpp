// san_main.cpp — runs the CPU oracle and the are:: file formats under
// AddressSanitizer and UndefinedBehaviorSanitizer (test infrastructure: built
// by oracle/san/Makefile with g++, never part of libaurora).  Every entry
// point of the oracle is called on a scene with every primitive, material
// and texture kind, at a pairwise-collapsed (small) and a greedy-collapsed
// (>= 4096 primitives) tree size, plus the error paths; the PFM and
// checkpoint formats are round-tripped and fed damaged files.  Exit status 0
// and the line "san: ok" mean no report (reports abort: -fno-sanitize-recover).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "../oracle.h"
#include "are/aurora.hpp"

namespace {

int failures = 0;

void check(bool ok, const char* what) {
    if (!ok) {
        std::fprintf(stderr, "san: check failed: %s\n", what);
        ++failures;
    }
}

struct Scene {
    std::vector<float> spheres, quads, texels;
    std::vector<uint32_t> smat, qmat, qkind;
    std::vector<aurora_material> mats;
    std::vector<aurora_texture> texs;
    aurora_scene_desc desc() const {
        aurora_scene_desc d{};
        d.n_spheres = uint32_t(smat.size());
        d.spheres = spheres.data();
        d.sphere_material = smat.data();
        d.n_quads = uint32_t(qmat.size());
        d.quads = quads.data();
        d.quad_material = qmat.data();
        d.quad_kind = qkind.data();
        d.n_materials = uint32_t(mats.size());
        d.materials = mats.data();
        d.n_textures = uint32_t(texs.size());
        d.textures = texs.data();
        d.n_texels = texels.size() / 3;
        d.texels = texels.data();
        return d;
    }
};

Scene make_scene(int n_spheres, int n_quads, uint32_t seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    Scene s;
    aurora_texture t{};
    t.type = AURORA_TEX_SOLID;
    t.color[0] = 0.7f, t.color[1] = 0.5f, t.color[2] = 0.3f;
    s.texs.push_back(t);  // 0 solid
    t = {};
    t.type = AURORA_TEX_NOISE;
    t.scale = 4.0f;
    t.seed = 7;
    s.texs.push_back(t);  // 1 noise
    t = {};
    t.type = AURORA_TEX_IMAGE;
    t.width = 5, t.height = 3, t.texel_offset = 0;
    s.texs.push_back(t);  // 2 image
    for (int i = 0; i < 15 * 3; ++i) s.texels.push_back(u(g));
    t = {};
    t.type = AURORA_TEX_CHECKER;
    t.even = 0, t.odd = 2, t.scale = 0.3f;
    s.texs.push_back(t);  // 3 checker
    const uint32_t kinds[4] = {AURORA_MAT_LAMBERTIAN, AURORA_MAT_METAL, AURORA_MAT_DIELECTRIC,
                               AURORA_MAT_DIFFUSE_LIGHT};
    for (uint32_t i = 0; i < 8; ++i) {
        aurora_material m{};
        m.type = kinds[i % 4];
        m.texture = i % 4;
        m.fuzz = 0.1f * float(i);
        m.ior = 1.5f;
        s.mats.push_back(m);
    }
    for (int i = 0; i < n_spheres; ++i) {
        s.spheres.insert(s.spheres.end(), {4 * u(g) - 2, 4 * u(g) - 2, 4 * u(g) - 2, 0.02f + 0.2f * u(g)});
        s.smat.push_back(g() % 8);
    }
    for (int i = 0; i < n_quads; ++i) {
        for (int k = 0; k < 3; ++k) s.quads.push_back(4 * u(g) - 2);
        for (int k = 0; k < 6; ++k) s.quads.push_back(0.6f * u(g) - 0.3f);
        s.qmat.push_back(g() % 8);
        s.qkind.push_back(i % 2 ? AURORA_QUAD_TRIANGLE : AURORA_QUAD_PARALLELOGRAM);
    }
    return s;
}

void run_scene(int n_spheres, int n_quads, uint32_t seed) {
    Scene sc = make_scene(n_spheres, n_quads, seed);
    const aurora_scene_desc d = sc.desc();
    orc_scene* s = nullptr;
    check(orc_scene_create(&d, &s) == AURORA_OK && s, "scene create");
    aurora_bvh_stats st{};
    check(orc_bvh_build(s, &st) == AURORA_OK, "bvh build");
    const uint32_t n = st.n_prims;
    std::vector<aurora_bvh_node> nodes(st.n_nodes);
    check(orc_bvh_nodes(s, nodes.data(), st.n_nodes) == AURORA_OK, "bvh nodes");
    std::vector<uint32_t> morton(n), order(n);
    check(orc_bvh_order(s, morton.data(), order.data(), n) == AURORA_OK, "bvh order");
    std::vector<aurora_bvh4_node> wide(st.n_wide);
    check(orc_bvh_wide(s, wide.data(), st.n_wide) == AURORA_OK, "bvh wide");
    check(orc_bvh_nodes(s, nodes.data(), 0) != AURORA_OK, "short node buffer is an error");

    const uint32_t rays = 4096;
    std::mt19937 g(seed + 1);
    std::uniform_real_distribution<float> u(-1.0f, 1.0f);
    std::vector<float> o(3 * rays), dir(3 * rays);
    for (uint32_t i = 0; i < rays; ++i)
        for (int k = 0; k < 3; ++k) o[3 * i + k] = 3 * u(g), dir[3 * i + k] = u(g);
    dir[0] = 0.0f, dir[1] = 0.0f, dir[2] = 1.0f;  // axis-aligned: infinite reciprocals
    std::vector<int32_t> prim[3];
    std::vector<float> t[3];
    std::vector<uint32_t> counts(3 * rays);
    for (int b = 0; b < 3; ++b) {
        prim[b].resize(rays), t[b].resize(rays);
        check(orc_trace(s, o.data(), dir.data(), rays, 0.001f, 1e30f, b, prim[b].data(), t[b].data(),
                        b == 1 ? nullptr : counts.data(), 2) == AURORA_OK,
              "trace");
    }
    check(prim[0] == prim[1] && prim[2] == prim[1], "tree hits equal brute force");
    check(std::memcmp(t[0].data(), t[1].data(), 4 * rays) == 0, "tree t equals brute force");

    aurora_camera_desc cam{};
    cam.image_width = 24, cam.image_height = 16, cam.vfov = 50.0f;
    cam.lookfrom[2] = 6.0f, cam.vup[1] = 1.0f;
    cam.defocus_angle = 0.5f, cam.focus_dist = 6.0f;
    aurora_render_desc rd{};
    rd.spp = 3, rd.max_depth = 12, rd.seed = seed;
    rd.background[0] = rd.background[1] = rd.background[2] = 0.5f;
    const uint32_t np = cam.image_width * cam.image_height;
    std::vector<float> rgb(3 * np), alb(3 * np), nrm(3 * np);
    uint64_t segs = 0;
    check(orc_render_pixels(s, &cam, &rd, nullptr, np, 2, rgb.data(), alb.data(), nrm.data(), &segs) == AURORA_OK,
          "render f32");
    std::vector<double> rgb64(3 * 16);
    std::vector<uint32_t> pix(16);
    for (uint32_t i = 0; i < 16; ++i) pix[i] = (i * 37) % np;
    check(orc_render_pixels_f64(s, &cam, &rd, pix.data(), 16, 1, rgb64.data(), nullptr, nullptr, &segs) ==
              AURORA_OK,
          "render f64");
    std::vector<float> seg(8 * 64);
    check(orc_path_segments(s, &cam, &rd, np / 2, 1, seg.data(), 64) >= 1, "path segments");
    orc_set_render_tree(1);
    check(orc_render_pixels(s, &cam, &rd, pix.data(), 16, 1, rgb.data(), nullptr, nullptr, &segs) == AURORA_OK,
          "render over the four-wide tree");
    orc_set_render_tree(0);
    std::vector<float> co(3 * 16), cd(3 * 16);
    std::vector<uint32_t> smp(16, 2);
    check(orc_camera_rays(&cam, seed, pix.data(), smp.data(), 16, co.data(), cd.data()) == AURORA_OK, "camera rays");
    float out18[18];
    check(orc_camera(&cam, out18) == AURORA_OK, "camera frame");
    aurora_camera_desc bad = cam;
    bad.lookat[2] = bad.lookfrom[2];
    check(orc_camera(&bad, out18) != AURORA_OK, "coincident eye and target is an error");
    float tv[3], p[3] = {0.3f, -0.2f, 0.9f};
    for (uint32_t k = 0; k < 4; ++k) orc_texture_value(s, k, 0.25f, 0.75f, p, tv);
    orc_scene_destroy(s);
}

void run_kernels() {
    uint32_t ctr[4] = {1, 2, 3, 4}, key[2] = {5, 6}, r[4];
    orc_philox(ctr, key, r);
    std::vector<uint32_t> batch(4 * 64), outb(4 * 64);
    for (uint32_t i = 0; i < batch.size(); ++i) batch[i] = i * 2654435761u;
    orc_philox_batch(batch.data(), 64, 9, 10, outb.data());
    std::vector<float> ranvec(3 * 256);
    std::vector<int32_t> perm(3 * 256);
    orc_perlin(3, ranvec.data(), perm.data());
    for (double x = -7.0; x < 7.0; x += 0.37) {
        (void)orc_sin(x);
        (void)orc_atan2(x, 1.0 - x);
        (void)orc_acos(std::fmod(x, 1.0));
    }
    float v[3] = {0.6f, -0.8f, 0.0f}, nn[3] = {0.0f, 1.0f, 0.0f}, o3[3];
    orc_reflect(v, nn, o3);
    orc_refract(v, nn, 1.0f / 1.5f, o3);
    (void)orc_schlick(0.3f, 1.5f);
    orc_random_unit(1, 2, 3, 4, o3);
    float tw[4] = {0.5f, -0.0f, 2.0f, 0.0f};
    int32_t hw[4] = {1, 0, 1, 1}, ord[4];
    check(orc_wide_order(tw, hw, ord) == 3, "wide order");
    float sph[4] = {0, 0, -3, 1}, org[3] = {0, 0, 0}, dd[3] = {0, 0, -1}, th, tab[3];
    check(orc_hit_sphere(sph, org, dd, 0.001f, 1e30f, &th) == 1, "sphere hit");
    float quad[9] = {-1, -1, -2, 2, 0, 0, 0, 2, 0};
    check(orc_hit_quad(quad, AURORA_QUAD_PARALLELOGRAM, org, dd, 0.001f, 1e30f, tab) == 1, "quad hit");
    float uv[2];
    orc_sphere_uv(nn, uv);

    // paste: random corners around and outside the target, degenerate ones
    std::mt19937 g(11);
    std::uniform_int_distribution<int> c(-40, 90);
    std::vector<float> src(7 * 5 * 3), dst(48 * 32 * 3);
    for (auto& x : src) x = std::generate_canonical<float, 24>(g);
    for (int k = 0; k < 64; ++k) {
        int32_t corners[8];
        for (int& x : corners) x = c(g);
        (void)orc_paste(src.data(), 7, 5, dst.data(), 48, 32, corners);
    }
    const int32_t flat[8] = {3, 3, 3, 3, 3, 3, 3, 3};
    check(orc_paste(src.data(), 7, 5, dst.data(), 48, 32, flat) == 0, "degenerate paste paints nothing");
    std::vector<float> sums(3 * 40);
    for (size_t i = 0; i < sums.size(); ++i) sums[i] = 0.25f * float(i) - 3.0f;
    std::vector<uint8_t> bytes(sums.size());
    for (uint32_t tr = 0; tr < 3; ++tr)
        check(orc_resolve_rgb8(sums.data(), sums.size(), 4, tr, bytes.data()) == AURORA_OK, "resolve");
    check(orc_resolve_rgb8(sums.data(), sums.size(), 4, 9u, bytes.data()) != AURORA_OK, "unknown transfer");
}

void run_formats(const std::string& dir) {
    const int w = 5, h = 3;
    std::vector<float> rgb(3 * w * h);
    for (size_t i = 0; i < rgb.size(); ++i) rgb[i] = 0.1f * float(i) - 0.7f;
    const std::string pfm = dir + "/san.pfm";
    are::write_pfm(pfm, w, h, rgb);
    int rw = 0, rh = 0;
    check(are::read_pfm(pfm, rw, rh) == rgb && rw == w && rh == h, "pfm round trip");
    // damaged files: every truncation of the PFM must throw, never read past the end
    std::vector<unsigned char> bytes;
    {
        FILE* f = std::fopen(pfm.c_str(), "rb");
        int ch;
        while ((ch = std::fgetc(f)) != EOF) bytes.push_back((unsigned char)ch);
        std::fclose(f);
    }
    const std::string cut = dir + "/san_cut.bin";
    int thrown = 0;
    for (size_t len = 0; len < bytes.size(); len += 3) {
        FILE* f = std::fopen(cut.c_str(), "wb");
        std::fwrite(bytes.data(), 1, len, f);
        std::fclose(f);
        try {
            (void)are::read_pfm(cut, rw, rh);
        } catch (const std::runtime_error&) {
            ++thrown;
        }
    }
    check(thrown == int((bytes.size() + 2) / 3), "every truncated PFM throws");

    are::accumulation a;
    a.camera.image_width = w, a.camera.image_height = h, a.camera.vfov = 40.0f;
    a.render.max_depth = 9, a.render.seed = 3;
    a.spp = 17;
    a.rgb_sum = rgb, a.albedo_sum = rgb, a.normal_sum = rgb;
    const std::string ck = dir + "/san.acc";
    a.save(ck);
    are::accumulation b = are::accumulation::load(ck);
    check(b.spp == 17 && b.rgb_sum == rgb && b.normal_sum == rgb, "checkpoint round trip");
    bytes.clear();
    {
        FILE* f = std::fopen(ck.c_str(), "rb");
        int ch;
        while ((ch = std::fgetc(f)) != EOF) bytes.push_back((unsigned char)ch);
        std::fclose(f);
    }
    thrown = 0;
    int tries = 0;
    for (size_t pos = 0; pos < bytes.size(); pos += 7, ++tries) {
        std::vector<unsigned char> bad = bytes;
        bad[pos] ^= 0x40;
        FILE* f = std::fopen(cut.c_str(), "wb");
        std::fwrite(bad.data(), 1, bad.size(), f);
        std::fclose(f);
        try {
            (void)are::accumulation::load(cut);
        } catch (const std::runtime_error&) {
            ++thrown;
        }
    }
    // a flipped bit in the header's padding is not checked; everything else is
    check(thrown >= tries * 3 / 4, "damaged checkpoints throw");
    for (size_t len = 0; len < bytes.size(); len += 11) {
        FILE* f = std::fopen(cut.c_str(), "wb");
        std::fwrite(bytes.data(), 1, len, f);
        std::fclose(f);
        try {
            (void)are::accumulation::load(cut);
            check(false, "truncated checkpoint loads");
        } catch (const std::runtime_error&) {
        }
    }
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    run_kernels();
    run_scene(40, 60, 1);      // pairwise four-wide collapse
    run_scene(2500, 2500, 2);  // greedy collapse
    Scene empty = make_scene(0, 0, 3);
    aurora_scene_desc d = empty.desc();
    orc_scene* s = nullptr;
    if (orc_scene_create(&d, &s) == AURORA_OK) {
        aurora_bvh_stats st{};
        (void)orc_bvh_build(s, &st);
        orc_scene_destroy(s);
    }
    Scene broken = make_scene(3, 3, 4);
    broken.smat[1] = 99;  // material index out of range
    d = broken.desc();
    s = nullptr;
    check(orc_scene_create(&d, &s) != AURORA_OK, "bad material index is an error");
    if (s) orc_scene_destroy(s);
    run_formats(dir);
    if (failures) return 1;
    std::printf("san: ok\n");
    return 0;
}
