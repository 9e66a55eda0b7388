// ingest_triangles.cpp — are::from_object_set on triangle objects of the
// public shape of the reference's are::Triangle / are::ObjectSet
// (get_vertices() of Q, Q+u, Q+v: include/object/triangle.h:34,
// src/object/triangle.cpp:42-45; a vector of triangle pointers:
// include/object/object_set.h:10-12), with materials of the shape of
// are::Diffuse / are::Reflective.  The scene: a 555-unit Cornell room, its
// light and two axis-aligned boxes (one a mirror), every quad given as two
// triangles (Q, Q+u, Q+v) and (Q+u+v, Q+v, Q+u).
//   ingest_triangles --dump <quads.f32>        flattened records (no GPU)
//   ingest_triangles <width> <spp> <out.f32>   radiance means, row 0 = top
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "are/from_reference.hpp"

namespace shape {  // types with the reference's public interface
struct Point3 {
    double e[3];
    double x() const { return e[0]; }
    double y() const { return e[1]; }
    double z() const { return e[2]; }
};
struct Material {
    virtual ~Material() = default;
};
struct Diffuse : Material {};
struct Reflective : Material {
    explicit Reflective(double r) : reflectivity_(r) {}
    double reflectivity_;
};
struct Triangle {
    Triangle(const Point3& Q, const Point3& u, const Point3& v) {
        vertices_ = {Q, {{Q.e[0] + u.e[0], Q.e[1] + u.e[1], Q.e[2] + u.e[2]}},
                     {{Q.e[0] + v.e[0], Q.e[1] + v.e[1], Q.e[2] + v.e[2]}}};
    }
    const std::vector<Point3>& get_vertices() const { return vertices_; }
    std::vector<Point3> vertices_;
};
struct ObjectSet {
    std::vector<Triangle*> triangles;
};
}  // namespace shape

int main(int argc, char** argv) {
    using shape::Point3;
    const bool dump = argc == 3 && std::string(argv[1]) == "--dump";
    if (!dump && argc != 4) {
        fprintf(stderr, "usage: ingest_triangles --dump <quads.f32> | <width> <spp> <out.f32>\n");
        return 1;
    }
    // quads (Q, u, v) and the material each takes: 0 red, 1 green, 2 white,
    // 3 light, 4 mirror (Reflective 0.9), 5 diffuse box (Diffuse)
    struct Q9 {
        double q[9];
        int m;
    };
    std::vector<Q9> quads = {
        {{555, 0, 0, 0, 555, 0, 0, 0, 555}, 1},       {{0, 0, 0, 0, 555, 0, 0, 0, 555}, 0},
        {{343, 554, 332, -130, 0, 0, 0, 0, -105}, 3}, {{0, 0, 0, 555, 0, 0, 0, 0, 555}, 2},
        {{555, 555, 555, -555, 0, 0, 0, 0, -555}, 2}, {{0, 0, 555, 555, 0, 0, 0, 555, 0}, 2},
    };
    auto box = [&](const double a[3], const double b[3], int m) {
        const double dx = b[0] - a[0], dy = b[1] - a[1], dz = b[2] - a[2];
        quads.push_back({{a[0], a[1], b[2], dx, 0, 0, 0, dy, 0}, m});
        quads.push_back({{b[0], a[1], b[2], 0, 0, -dz, 0, dy, 0}, m});
        quads.push_back({{b[0], a[1], a[2], -dx, 0, 0, 0, dy, 0}, m});
        quads.push_back({{a[0], a[1], a[2], 0, 0, dz, 0, dy, 0}, m});
        quads.push_back({{a[0], b[1], b[2], dx, 0, 0, 0, 0, -dz}, m});
        quads.push_back({{a[0], a[1], a[2], dx, 0, 0, 0, 0, dz}, m});
    };
    const double a1[3] = {130, 0, 65}, b1[3] = {295, 165, 230}, a2[3] = {265, 0, 295}, b2[3] = {430, 330, 460};
    box(a1, b1, 4);
    box(a2, b2, 5);

    shape::Diffuse diffuse;
    shape::Reflective reflective(0.9);
    std::vector<std::unique_ptr<shape::Triangle>> owned;
    std::map<const shape::Triangle*, int> role;
    shape::ObjectSet set;
    for (const Q9& x : quads) {
        const Point3 Q{{x.q[0], x.q[1], x.q[2]}}, u{{x.q[3], x.q[4], x.q[5]}}, v{{x.q[6], x.q[7], x.q[8]}};
        const Point3 far{{Q.e[0] + u.e[0] + v.e[0], Q.e[1] + u.e[1] + v.e[1], Q.e[2] + u.e[2] + v.e[2]}};
        const Point3 mu{{-u.e[0], -u.e[1], -u.e[2]}}, mv{{-v.e[0], -v.e[1], -v.e[2]}};
        owned.push_back(std::make_unique<shape::Triangle>(Q, u, v));
        owned.push_back(std::make_unique<shape::Triangle>(far, mu, mv));
        for (int k = 0; k < 2; ++k) {
            set.triangles.push_back(owned[owned.size() - 2 + k].get());
            role[set.triangles.back()] = x.m;
        }
    }
    auto red = std::make_shared<are::lambertian>(are::color(.65, .05, .05));
    auto green = std::make_shared<are::lambertian>(are::color(.12, .45, .15));
    auto white = std::make_shared<are::lambertian>(are::color(.73, .73, .73));
    auto light = std::make_shared<are::diffuse_light>(are::color(15, 15, 15));
    auto world = are::from_object_set(set, [&](const shape::Triangle* t) -> std::shared_ptr<are::material> {
        switch (role.at(t)) {
            case 0: return red;
            case 1: return green;
            case 2: return white;
            case 3: return light;
            case 4: return are::material_from_reference<shape::Reflective>(
                        static_cast<const shape::Material*>(&reflective), are::color(.73, .73, .73));
            default: return are::material_from_reference<shape::Reflective>(
                         static_cast<const shape::Material*>(&diffuse), are::color(.73, .73, .73));
        }
    });
    if (dump) {
        are::scene_builder sb;
        world->flatten(sb, are::xform());
        FILE* f = fopen(argv[2], "wb");
        if (!f || fwrite(sb.quads.data(), sizeof(float), sb.quads.size(), f) != sb.quads.size()) return 1;
        fclose(f);
        return 0;
    }
    are::camera cam;
    cam.image_width = atoi(argv[1]);
    cam.samples_per_pixel = atoi(argv[2]);
    cam.max_depth = 50;
    cam.background = are::color(0, 0, 0);
    cam.vfov = 40;
    cam.lookfrom = are::point3(278, 278, -800);
    cam.lookat = are::point3(278, 278, 0);
    std::vector<float> img;
    cam.render(*world, img);
    FILE* f = fopen(argv[3], "wb");
    if (!f || fwrite(img.data(), sizeof(float), img.size(), f) != img.size()) return 1;
    fclose(f);
    printf("rendered %zu triangles at %dx%d, %d spp\n", set.triangles.size(), cam.image_width, cam.image_height(),
           cam.samples_per_pixel);
    return 0;
}
