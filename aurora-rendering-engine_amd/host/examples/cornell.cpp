// cornell.cpp — the Cornell box of "Ray Tracing: The Next Week" written
// against the are:: host API, rendered on the GPU.
//   cornell <width> <spp> <out.f32> [out.ppm]
// writes the linear RGB means (float32, row 0 = top) to out.f32;
//   cornell --dump <quads.f32>
// writes the flattened quad records (9 floats each) without touching a GPU.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <vector>

#include "are/aurora.hpp"

using namespace are;

int main(int argc, char** argv) {
    const bool dump = argc > 2 && std::string(argv[1]) == "--dump";
    const int width = argc > 1 && !dump ? atoi(argv[1]) : 600;
    const int spp = argc > 2 && !dump ? atoi(argv[2]) : 200;
    hittable_list world;
    auto red = std::make_shared<lambertian>(color(.65, .05, .05));
    auto white = std::make_shared<lambertian>(color(.73, .73, .73));
    auto green = std::make_shared<lambertian>(color(.12, .45, .15));
    auto light = std::make_shared<diffuse_light>(color(15, 15, 15));
    world.add(std::make_shared<quad>(point3(555, 0, 0), vec3(0, 555, 0), vec3(0, 0, 555), green));
    world.add(std::make_shared<quad>(point3(0, 0, 0), vec3(0, 555, 0), vec3(0, 0, 555), red));
    world.add(std::make_shared<quad>(point3(343, 554, 332), vec3(-130, 0, 0), vec3(0, 0, -105), light));
    world.add(std::make_shared<quad>(point3(0, 0, 0), vec3(555, 0, 0), vec3(0, 0, 555), white));
    world.add(std::make_shared<quad>(point3(555, 555, 555), vec3(-555, 0, 0), vec3(0, 0, -555), white));
    world.add(std::make_shared<quad>(point3(0, 0, 555), vec3(555, 0, 0), vec3(0, 555, 0), white));
    std::shared_ptr<hittable> box1 = box(point3(0, 0, 0), point3(165, 330, 165), white);
    box1 = std::make_shared<rotate_y>(box1, 15);
    box1 = std::make_shared<translate>(box1, vec3(265, 0, 295));
    world.add(box1);
    std::shared_ptr<hittable> box2 = box(point3(0, 0, 0), point3(165, 165, 165), white);
    box2 = std::make_shared<rotate_y>(box2, -18);
    box2 = std::make_shared<translate>(box2, vec3(130, 0, 65));
    world.add(box2);

    if (dump) {
        scene_builder sb;
        world.flatten(sb, xform());
        FILE* f = fopen(argv[2], "wb");
        if (!f || fwrite(sb.quads.data(), sizeof(float), sb.quads.size(), f) != sb.quads.size()) return 1;
        fclose(f);
        return 0;
    }

    camera cam;
    cam.aspect_ratio = 1.0;
    cam.image_width = width;
    cam.samples_per_pixel = spp;
    cam.max_depth = 50;
    cam.background = color(0, 0, 0);
    cam.vfov = 40;
    cam.lookfrom = point3(278, 278, -800);
    cam.lookat = point3(278, 278, 0);
    cam.vup = vec3(0, 1, 0);
    cam.defocus_angle = 0;

    // a ray query through the same API: straight at the back wall's centre
    hit_record rec;
    if (!world.hit(ray(point3(278, 278, -800), vec3(0, 0, 1)), interval(0.001, INFINITY), rec)) {
        fprintf(stderr, "expected a hit\n");
        return 1;
    }
    printf("hit prim %d t %.6f normal (%.3f %.3f %.3f)\n", rec.prim, rec.t, rec.normal.x(), rec.normal.y(),
           rec.normal.z());

    std::vector<float> img;
    cam.render(world, img);
    printf("rendered %dx%d at %d spp: %llu paths, %llu segments\n", width, cam.image_height(), spp,
           (unsigned long long)cam.last_stats.paths, (unsigned long long)cam.last_stats.segments);
    if (argc > 3) {
        FILE* f = fopen(argv[3], "wb");
        if (!f || fwrite(img.data(), sizeof(float), img.size(), f) != img.size()) return 1;
        fclose(f);
    }
    if (argc > 4) cam.render_ppm(world, argv[4]);
    return 0;
}
