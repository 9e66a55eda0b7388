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
#include "cornell_scene.hpp"

using namespace are;

int main(int argc, char** argv) {
    const bool dump = argc > 2 && std::string(argv[1]) == "--dump";
    const int width = argc > 1 && !dump ? atoi(argv[1]) : 600;
    const int spp = argc > 2 && !dump ? atoi(argv[2]) : 200;
    hittable_list world;
    add_cornell_box(world);

    if (dump) {
        scene_builder sb;
        world.flatten(sb, xform());
        FILE* f = fopen(argv[2], "wb");
        if (!f || fwrite(sb.quads.data(), sizeof(float), sb.quads.size(), f) != sb.quads.size()) return 1;
        fclose(f);
        return 0;
    }

    camera cam = cornell_camera(width, spp);

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
