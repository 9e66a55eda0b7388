// progressive.cpp — progressive rendering of the Cornell box with an
// accumulation checkpoint, through the are:: host API:
//   progressive start  <width> <spp> <ckpt> [aov]  render spp samples, save
//   progressive resume <ckpt> <more> <ckpt_out> <out.pfm>
//       load, add `more` samples, save, and write the radiance mean as PFM
//   progressive pfm <in.pfm> <out.f32> <out.pfm>
//       read a PFM, write its floats (rows top-down) and write it again
//   progressive copy <ckpt> <ckpt_out>   load a checkpoint and save it again
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "are/aurora.hpp"
#include "cornell_scene.hpp"

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "";
    try {
        if (mode == "pfm" && argc == 5) {
            int w = 0, h = 0;
            const std::vector<float> rgb = are::read_pfm(argv[2], w, h);
            FILE* f = fopen(argv[3], "wb");
            if (!f || fwrite(rgb.data(), sizeof(float), rgb.size(), f) != rgb.size()) return 1;
            fclose(f);
            are::write_pfm(argv[4], w, h, rgb);
            printf("ok %d %d\n", w, h);
            return 0;
        }
        if (mode == "copy" && argc == 4) {
            const are::accumulation acc = are::accumulation::load(argv[2]);
            acc.save(argv[3]);
            printf("ok %u samples\n", acc.spp);
            return 0;
        }
        are::hittable_list world;
        add_cornell_box(world);
        if (mode == "start" && (argc == 5 || argc == 6)) {
            are::camera cam = cornell_camera(atoi(argv[2]), 1);
            are::accumulation acc;
            cam.accumulate(world, acc, atoi(argv[3]), argc == 6);
            acc.save(argv[4]);
            printf("ok %u samples\n", acc.spp);
            return 0;
        }
        if (mode == "resume" && argc == 6) {
            are::accumulation acc = are::accumulation::load(argv[2]);
            are::camera cam = cornell_camera((int)acc.camera.image_width, 1);
            cam.accumulate(world, acc, atoi(argv[3]));
            acc.save(argv[4]);
            std::vector<float> mean(acc.rgb_sum);
            for (float& v : mean) v /= (float)acc.spp;
            are::write_pfm(argv[5], (int)acc.camera.image_width, (int)acc.camera.image_height, mean);
            printf("ok %u samples\n", acc.spp);
            return 0;
        }
    } catch (const std::exception& e) {
        printf("error: %s\n", e.what());
        return 2;
    }
    fprintf(stderr, "usage: progressive start <width> <spp> <ckpt> [aov] | resume <ckpt> <more> <ckpt_out> <out.pfm>"
                    " | pfm <in.pfm> <out.f32> <out.pfm> | copy <ckpt> <ckpt_out>\n");
    return 1;
}
