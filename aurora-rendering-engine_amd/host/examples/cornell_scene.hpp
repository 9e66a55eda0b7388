// cornell_scene.hpp — the Cornell box of "Ray Tracing: The Next Week" (the
// reference's materials of experiments/rt9.cpp:1446-1451 in the 555-unit
// room of experiments/rt12.cpp:508-601) and its camera, shared by the
// example programs.
#pragma once

#include <memory>

#include "are/aurora.hpp"

inline void add_cornell_box(are::hittable_list& world) {
    using namespace are;
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
}

inline are::camera cornell_camera(int width, int spp) {
    are::camera cam;
    cam.aspect_ratio = 1.0;
    cam.image_width = width;
    cam.samples_per_pixel = spp;
    cam.max_depth = 50;
    cam.background = are::color(0, 0, 0);
    cam.vfov = 40;
    cam.lookfrom = are::point3(278, 278, -800);
    cam.lookat = are::point3(278, 278, 0);
    cam.vup = are::vec3(0, 1, 0);
    cam.defocus_angle = 0;
    return cam;
}
