// ingest_reference.cpp — how code written against the reference renderer's
// scene types hands its scene to the path tracer: an are::ObjectSet of
// are::Triangle (include/object/object_set.h:10-12) whose materials are
// are::Diffuse / are::Reflective becomes an are::hittable_list through
// are::from_object_set and renders on the GPU.
//
// Compiled against the reference's headers (-I <reference>/include).  A
// program that also constructs are::Triangle objects links the reference's
// sources; are::Triangle::trace_texture is declared there
// (include/object/triangle.h:31) but not defined, which such a program has
// to supply.
#include <map>
#include <memory>
#include <vector>

#include <material/diffuse.h>
#include <material/reflective.h>
#include <object/object_set.h>

#include "are/from_reference.hpp"

// Radiance means (height x width x 3, row 0 = top) of an object set whose
// triangles' materials the caller recorded when it built them.  Diffuse
// triangles take `albedo`, reflective ones become mirrors.
std::vector<float> render_object_set(const are::ObjectSet& set,
                                     const std::map<const are::Triangle*, const are::Material*>& material_of,
                                     const are::color& albedo, are::camera& cam) {
    auto world = are::from_object_set(set, [&](const are::Triangle* t) {
        const auto it = material_of.find(t);
        return are::material_from_reference<are::Reflective>(it == material_of.end() ? nullptr : it->second, albedo);
    });
    std::vector<float> img;
    cam.render(*world, img);
    return img;
}
