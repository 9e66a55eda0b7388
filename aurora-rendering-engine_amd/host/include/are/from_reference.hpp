// are/from_reference.hpp — scenes built with the reference renderer's own
// types, ingested into the path tracer's host API.
//
// The reference keeps a scene as are::ObjectSet { std::vector<Triangle*>
// triangles } (include/object/object_set.h:10-12).  Each are::Triangle(Q, u,
// v, Material*, Texture*) stores the vertices Q, Q+u, Q+v
// (src/object/triangle.cpp:42-45) and exposes them through get_vertices()
// (include/object/triangle.h:34); its Material* and Texture* are private
// and have no accessor (include/object/triangle.h:38-39), so the caller says
// which material each triangle takes (it chose them when it built the
// triangles).
//
// Everything here is a template over the reference's types, so this header
// includes none of the reference's headers: it takes are::ObjectSet and any
// type of the same shape (a `triangles` range of pointers to objects whose
// get_vertices() returns three points with x(), y(), z()).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <utility>

#include "are/aurora.hpp"

namespace are {

// One triangle (kind AURORA_QUAD_TRIANGLE) per entry of set.triangles, in
// order: Q = v0, u = v1 - v0, v = v2 - v0 from get_vertices(), in double (the
// flattening rounds them to float), with material_of(triangle pointer) as its
// material.  The list renders with camera::render and answers
// hittable::hit.  Throws std::invalid_argument for a null triangle, a
// triangle without three vertices or a null material, and for degenerate
// edges as are::Triangle's constructor does (src/object/triangle.cpp:25-36).
template <class ObjectSetT, class MaterialOf>
std::shared_ptr<hittable_list> from_object_set(const ObjectSetT& set, MaterialOf&& material_of) {
    auto list = std::make_shared<hittable_list>();
    size_t k = 0;
    for (const auto* tri : set.triangles) {
        const std::string at = "object set triangle " + std::to_string(k);
        if (!tri) throw std::invalid_argument(at + " is null");
        const auto& v = tri->get_vertices();
        if (v.size() != 3) throw std::invalid_argument(at + " does not have three vertices");
        const point3 q(v[0].x(), v[0].y(), v[0].z());
        const vec3 eu = point3(v[1].x(), v[1].y(), v[1].z()) - q;
        const vec3 ev = point3(v[2].x(), v[2].y(), v[2].z()) - q;
        const double eps = 1e-8;  // the near-zero test of are::Vec3::near_zero
        auto near_zero = [&](const vec3& e) {
            return std::fabs(e[0]) < eps && std::fabs(e[1]) < eps && std::fabs(e[2]) < eps;
        };
        if (near_zero(eu) || near_zero(ev)) throw std::invalid_argument(at + ": edge vector cannot be zero");
        if (near_zero(cross(eu, ev))) throw std::invalid_argument(at + ": edge vectors cannot be collinear");
        std::shared_ptr<material> m = material_of(tri);
        if (!m) throw std::invalid_argument(at + ": no material");
        list->add(std::make_shared<triangle>(q, eu, ev, std::move(m)));
        ++k;
    }
    return list;
}

// are::Reflective(reflectivity) (include/material/reflective.h:13-27) mirrors
// the viewpoint across the surface (src/material/reflective.cpp:9-26): here a
// metal of albedo `reflectivity` in every channel and fuzz 0.
inline std::shared_ptr<material> mirror(double reflectivity) {
    return std::make_shared<metal>(color(reflectivity, reflectivity, reflectivity), 0.0);
}

// The reference's material of a triangle as a path-tracer material:
// ReflectiveT (are::Reflective) becomes mirror(reflectivity_); anything
// else (are::Diffuse, include/material/diffuse.h, which carries no colour)
// becomes lambertian(diffuse_albedo).  Call as
// material_from_reference<are::Reflective>(m, albedo).
template <class ReflectiveT, class MaterialT>
std::shared_ptr<material> material_from_reference(const MaterialT* m, const color& diffuse_albedo) {
    if (!m) throw std::invalid_argument("Material pointer cannot be null");
    if (const auto* r = dynamic_cast<const ReflectiveT*>(m)) return mirror(r->reflectivity_);
    return std::make_shared<lambertian>(diffuse_albedo);
}

}  // namespace are
