// oracle/orc_scene.hpp — CPU oracle: scene, textures, LBVH, closest hit.
//
// TEST INFRASTRUCTURE ONLY (see orc_core.hpp).  An independent restatement
// of the algorithms libaurora runs on the GPU; written from the algorithm
// descriptions, not from the GPU sources.
#pragma once

#include <string>
#include <vector>

#include "../include/aurora.h"
#include "orc_core.hpp"

namespace orc {

struct Sphere { float c[3]; float r; };
struct Quad { float Q[3], u[3], v[3]; uint32_t kind; };

struct Perlin {
    float ranvec[256][3];
    int perm[3][256];
};

struct Scene {
    std::vector<Sphere> spheres;
    std::vector<uint32_t> sphere_mat;
    std::vector<Quad> quads;
    std::vector<uint32_t> quad_mat;
    std::vector<aurora_material> mats;
    std::vector<aurora_texture> texs;
    std::vector<float> texels;
    std::vector<Perlin> perlin;        // one per texture (only noise ones used)
    // LBVH
    std::vector<aurora_bvh_node> nodes;
    std::vector<aurora_bvh4_node> wide;  // four-wide traversal tree
    std::vector<uint32_t> sorted_morton, order;
    aurora_bvh_stats stats{};
    bool built = false;
    uint32_t n_prims() const { return uint32_t(spheres.size() + quads.size()); }
};

// Validation mirrors libaurora's documented contract; returns "" when valid.
std::string validate(const aurora_scene_desc& d);
void load(Scene& s, const aurora_scene_desc& d);
void perlin_tables(uint32_t seed, Perlin& p);

// Primitive bounds (padded) used by the LBVH.
void prim_box(const Scene& s, uint32_t id, float lo[3], float hi[3]);
void build_lbvh(Scene& s);

struct Counts { uint32_t nodes = 0, spheres = 0, quads = 0; };

template <class R>
struct Hit {
    R t;
    int32_t prim;
};

// closest hit of (o, d) in (tmin, tmax); ties go to the lower id.
template <class R>
Hit<R> closest_bvh(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax, Counts* cnt);
// closest hit over the four-wide tree (counts: 4 node records per wide node)
template <class R>
Hit<R> closest_bvh4(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax, Counts* cnt);
template <class R>
int wide_order(const bool hit[4], const R tn[4], int ord[4]);
template <class R>
Hit<R> closest_brute(const Scene& s, V3<R> o, V3<R> d, R tmin, R tmax);

// single-primitive tests; return true and t when hit in (tmin, tmax)
template <class R>
bool hit_sphere(const Sphere& sp, V3<R> o, V3<R> d, R tmin, R tmax, R& t);
template <class R>
bool hit_quad(const Quad& q, V3<R> o, V3<R> d, R tmin, R tmax, R& t, R& a, R& b);

template <class R>
V3<R> texture_value(const Scene& s, uint32_t tex, R u, R v, V3<R> p);
template <class R>
R perlin_noise(const Perlin& pl, V3<R> p);
template <class R>
R perlin_turb(const Perlin& pl, V3<R> p, int depth);

}  // namespace orc
