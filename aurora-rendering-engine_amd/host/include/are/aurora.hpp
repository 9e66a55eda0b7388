// are/aurora.hpp — host C++ API of the path tracer (namespace are).
//
// The classes keep the hittable / material / texture / camera shape that
// scene code written for a "Ray Tracing in One Weekend"-style renderer
// expects: build a world from spheres, quads, boxes, rotate_y / translate
// wrappers and bvh_node, give objects lambertian / metal / dielectric /
// diffuse_light materials with solid / checker / noise / image textures,
// then call camera::render(world).  Rendering does not run on the host: the
// world is flattened into structure-of-arrays buffers (aurora_scene_desc)
// and traced by libaurora's HIP kernels on the GPU through the C ABI of
// include/aurora.h.  hittable::hit is answered by the same GPU traversal.
//
// Reference anchors: the reference's scene classes are are::Object /
// are::Triangle / are::ObjectSet with raw Material* / Texture* pointers
// (include/object/object.h:14-39, include/object/triangle.h:19,38-39,
// include/object/object_set.h:10-12) and its only renderer loop is
// experiments/rt.cpp:337-374.  Errors surface as std::invalid_argument
// (bad input) and std::runtime_error (device failures), the exception types
// of the reference's constructors (src/object/triangle.cpp:13-36,
// src/texture.cpp:11-43).
#pragma once

#include <cmath>
#include <cstdint>
#include <map>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "../../../../include/aurora.h"

namespace are {

// ------------------------------------------------------------- vec3 -----
struct vec3 {
    double e[3] = {0, 0, 0};
    vec3() = default;
    vec3(double x, double y, double z) : e{x, y, z} {}
    double x() const { return e[0]; }
    double y() const { return e[1]; }
    double z() const { return e[2]; }
    double operator[](int i) const { return e[i]; }
    double& operator[](int i) { return e[i]; }
    vec3 operator-() const { return vec3(-e[0], -e[1], -e[2]); }
    double length_squared() const { return e[0] * e[0] + e[1] * e[1] + e[2] * e[2]; }
    double length() const { return std::sqrt(length_squared()); }
};
using point3 = vec3;
using color = vec3;
inline vec3 operator+(const vec3& a, const vec3& b) { return vec3(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
inline vec3 operator-(const vec3& a, const vec3& b) { return vec3(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
inline vec3 operator*(double t, const vec3& v) { return vec3(t * v[0], t * v[1], t * v[2]); }
inline vec3 operator*(const vec3& v, double t) { return t * v; }
inline vec3 operator/(const vec3& v, double t) { return (1 / t) * v; }
inline double dot(const vec3& a, const vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline vec3 cross(const vec3& a, const vec3& b) {
    return vec3(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
inline vec3 unit_vector(const vec3& v) { return v / v.length(); }

class ray {
   public:
    ray() = default;
    ray(const point3& origin, const vec3& direction) : orig(origin), dir(direction) {}
    const point3& origin() const { return orig; }
    const vec3& direction() const { return dir; }
    point3 at(double t) const { return orig + t * dir; }

   private:
    point3 orig;
    vec3 dir;
};

struct interval {
    double min = +INFINITY, max = -INFINITY;
    interval() = default;
    interval(double lo, double hi) : min(lo), max(hi) {}
    bool contains(double x) const { return min <= x && x <= max; }
    bool surrounds(double x) const { return min < x && x < max; }
};

// --------------------------------------------------------- textures -----
class texture {
   public:
    virtual ~texture() = default;
    virtual aurora_texture describe(class scene_builder& sb) const = 0;
};

class solid_color : public texture {
   public:
    explicit solid_color(const color& albedo) : albedo(albedo) {}
    solid_color(double r, double g, double b) : albedo(r, g, b) {}
    aurora_texture describe(scene_builder& sb) const override;
    color albedo;
};

class checker_texture : public texture {
   public:
    checker_texture(double scale, std::shared_ptr<texture> even, std::shared_ptr<texture> odd)
        : scale(scale), even(std::move(even)), odd(std::move(odd)) {}
    checker_texture(double scale, const color& c1, const color& c2)
        : checker_texture(scale, std::make_shared<solid_color>(c1), std::make_shared<solid_color>(c2)) {}
    aurora_texture describe(scene_builder& sb) const override;
    double scale;
    std::shared_ptr<texture> even, odd;
};

class noise_texture : public texture {
   public:
    explicit noise_texture(double scale, uint32_t seed = 0) : scale(scale), seed(seed) {}
    aurora_texture describe(scene_builder& sb) const override;
    double scale;
    uint32_t seed;
};

// RGB image sampled bilinearly (the sampling of src/texture.cpp:310-330);
// loads binary PPM (P6, maxval 255) like are::Texture (src/texture.cpp:9-50).
class image_texture : public texture {
   public:
    explicit image_texture(const std::string& ppm_path);
    image_texture(int width, int height, std::vector<float> rgb_linear);
    aurora_texture describe(scene_builder& sb) const override;
    int width = 0, height = 0;
    std::vector<float> rgb;  // linear RGB, row 0 = top
};

// -------------------------------------------------------- materials -----
class material {
   public:
    virtual ~material() = default;
    virtual aurora_material describe(scene_builder& sb) const = 0;
};

class lambertian : public material {
   public:
    explicit lambertian(const color& albedo) : tex(std::make_shared<solid_color>(albedo)) {}
    explicit lambertian(std::shared_ptr<texture> tex) : tex(std::move(tex)) {}
    aurora_material describe(scene_builder& sb) const override;
    std::shared_ptr<texture> tex;
};

class metal : public material {
   public:
    metal(const color& albedo, double fuzz) : tex(std::make_shared<solid_color>(albedo)), fuzz(fuzz < 1 ? fuzz : 1) {}
    aurora_material describe(scene_builder& sb) const override;
    std::shared_ptr<texture> tex;
    double fuzz;
};

class dielectric : public material {
   public:
    explicit dielectric(double refraction_index) : refraction_index(refraction_index) {}
    aurora_material describe(scene_builder& sb) const override;
    double refraction_index;
};

class diffuse_light : public material {
   public:
    explicit diffuse_light(const color& emit) : tex(std::make_shared<solid_color>(emit)) {}
    explicit diffuse_light(std::shared_ptr<texture> tex) : tex(std::move(tex)) {}
    aurora_material describe(scene_builder& sb) const override;
    std::shared_ptr<texture> tex;
};

// ------------------------------------------------------------ scene -----
// Affine map applied while flattening: rotation about +y then translation.
struct xform {
    double m[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    vec3 t;
    vec3 point(const vec3& p) const { return vector(p) + t; }
    vec3 vector(const vec3& v) const {
        return vec3(m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2], m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2],
                    m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2]);
    }
    xform then(const xform& outer) const;  // outer ∘ this
};

// Collects the flattened arrays of aurora_scene_desc.
class scene_builder {
   public:
    uint32_t material_id(const std::shared_ptr<material>& m);
    uint32_t texture_id(const texture* t);
    void add_sphere(const point3& c, double r, uint32_t mat);
    void add_quad(const point3& Q, const vec3& u, const vec3& v, uint32_t mat, uint32_t kind);
    uint32_t add_texels(const std::vector<float>& rgb);  // returns first texel
    aurora_scene_desc desc() const;

    std::vector<float> spheres, quads, texels;
    std::vector<uint32_t> sphere_mat, quad_mat, quad_kind;
    std::vector<aurora_material> mats;
    std::vector<aurora_texture> texs;

   private:
    std::map<const material*, uint32_t> mat_ids_;
    std::map<const texture*, uint32_t> tex_ids_;
    std::vector<std::shared_ptr<material>> keep_;
};

struct hit_record {
    point3 p;
    vec3 normal;
    std::shared_ptr<material> mat;
    double t = 0, u = 0, v = 0;
    bool front_face = false;
    int32_t prim = AURORA_MISS;
};

class hittable {
   public:
    virtual ~hittable() = default;
    virtual void flatten(scene_builder& sb, const xform& x) const = 0;
    // Closest hit of r in (ray_t.min, ray_t.max), traced on the GPU against
    // this object's flattened scene (cached after the first call).
    bool hit(const ray& r, interval ray_t, hit_record& rec) const;

   private:
    mutable std::shared_ptr<struct device_scene> cache_;
    friend class camera;
    std::shared_ptr<struct device_scene> device() const;
};

class hittable_list : public hittable {
   public:
    hittable_list() = default;
    explicit hittable_list(std::shared_ptr<hittable> object) { add(std::move(object)); }
    void add(std::shared_ptr<hittable> object) { objects.push_back(std::move(object)); }
    void clear() { objects.clear(); }
    void flatten(scene_builder& sb, const xform& x) const override;
    std::vector<std::shared_ptr<hittable>> objects;
};

class sphere : public hittable {
   public:
    sphere(const point3& center, double radius, std::shared_ptr<material> mat)
        : center(center), radius(std::fmax(0, radius)), mat(std::move(mat)) {}
    void flatten(scene_builder& sb, const xform& x) const override;
    point3 center;
    double radius;
    std::shared_ptr<material> mat;
};

class quad : public hittable {
   public:
    quad(const point3& Q, const vec3& u, const vec3& v, std::shared_ptr<material> mat)
        : Q(Q), u(u), v(v), mat(std::move(mat)) {}
    void flatten(scene_builder& sb, const xform& x) const override;
    point3 Q;
    vec3 u, v;
    std::shared_ptr<material> mat;
    uint32_t kind = AURORA_QUAD_PARALLELOGRAM;
};

// Triangle with vertices Q, Q+u, Q+v (are::Triangle, src/object/triangle.cpp:42-45).
class triangle : public quad {
   public:
    triangle(const point3& Q, const vec3& u, const vec3& v, std::shared_ptr<material> mat)
        : quad(Q, u, v, std::move(mat)) {
        kind = AURORA_QUAD_TRIANGLE;
    }
};

// The six sides of the box with opposite vertices a and b.
std::shared_ptr<hittable_list> box(const point3& a, const point3& b, std::shared_ptr<material> mat);

class translate : public hittable {
   public:
    translate(std::shared_ptr<hittable> object, const vec3& offset) : object(std::move(object)), offset(offset) {}
    void flatten(scene_builder& sb, const xform& x) const override;
    std::shared_ptr<hittable> object;
    vec3 offset;
};

class rotate_y : public hittable {
   public:
    rotate_y(std::shared_ptr<hittable> object, double angle_degrees);
    void flatten(scene_builder& sb, const xform& x) const override;
    std::shared_ptr<hittable> object;
    double sin_theta, cos_theta;
};

// The device always builds an LBVH over the flattened primitives; bvh_node
// marks the grouping and flattens its children.
class bvh_node : public hittable {
   public:
    explicit bvh_node(const hittable_list& list) : objects(list.objects) {}
    void flatten(scene_builder& sb, const xform& x) const override;
    std::vector<std::shared_ptr<hittable>> objects;
};

// ------------------------------------------------------- file formats -----
// Portable FloatMap (PFM): FP32 radiance or AOV images without loss.  Header
// "PF\n<w> <h>\n-1.0\n" (colour, negative scale = little-endian floats),
// then the scanlines from the bottom row up; in memory rows run top-down
// (width*height*3 floats, the layout of aurora_render's sums).  read_pfm
// also takes big-endian files (positive scale).  Errors throw
// std::runtime_error naming the file, as are::Texture's PPM loader does
// (src/texture.cpp:11-43).
void write_pfm(const std::string& path, int width, int height, const std::vector<float>& rgb);
std::vector<float> read_pfm(const std::string& path, int& width, int& height);

// Checkpoint of a progressive render: the camera and render settings, the
// number of samples accumulated (samples [render.sample_offset,
// render.sample_offset + spp) of every pixel) and the per-pixel sums
// (width*height*3 floats each, rows top-down; albedo/normal empty without
// AOVs).  Resuming adds the following samples in sample order, so stop ->
// save -> load -> resume gives the sums of an uninterrupted render bit for
// bit.  File: magic "AURACC01", a 128-byte little-endian header (u32 header
// size, width, height, flags (1: AOV sums), camera vfov, lookfrom, lookat,
// vup, defocus_angle, focus_dist (f32), u32 spp, sample_offset, max_depth,
// seed, f32 background[3], tmin, zero padding, u32 CRC-32 (zlib polynomial)
// of the payload, u32 0), then the f32 sums: rgb, then albedo and normal.  load() checks
// the magic, sizes, trailing bytes and checksum (std::runtime_error).
struct accumulation {
    aurora_camera_desc camera{};
    aurora_render_desc render{};  // max_depth, seed, background, tmin, sample_offset
    uint32_t spp = 0;             // samples accumulated so far
    std::vector<float> rgb_sum, albedo_sum, normal_sum;
    void save(const std::string& path) const;
    static accumulation load(const std::string& path);
};

// ----------------------------------------------------------- camera -----
class camera {
   public:
    double aspect_ratio = 1.0;
    int image_width = 100;
    int samples_per_pixel = 10;
    int max_depth = 10;
    color background;
    double vfov = 90;
    point3 lookfrom = point3(0, 0, 0);
    point3 lookat = point3(0, 0, -1);
    vec3 vup = vec3(0, 1, 0);
    double defocus_angle = 0;
    double focus_dist = 10;
    uint32_t seed = 0;
    // transfer curve of render_ppm (AURORA_TRANSFER_*): by default the
    // writer of experiments/rt.cpp:377-392 (pow(c, 1/2.2f) x 255)
    uint32_t transfer = AURORA_TRANSFER_GAMMA22;

    // Render and write the image to std::cout as ASCII PPM (P3), gamma 2
    // ("Ray Tracing in One Weekend" write_color).
    void render(const hittable& world);
    // Render into `out`: linear RGB means, row 0 = top (height x width x 3).
    void render(const hittable& world, std::vector<float>& out);
    // Render and write a binary P6 PPM (layout of src/texture.cpp:370-391).
    void render_ppm(const hittable& world, const std::string& path);
    // Progressive rendering: add `more` samples to acc.  An empty acc
    // (spp 0, no sums) starts from this camera's settings (aovs: also sum the
    // first-hit albedo and normal); a loaded one must have been made with
    // the same settings (std::invalid_argument otherwise).  Samples
    // [sample_offset + spp, sample_offset + spp + more) are added.
    void accumulate(const hittable& world, accumulation& acc, int more, bool aovs = false);
    int image_height() const;
    aurora_render_stats last_stats{};

   private:
    void render_sum(const hittable& world, std::vector<float>& sum);
};

// Write 8-bit RGB bytes (width*height*3, rows top-down) as a binary P6 PPM,
// the layout of are::Texture::save_texture (src/texture.cpp:372-391):
// header "P6\n<w> <h>\n255\n", then the bytes.  Throws std::runtime_error
// when the file cannot be written.
void write_ppm(const std::string& path, int width, int height, const std::vector<uint8_t>& rgb8);

// Process-wide libaurora context on `device` (created on first use).
aurora_ctx* default_context(int device = 0);

}  // namespace are
