// aurora_api.cpp — implementation of the are:: host API (are/aurora.hpp):
// flattening of hittable trees into aurora_scene_desc and the calls into
// libaurora's C ABI.  No host-side ray tracing: render and hit both run the
// GPU kernels.
#include "are/aurora.hpp"

#include <cstdio>
#include <cstring>
#include <iostream>
#include <mutex>
#include <stdexcept>

namespace are {

namespace {

[[noreturn]] void raise(int code, const char* what) {
    if (code == AURORA_ERR_INVALID) throw std::invalid_argument(what);
    throw std::runtime_error(what);
}

void check(aurora_ctx* ctx, int code) {
    if (code != AURORA_OK) raise(code, aurora_last_error(ctx));
}

void put3(float* dst, const vec3& v) {
    dst[0] = (float)v[0];
    dst[1] = (float)v[1];
    dst[2] = (float)v[2];
}

}  // namespace

aurora_ctx* default_context(int device) {
    static std::mutex mu;
    static std::map<int, aurora_ctx*> ctxs;
    std::lock_guard<std::mutex> lock(mu);
    auto it = ctxs.find(device);
    if (it != ctxs.end()) return it->second;
    aurora_ctx* c = nullptr;
    int rc = aurora_create(device, &c);
    if (rc != AURORA_OK) raise(rc, aurora_last_error(nullptr));
    ctxs[device] = c;
    return c;
}

// ------------------------------------------------------------ xform -----
xform xform::then(const xform& o) const {
    xform r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = o.m[i][0] * m[0][j] + o.m[i][1] * m[1][j] + o.m[i][2] * m[2][j];
    r.t = o.point(t);
    return r;
}

// ---------------------------------------------------- scene_builder -----
uint32_t scene_builder::texture_id(const texture* t) {
    if (!t) throw std::invalid_argument("texture pointer cannot be null");
    auto it = tex_ids_.find(t);
    if (it != tex_ids_.end()) return it->second;
    aurora_texture d = t->describe(*this);
    uint32_t id = (uint32_t)texs.size();
    texs.push_back(d);
    tex_ids_[t] = id;
    return id;
}

uint32_t scene_builder::material_id(const std::shared_ptr<material>& m) {
    if (!m) throw std::invalid_argument("material pointer cannot be null");
    auto it = mat_ids_.find(m.get());
    if (it != mat_ids_.end()) return it->second;
    aurora_material d = m->describe(*this);
    uint32_t id = (uint32_t)mats.size();
    mats.push_back(d);
    mat_ids_[m.get()] = id;
    keep_.push_back(m);
    return id;
}

void scene_builder::add_sphere(const point3& c, double r, uint32_t mat) {
    spheres.insert(spheres.end(), {(float)c[0], (float)c[1], (float)c[2], (float)r});
    sphere_mat.push_back(mat);
}

void scene_builder::add_quad(const point3& Q, const vec3& u, const vec3& v, uint32_t mat, uint32_t kind) {
    float q[9];
    put3(q, Q);
    put3(q + 3, u);
    put3(q + 6, v);
    quads.insert(quads.end(), q, q + 9);
    quad_mat.push_back(mat);
    quad_kind.push_back(kind);
}

uint32_t scene_builder::add_texels(const std::vector<float>& rgb) {
    uint32_t first = (uint32_t)(texels.size() / 3);
    texels.insert(texels.end(), rgb.begin(), rgb.end());
    return first;
}

aurora_scene_desc scene_builder::desc() const {
    aurora_scene_desc d{};
    d.n_spheres = (uint32_t)sphere_mat.size();
    d.spheres = spheres.data();
    d.sphere_material = sphere_mat.data();
    d.n_quads = (uint32_t)quad_mat.size();
    d.quads = quads.data();
    d.quad_material = quad_mat.data();
    d.quad_kind = quad_kind.data();
    d.n_materials = (uint32_t)mats.size();
    d.materials = mats.data();
    d.n_textures = (uint32_t)texs.size();
    d.textures = texs.data();
    d.n_texels = texels.size() / 3;
    d.texels = texels.data();
    return d;
}

// --------------------------------------------------------- textures -----
aurora_texture solid_color::describe(scene_builder&) const {
    aurora_texture t{};
    t.type = AURORA_TEX_SOLID;
    put3(t.color, albedo);
    return t;
}

aurora_texture checker_texture::describe(scene_builder& sb) const {
    aurora_texture t{};
    t.type = AURORA_TEX_CHECKER;
    t.scale = (float)scale;
    t.even = sb.texture_id(even.get());
    t.odd = sb.texture_id(odd.get());
    return t;
}

aurora_texture noise_texture::describe(scene_builder&) const {
    aurora_texture t{};
    t.type = AURORA_TEX_NOISE;
    t.scale = (float)scale;
    t.seed = seed;
    return t;
}

image_texture::image_texture(int w, int h, std::vector<float> rgb_linear)
    : width(w), height(h), rgb(std::move(rgb_linear)) {
    if (w <= 0 || h <= 0) throw std::runtime_error("Texture width and height must be positive.");
    if (rgb.size() != (size_t)w * h * 3) throw std::invalid_argument("image size does not match width*height*3");
}

image_texture::image_texture(const std::string& path) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) throw std::runtime_error("Failed to open texture file: " + path);
    char fmt[3] = {0};
    int w = 0, h = 0, maxv = 0;
    if (fscanf(fp, "%2s", fmt) != 1 || std::string(fmt) != "P6") {
        fclose(fp);
        throw std::runtime_error("Unsupported PPM format in file: " + path);
    }
    if (fscanf(fp, "%d %d %d", &w, &h, &maxv) != 3 || maxv != 255 || w <= 0 || h <= 0) {
        fclose(fp);
        throw std::runtime_error("Invalid PPM header in file: " + path);
    }
    fgetc(fp);
    std::vector<unsigned char> bytes((size_t)w * h * 3);
    if (fread(bytes.data(), 1, bytes.size(), fp) != bytes.size()) {
        fclose(fp);
        throw std::runtime_error("Unexpected end of file while reading pixel data in file: " + path);
    }
    fclose(fp);
    width = w;
    height = h;
    rgb.resize(bytes.size());
    for (size_t i = 0; i < bytes.size(); ++i) rgb[i] = (float)((double)bytes[i] / 255.0);
}

aurora_texture image_texture::describe(scene_builder& sb) const {
    aurora_texture t{};
    t.type = AURORA_TEX_IMAGE;
    t.width = (uint32_t)width;
    t.height = (uint32_t)height;
    t.texel_offset = sb.add_texels(rgb);
    return t;
}

// -------------------------------------------------------- materials -----
aurora_material lambertian::describe(scene_builder& sb) const {
    return aurora_material{AURORA_MAT_LAMBERTIAN, sb.texture_id(tex.get()), 0.0f, 1.0f};
}
aurora_material metal::describe(scene_builder& sb) const {
    return aurora_material{AURORA_MAT_METAL, sb.texture_id(tex.get()), (float)fuzz, 1.0f};
}
aurora_material dielectric::describe(scene_builder&) const {
    return aurora_material{AURORA_MAT_DIELECTRIC, 0u, 0.0f, (float)refraction_index};
}
aurora_material diffuse_light::describe(scene_builder& sb) const {
    return aurora_material{AURORA_MAT_DIFFUSE_LIGHT, sb.texture_id(tex.get()), 0.0f, 1.0f};
}

// -------------------------------------------------------- hittables -----
void hittable_list::flatten(scene_builder& sb, const xform& x) const {
    for (const auto& o : objects) o->flatten(sb, x);
}
void bvh_node::flatten(scene_builder& sb, const xform& x) const {
    for (const auto& o : objects) o->flatten(sb, x);
}
void sphere::flatten(scene_builder& sb, const xform& x) const { sb.add_sphere(x.point(center), radius, sb.material_id(mat)); }
void quad::flatten(scene_builder& sb, const xform& x) const {
    sb.add_quad(x.point(Q), x.vector(u), x.vector(v), sb.material_id(mat), kind);
}

void translate::flatten(scene_builder& sb, const xform& x) const {
    xform local;
    local.t = offset;
    object->flatten(sb, local.then(x));
}

rotate_y::rotate_y(std::shared_ptr<hittable> object, double angle_degrees) : object(std::move(object)) {
    const double r = angle_degrees * (3.1415926535897932385 / 180.0);
    sin_theta = std::sin(r);
    cos_theta = std::cos(r);
}

void rotate_y::flatten(scene_builder& sb, const xform& x) const {
    // object space -> world space: x' = cos x + sin z, z' = -sin x + cos z
    xform local;
    local.m[0][0] = cos_theta;
    local.m[0][2] = sin_theta;
    local.m[2][0] = -sin_theta;
    local.m[2][2] = cos_theta;
    object->flatten(sb, local.then(x));
}

std::shared_ptr<hittable_list> box(const point3& a, const point3& b, std::shared_ptr<material> mat) {
    auto sides = std::make_shared<hittable_list>();
    const point3 lo(std::fmin(a.x(), b.x()), std::fmin(a.y(), b.y()), std::fmin(a.z(), b.z()));
    const point3 hi(std::fmax(a.x(), b.x()), std::fmax(a.y(), b.y()), std::fmax(a.z(), b.z()));
    const vec3 dx(hi.x() - lo.x(), 0, 0), dy(0, hi.y() - lo.y(), 0), dz(0, 0, hi.z() - lo.z());
    sides->add(std::make_shared<quad>(point3(lo.x(), lo.y(), hi.z()), dx, dy, mat));
    sides->add(std::make_shared<quad>(point3(hi.x(), lo.y(), hi.z()), -dz, dy, mat));
    sides->add(std::make_shared<quad>(point3(hi.x(), lo.y(), lo.z()), -dx, dy, mat));
    sides->add(std::make_shared<quad>(point3(lo.x(), lo.y(), lo.z()), dz, dy, mat));
    sides->add(std::make_shared<quad>(point3(lo.x(), hi.y(), hi.z()), dx, -dz, mat));
    sides->add(std::make_shared<quad>(point3(lo.x(), lo.y(), lo.z()), dx, dz, mat));
    return sides;
}

// A flattened, uploaded scene with its LBVH.
struct device_scene {
    aurora_ctx* ctx = nullptr;
    aurora_scene* scene = nullptr;
    scene_builder sb;
    ~device_scene() {
        if (scene) aurora_scene_destroy(scene);
    }
};

std::shared_ptr<device_scene> hittable::device() const {
    if (cache_) return cache_;
    auto d = std::make_shared<device_scene>();
    d->ctx = default_context();
    flatten(d->sb, xform());
    aurora_scene_desc desc = d->sb.desc();
    check(d->ctx, aurora_scene_upload(d->ctx, &desc, &d->scene));
    check(d->ctx, aurora_bvh_build(d->scene, nullptr));
    cache_ = d;
    return d;
}

bool hittable::hit(const ray& r, interval ray_t, hit_record& rec) const {
    auto d = device();
    float o[3], dir[3];
    put3(o, r.origin());
    put3(dir, r.direction());
    int32_t prim = AURORA_MISS;
    float t = 0;
    check(d->ctx, aurora_trace_closest(d->scene, o, dir, 1, (float)ray_t.min, (float)ray_t.max, &prim, &t));
    if (prim < 0) return false;
    rec.prim = prim;
    rec.t = t;
    rec.p = r.at(t);
    const uint32_t ns = (uint32_t)d->sb.sphere_mat.size();
    vec3 outward;
    if ((uint32_t)prim < ns) {
        const float* s = &d->sb.spheres[4 * (size_t)prim];
        outward = (rec.p - point3(s[0], s[1], s[2])) / (double)s[3];
    } else {
        const float* q = &d->sb.quads[9 * (size_t)(prim - ns)];
        outward = unit_vector(cross(vec3(q[3], q[4], q[5]), vec3(q[6], q[7], q[8])));
    }
    rec.front_face = dot(r.direction(), outward) < 0;
    rec.normal = rec.front_face ? outward : -outward;
    return true;
}

// ----------------------------------------------------------- camera -----
int camera::image_height() const {
    int h = (int)(image_width / aspect_ratio);
    return h < 1 ? 1 : h;
}

void camera::accumulate(const hittable& world, accumulation& acc, int more, bool aovs) {
    if (more < 0) throw std::invalid_argument("sample count must be non-negative");
    const int h = image_height();
    aurora_camera_desc cd{};
    cd.image_width = (uint32_t)image_width;
    cd.image_height = (uint32_t)h;
    cd.vfov = (float)vfov;
    put3(cd.lookfrom, lookfrom);
    put3(cd.lookat, lookat);
    put3(cd.vup, vup);
    cd.defocus_angle = (float)defocus_angle;
    cd.focus_dist = (float)focus_dist;
    aurora_render_desc rd{};
    rd.max_depth = (uint32_t)max_depth;
    rd.seed = seed;
    put3(rd.background, background);
    const size_t n = (size_t)image_width * h * 3;
    if (acc.spp == 0 && acc.rgb_sum.empty()) {
        acc.camera = cd;
        acc.render = rd;
        acc.rgb_sum.assign(n, 0.0f);
        if (aovs) {
            acc.albedo_sum.assign(n, 0.0f);
            acc.normal_sum.assign(n, 0.0f);
        }
    } else {
        if (std::memcmp(&acc.camera, &cd, sizeof cd) != 0)
            throw std::invalid_argument("accumulation: made with another camera");
        const aurora_render_desc& r = acc.render;
        if (r.max_depth != rd.max_depth || r.seed != rd.seed ||
            std::memcmp(r.background, rd.background, sizeof rd.background) != 0)
            throw std::invalid_argument("accumulation: made with other render settings");
        if (acc.rgb_sum.size() != n) throw std::invalid_argument("accumulation: sums do not match the image");
    }
    const bool aov = !acc.albedo_sum.empty();
    if (aov && (acc.albedo_sum.size() != n || acc.normal_sum.size() != n))
        throw std::invalid_argument("accumulation: AOV sums do not match the image");
    if ((uint64_t)acc.render.sample_offset + acc.spp + (uint64_t)more > 0xffffffffull)
        throw std::invalid_argument("accumulation: sample index overflow");
    if (more == 0) return;
    auto d = world.device();
    aurora_render_desc go = acc.render;
    go.spp = (uint32_t)more;
    go.sample_offset = acc.render.sample_offset + acc.spp;
    go.flags = aov ? AURORA_RENDER_AOV : 0u;
    go.spp_per_pass = 0;
    check(d->ctx, aurora_render(d->scene, &acc.camera, &go, acc.rgb_sum.data(), aov ? acc.albedo_sum.data() : nullptr,
                                aov ? acc.normal_sum.data() : nullptr, &last_stats));
    acc.spp += (uint32_t)more;
}

void camera::render_sum(const hittable& world, std::vector<float>& sum) {
    auto d = world.device();
    const int h = image_height();
    aurora_camera_desc cd{};
    cd.image_width = (uint32_t)image_width;
    cd.image_height = (uint32_t)h;
    cd.vfov = (float)vfov;
    put3(cd.lookfrom, lookfrom);
    put3(cd.lookat, lookat);
    put3(cd.vup, vup);
    cd.defocus_angle = (float)defocus_angle;
    cd.focus_dist = (float)focus_dist;
    aurora_render_desc rd{};
    rd.spp = (uint32_t)samples_per_pixel;
    rd.max_depth = (uint32_t)max_depth;
    rd.seed = seed;
    put3(rd.background, background);
    sum.assign((size_t)image_width * h * 3, 0.0f);
    check(d->ctx, aurora_render(d->scene, &cd, &rd, sum.data(), nullptr, nullptr, &last_stats));
}

void camera::render(const hittable& world, std::vector<float>& out) {
    render_sum(world, out);
    const float s = 1.0f / (float)samples_per_pixel;
    for (float& v : out) v *= s;
}

void camera::render(const hittable& world) {
    std::vector<float> sum;
    render_sum(world, sum);
    const int h = image_height();
    std::vector<uint8_t> px(sum.size());
    check(nullptr, aurora_resolve_rgb8(sum.data(), (uint32_t)image_width, (uint32_t)h, (uint32_t)samples_per_pixel,
                                       AURORA_TRANSFER_GAMMA2, px.data()));
    std::cout << "P3\n" << image_width << ' ' << h << "\n255\n";
    for (size_t i = 0; i < px.size(); i += 3) std::cout << (int)px[i] << ' ' << (int)px[i + 1] << ' ' << (int)px[i + 2] << '\n';
}

void camera::render_ppm(const hittable& world, const std::string& path) {
    std::vector<float> sum;
    render_sum(world, sum);
    const int h = image_height();
    std::vector<uint8_t> px(sum.size());
    check(nullptr, aurora_resolve_rgb8(sum.data(), (uint32_t)image_width, (uint32_t)h, (uint32_t)samples_per_pixel,
                                       transfer, px.data()));
    write_ppm(path, image_width, h, px);
}

void write_ppm(const std::string& path, int width, int height, const std::vector<uint8_t>& rgb8) {
    if (width <= 0 || height <= 0 || rgb8.size() != (size_t)width * height * 3)
        throw std::invalid_argument("image size does not match width*height*3");
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) throw std::runtime_error("Failed to open file for writing: " + path);
    const bool ok = fprintf(fp, "P6\n%d %d\n255\n", width, height) > 0 &&
                    fwrite(rgb8.data(), 1, rgb8.size(), fp) == rgb8.size();
    if (fclose(fp) != 0 || !ok) throw std::runtime_error("Failed to write file: " + path);
}

}  // namespace are
