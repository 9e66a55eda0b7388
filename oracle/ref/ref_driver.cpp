// ref_driver.cpp — C entry points over routines of the reference renderer,
// compiled from the reference's own sources (see Makefile) so the oracle's
// restatements can be checked against them.  TEST INFRASTRUCTURE ONLY.
//
// Wrapped routines:
//   are::reflect / are::refract       src/basic/vec3.cpp:182-199
//   are::Plane::intersect_ray         src/basic/plane.cpp:13-27
//   are::Texture fill ctor / pixel / paste (bilinear)  src/texture.cpp:52-83, 85-360
//   are::Texture(path) P6 loader and save_texture       src/texture.cpp:9-50, 362-395
// are::Triangle is not wrapped: its vtable needs Triangle::trace_texture,
// which the reference declares (include/object/triangle.h:31) but never
// defines, so it cannot be linked without a stand-in.
#include <basic/plane.h>
#include <basic/ray.h>
#include <basic/vec3.h>
#include <texture.h>

#include <cmath>
#include <cstring>
#include <stdexcept>

extern "C" {

void ref_reflect(const double* v, const double* n, double* out) {
    are::Vec3 r = are::reflect(are::Vec3(v[0], v[1], v[2]), are::Vec3(n[0], n[1], n[2]));
    out[0] = r.x();
    out[1] = r.y();
    out[2] = r.z();
}

void ref_refract(const double* v, const double* n, double eta, double* out) {
    are::Vec3 r = are::refract(are::Vec3(v[0], v[1], v[2]), are::Vec3(n[0], n[1], n[2]), eta);
    out[0] = r.x();
    out[1] = r.y();
    out[2] = r.z();
}

// plane through p with normal n; ray (o, d) (the reference normalises d)
int ref_plane_hit(const double* p, const double* n, const double* o, const double* d, double* hit) {
    are::Plane pl(are::Point3(p[0], p[1], p[2]), are::Vec3(n[0], n[1], n[2]));
    are::Ray r(are::Point3(o[0], o[1], o[2]), are::Vec3(d[0], d[1], d[2]));
    are::Point3 h;
    if (!pl.intersect_ray(r, h)) return 0;
    hit[0] = h.x();
    hit[1] = h.y();
    hit[2] = h.z();
    return 1;
}

// paste a sw x sh RGB image (row-major, rows top-down) into a dw x dh image
// filled with `fill`, with the source corners (0,0), (sw-1,0), (0,sh-1),
// (sw-1,sh-1) sent to corners[0..7] = LT, RT, LB, RB (x, y pairs)
void ref_paste(int sw, int sh, const double* src, int dw, int dh, const int* corners, double fill, double* dst) {
    are::Texture s(sw, sh, are::Color3(0, 0, 0));
    for (int y = 0; y < sh; ++y)
        for (int x = 0; x < sw; ++x) {
            const double* c = src + 3 * ((size_t)y * sw + x);
            s.pixel(x, y) = are::Color3(c[0], c[1], c[2]);
        }
    are::Texture d(dw, dh, are::Color3(fill, fill, fill));
    d.paste(s, {corners[0], corners[1]}, {corners[2], corners[3]}, {corners[4], corners[5]}, {corners[6], corners[7]});
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            const are::Color3& c = d.pixel(x, y);
            double* o = dst + 3 * ((size_t)y * dw + x);
            o[0] = c.x();
            o[1] = c.y();
            o[2] = c.z();
        }
}

// load a P6 file with are::Texture(path): returns 0 and w, h and up to cap
// texels (3 doubles each, rows top-down), or 1 and the exception's message
int ref_load_ppm(const char* path, int* w, int* h, double* out, long cap, char* err, int errcap) {
    try {
        are::Texture t{std::string(path)};
        *w = t.width_;
        *h = t.height_;
        long k = 0;
        for (int y = 0; y < t.height_; ++y)
            for (int x = 0; x < t.width_; ++x, ++k) {
                if (k >= cap) return 0;
                const are::Color3& c = t.pixel(x, y);
                out[3 * k] = c.x();
                out[3 * k + 1] = c.y();
                out[3 * k + 2] = c.z();
            }
        return 0;
    } catch (const std::exception& e) {
        std::strncpy(err, e.what(), (size_t)errcap - 1);
        err[errcap - 1] = 0;
        return 1;
    }
}

// write a w x h image (3 doubles per texel, rows top-down) with
// are::Texture::save_texture; returns its result (1 written, 0 refused)
int ref_save_ppm(int w, int h, const double* rgb, const char* path) {
    are::Texture t(w, h, are::Color3(0, 0, 0));
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const double* c = rgb + 3 * ((size_t)y * w + x);
            t.pixel(x, y) = are::Color3(c[0], c[1], c[2]);
        }
    return t.save_texture(path) ? 1 : 0;
}

}  // extern "C"
