/* oracle.h — C ABI of the CPU oracle (liborc.so).
 *
 * TEST INFRASTRUCTURE ONLY: the CPU restatement of libaurora's path-tracing
 * hot path that the tests check the GPU against.  Loaded by tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg; never linked into
 * or called by libaurora.  Scene, node and camera structs are the ones of
 * include/aurora.h so both sides read identical inputs. */
#ifndef AURORA_ORACLE_H
#define AURORA_ORACLE_H

#include "../include/aurora.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

const char* orc_last_error(void);
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out);
void orc_philox_batch(const uint32_t* ctr, uint32_t n, uint32_t k0, uint32_t k1, uint32_t* out);
void orc_perlin(uint32_t seed, float* ranvec, int32_t* perm);
double orc_sin(double x);
double orc_atan2(double y, double x);
double orc_acos(double x);
void orc_reflect(const float* v, const float* n, float* out);
void orc_refract(const float* v, const float* n, float eta, float* out);
float orc_schlick(float cosine, float ri);
/* the rejection-sampled unit vector of scatter draw (pixel, sample, bounce) */
void orc_random_unit(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t seed, float* out);

int orc_scene_create(const aurora_scene_desc* d, orc_scene** out);
void orc_scene_destroy(orc_scene* s);
int orc_bvh_build(orc_scene* s, aurora_bvh_stats* st);
int orc_bvh_nodes(orc_scene* s, aurora_bvh_node* nodes, uint32_t cap);
int orc_bvh_order(orc_scene* s, uint32_t* morton, uint32_t* order, uint32_t cap);
int orc_bvh_wide(orc_scene* s, aurora_bvh4_node* nodes, uint32_t cap);
void orc_prim_box(orc_scene* s, uint32_t id, float* lo, float* hi);
/* brute: 0 = the four-wide tree, 1 = every primitive, 2 = the binary tree.
 * counts: 3 per ray (node records, spheres, quads fetched; a four-wide node
 * is 4 node records), tree modes only. */
int orc_trace(orc_scene* s, const float* o, const float* d, uint32_t n, float tmin, float tmax, int brute,
              int32_t* prim, float* t, uint32_t* counts, int threads);
/* the segments of one path (pixel, sample): per segment origin (3),
 * direction (3), hit primitive (int32 bits) and hit t, at most cap segments;
 * returns the segment count, or minus a status code */
int orc_path_segments(orc_scene* s, const aurora_camera_desc* c, const aurora_render_desc* rd, uint32_t pixel,
                      uint32_t sample, float* out, uint32_t cap);
/* the tree later renders walk: 0 the binary LBVH (default), 1 the four-wide
 * tree; process-wide */
void orc_set_render_tree(int wide);
/* the order in which a four-wide step descends into its hit internal slots:
 * t[4] entry distances, hit[4] flags; ord[0..m) receives the hit slots
 * nearest first, the rest the others; returns m */
int orc_wide_order(const float* t, const int32_t* hit, int32_t* ord);
int orc_hit_sphere(const float* sphere, const float* o, const float* d, float tmin, float tmax, float* t);
/* tab receives t, alpha, beta */
int orc_hit_quad(const float* quad, uint32_t kind, const float* o, const float* d, float tmin, float tmax,
                 float* tab);
void orc_texture_value(orc_scene* s, uint32_t tex, float u, float v, const float* p, float* out);
/* (u, v) of the outward unit normal n on a sphere, as shading computes it */
void orc_sphere_uv(const float* n, float* uv);
/* out18: center, pixel00, delta_u, delta_v, disk_u, disk_v */
int orc_camera(const aurora_camera_desc* c, float* out18);
/* first ray of sample samples[k] through pixel pixels[k] (row-major), as
 * aurora_camera_rays: 3*n origins and unit directions */
int orc_camera_rays(const aurora_camera_desc* c, uint32_t seed, const uint32_t* pixels, const uint32_t* samples,
                    uint32_t n, float* org, float* dir);
/* are::Texture::paste restated (src/texture.cpp:85-360): src (sw x sh RGB
 * floats) warped into dst (dw x dh) by the homography of the source corners
 * onto corners[0..7] = LT, RT, LB, RB (x, y); returns the pixels painted */
int orc_paste(const float* src, int sw, int sh, float* dst, int dw, int dh, const int32_t* corners);
/* 8-bit resolve of n floats (sum over spp samples), AURORA_TRANSFER_* */
int orc_resolve_rgb8(const float* sum, uint64_t n, uint32_t spp, uint32_t transfer, uint8_t* out);
/* pixels NULL = every pixel in row-major order; rgb/albedo/normal are 3 per
 * pixel and are added to (sum over samples), as libaurora does */
int orc_render_pixels(orc_scene* s, const aurora_camera_desc* c, const aurora_render_desc* rd,
                      const uint32_t* pixels, uint32_t n, int threads, float* rgb, float* albedo, float* normal,
                      uint64_t* segments);
/* experiments/rt10.cpp's patch-projection renderer restated in double:
 * rgb receives height x width x 3 (row 0 at the top); textures (optional)
 * the number of triangle textures of the recursion tree */
int orc_patch_render(const aurora_patch_triangle* tris, uint32_t n_tris, const aurora_patch_material* mats,
                     uint32_t n_mats, const aurora_patch_camera* cam, const aurora_patch_config* cfg, double* rgb,
                     uint64_t* textures);
const char* orc_patch_last_error(void);
int orc_render_pixels_f64(orc_scene* s, const aurora_camera_desc* c, const aurora_render_desc* rd,
                          const uint32_t* pixels, uint32_t n, int threads, double* rgb, double* albedo,
                          double* normal, uint64_t* segments);

#ifdef __cplusplus
}
#endif
#endif
