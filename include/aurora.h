/*
 * aurora.h — C ABI of libaurora, the MI355X (gfx950) path-tracing hot path.
 *
 * Every entry point takes plain pointers and sizes, returns a status code and
 * never throws.  The host C++ API (namespace are: hittable / material /
 * texture / camera) flattens a scene into the descriptors below and calls
 * these functions; ctypes, cgo or JNI bindings can do the same.
 *
 * The reference renderer (ARE) has no path tracer, BVH, sphere, quad,
 * scatter or camera; each entry point names the reference code whose role
 * it takes over:
 *   - scene objects      : are::Object / are::Triangle / are::ObjectSet
 *                          (include/object/object.h:14-39,
 *                           include/object/triangle.h:19,
 *                           include/object/object_set.h:10-12)
 *   - ray query          : are::Triangle::intersect_ray
 *                          (src/object/triangle.cpp:82-121) and the linear
 *                          closest-hit scan experiments/rt.cpp:209-218
 *   - per-pixel render   : experiments/rt.cpp:337-374 (render) and
 *                          experiments/rt.cpp:251-334 (trace)
 *   - image output       : are::Texture::save_texture (src/texture.cpp:362-395)
 *
 * Conventions
 *   - status 0 = AURORA_OK.  On failure, aurora_last_error(ctx) holds a
 *     message; the host C++ wrapper rethrows it as std::invalid_argument
 *     (AURORA_ERR_INVALID) or std::runtime_error (everything else), the
 *     exception types the reference's constructors throw
 *     (src/object/triangle.cpp:13-36, src/texture.cpp:11-43).
 *   - the caller owns host buffers; the library owns device buffers behind
 *     the opaque handles.  Entry points with a _device suffix take device
 *     pointers that the caller allocated on the context's device.
 *   - one HIP stream per context; a context is not thread-safe, separate
 *     contexts are independent.
 */
#ifndef AURORA_H
#define AURORA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AURORA_ABI_VERSION 6

/* ---- status codes ---------------------------------------------------- */
#define AURORA_OK 0
#define AURORA_ERR_INVALID 1   /* bad argument or scene descriptor        */
#define AURORA_ERR_RUNTIME 2   /* HIP runtime / kernel launch failure      */
#define AURORA_ERR_NOMEM 3     /* device allocation failed                 */
#define AURORA_ERR_STATE 4     /* call out of order (e.g. no BVH built)    */

/* ---- primitive kinds -------------------------------------------------- */
/* A quad record (Q, u, v) spans the parallelogram Q + a*u + b*v with
 * 0<=a,b<=1; the triangle kind keeps a,b>=0, a+b<=1 (vertices Q, Q+u, Q+v as
 * in are::Triangle, src/object/triangle.cpp:42-45). */
#define AURORA_QUAD_PARALLELOGRAM 0u
#define AURORA_QUAD_TRIANGLE 1u

/* ---- materials -------------------------------------------------------- */
#define AURORA_MAT_LAMBERTIAN 0u     /* albedo = texture                    */
#define AURORA_MAT_METAL 1u          /* albedo = texture, fuzz              */
#define AURORA_MAT_DIELECTRIC 2u     /* ior                                 */
#define AURORA_MAT_DIFFUSE_LIGHT 3u  /* emit = texture, no scatter          */

typedef struct aurora_material {
    uint32_t type;     /* AURORA_MAT_*                                      */
    uint32_t texture;  /* index into the texture table (unused: dielectric) */
    float fuzz;        /* metal: clamped to [0,1] at upload                  */
    float ior;         /* dielectric: refraction index                       */
} aurora_material;

/* ---- textures --------------------------------------------------------- */
#define AURORA_TEX_SOLID 0u    /* color                                     */
#define AURORA_TEX_CHECKER 1u  /* spatial checker: even/odd texture, scale  */
#define AURORA_TEX_NOISE 2u    /* Perlin marble: scale, seed                */
#define AURORA_TEX_IMAGE 3u    /* bilinear RGB image: width, height, offset */

typedef struct aurora_texture {
    uint32_t type;       /* AURORA_TEX_*                                    */
    uint32_t even, odd;  /* checker: texture indices (must not be checkers) */
    uint32_t seed;       /* noise: Philox seed of the Perlin tables         */
    float color[3];      /* solid                                           */
    float scale;         /* checker: cell size; noise: frequency            */
    uint32_t width, height;   /* image                                      */
    uint64_t texel_offset;    /* image: first texel (RGB float triple) in
                                 aurora_scene_desc.texels                   */
} aurora_texture;

/* ---- scene ------------------------------------------------------------ */
typedef struct aurora_scene_desc {
    uint32_t n_spheres;
    const float* spheres;           /* n_spheres x {cx, cy, cz, radius}     */
    const uint32_t* sphere_material;/* n_spheres                            */
    uint32_t n_quads;
    const float* quads;             /* n_quads x {Qx,Qy,Qz, ux,uy,uz, vx,vy,vz} */
    const uint32_t* quad_material;  /* n_quads                              */
    const uint32_t* quad_kind;      /* n_quads, AURORA_QUAD_* (NULL = all
                                       parallelograms)                      */
    uint32_t n_materials;
    const aurora_material* materials;
    uint32_t n_textures;
    const aurora_texture* textures;
    uint64_t n_texels;              /* RGB triples                          */
    const float* texels;            /* 3*n_texels floats, linear RGB        */
} aurora_scene_desc;

/* Primitive ids used by every query: spheres are [0, n_spheres), quads are
 * [n_spheres, n_spheres + n_quads).  AURORA_MISS marks no hit. */
#define AURORA_MISS (-1)

/* ---- LBVH ------------------------------------------------------------- */
/* 32-byte node.  Node 0 is the root and node 1 is unused padding; every
 * internal node's two children sit side by side at nodes[child] and
 * nodes[child+1] with `child` even, so one 64-byte-aligned fetch reads both
 * child boxes.  The tree of n primitives has 2n nodes (n = 1: one leaf).
 * A leaf has AURORA_NODE_LEAF set in `child` and holds the primitive id in
 * the low 31 bits. */
#define AURORA_NODE_LEAF 0x80000000u
typedef struct aurora_bvh_node {
    float lo[3];
    uint32_t child;
    float hi[3];
    uint32_t parent;  /* index of the parent node (root: 0xffffffff)     */
} aurora_bvh_node;

/* Four-wide node of the traversal tree (128 bytes, one cache line): the
 * binary LBVH collapsed so that each node holds a cut of up to four nodes
 * below a binary node.  Scenes of at least 4096 primitives: greedy (the
 * internal slot of largest surface area opened first); smaller scenes:
 * pairwise (each internal child opened once, so the wide nodes are the
 * even-depth binary nodes).  ref is AURORA_NODE_LEAF | primitive, the index
 * of another four-wide node, or AURORA_BVH4_EMPTY.  Node 0 is the root. */
#define AURORA_BVH4_EMPTY 0xffffffffu
typedef struct aurora_bvh4_slot {
    float lo[3];
    uint32_t ref;
    float hi[3];
    uint32_t pad;           /* 0                                            */
} aurora_bvh4_slot;
typedef struct aurora_bvh4_node {
    aurora_bvh4_slot slot[4];
} aurora_bvh4_node;

#define AURORA_BVH4_PAIRWISE 0u
#define AURORA_BVH4_GREEDY 1u
typedef struct aurora_bvh_stats {
    uint32_t n_prims;
    uint32_t n_nodes;       /* 2*n_prims (1 for a single primitive)         */
    uint32_t depth;         /* longest root-to-leaf path, in edges          */
    uint32_t n_wide;        /* four-wide traversal nodes                    */
    uint32_t wide_depth;    /* four-wide nodes on the longest root path     */
    uint32_t wide_form;     /* AURORA_BVH4_GREEDY from 4096 primitives while
                               its walk fits the 96-entry traversal stack
                               (three pushes per level above the deepest),
                               else AURORA_BVH4_PAIRWISE                    */
    float scene_lo[3], scene_hi[3];     /* union of primitive boxes         */
    float centroid_lo[3], centroid_hi[3];/* Morton quantisation frame        */
    double build_ms;        /* device time of the build kernels            */
} aurora_bvh_stats;

/* ---- camera and render ------------------------------------------------ */
typedef struct aurora_camera_desc {
    uint32_t image_width, image_height;
    float vfov;            /* vertical field of view, degrees              */
    float lookfrom[3], lookat[3], vup[3];
    float defocus_angle;   /* degrees; <= 0 disables depth of field        */
    float focus_dist;
} aurora_camera_desc;

#define AURORA_RENDER_AOV 1u   /* also write first-hit albedo and normal */
#define AURORA_RENDER_COUNT_FETCHES 2u  /* count BVH node and primitive
                                           records the traversal fetches
                                           (slower; for roofline accounting;
                                           every segment then runs in the
                                           wavefront kernels, none in the
                                           fused tail) */
#define AURORA_RENDER_GRAPH 4u  /* capture one pass in a hipGraph and replay
                                   it for every pass of the same shape
                                   (fixed launch shapes; same results) */

typedef struct aurora_render_desc {
    uint32_t spp;            /* samples this call adds                       */
    uint32_t sample_offset;  /* first sample index (progressive rendering)   */
    uint32_t max_depth;      /* ray segments per path (>= 1)                 */
    uint32_t seed;           /* Philox key word 0                            */
    float background[3];     /* radiance of rays that leave the scene        */
    uint32_t flags;          /* AURORA_RENDER_*                              */
    uint32_t spp_per_pass;   /* samples per wavefront pass; 0 = automatic:
                              * the fewest passes of at most 2^27 paths   */
    float tmin;              /* ray offset; <= 0 selects 0.001               */
} aurora_render_desc;

typedef struct aurora_render_stats {
    uint64_t paths;          /* camera samples traced                        */
    uint64_t segments;       /* ray segments traversed (extension rays)      */
    double kernel_ms;        /* device time from first to last render kernel */
    double extend_ms;        /* summed device time of the traversal launches */
    uint32_t passes;
    uint32_t max_live;       /* largest wavefront queue                      */
    uint32_t extend_launches;
    uint32_t tail_launches;  /* fused-tail launches (eager passes)           */
    uint64_t nodes_fetched;  /* AURORA_RENDER_COUNT_FETCHES: 32-byte nodes   */
    uint64_t spheres_fetched;/*   16-byte sphere records                     */
    uint64_t quads_fetched;  /*   48-byte quad records                       */
    double tail_ms;          /* summed device time of the fused-tail launches */
} aurora_render_stats;

/* ---- context ---------------------------------------------------------- */
typedef struct aurora_ctx aurora_ctx;
typedef struct aurora_scene aurora_scene;

int aurora_abi_version(void);

/* Create a context on HIP device `device` (takes over from nothing in the
 * reference: ARE is single-threaded CPU code without a device). */
int aurora_create(int device, aurora_ctx** out);
void aurora_destroy(aurora_ctx* ctx);
/* Message of the last failure on ctx (ctx NULL: last failure of
 * aurora_create on this thread).  Never NULL. */
const char* aurora_last_error(const aurora_ctx* ctx);
/* Synchronise the context's stream. */
int aurora_synchronize(aurora_ctx* ctx);
/* The context's HIP stream, as a void* (hipStream_t). */
void* aurora_stream(aurora_ctx* ctx);
/* Launch shape: persistent grid sizes of the traversal and shade kernels in
 * blocks per compute unit (default 8 each; both in [1, 64]), and the live
 * path count at or below which the remaining bounces of a pass run in the
 * persistent fused-tail kernel, each lane carrying a path to termination
 * and then taking the next (default 0xffffffff: the whole pass; 0 = never,
 * the wavefront extend/shade kernels carry every bounce).  Changes speed
 * only, never results. */
int aurora_set_launch(aurora_ctx* ctx, int extend_blocks_per_cu,
                      int shade_blocks_per_cu, uint32_t tail_paths);
/* Queue sizes: the most paths an automatic pass holds (spp_per_pass 0;
 * default 0 = 2^27, else in [2^16, 2^28]; the path queues take 152 bytes per
 * path), and the paths a wave of the fused tail claims per work-pool atomic
 * (default 0 = by pass size, 64 to 128; else in [16, 1024]).  Changes speed
 * only, never results: passes add their samples in order, and claims only
 * decide which lane carries which path. */
int aurora_set_queue(aurora_ctx* ctx, uint32_t pass_paths, uint32_t claim_paths);

/* Tree walked by the render kernels: 0 = the binary LBVH (default; one
 * 64-byte child pair per step), 1 = the four-wide tree (one 128-byte node
 * per step), 2 = flat: for single-scene renders of scenes of at most 256
 * primitives, the persistent path kernel tests every leaf in Morton order,
 * the same sequence for every ray (no stack, no divergence); elsewhere 2
 * walks the four-wide tree.  Which is fastest depends on the scene (the
 * autotuner tries them); the images are the same.  No reference
 * counterpart. */
int aurora_set_render_tree(aurora_ctx* ctx, int tree);

/* Order of the rays of aurora_trace_closest*: 0 = as given, 1 = sorted by
 * origin (Morton code in the root box) and direction octant when the query
 * has at least 2^20 rays and the tree at least 2^21 nodes (default), 2 =
 * always sorted.  The hit of every ray is the same in every mode; sorting
 * makes the rays traced together on one XCD share the nodes its L2 caches.
 * No reference counterpart. */
int aurora_set_ray_sort(aurora_ctx* ctx, int mode);

/* Validate and upload a flattened scene (replaces are::ObjectSet,
 * include/object/object_set.h:10-12, with its Triangle, Material and Texture
 * pointers, include/object/triangle.h:38-39).  Rejects degenerate quads the
 * way are::Triangle's constructor does (src/object/triangle.cpp:25-36). */
int aurora_scene_upload(aurora_ctx* ctx, const aurora_scene_desc* desc,
                        aurora_scene** out);
void aurora_scene_destroy(aurora_scene* scene);

/* Build the LBVH on the device: Morton codes, LDS radix sort, Karras radix
 * tree, bottom-up refit into 32-byte nodes, then the four-wide traversal
 * tree collapsed from it.  No reference counterpart (ARE scans linearly,
 * experiments/rt.cpp:209-218). */
int aurora_bvh_build(aurora_scene* scene, aurora_bvh_stats* stats);
/* Copy the node array (stats.n_nodes nodes) to host memory. */
int aurora_bvh_download(aurora_scene* scene, aurora_bvh_node* nodes,
                        uint32_t capacity);
/* Copy the four-wide traversal nodes (stats.n_wide of them) to host memory. */
int aurora_bvh_download_wide(aurora_scene* scene, aurora_bvh4_node* nodes,
                             uint32_t capacity);
/* Copy the sorted 30-bit Morton codes and primitive order to host memory. */
int aurora_bvh_download_order(aurora_scene* scene, uint32_t* morton,
                              uint32_t* prim_order, uint32_t capacity);

/* Closest hit of n rays in (tmin, tmax): the role of
 * are::Triangle::intersect_ray (src/object/triangle.cpp:82-121) over the
 * linear scan of experiments/rt.cpp:209-218, on spheres and quads through
 * the LBVH.  Directions need not be unit length; t is in units of the given
 * direction.  Ties in t go to the lower primitive id.  origins/dirs are
 * 3*n floats; prim receives the primitive id or AURORA_MISS, t the hit
 * distance (tmax on a miss).  Host buffers. */
int aurora_trace_closest(aurora_scene* scene, const float* origins,
                         const float* dirs, uint32_t n, float tmin,
                         float tmax, int32_t* prim, float* t);
/* Same on device buffers; counts (optional, device, 2*n u32) receives the
 * number of BVH nodes and primitive records fetched per ray. */
int aurora_trace_closest_device(aurora_scene* scene, const float* d_origins,
                                const float* d_dirs, uint32_t n, float tmin,
                                float tmax, int32_t* d_prim, float* d_t,
                                uint32_t* d_counts);
/* Device time and fetch totals of one traversal launch. */
typedef struct aurora_trace_stats {
    double extend_ms;          /* HIP-event time of the traversal kernel     */
    uint64_t nodes_fetched;    /* 32-byte BVH nodes (count_fetches only)     */
    uint64_t spheres_fetched;  /* 16-byte sphere records                     */
    uint64_t quads_fetched;    /* 48-byte quad records                       */
} aurora_trace_stats;
/* aurora_trace_closest_device with the traversal launch timed on the
 * context stream (synchronous on return); with count_fetches != 0 the
 * counting kernel variant also sums the records fetched. */
int aurora_trace_closest_device_stats(aurora_scene* scene,
                                      const float* d_origins,
                                      const float* d_dirs, uint32_t n,
                                      float tmin, float tmax, int32_t* d_prim,
                                      float* d_t, int count_fetches,
                                      aurora_trace_stats* stats);
/* Allocation-free form of aurora_trace_closest_device (same arguments and
 * hits).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no allocation, copy or
 * synchronisation; every buffer beyond the caller's rays, hits and counts
 * lives in a static device array of AURORA_NOALLOC_SCRATCH_BYTES (50 MiB)
 * that each GPU the library runs on holds until the process ends (packed
 * rays and hits, work counters, the scene table, the traversal stack
 * spill).  At most AURORA_NOALLOC_MAX_RAYS rays per call (more:
 * AURORA_ERR_INVALID); rays are traced in the given order (the
 * aurora_set_ray_sort mode does not apply).  Asynchronous on the context
 * stream.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (other contexts or streams, other threads) race. */
#define AURORA_NOALLOC_MAX_RAYS 262144u
#define AURORA_NOALLOC_SCRATCH_BYTES 52430848u
int aurora_trace_closest_device_noalloc(aurora_scene* scene,
                                        const float* d_origins,
                                        const float* d_dirs, uint32_t n,
                                        float tmin, float tmax,
                                        int32_t* d_prim, float* d_t,
                                        uint32_t* d_counts);
/* Brute-force closest hit (every primitive tested, no BVH): the semantics
 * of experiments/rt.cpp:209-218.  Parity aid. Device buffers. */
int aurora_trace_brute_device(aurora_scene* scene, const float* d_origins,
                              const float* d_dirs, uint32_t n, float tmin,
                              float tmax, int32_t* d_prim, float* d_t);

/* Render: the role of experiments/rt.cpp:337-374 (per-pixel loop) and
 * rt.cpp:251-334 (trace).  Adds the radiance of samples
 * [sample_offset, sample_offset+spp) of every pixel to rgb_sum (W*H*3
 * floats, row 0 = top; caller zeroes it before the first call).  With
 * AURORA_RENDER_AOV, albedo and normal (W*H*3 each) receive the sum over the
 * same samples of the first hit's albedo and shading normal.  Host buffers;
 * aovs may be NULL. */
int aurora_render(aurora_scene* scene, const aurora_camera_desc* cam,
                  const aurora_render_desc* rd, float* rgb_sum,
                  float* albedo_sum, float* normal_sum,
                  aurora_render_stats* stats);
/* Same on device buffers (no PCIe traffic). */
int aurora_render_device(aurora_scene* scene, const aurora_camera_desc* cam,
                         const aurora_render_desc* rd, float* d_rgb_sum,
                         float* d_albedo_sum, float* d_normal_sum,
                         aurora_render_stats* stats);

/* One image of a batched render. */
typedef struct aurora_batch_item {
    aurora_scene* scene;
    aurora_camera_desc camera;
    aurora_render_desc render;
    float* d_rgb_sum;        /* device, W*H*3                                */
    float* d_albedo_sum;     /* device, with AURORA_RENDER_AOV               */
    float* d_normal_sum;
} aurora_batch_item;
/* Render many scenes (uneven sizes, resolutions and sample counts) in
 * shared wavefront passes: every launch of a pass (raygen, traversal,
 * shading, fused tail, accumulation) serves all of them, each path carrying
 * its scene's index.  Every image equals its own aurora_render_device
 * result bit for bit.  flags: AURORA_RENDER_GRAPH.  No reference
 * counterpart (ARE renders one fixed scene per program). */
int aurora_render_batch_device(aurora_ctx* ctx, const aurora_batch_item* items,
                               uint32_t n, uint32_t flags,
                               aurora_render_stats* stats);

/* Camera rays of n (pixel, sample) pairs: ray k is the first ray that
 * aurora_render traces for sample samples[k] of pixel pixels[k] (pixel =
 * row * image_width + column, row 0 at the top) under Philox key word
 * `seed` (aurora_render_desc.seed): the pixel centre jittered within its
 * pixel, from the centre of projection or, with defocus_angle > 0, from a
 * point of the lens disk of radius focus_dist * tan(defocus_angle / 2).
 * The role of experiments/rt.cpp:337-343 (camera basis) and :364-366
 * (pixel -> direction), on the "Ray Tracing in One Weekend" viewport
 * (vertical field of view vfov, aspect width/height).  origins and dirs
 * receive 3*n floats (unit directions).  Host buffers; pixels must be below
 * width*height. */
int aurora_camera_rays(aurora_ctx* ctx, const aurora_camera_desc* cam,
                       uint32_t seed, const uint32_t* pixels,
                       const uint32_t* samples, uint32_t n, float* origins,
                       float* dirs);
/* Same on device buffers (a pixel beyond width*height gives the ray through
 * that position past the image's last row). */
int aurora_camera_rays_device(aurora_ctx* ctx, const aurora_camera_desc* cam,
                              uint32_t seed, const uint32_t* d_pixels,
                              const uint32_t* d_samples, uint32_t n,
                              float* d_origins, float* d_dirs);

/* Transfer curves of the 8-bit resolve.  The mean is sum / spp; NaN
 * counts as 0; each result is truncated to a byte. */
#define AURORA_TRANSFER_GAMMA2 0u   /* "Ray Tracing in One Weekend":
                                       256 * clamp(sqrt(x), 0, 0.999)       */
#define AURORA_TRANSFER_GAMMA22 1u  /* experiments/rt.cpp:377-392:
                                       255 * pow(clamp(x, 0, 1), 1/2.2f)    */
#define AURORA_TRANSFER_LINEAR 2u   /* are::Texture::save_texture
                                       (src/texture.cpp:378-391):
                                       clamp(255 * x, 0, 255)               */

/* Resolve a radiance sum (W*H*3 floats, rows top-down) to bytes in the
 * layout of a P6 PPM body (src/texture.cpp:378-391; header
 * "P6\n<w> <h>\n255\n" as at src/texture.cpp:375).  Host buffers. */
int aurora_resolve_rgb8(const float* rgb_sum, uint32_t width,
                        uint32_t height, uint32_t spp, uint32_t transfer,
                        uint8_t* out);
/* Same on device buffers, on the context's stream (asynchronous). */
int aurora_resolve_rgb8_device(aurora_ctx* ctx, const float* d_rgb_sum,
                               uint32_t width, uint32_t height, uint32_t spp,
                               uint32_t transfer, uint8_t* d_out);

/* are::Texture::paste (src/texture.cpp:85-360) on the GPU: warp src
 * (src_w x src_h RGB floats, rows top-down) into dst (dst_w x dst_h) by the
 * homography that sends the source corner pixels (0,0), (w-1,0), (0,h-1),
 * (w-1,h-1) to corners[0..7] = left_top, right_top, left_bottom,
 * right_bottom (x, y pixel positions).  A destination pixel whose centre
 * lies inside the quad and maps inside the source takes the source's
 * bilinear sample (clamped neighbours); every other pixel keeps its value.
 * An empty source, a quad outside the destination or a degenerate mapping
 * paste nothing and return AURORA_OK, as the reference returns.  The
 * homography is solved on the host and the warp runs in double on the
 * device, so each painted value is the reference's double result rounded
 * to float.  Device buffers, asynchronous on the context stream. */
int aurora_paste_device(aurora_ctx* ctx, const float* d_src, uint32_t src_w,
                        uint32_t src_h, float* d_dst, uint32_t dst_w,
                        uint32_t dst_h, const int32_t* corners);
/* Same on host buffers (dst is read and written back). */
int aurora_paste(aurora_ctx* ctx, const float* src, uint32_t src_w,
                 uint32_t src_h, float* dst, uint32_t dst_w, uint32_t dst_h,
                 const int32_t* corners);

/* ---- patch-projection renderer ------------------------------------------
 * The recursive mirror renderer of experiments/rt10.cpp:551-772, all in
 * double as there.  Each viewport triangle paints, far to near (triangle
 * centroid distance from the eye), every scene triangle with a vertex that
 * projects inside it, as that triangle's texture warped through the
 * Sutherland-Hodgman clip of its projection; a reflective triangle's
 * texture is rendered by the same procedure from the eye mirrored across
 * its plane, onto the triangle itself, and mixed with its albedo.  The
 * recursion stops at max_depth, below min_area_px estimated texels, and at
 * a triangle id already on the recursion path.  The recursion tree and the
 * clipped warp triangles are planned on the host; every texel of every
 * texture and of the image is rasterised on the device, one launch per
 * recursion level, deepest first. */
#define AURORA_PATCH_DIFFUSE 0u
#define AURORA_PATCH_REFLECTIVE 1u
#define AURORA_PATCH_NO_MATERIAL 0xffffffffu  /* white diffuse (rt10.cpp:570-574) */
#define AURORA_PATCH_MAX_DEPTH 64
#define AURORA_PATCH_MAX_TEX_RES 4096

/* rt10.cpp:145-155 MaterialType, Material */
typedef struct aurora_patch_material {
    uint32_t type;         /* AURORA_PATCH_DIFFUSE / AURORA_PATCH_REFLECTIVE */
    uint32_t reserved;     /* 0                                            */
    double albedo[3];
    double metalness;      /* reflective: weight of the albedo-tinted
                              reflection, clamped to [0, 1]                */
} aurora_patch_material;

/* rt10.cpp:157-169 Triangle */
typedef struct aurora_patch_triangle {
    double p[3][3];        /* vertices                                     */
    double uv[3][2];       /* texture coordinates, v down                  */
    int32_t id;            /* recursion-guard identity                     */
    uint32_t material;     /* material index or AURORA_PATCH_NO_MATERIAL   */
} aurora_patch_triangle;

/* rt10.cpp:670-676 Camera: an eye and two viewport triangles whose uv
 * (in [0,1]^2, v down) place them in the image */
typedef struct aurora_patch_camera {
    double origin[3];
    aurora_patch_triangle viewport[2];  /* id and material unused        */
    uint32_t width, height;             /* >= 2                          */
} aurora_patch_camera;

/* rt10.cpp:543-549 RenderConfig */
typedef struct aurora_patch_config {
    int32_t max_depth;     /* 0 .. AURORA_PATCH_MAX_DEPTH                  */
    int32_t max_tex_res;   /* texture side clamp, 1 <= min <= max <=      */
    int32_t min_tex_res;   /*   AURORA_PATCH_MAX_TEX_RES                   */
    uint32_t reserved;     /* 0                                            */
    double min_area_px;    /* no recursion below this many texels          */
    double env[3];         /* colour where a reflection sees nothing       */
} aurora_patch_config;

typedef struct aurora_patch_stats {
    uint32_t textures;         /* triangle textures in the recursion tree  */
    uint32_t raster_textures;  /* of them rasterised (reflective and
                                  recursed; the others are constant)       */
    uint32_t levels;           /* recursion levels with rasterised textures */
    uint32_t warps;            /* warp triangles (clipped-polygon fans)     */
    uint64_t texels;           /* texels rasterised, viewports included     */
    uint64_t warp_texels;      /* texel x warp-triangle pairs in the warp's
                                  clamped pixel box (the reference's
                                  per-pixel loop iterations)                */
    double plan_ms;            /* host: visibility, clipping, recursion     */
    double kernel_ms;          /* device: the raster launches               */
} aurora_patch_stats;

/* Render into rgb (height x width x 3 doubles, row 0 at the top, clamped
 * to [0,1] as rt10.cpp:775-784).  Host buffer; synchronous.  The image is
 * the reference's bit for bit.  Errors: AURORA_ERR_INVALID for bad sizes,
 * material indices or configuration; AURORA_ERR_NOMEM when the recursion
 * tree's textures exceed the device (stats then hold the plan's size).
 * stats may be NULL. */
int aurora_patch_render(aurora_ctx* ctx, const aurora_patch_triangle* tris,
                        uint32_t n_tris, const aurora_patch_material* mats,
                        uint32_t n_mats, const aurora_patch_camera* cam,
                        const aurora_patch_config* cfg, double* rgb,
                        aurora_patch_stats* stats);
/* Same with a device output buffer (waits for the raster launches). */
int aurora_patch_render_device(aurora_ctx* ctx,
                               const aurora_patch_triangle* tris,
                               uint32_t n_tris,
                               const aurora_patch_material* mats,
                               uint32_t n_mats, const aurora_patch_camera* cam,
                               const aurora_patch_config* cfg, double* d_rgb,
                               aurora_patch_stats* stats);

/* ---- learned denoiser ---------------------------------------------------
 * A 4-level U-Net (32/48/64/64 channels, 15 3x3 convolutions) on BF16 MFMA
 * with FP32 accumulation, fed the noisy radiance mean and the albedo and
 * normal AOVs of aurora_render (AURORA_RENDER_AOV) and predicting a residual
 * in log(1 + radiance).  No reference counterpart (ARE has no learned
 * component).  Parameters are a flat FP32 vector: per layer the weights
 * [Cout][3x3 tap][Cin] then the biases. */
typedef struct aurora_denoiser aurora_denoiser;
/* He-normal initialisation from Philox(seed); output layer zero. */
int aurora_denoiser_create(aurora_ctx* ctx, uint32_t seed,
                           aurora_denoiser** out);
void aurora_denoiser_destroy(aurora_denoiser* dn);
uint64_t aurora_denoiser_param_count(const aurora_denoiser* dn);
int aurora_denoiser_get_params(aurora_denoiser* dn, float* host, uint64_t n);
int aurora_denoiser_set_params(aurora_denoiser* dn, const float* host,
                               uint64_t n);
/* Denoise one image: inputs are device sums over `spp` samples (W*H*3 each,
 * as aurora_render_device writes them); out receives linear RGB means.
 * width and height must be multiples of 8. */
int aurora_denoise(aurora_denoiser* dn, const float* d_rgb_sum,
                   const float* d_albedo_sum, const float* d_normal_sum,
                   uint32_t spp, uint32_t width, uint32_t height,
                   float* d_out);
/* Training loss against d_target (linear RGB means, batch x H x W x 3).
 * AURORA_LOSS_LOG_L2 (default): mean squared error in log(1 + radiance).
 * AURORA_LOSS_REL_L2: mean of (D - T)^2 / (D^2 + 0.01) over pixels and
 * channels, D the denoised radiance, with the denominator held constant in
 * the gradient (the relative MSE of the error metric; unbiased with noisy
 * targets). */
#define AURORA_LOSS_LOG_L2 0
#define AURORA_LOSS_REL_L2 1
int aurora_denoiser_set_loss(aurora_denoiser* dn, int kind);
/* Loss (as set by aurora_denoiser_set_loss) and its gradient; host_grads
 * (optional) receives the flat FP32 gradient. */
int aurora_denoiser_loss_grad(aurora_denoiser* dn, const float* d_rgb_sum,
                              const float* d_albedo_sum,
                              const float* d_normal_sum, const float* d_target,
                              uint32_t batch, uint32_t spp, uint32_t width,
                              uint32_t height, float* loss, float* host_grads);
/* loss_grad followed by one Adam step (beta 0.9 / 0.999, eps 1e-8). */
int aurora_denoiser_train_step(aurora_denoiser* dn, const float* d_rgb_sum,
                               const float* d_albedo_sum,
                               const float* d_normal_sum,
                               const float* d_target, uint32_t batch,
                               uint32_t spp, uint32_t width, uint32_t height,
                               float lr, float* loss);
/* Self-test of the MFMA lane layout: D (16x16 FP32) = A (16x32 BF16 bits,
 * row-major) x B (32x16). Host buffers. */
int aurora_selftest_mfma(aurora_ctx* ctx, const uint16_t* A, const uint16_t* B,
                         float* D);
/* Self-test of the fast reciprocal the path kernels use: the float32 bit
 * patterns in [lo, hi) whose hardware-reciprocal-plus-Newton-step result
 * differs from IEEE division 1.0f / x are counted into *mismatches, and the
 * first (up to 16, in no particular order) written to first (may be NULL).
 * The kernels use that form only for 2^-126 <= |x| < 2^126, the patterns
 * [0x00800000, 0x7e800000) and [0x80800000, 0xfe800000), where the count
 * is zero.  No counterpart in the reference (its divisions run on the CPU). */
int aurora_selftest_rcp(aurora_ctx* ctx, uint32_t lo, uint32_t hi,
                        uint64_t* mismatches, uint32_t* first);

#ifdef __cplusplus
}
#endif

#endif /* AURORA_H */
