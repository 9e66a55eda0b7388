// internal.hpp — libaurora internals shared by the host-side C ABI and the
// kernel translation units.  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "dev_common.hpp"

namespace aur {

// extend work pools: one counter per XCD, each on its own 128-byte line
constexpr int kPools = 8;
constexpr uint32_t kPoolStride = 32;
// render counters: [0], [1] queue counts, pools from kPoolBase
constexpr uint32_t kPoolBase = 32;
// 64-bit count of ray segments traced in the pass (shade and tail add)
constexpr uint32_t kTailSegs = kPoolBase + kPools * kPoolStride;
constexpr size_t kCounterBytes = (kTailSegs + 32) * 4;

struct Status {
    int code = AURORA_OK;
    std::string msg;
    static Status ok() { return Status(); }
    bool bad() const { return code != AURORA_OK; }
};

#define AUR_HIP(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return Status{e_ == hipErrorOutOfMemory ? AURORA_ERR_NOMEM : AURORA_ERR_RUNTIME,      \
                          std::string(#expr) + ": " + hipGetErrorString(e_)};                      \
    } while (0)

#define AUR_TRY(expr)             \
    do {                          \
        Status aur_st_ = (expr);  \
        if (aur_st_.bad()) return aur_st_; \
    } while (0)

// Device buffer owned by the library.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    Status alloc(size_t n) {
        if (n <= bytes && p) return Status::ok();
        if (p) { (void)hipFree(p); p = nullptr; bytes = 0; }
        if (n == 0) return Status::ok();
        AUR_HIP(hipMalloc(&p, n));
        bytes = n;
        return Status::ok();
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Two timing events, destroyed on every return path.
struct EventPair {
    hipEvent_t e[2] = {nullptr, nullptr};
    Status create() {
        for (hipEvent_t& x : e) AUR_HIP(hipEventCreate(&x));
        return Status::ok();
    }
    ~EventPair() {
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
    }
};

// Scratch of the LBVH build.
struct BvhScratch {
    DevBuf box_lo, box_hi;      // float4 per prim (padded boxes)
    DevBuf bounds;              // 12 ordered-u32: scene lo/hi, centroid lo/hi
    DevBuf bpart;               // 12 ordered-u32 per block of the bounds reduction
    DevBuf keys[2], vals[2];    // radix sort ping-pong
    DevBuf hist;                // 256 * blocks
    DevBuf gamma, flags, int_parent, leaf_parent, int_pos, leaf_pos;
    DevBuf jump[2];             // (ancestor, distance) per internal node, pointer-jumping ping-pong
    DevBuf lvl_key, lvl_node;   // internal nodes sorted by depth
    DevBuf lvl_start;           // first sorted index of each depth (257 u32)
    DevBuf widx;                // wide index of each internal node (exclusive scan of even-depth flags)
    void release() {
        DevBuf* all[] = {&box_lo, &box_hi, &bounds, &bpart, &keys[0], &keys[1], &vals[0], &vals[1], &hist, &gamma,
                         &flags, &int_parent, &leaf_parent, &int_pos, &leaf_pos, &jump[0], &jump[1], &lvl_key,
                         &lvl_node, &lvl_start, &widx};
        for (DevBuf* b : all) b->release();
    }
};

// Build the LBVH into `nodes` (2n float4 pairs) and the four-wide traversal
// tree collapsed from it into `wide` (8 float4 per node) on `stream`.
// morton/order receive the sorted codes and primitive order.
Status lbvh_build(const DevScene& sc, uint32_t n_prims, DevBuf& nodes, DevBuf& wide, DevBuf& morton, DevBuf& order,
                  BvhScratch& scr, hipStream_t stream, aurora_bvh_stats* stats);

// Scenes of at most kFlatMaxPrims primitives also keep their leaf records in
// Morton order (32 bytes each) for the flat traversal; flat_build fills them
// from the LBVH that lbvh_build left in `nodes` and `scr`.
constexpr uint32_t kFlatMaxPrims = 256;
Status flat_build(uint32_t n_prims, const DevBuf& nodes, const BvhScratch& scr, DevBuf& flat, hipStream_t stream);

// Stable LSD radix sort of (key, value) pairs, 8 bits per pass from bit 0.
// Sorted pairs end in (k0, v0) for an even number of passes, else in
// (k1, v1).  hist holds radix_hist_bytes(n); after the call it holds the
// last pass's exclusive scan in digit-major order (entry d * blocks + b is
// where block b's keys of digit d start, blocks = ceil(n / 2048)), so
// hist[d * blocks] is the first sorted index of digit d (the LBVH build reads
// its per-depth node ranges there).
size_t radix_hist_bytes(uint32_t n);
Status radix_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, uint32_t* hist, uint32_t n,
                        int passes, hipStream_t stream);

// Extend (closest hit) launch parameters.
struct ExtendArgs {
    const float4* ray_o;      // (origin, any)
    const float4* ray_d;      // (direction, any)
    const uint32_t* count;    // device count of rays (or nullptr: use n)
    uint32_t n;
    uint32_t* pool;           // kPools work counters kPoolStride u32 apart, 0 at launch
    uint32_t* zero_after;     // optional: written 0 by the kernel (next queue count)
    float tmax;               // ray interval (SceneRec::tmin, tmax)
    float2* hit;              // (t, prim bits)
    const struct SceneRec* recs;  // scene table; a ray's scene is the bits of ray_d.w
    uint32_t one_scene;       // nonzero: every ray belongs to recs[0]
    uint32_t wide;            // nonzero: walk the four-wide tree, else the binary one
    const uint32_t* order;    // optional: claim order of the rays (slot -> ray index)
    uint32_t* counts;         // optional: 3 per ray (nodes, spheres, quads fetched)
    unsigned long long* totals; // optional: [nodes, spheres, quads] summed
    uint32_t* spill;          // stack spill, spill_stride * extend_spill_levels() u32
    uint32_t spill_stride;    // threads in the persistent grid
};
Status launch_extend(const ExtendArgs& a, int grid, hipStream_t stream);
Status launch_brute(const DevScene& sc, const float4* o, const float4* d, uint32_t n, float tmin, float tmax,
                    float2* hit, hipStream_t stream);
// Sort keys of n rays for traversal order: 27-bit Morton code of the origin
// within the box of the scene's root node (origins outside are clamped),
// then the 3-bit direction octant; values are the ray indices.
Status launch_ray_keys(const float4* ray_o, const float4* ray_d, uint32_t n, const float4* nodes, uint32_t* keys,
                       uint32_t* vals, hipStream_t stream);
int extend_block();         // threads per block of the extend kernel
int extend_spill_levels();  // u32 stack entries per thread in ExtendArgs::spill
constexpr int kExtendBlock = 256;        // extend_block()
constexpr int kExtendSpillLevels = 80;   // extend_spill_levels()

// Wavefront render.  A pass renders spp_pass samples of every pixel of one
// or more scenes; each scene of the pass has a record in a device table and
// every path carries its scene's index.
struct CamF {
    v3 center, p00, du, dv, disk_u, disk_v;
    uint32_t W, H;
    uint32_t defocus;
};
struct SceneRec {
    DevScene sc;
    CamF cam;
    uint32_t seed, sample_base, spp_pass, max_depth;
    float tmin;
    float bg[3];
    uint32_t aov;                 // write first-hit albedo / normal
    uint32_t WH;                  // pixels
    uint32_t path_off, pix_off;   // first path id / pixel of the scene in the pass
    float *rgb, *alb, *nrm;       // per-pixel sums (device)
};
struct Pass {
    const SceneRec* recs;  // device table
    uint32_t n_recs;
    uint32_t n_paths;      // sum of WH * spp_pass
    uint32_t n_pixels;     // sum of WH over scenes with spp_pass > 0
    uint32_t any_aov;
    uint32_t wide;         // render traversal over the four-wide tree (else the binary one)
    uint32_t flat;         // the fused tail walks the scene's flat leaf array (one scene with one)
    uint32_t claim;        // paths per work-pool claim in the fused tail (0: by pass size)
    uint32_t lean_queue;   // the fused tail carries the pass from the first bounce: raygen writes
                           // directions, and origins for a thin lens only (throughput 1, radiance
                           // set by the tail; a queue position is its path)
};
struct Workspace {
    DevBuf q_o[2], q_d[2], q_b[2];   // path queues (float4 each)
    DevBuf hit;                       // float2 per slot
    DevBuf L;                         // float4 per path id (radiance)
    DevBuf aov_a, aov_n;              // 3 floats per path id
    DevBuf counters;                  // queue counts [0], [1]; pools from kPoolBase
    DevBuf spill;
    DevBuf recs;                      // SceneRec table of the current pass
    uint32_t capacity = 0;
    uint32_t* host_counts = nullptr;  // pinned, 256 slots
};
Status launch_raygen(const Pass& p, Workspace& ws, hipStream_t s);
// camera rays of n (pixel, sample) pairs into 3*n float origins/directions
// (pixel = j*W + i < W*H, checked by the caller)
Status launch_camera_rays(const CamF& cam, uint32_t seed, const uint32_t* pixels, const uint32_t* samples,
                          uint32_t n, float* org, float* dir, hipStream_t s);
// are::Texture::paste on the device: src (sw x sh RGB floats) warped into the
// quad corners[0..7] = LT, RT, LB, RB (x, y) of dst (dw x dh); *painted is
// false when there is nothing to paste (the reference's early returns)
Status paste_device(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                    const int32_t* corners, hipStream_t s, bool* painted);

// Patch-projection renderer (experiments/rt10.cpp): device tables of one
// render (texture records, warp triangles, tiles, the texel pool of the
// recursed textures) and the host-output staging image, kept by the context
// and grown as needed.
struct PatchWork {
    DevBuf texs, warps, tiles, pool, out;
    void release() {
        DevBuf* all[] = {&texs, &warps, &tiles, &pool, &out};
        for (DevBuf* b : all) b->release();
    }
};
Status patch_check(const aurora_patch_triangle* tris, uint32_t n_tris, const aurora_patch_material* mats,
                   uint32_t n_mats, const aurora_patch_camera* cam, const aurora_patch_config* cfg);
// plan on the host, rasterise on stream s into d_out (H x W x 3 doubles);
// returns after the raster launches completed (stats->kernel_ms)
Status patch_render_device(PatchWork& pw, const aurora_patch_triangle* tris, uint32_t n_tris,
                           const aurora_patch_material* mats, const aurora_patch_camera* cam,
                           const aurora_patch_config* cfg, double* d_out, hipStream_t s, aurora_patch_stats* stats);
// 8-bit resolve of n floats (sum and out 16-byte and 4-byte aligned)
Status launch_resolve_rgb8(const float* sum, size_t n, uint32_t spp, uint32_t transfer, uint8_t* out, hipStream_t s);
Status launch_shade(const Pass& p, Workspace& ws, int cur, uint32_t depth, int grid, hipStream_t s);
Status launch_tail(const Pass& p, Workspace& ws, int cur, uint32_t depth, int grid, hipStream_t s);
Status launch_accumulate(const Pass& p, Workspace& ws, hipStream_t s);
// sample_base += spp_pass for every record (lets a captured pass be replayed)
Status launch_advance(SceneRec* recs, uint32_t n, hipStream_t s);

}  // namespace aur
