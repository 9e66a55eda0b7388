// capi.hip — host side of libaurora's C ABI (include/aurora.h).
//
// Owns contexts (device + stream), uploads and validates flattened scenes,
// builds the LBVH, and drives the wavefront passes (raygen, then per bounce
// extend + shade, then accumulate).  Never throws across the ABI: every
// failure becomes a status code plus a message in aurora_last_error.
//
// Reference counterparts: scene containers are::ObjectSet / are::Triangle
// (include/object/object_set.h:10-12, include/object/triangle.h:19,38-39);
// validation rules of the Triangle constructor (src/object/triangle.cpp:
// 13-36); the render loop of experiments/rt.cpp:337-374; the P6 byte layout
// of are::Texture::save_texture (src/texture.cpp:378-391).
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"

using namespace aur;

// automatic ray sorting of ray queries: at least this many rays on a tree of
// at least this many nodes (32 bytes each: 64 MiB, beyond what the L2s hold)
constexpr uint32_t kSortMinRays = 1u << 20;
// resident 256-thread blocks of the fused tail kernel per CU: it is built for
// five waves per SIMD for the tree walks, six for the four-wide walk of a
// single scene, eight for the flat walk
constexpr int kTailBlocksPerCu = 5;
constexpr int kTailWideOneBlocksPerCu = 6;
constexpr int kTailFlatBlocksPerCu = 8;
constexpr uint32_t kSortMinNodes = 1u << 21;

struct aurora_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int cus = 0;
    int ext_grid = 0;    // resident blocks of the extend kernel
    int shade_grid = 0;
    uint32_t tail_paths = 0xffffffffu;  // fused-tail switch: live paths at or below (default: always)
    uint32_t graph_tail_depth = 8;   // captured passes switch to the fused tail here
    int ray_sort = 1;                // ray queries: 0 never sort, 1 large scenes, 2 always
    uint32_t render_tree = 0;        // render bounces: 0 binary LBVH, 1 four-wide tree, 2 flat (small scenes)
    uint64_t pass_paths = 0;         // automatic passes: at most this many paths (0: kPassPaths)
    uint32_t claim_paths = 0;        // fused-tail paths per pool claim (0: by pass size)
    Workspace ws;                    // path queues shared by the context's renders
    // pinned staging of the passes' scene tables, two in turn: a table is
    // rewritten only after the copy that read it two passes ago has run
    // (recs_ev), so passes and calls are enqueued without waiting for the GPU
    SceneRec* host_recs[2] = {nullptr, nullptr};
    hipEvent_t recs_ev[2] = {nullptr, nullptr};
    int recs_turn = 0;
    uint32_t host_recs_cap = 0;
    unsigned char* scratch = nullptr;  // this device's copy of the allocation-free scratch array
    PatchWork patch;                   // patch-projection renderer tables
};

struct aurora_scene {
    aurora_ctx* ctx = nullptr;
    DevScene view{};
    uint32_t n_prims = 0;
    DevBuf spheres, sphere_mat, quads, quad_n, mats, texs, texels, perlin_vec, perlin_perm;
    DevBuf nodes, wide, morton, order, flat;
    BvhScratch scr;
    bool built = false;
    aurora_bvh_stats stats{};
    DevBuf rec1;  // one-entry scene table for ray queries
    DevBuf tr_o, tr_d, tr_hit, tr_spill, tr_pool, totals;
    DevBuf tr_key[2], tr_val[2], tr_hist;  // ray sort of large queries
};

// ---- allocation-free ray queries ----------------------------------------
// aurora_trace_closest_device_noalloc keeps every buffer beyond the caller's
// rays and hits in `scratch`, a static device array of which each GPU holds
// its own copy until the process ends: the packed rays and hits of up to
// kNoAllocMaxRays rays, the work-pool counters, the one-entry scene table and
// the traversal stack spill of a grid of at most kNoAllocGrid blocks.  Host
// code never passes the array's name to a kernel (on the host it is only a
// stand-in): its device address is taken once per device with
// hipGetSymbolAddress (scratch_address, called by aurora_create) and kept
// in the context.
constexpr uint32_t kNoAllocMaxRays = AURORA_NOALLOC_MAX_RAYS;
constexpr int kNoAllocGrid = 512;  // two 256-thread blocks per CU
constexpr size_t kNaRayBytes = (size_t)kNoAllocMaxRays * 16;  // float4 per ray, origins then directions
constexpr size_t kNaHitBytes = (size_t)kNoAllocMaxRays * 8;   // float2 (t, prim) per ray
constexpr size_t kNaPoolBytes = (size_t)kPools * kPoolStride * 4;
constexpr size_t kNaRecBytes = 1024;
constexpr size_t kNaSpillBytes = (size_t)kNoAllocGrid * kExtendBlock * kExtendSpillLevels * 4;
constexpr size_t kNaOffDir = kNaRayBytes;
constexpr size_t kNaOffHit = 2 * kNaRayBytes;
constexpr size_t kNaOffPool = kNaOffHit + kNaHitBytes;
constexpr size_t kNaOffRec = kNaOffPool + kNaPoolBytes;
constexpr size_t kNaOffSpill = kNaOffRec + kNaRecBytes;
constexpr size_t kNoAllocScratchBytes = kNaOffSpill + kNaSpillBytes;
static_assert(kNoAllocScratchBytes == AURORA_NOALLOC_SCRATCH_BYTES, "scratch size stated in aurora.h");
static_assert(sizeof(SceneRec) <= kNaRecBytes, "scene record fits its slot");
static_assert(kNaOffDir % 256 == 0 && kNaOffHit % 256 == 0 && kNaOffPool % 256 == 0 && kNaOffRec % 256 == 0 &&
                  kNaOffSpill % 256 == 0,
              "scratch regions 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
unsigned char* g_scratch_dev[kMaxDevices] = {};

// the device address of this device's scratch array; `device` must be the
// current device
Status scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= kMaxDevices) return Status{AURORA_ERR_INVALID, "device ordinal beyond 63"};
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        AUR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_dev[device];
    return Status::ok();
}

// clears the work-pool counters and writes the one-entry scene table of an
// allocation-free query (a kernel, not a copy or a memset, so the call stays
// capturable)
__global__ void k_noalloc_setup(uint32_t* pool, SceneRec* rec, SceneRec value) {
    const uint32_t t = threadIdx.x;
    if (t < (uint32_t)(kPools * kPoolStride)) pool[t] = 0u;
    if (t == 0) *rec = value;
}

thread_local std::string g_create_err;

int fail(aurora_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_err = msg;
    return code;
}
int fail(aurora_ctx* c, const Status& s) { return fail(c, s.code, s.msg); }

#define CK(ctx, expr)                         \
    do {                                      \
        Status st_ = (expr);                  \
        if (st_.bad()) return fail(ctx, st_); \
    } while (0)

Status hip_status(hipError_t e, const char* what) {
    if (e == hipSuccess) return Status::ok();
    return Status{e == hipErrorOutOfMemory ? AURORA_ERR_NOMEM : AURORA_ERR_RUNTIME,
                  std::string(what) + ": " + hipGetErrorString(e)};
}

Status upload(DevBuf& b, const void* src, size_t bytes, hipStream_t s) {
    AUR_TRY(b.alloc(bytes ? bytes : 16));
    if (bytes) AUR_HIP(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, s));
    return Status::ok();
}

bool finite3(const float* v) { return std::isfinite(v[0]) && std::isfinite(v[1]) && std::isfinite(v[2]); }
bool tiny3(double x, double y, double z) { return std::fabs(x) < 1e-8 && std::fabs(y) < 1e-8 && std::fabs(z) < 1e-8; }

std::string check_scene(const aurora_scene_desc& d) {
    char b[200];
    if (d.n_spheres && (!d.spheres || !d.sphere_material)) return "sphere arrays are NULL";
    if (d.n_quads && (!d.quads || !d.quad_material)) return "quad arrays are NULL";
    if (d.n_materials && !d.materials) return "material table is NULL";
    if (d.n_textures && !d.textures) return "texture table is NULL";
    if (d.n_texels && !d.texels) return "texel array is NULL";
    if ((uint64_t)d.n_spheres + d.n_quads > 0x7fffffffull) return "too many primitives";
    for (uint32_t i = 0; i < d.n_spheres; ++i) {
        const float* s = d.spheres + 4 * (size_t)i;
        if (!finite3(s) || !std::isfinite(s[3]) || !(s[3] > 0.0f))
            return (snprintf(b, sizeof b, "sphere %u: non-finite centre or radius <= 0", i), b);
        if (d.sphere_material[i] >= d.n_materials)
            return (snprintf(b, sizeof b, "sphere %u: material index out of range", i), b);
    }
    for (uint32_t i = 0; i < d.n_quads; ++i) {
        const float* q = d.quads + 9 * (size_t)i;
        if (!finite3(q) || !finite3(q + 3) || !finite3(q + 6))
            return (snprintf(b, sizeof b, "quad %u: non-finite vertex or edge", i), b);
        const double ux = q[3], uy = q[4], uz = q[5], vx = q[6], vy = q[7], vz = q[8];
        if (tiny3(ux, uy, uz)) return (snprintf(b, sizeof b, "quad %u: edge vector u cannot be zero vector", i), b);
        if (tiny3(vx, vy, vz)) return (snprintf(b, sizeof b, "quad %u: edge vector v cannot be zero vector", i), b);
        if (tiny3(uy * vz - uz * vy, uz * vx - ux * vz, ux * vy - uy * vx))
            return (snprintf(b, sizeof b, "quad %u: edge vectors u and v cannot be collinear", i), b);
        if (d.quad_material[i] >= d.n_materials)
            return (snprintf(b, sizeof b, "quad %u: material index out of range", i), b);
        if (d.quad_kind && d.quad_kind[i] > AURORA_QUAD_TRIANGLE)
            return (snprintf(b, sizeof b, "quad %u: unknown kind", i), b);
    }
    for (uint32_t i = 0; i < d.n_materials; ++i) {
        const aurora_material& m = d.materials[i];
        if (m.type > AURORA_MAT_DIFFUSE_LIGHT) return (snprintf(b, sizeof b, "material %u: unknown type", i), b);
        if (m.type != AURORA_MAT_DIELECTRIC && m.texture >= d.n_textures)
            return (snprintf(b, sizeof b, "material %u: texture index out of range", i), b);
        if (m.type == AURORA_MAT_DIELECTRIC && !(m.ior > 0.0f && std::isfinite(m.ior)))
            return (snprintf(b, sizeof b, "material %u: ior must be positive", i), b);
        if (m.type == AURORA_MAT_METAL && !(m.fuzz >= 0.0f))
            return (snprintf(b, sizeof b, "material %u: fuzz must be >= 0", i), b);
    }
    for (uint32_t i = 0; i < d.n_textures; ++i) {
        const aurora_texture& t = d.textures[i];
        if (t.type > AURORA_TEX_IMAGE) return (snprintf(b, sizeof b, "texture %u: unknown type", i), b);
        if (t.type == AURORA_TEX_CHECKER) {
            if (t.even >= d.n_textures || t.odd >= d.n_textures)
                return (snprintf(b, sizeof b, "texture %u: checker child out of range", i), b);
            if (d.textures[t.even].type == AURORA_TEX_CHECKER || d.textures[t.odd].type == AURORA_TEX_CHECKER)
                return (snprintf(b, sizeof b, "texture %u: checker children cannot be checkers", i), b);
            if (!(t.scale > 0.0f)) return (snprintf(b, sizeof b, "texture %u: checker scale must be positive", i), b);
        }
        if (t.type == AURORA_TEX_IMAGE) {
            if (t.width == 0 || t.height == 0)
                return (snprintf(b, sizeof b, "texture %u: image width and height must be positive", i), b);
            if (t.texel_offset + (uint64_t)t.width * t.height > d.n_texels)
                return (snprintf(b, sizeof b, "texture %u: image texels out of range", i), b);
        }
    }
    return "";
}

// Perlin tables from Philox(seed, STREAM_PERLIN): 256 unit gradients from one
// draw each, then three Fisher–Yates shuffles (index j = floor(x*(i+1)/2^32)).
void perlin_fill(uint32_t seed, float* vec, int* perm) {
    for (uint32_t i = 0; i < 256; ++i) {
        const u4 r = philox(u4{i, 0u, 0u, 0u}, seed, STREAM_PERLIN);
        float x = 2.0f * u01(r.x) - 1.0f, y = 2.0f * u01(r.y) - 1.0f, z = 2.0f * u01(r.z) - 1.0f;
        if (x * x + y * y + z * z == 0.0f) { x = 1.0f; y = 0.0f; z = 0.0f; }
        const float l = sqrtf(x * x + y * y + z * z);
        vec[3 * i] = x / l;
        vec[3 * i + 1] = y / l;
        vec[3 * i + 2] = z / l;
    }
    for (uint32_t a = 0; a < 3; ++a) {
        int* p = perm + 256 * a;
        for (int i = 0; i < 256; ++i) p[i] = i;
        for (uint32_t i = 255; i >= 1; --i) {
            const u4 r = philox(u4{i, a + 1u, 0u, 0u}, seed, STREAM_PERLIN);
            const uint32_t j = (uint32_t)(((uint64_t)r.x * (i + 1u)) >> 32);
            const int t = p[i];
            p[i] = p[j];
            p[j] = t;
        }
    }
}

struct D3 {
    double x, y, z;
};
D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator*(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
D3 operator/(D3 a, double s) { return D3{a.x / s, a.y / s, a.z / s}; }
double len(D3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
D3 crs(D3 a, D3 b) { return D3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
v3 tof(D3 a) { return V((float)a.x, (float)a.y, (float)a.z); }

// Camera frame in double, then rounded once to float ("Ray Tracing in One
// Weekend" camera: viewport at focus_dist, pixel centres at +0.5 as in
// experiments/rt.cpp:364-366).
std::string camera_frame(const aurora_camera_desc& c, CamF& out) {
    if (c.image_width == 0 || c.image_height == 0) return "image width and height must be positive";
    if ((uint64_t)c.image_width * c.image_height > (1ull << 31)) return "image too large";
    if (!(c.vfov > 0.0f && c.vfov < 180.0f)) return "vfov must be in (0, 180) degrees";
    if (!(c.focus_dist > 0.0f)) return "focus_dist must be positive";
    const double PI = 3.1415926535897932385;
    const D3 lf{c.lookfrom[0], c.lookfrom[1], c.lookfrom[2]}, la{c.lookat[0], c.lookat[1], c.lookat[2]},
        up{c.vup[0], c.vup[1], c.vup[2]};
    if (len(lf - la) == 0.0) return "lookfrom and lookat coincide";
    const double theta = (double)c.vfov * PI / 180.0;
    const double h = std::tan(theta / 2.0);
    const double fd = c.focus_dist;
    const double vh = 2.0 * h * fd;
    const double vw = vh * ((double)c.image_width / (double)c.image_height);
    const D3 w = (lf - la) / len(lf - la);
    const D3 cu = crs(up, w);
    if (len(cu) == 0.0) return "vup is parallel to the view direction";
    const D3 u = cu / len(cu);
    const D3 v = crs(w, u);
    const D3 vu = vw * u, vv = vh * D3{-v.x, -v.y, -v.z};
    const D3 du = vu / (double)c.image_width, dv = vv / (double)c.image_height;
    const D3 ul = ((lf - fd * w) - vu / 2.0) - vv / 2.0;
    const D3 p00 = ul + 0.5 * (du + dv);
    const double rad = fd * std::tan(((double)c.defocus_angle / 2.0) * PI / 180.0);
    out.center = tof(lf);
    out.p00 = tof(p00);
    out.du = tof(du);
    out.dv = tof(dv);
    out.disk_u = tof(rad * u);
    out.disk_v = tof(rad * v);
    out.defocus = c.defocus_angle > 0.0f ? 1u : 0u;
    out.W = c.image_width;
    out.H = c.image_height;
    return "";
}

// the unit normal of each quad, unitv(u x v): the same operations shading
// would run per hit (each correctly rounded, no contraction), done once
__global__ void k_quad_normals(const float4* quads, uint32_t n, float4* nrm) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const v3 w = unitv(cross(xyz(quads[3 * (size_t)i + 1]), xyz(quads[3 * (size_t)i + 2])));
    nrm[i] = make_float4(w.x, w.y, w.z, 0.0f);
}

__global__ void k_pack3(const float* src, float4* dst, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = make_float4(src[3 * (size_t)i], src[3 * (size_t)i + 1], src[3 * (size_t)i + 2], 0.0f);
}
__global__ void k_unpack_hits(const float2* h, int32_t* prim, float* t, uint32_t n) {
    const uint32_t i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const float2 v = h[i];
        prim[i] = __float_as_int(v.y);
        t[i] = v.x;
    }
}

int grid_for_extend(aurora_ctx* c, uint64_t n) {
    const uint64_t need = (n + extend_block() - 1) / extend_block();
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(need, (uint64_t)c->ext_grid));
}

Status ensure_spill(DevBuf& spill, int grid) {
    return spill.alloc((size_t)grid * extend_block() * extend_spill_levels() * 4);
}

Status trace_device(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                    int32_t* prim, float* t, uint32_t* counts, bool brute, aurora_trace_stats* ts = nullptr,
                    bool count_fetches = false) {
    aurora_ctx* c = s->ctx;
    if (!brute && !s->built) return Status{AURORA_ERR_STATE, "BVH not built: call aurora_bvh_build first"};
    if (n == 0) return Status::ok();
    AUR_TRY(s->tr_o.alloc((size_t)n * 16));
    AUR_TRY(s->tr_d.alloc((size_t)n * 16));
    AUR_TRY(s->tr_hit.alloc((size_t)n * 8));
    const uint32_t blocks = (n + 255) / 256;
    k_pack3<<<blocks, 256, 0, c->stream>>>(o, s->tr_o.as<float4>(), n);
    k_pack3<<<blocks, 256, 0, c->stream>>>(d, s->tr_d.as<float4>(), n);
    AUR_HIP(hipGetLastError());
    if (brute) {
        AUR_TRY(launch_brute(s->view, s->tr_o.as<float4>(), s->tr_d.as<float4>(), n, tmin, tmax,
                             s->tr_hit.as<float2>(), c->stream));
    } else {
        const int grid = grid_for_extend(c, n);
        AUR_TRY(ensure_spill(s->tr_spill, grid));
        // the query's ray offset travels in the one-entry scene table
        SceneRec rec{};
        rec.sc = s->view;
        rec.tmin = tmin;
        AUR_TRY(s->rec1.alloc(sizeof(SceneRec)));
        AUR_HIP(hipStreamSynchronize(c->stream));
        AUR_HIP(hipMemcpy(s->rec1.p, &rec, sizeof(SceneRec), hipMemcpyHostToDevice));
        AUR_TRY(s->tr_pool.alloc(kPools * kPoolStride * 4));
        AUR_HIP(hipMemsetAsync(s->tr_pool.p, 0, kPools * kPoolStride * 4, c->stream));
        ExtendArgs a{};
        a.ray_o = s->tr_o.as<float4>();
        a.ray_d = s->tr_d.as<float4>();
        a.count = nullptr;
        a.n = n;
        a.pool = s->tr_pool.as<uint32_t>();
        a.tmax = tmax;
        a.hit = s->tr_hit.as<float2>();
        a.counts = counts;
        if (count_fetches) {
            AUR_TRY(s->totals.alloc(64));
            AUR_HIP(hipMemsetAsync(s->totals.p, 0, 24, c->stream));
            a.totals = s->totals.as<unsigned long long>();
        }
        a.spill = s->tr_spill.as<uint32_t>();
        a.spill_stride = (uint32_t)(grid * extend_block());
        a.recs = s->rec1.as<SceneRec>();
        a.one_scene = 1;
        a.wide = 1;  // ray queries: memory bound, the four-wide tree wins
        // Large queries on large trees: claim rays in (origin Morton code,
        // octant) order, so the rays in flight on one XCD share the nodes
        // its L2 holds.  The hit of each ray is unchanged.
        const bool sort = s->stats.n_nodes > 0 &&
                          (c->ray_sort == 2 ||
                           (c->ray_sort == 1 && n >= kSortMinRays && s->stats.n_nodes >= kSortMinNodes));
        if (sort) {
            for (int k = 0; k < 2; ++k) {
                AUR_TRY(s->tr_key[k].alloc((size_t)n * 4));
                AUR_TRY(s->tr_val[k].alloc((size_t)n * 4));
            }
            AUR_TRY(s->tr_hist.alloc(radix_hist_bytes(n)));
            a.order = s->tr_val[0].as<uint32_t>();
        }
        EventPair ev;
        if (ts) {
            AUR_TRY(ev.create());
            AUR_HIP(hipEventRecord(ev.e[0], c->stream));
        }
        hipEvent_t e0 = ev.e[0], e1 = ev.e[1];
        Status ls;
        if (sort) {
            ls = launch_ray_keys(a.ray_o, a.ray_d, n, s->view.nodes, s->tr_key[0].as<uint32_t>(),
                                 s->tr_val[0].as<uint32_t>(), c->stream);
            if (!ls.bad())
                ls = radix_sort_pairs(s->tr_key[0].as<uint32_t>(), s->tr_val[0].as<uint32_t>(),
                                      s->tr_key[1].as<uint32_t>(), s->tr_val[1].as<uint32_t>(),
                                      s->tr_hist.as<uint32_t>(), n, 4, c->stream);
        }
        if (!ls.bad()) ls = launch_extend(a, grid, c->stream);
        if (ts) {
            Status es = hip_status(hipEventRecord(e1, c->stream), "hipEventRecord");
            if (!es.bad()) es = hip_status(hipEventSynchronize(e1), "hipEventSynchronize");
            float ms = 0.0f;
            if (!es.bad()) es = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
            AUR_TRY(ls);
            AUR_TRY(es);
            *ts = aurora_trace_stats{};
            ts->extend_ms = ms;
            if (count_fetches) {
                unsigned long long tot[3];
                AUR_HIP(hipMemcpy(tot, s->totals.p, 24, hipMemcpyDeviceToHost));
                ts->nodes_fetched = tot[0];
                ts->spheres_fetched = tot[1];
                ts->quads_fetched = tot[2];
            }
        }
        AUR_TRY(ls);
    }
    k_unpack_hits<<<blocks, 256, 0, c->stream>>>(s->tr_hit.as<float2>(), prim, t, n);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status ensure_workspace(aurora_ctx* c, uint32_t cap, int grid) {
    Workspace& w = c->ws;
    if (cap > w.capacity) {
        for (int k = 0; k < 2; ++k) {
            AUR_TRY(w.q_o[k].alloc((size_t)cap * 16));
            AUR_TRY(w.q_d[k].alloc((size_t)cap * 16));
            AUR_TRY(w.q_b[k].alloc((size_t)cap * 16));
        }
        AUR_TRY(w.hit.alloc((size_t)cap * 8));
        AUR_TRY(w.L.alloc((size_t)cap * 16));
        AUR_TRY(w.aov_a.alloc((size_t)cap * 12));
        AUR_TRY(w.aov_n.alloc((size_t)cap * 12));
        w.capacity = cap;
    }
    AUR_TRY(w.counters.alloc(kCounterBytes));
    AUR_TRY(ensure_spill(w.spill, grid));
    if (!w.host_counts) AUR_HIP(hipHostMalloc((void**)&w.host_counts, 256 * sizeof(uint32_t), hipHostMallocDefault));
    return Status::ok();
}

// Paths per pass when the caller leaves spp_per_pass at 0.  A pass holds
// 144 bytes of queues per path (two ping-pong queues, hit, radiance and the
// two AOV planes), so 2^27 paths take 19 GB of the 288 GB of HBM: few, long
// passes amortise each launch's ramp and drain.
constexpr uint64_t kPassPaths = 1ull << 27;

// One image of a render call.
struct Job {
    aurora_scene* scene;
    CamF cam;
    aurora_render_desc rd;
    float *rgb, *alb, *nrm;
    uint32_t S;     // samples per pass
    uint32_t done;  // samples rendered so far
};

Status validate_job(aurora_scene* s, const aurora_camera_desc* cam, const aurora_render_desc* rd, float* rgb,
                    float* alb, float* nrm, Job& j) {
    if (!s) return Status{AURORA_ERR_INVALID, "scene is NULL"};
    if (!cam || !rd) return Status{AURORA_ERR_INVALID, "null camera or render descriptor"};
    if (!rgb) return Status{AURORA_ERR_INVALID, "rgb_sum is NULL"};
    const std::string e = camera_frame(*cam, j.cam);
    if (!e.empty()) return Status{AURORA_ERR_INVALID, e};
    if (rd->max_depth == 0) return Status{AURORA_ERR_INVALID, "max_depth must be >= 1"};
    if (rd->max_depth > 250) return Status{AURORA_ERR_INVALID, "max_depth must be <= 250"};
    if (!finite3(rd->background)) return Status{AURORA_ERR_INVALID, "background must be finite"};
    if ((rd->flags & AURORA_RENDER_AOV) && (!alb || !nrm))
        return Status{AURORA_ERR_INVALID, "AOV flag set but albedo/normal buffers are NULL"};
    if (!s->built) return Status{AURORA_ERR_STATE, "BVH not built: call aurora_bvh_build first"};
    if ((uint64_t)rd->sample_offset + rd->spp > 0xffffffffull) return Status{AURORA_ERR_INVALID, "sample index overflow"};
    j.scene = s;
    j.rd = *rd;
    j.rgb = rgb;
    j.alb = alb;
    j.nrm = nrm;
    j.done = 0;
    return Status::ok();
}

// Render every job's samples in wavefront passes that share launches.
Status render_jobs(aurora_ctx* c, std::vector<Job>& jobs, aurora_render_stats* stats, bool count_fetches,
                   bool graph) {
    aurora_render_stats st{};
    // samples per pass: explicit (spp_per_pass), or automatic: the fewest
    // passes P that keep a pass within kPassPaths paths next to the explicit
    // jobs, each automatic job rendering ceil(spp / P) samples per pass, so
    // all of them finish in the same pass
    uint64_t fixed = 0;
    uint32_t max_spp = 1;
    for (const Job& j : jobs) {
        const uint64_t wh = (uint64_t)j.cam.W * j.cam.H;
        if (j.rd.spp_per_pass) fixed += std::min<uint64_t>(j.rd.spp_per_pass, std::max<uint32_t>(1, j.rd.spp)) * wh;
        else max_spp = std::max(max_spp, j.rd.spp);
    }
    const uint64_t pass_cap = c->pass_paths ? c->pass_paths : kPassPaths;
    const uint64_t room = pass_cap > fixed ? pass_cap - fixed : 0;
    auto auto_paths = [&](uint64_t P) {
        uint64_t n = 0;
        for (const Job& j : jobs)
            if (!j.rd.spp_per_pass)
                n += std::max<uint64_t>(1, (j.rd.spp + P - 1) / P) * ((uint64_t)j.cam.W * j.cam.H);
        return n;
    };
    uint64_t P_lo = 1, P_hi = max_spp;  // auto_paths is non-increasing in P; at max_spp every S is 1
    while (P_lo < P_hi) {
        const uint64_t mid = (P_lo + P_hi) / 2;
        if (auto_paths(mid) <= room) P_hi = mid;
        else P_lo = mid + 1;
    }
    for (Job& j : jobs) {
        const uint64_t wh = (uint64_t)j.cam.W * j.cam.H;
        uint64_t S = j.rd.spp_per_pass ? j.rd.spp_per_pass : (j.rd.spp + P_lo - 1) / P_lo;
        S = std::max<uint64_t>(1, std::min<uint64_t>(S, std::max<uint32_t>(1, j.rd.spp)));
        if (S * wh > 0xffffffffull) return Status{AURORA_ERR_INVALID, "spp_per_pass * pixels exceeds 2^32"};
        j.S = (uint32_t)S;
    }
    uint64_t cap = 0;
    for (const Job& j : jobs) cap += (uint64_t)j.S * j.cam.W * j.cam.H;
    if (cap > 0x7fffffffull) return Status{AURORA_ERR_INVALID, "a pass would exceed 2^31 paths"};
    uint32_t max_depth = 1;
    bool any_work = false;
    for (const Job& j : jobs) {
        max_depth = std::max(max_depth, j.rd.max_depth);
        any_work |= j.rd.spp > 0;
    }
    if (!any_work) {
        if (stats) *stats = st;
        return Status::ok();
    }
    const int ext_grid = c->ext_grid;
    // the fused tail is persistent: its grid is the resident capacity
    // the passes walk the flat leaf array under the condition render_jobs
    // sets pa.flat by (one scene, small enough to have the array)
    const bool flat = c->render_tree == 2 && jobs.size() == 1 && jobs[0].scene->view.n_flat > 0;
    const bool wide_one = !flat && c->render_tree != 0 && jobs.size() == 1;
    const int tail_grid =
        c->cus * (flat ? kTailFlatBlocksPerCu : wide_one ? kTailWideOneBlocksPerCu : kTailBlocksPerCu);
    const int graph_tail_grid = tail_grid;
    AUR_TRY(ensure_workspace(c, (uint32_t)cap, std::max(ext_grid, graph ? graph_tail_grid : tail_grid)));
    Workspace& w = c->ws;
    AUR_TRY(w.recs.alloc(jobs.size() * sizeof(SceneRec)));
    if (c->host_recs_cap < jobs.size()) {
        for (int k = 0; k < 2; ++k) {
            AUR_HIP(hipEventSynchronize(c->recs_ev[k]));  // no copy still reads the old table
            if (c->host_recs[k]) (void)hipHostFree(c->host_recs[k]);
            c->host_recs[k] = nullptr;
        }
        c->host_recs_cap = 0;
        for (int k = 0; k < 2; ++k)
            AUR_HIP(hipHostMalloc((void**)&c->host_recs[k], jobs.size() * sizeof(SceneRec), hipHostMallocDefault));
        c->host_recs_cap = (uint32_t)jobs.size();
    }
    if (count_fetches) {
        AUR_TRY(jobs[0].scene->totals.alloc(64));
        AUR_HIP(hipMemsetAsync(jobs[0].scene->totals.p, 0, 24, c->stream));
    }
    const bool timed = stats != nullptr;
    std::vector<hipEvent_t> evs;
    auto ev = [&](hipEvent_t& e) -> Status {
        AUR_HIP(hipEventCreate(&e));
        evs.push_back(e);
        return Status::ok();
    };
    struct Cleanup {
        std::vector<hipEvent_t>& e;
        ~Cleanup() {
            for (auto& x : e) (void)hipEventDestroy(x);
        }
    } cleanup{evs};
    hipEvent_t t_begin = nullptr, t_end = nullptr;
    if (timed) {
        AUR_TRY(ev(t_begin));
        AUR_TRY(ev(t_end));
        AUR_HIP(hipEventRecord(t_begin, c->stream));
    }
    std::vector<hipEvent_t> ext_ev, tail_ev;
    std::vector<hipEvent_t> cnt_ev(max_depth);
    // counting runs trace every segment in k_extend, whose counters they read
    const uint32_t tail_paths = count_fetches ? 0u : c->tail_paths;
    for (auto& e2 : cnt_ev) AUR_TRY(ev(e2));
    uint64_t total_segments = 0;

    // build the scene table of the next pass; returns its shape
    // the table staged for the next pass, and its copy to the device
    SceneRec* hrecs = nullptr;
    auto stage_recs = [&]() -> Status {
        AUR_HIP(hipMemcpyAsync(w.recs.p, hrecs, jobs.size() * sizeof(SceneRec), hipMemcpyHostToDevice, c->stream));
        AUR_HIP(hipEventRecord(c->recs_ev[c->recs_turn], c->stream));
        c->recs_turn ^= 1;
        return Status::ok();
    };
    auto make_pass = [&](Pass& pa, bool& same_shape, uint32_t& prev_sig, bool& any) -> Status {
        uint32_t path_off = 0, pix_off = 0, sig = 0;
        any = false;
        AUR_HIP(hipEventSynchronize(c->recs_ev[c->recs_turn]));
        hrecs = c->host_recs[c->recs_turn];
        for (size_t k = 0; k < jobs.size(); ++k) {
            Job& j = jobs[k];
            const uint32_t sp = j.done < j.rd.spp ? std::min(j.S, j.rd.spp - j.done) : 0u;
            SceneRec& r = hrecs[k];
            r.sc = j.scene->view;
            r.cam = j.cam;
            r.seed = j.rd.seed;
            r.sample_base = j.rd.sample_offset + j.done;
            r.spp_pass = sp;
            r.max_depth = j.rd.max_depth;
            r.tmin = j.rd.tmin > 0.0f ? j.rd.tmin : 0.001f;
            for (int q = 0; q < 3; ++q) r.bg[q] = j.rd.background[q];
            r.aov = (j.rd.flags & AURORA_RENDER_AOV) ? 1u : 0u;
            r.WH = j.cam.W * j.cam.H;
            r.path_off = path_off;
            r.pix_off = pix_off;
            r.rgb = j.rgb;
            r.alb = j.alb;
            r.nrm = j.nrm;
            path_off += r.WH * sp;
            if (sp) pix_off += r.WH;
            any |= sp > 0;
            sig = sig * 1000003u + sp;
        }
        pa.recs = w.recs.as<SceneRec>();
        pa.n_recs = (uint32_t)jobs.size();
        // the flat form needs one scene with a flat array; otherwise, and
        // in the wavefront traversal kernels, the four-wide tree stands in
        pa.wide = c->render_tree != 0 ? 1u : 0u;
        pa.flat = flat ? 1u : 0u;
        pa.claim = c->claim_paths;
        pa.n_paths = path_off;
        pa.n_pixels = pix_off;
        // the fused tail carries the pass from the first bounce (the pass
        // fits the tail switch): the raygen queue needs no throughput, and
        // the tail writes every path's radiance
        pa.lean_queue = path_off <= tail_paths ? 1u : 0u;
        same_shape = sig == prev_sig;
        prev_sig = sig;
        return Status::ok();
    };
    auto finish_pass = [&](const Pass& pa) {
        for (Job& j : jobs) j.done += j.done < j.rd.spp ? std::min(j.S, j.rd.spp - j.done) : 0u;
        st.passes++;
        st.paths += pa.n_paths;
        st.max_live = std::max(st.max_live, pa.n_paths);
    };
    // the kernels add the segments they trace to one device counter over
    // the whole call; it is read once, at the end, and only for statistics
    AUR_HIP(hipMemsetAsync(w.counters.as<uint32_t>() + kTailSegs, 0, 8, c->stream));
    auto extend_args = [&](int cur) {
        ExtendArgs a{};
        a.ray_o = w.q_o[cur].as<float4>();
        a.ray_d = w.q_d[cur].as<float4>();
        a.count = w.counters.as<uint32_t>() + cur;
        a.n = 0;
        a.pool = w.counters.as<uint32_t>() + kPoolBase;
        a.zero_after = w.counters.as<uint32_t>() + (cur ^ 1);
        a.tmax = INFINITY;
        a.hit = w.hit.as<float2>();
        a.recs = w.recs.as<SceneRec>();
        a.one_scene = jobs.size() == 1;
        a.wide = c->render_tree != 0 ? 1u : 0u;  // the counting runs trace as the timed ones do
        a.totals = count_fetches ? jobs[0].scene->totals.as<unsigned long long>() : nullptr;
        a.spill = w.spill.as<uint32_t>();
        a.spill_stride = (uint32_t)(std::max(ext_grid, graph ? graph_tail_grid : tail_grid) * extend_block());
        return a;
    };

    uint32_t prev_sig = 0xffffffffu;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t g = nullptr;
    for (;;) {
        Pass pa{};
        bool same = false;
        bool any = false;
        AUR_TRY(make_pass(pa, same, prev_sig, any));
        if (!any) break;
        if (graph && same && gexec) {
            // same pass shape as the captured one: replay (the graph's last
            // kernel advanced the device table's sample bases)
            AUR_HIP(hipGraphLaunch(gexec, c->stream));
            finish_pass(pa);
            continue;
        }
        AUR_TRY(stage_recs());
        if (graph) {
            // capture: fixed launch shapes, no host round trips inside
            if (gexec) {
                (void)hipGraphExecDestroy(gexec);
                gexec = nullptr;
            }
            if (g) {
                (void)hipGraphDestroy(g);
                g = nullptr;
            }
            AUR_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            Status cs = launch_raygen(pa, w, c->stream);
            // captured passes run a fixed number of wavefront bounces, then
            // the fused tail; none when the whole pass fits the tail switch
            const uint32_t dt = pa.n_paths <= tail_paths ? 0u : std::min(max_depth, c->graph_tail_depth);
            for (uint32_t depth = 0; depth < dt && !cs.bad(); ++depth) {
                const int cur = depth & 1;
                ExtendArgs a = extend_args(cur);
                a.spill_stride = (uint32_t)(std::max(ext_grid, graph_tail_grid) * extend_block());
                cs = launch_extend(a, ext_grid, c->stream);
                if (!cs.bad()) cs = launch_shade(pa, w, cur, depth, c->shade_grid, c->stream);
            }
            if (!cs.bad() && dt < max_depth) cs = launch_tail(pa, w, dt & 1, dt, graph_tail_grid, c->stream);
            if (!cs.bad()) cs = launch_accumulate(pa, w, c->stream);
            if (!cs.bad()) cs = launch_advance(w.recs.as<SceneRec>(), pa.n_recs, c->stream);
            hipError_t ce = hipStreamEndCapture(c->stream, &g);
            AUR_TRY(cs);
            AUR_HIP(ce);
            AUR_HIP(hipGraphInstantiate(&gexec, g, nullptr, nullptr, 0));
            AUR_HIP(hipGraphLaunch(gexec, c->stream));
            finish_pass(pa);
            continue;
        }
        AUR_TRY(launch_raygen(pa, w, c->stream));
        for (uint32_t depth = 0; depth < max_depth; ++depth) {
            const int cur = depth & 1;
            // queues only shrink, so the count two bounces back bounds this
            // bounce's queue: size both launches from it
            uint32_t ub = pa.n_paths;
            if (depth >= 2) {
                AUR_HIP(hipEventSynchronize(cnt_ev[depth - 2]));
                ub = w.host_counts[depth - 2];
                if (ub == 0) break;
            }
            // one item per thread while the queue is shorter than the chip
            const uint64_t blocks = std::max<uint64_t>(1, ((uint64_t)ub + 255) / 256);
            if (ub <= tail_paths) {
                const int tg = (int)std::min<uint64_t>((uint64_t)tail_grid, blocks);
                hipEvent_t e0 = nullptr, e1 = nullptr;
                if (timed) {
                    AUR_TRY(ev(e0));
                    AUR_TRY(ev(e1));
                    AUR_HIP(hipEventRecord(e0, c->stream));
                }
                AUR_TRY(launch_tail(pa, w, cur, depth, tg, c->stream));
                if (timed) {
                    AUR_HIP(hipEventRecord(e1, c->stream));
                    tail_ev.push_back(e0);
                    tail_ev.push_back(e1);
                }
                break;
            }
            const int eg = (int)std::min<uint64_t>((uint64_t)ext_grid, blocks);
            const int sg = (int)std::min<uint64_t>((uint64_t)c->shade_grid, blocks);
            ExtendArgs a = extend_args(cur);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (timed) {
                AUR_TRY(ev(e0));
                AUR_TRY(ev(e1));
                AUR_HIP(hipEventRecord(e0, c->stream));
            }
            AUR_TRY(launch_extend(a, eg, c->stream));
            if (timed) {
                AUR_HIP(hipEventRecord(e1, c->stream));
                ext_ev.push_back(e0);
                ext_ev.push_back(e1);
            }
            AUR_TRY(launch_shade(pa, w, cur, depth, sg, c->stream));
            AUR_HIP(hipMemcpyAsync(w.host_counts + depth, w.counters.as<uint32_t>() + (cur ^ 1), 4,
                                   hipMemcpyDeviceToHost, c->stream));
            AUR_HIP(hipEventRecord(cnt_ev[depth], c->stream));
        }
        AUR_TRY(launch_accumulate(pa, w, c->stream));
        finish_pass(pa);
    }
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (g) (void)hipGraphDestroy(g);
    if (stats) {
        unsigned long long v = 0;
        AUR_HIP(hipMemcpyAsync(&v, w.counters.as<uint32_t>() + kTailSegs, 8, hipMemcpyDeviceToHost, c->stream));
        AUR_HIP(hipStreamSynchronize(c->stream));
        total_segments = v;
    }
    st.segments = total_segments;
    if (timed) {
        AUR_HIP(hipEventRecord(t_end, c->stream));
        AUR_HIP(hipEventSynchronize(t_end));
        float ms = 0.0f;
        AUR_HIP(hipEventElapsedTime(&ms, t_begin, t_end));
        st.kernel_ms = ms;
        double ext = 0.0;
        for (size_t k = 0; k + 1 < ext_ev.size(); k += 2) {
            float m2 = 0.0f;
            AUR_HIP(hipEventElapsedTime(&m2, ext_ev[k], ext_ev[k + 1]));
            ext += m2;
        }
        st.extend_ms = ext;
        st.extend_launches = (uint32_t)(ext_ev.size() / 2);
        double tail = 0.0;
        for (size_t k = 0; k + 1 < tail_ev.size(); k += 2) {
            float m2 = 0.0f;
            AUR_HIP(hipEventElapsedTime(&m2, tail_ev[k], tail_ev[k + 1]));
            tail += m2;
        }
        st.tail_ms = tail;
        st.tail_launches = (uint32_t)(tail_ev.size() / 2);
    }
    if (count_fetches) {
        unsigned long long tot[3];
        AUR_HIP(hipMemcpyAsync(tot, jobs[0].scene->totals.p, 24, hipMemcpyDeviceToHost, c->stream));
        AUR_HIP(hipStreamSynchronize(c->stream));
        st.nodes_fetched = tot[0];
        st.spheres_fetched = tot[1];
        st.quads_fetched = tot[2];
    }
    if (stats) *stats = st;
    return Status::ok();
}

Status render_device(aurora_scene* s, const aurora_camera_desc* cam, const aurora_render_desc* rd, float* rgb,
                     float* alb, float* nrm, aurora_render_stats* stats, bool count_fetches) {
    std::vector<Job> jobs(1);
    AUR_TRY(validate_job(s, cam, rd, rgb, alb, nrm, jobs[0]));
    return render_jobs(s->ctx, jobs, stats, count_fetches, (rd->flags & AURORA_RENDER_GRAPH) != 0);
}

Status check_paste(const float* src, uint32_t sw, uint32_t sh, const float* dst, uint32_t dw, uint32_t dh,
                   const int32_t* corners) {
    if (!dst || !corners) return Status{AURORA_ERR_INVALID, "NULL destination or corners"};
    if (dw == 0 || dh == 0) return Status{AURORA_ERR_INVALID, "destination width and height must be positive"};
    if ((uint64_t)dw * dh > (1ull << 30) || (uint64_t)sw * sh > (1ull << 30))
        return Status{AURORA_ERR_INVALID, "image too large"};
    if (sw && sh && !src) return Status{AURORA_ERR_INVALID, "NULL source"};
    if (sw > 0x7fffffffu || sh > 0x7fffffffu || dw > 0x7fffffffu || dh > 0x7fffffffu)
        return Status{AURORA_ERR_INVALID, "image size beyond int range"};
    return Status::ok();
}

}  // namespace

// accessors for the other translation units of the C ABI (denoise.hip)
hipStream_t ctx_stream(aurora_ctx* c) { return c->stream; }
int ctx_device(aurora_ctx* c) { return c->device; }
int ctx_fail(aurora_ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }

extern "C" {

int aurora_abi_version(void) { return AURORA_ABI_VERSION; }

int aurora_create(int device, aurora_ctx** out) {
    if (!out) return fail(nullptr, AURORA_ERR_INVALID, "out is NULL");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0)
        return fail(nullptr, AURORA_ERR_RUNTIME, std::string("no HIP device: ") + hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(nullptr, AURORA_ERR_INVALID, "device ordinal out of range");
    auto* c = new aurora_ctx();
    c->device = device;
    Status s = hip_status(hipSetDevice(device), "hipSetDevice");
    hipDeviceProp_t prop{};
    if (!s.bad()) s = hip_status(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    // a blocking stream: ordered with work on the legacy default stream
    // (where callers such as PyTorch place their copies and fills)
    if (!s.bad()) s = hip_status(hipStreamCreate(&c->stream), "hipStreamCreate");
    for (int k = 0; k < 2 && !s.bad(); ++k)
        s = hip_status(hipEventCreateWithFlags(&c->recs_ev[k], hipEventDisableTiming), "hipEventCreate");
    if (s.bad()) {
        for (int k = 0; k < 2; ++k)
            if (c->recs_ev[k]) (void)hipEventDestroy(c->recs_ev[k]);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, s);
    }
    if (!s.bad()) s = scratch_address(device, &c->scratch);
    if (s.bad()) {
        for (int k = 0; k < 2; ++k)
            if (c->recs_ev[k]) (void)hipEventDestroy(c->recs_ev[k]);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, s);
    }
    c->cus = prop.multiProcessorCount;
    c->ext_grid = c->cus * 8;
    c->shade_grid = c->cus * 8;
    *out = c;
    return AURORA_OK;
}

void aurora_destroy(aurora_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    Workspace& w = c->ws;
    DevBuf* bufs[] = {&w.q_o[0], &w.q_o[1], &w.q_d[0], &w.q_d[1], &w.q_b[0], &w.q_b[1], &w.hit, &w.L,
                      &w.aov_a, &w.aov_n, &w.counters, &w.spill, &w.recs};
    for (DevBuf* b : bufs) b->release();
    c->patch.release();
    if (w.host_counts) (void)hipHostFree(w.host_counts);
    for (int k = 0; k < 2; ++k) {
        if (c->host_recs[k]) (void)hipHostFree(c->host_recs[k]);
        if (c->recs_ev[k]) (void)hipEventDestroy(c->recs_ev[k]);
    }
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* aurora_last_error(const aurora_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int aurora_synchronize(aurora_ctx* c) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    CK(c, hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return AURORA_OK;
}

void* aurora_stream(aurora_ctx* c) { return c ? (void*)c->stream : nullptr; }

int aurora_set_launch(aurora_ctx* c, int ext_blocks_per_cu, int shade_blocks_per_cu, uint32_t tail_paths) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (ext_blocks_per_cu < 1 || ext_blocks_per_cu > 64 || shade_blocks_per_cu < 1 || shade_blocks_per_cu > 64)
        return fail(c, AURORA_ERR_INVALID, "blocks per CU must be in [1, 64]");
    c->ext_grid = c->cus * ext_blocks_per_cu;
    c->shade_grid = c->cus * shade_blocks_per_cu;
    c->tail_paths = tail_paths;
    return AURORA_OK;
}

int aurora_set_queue(aurora_ctx* c, uint32_t pass_paths, uint32_t claim_paths) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (pass_paths && (pass_paths < (1u << 16) || pass_paths > (1u << 28)))
        return fail(c, AURORA_ERR_INVALID, "pass_paths must be 0 or in [2^16, 2^28]");
    if (claim_paths && (claim_paths < 16 || claim_paths > 1024))
        return fail(c, AURORA_ERR_INVALID, "claim_paths must be 0 or in [16, 1024]");
    c->pass_paths = pass_paths;
    c->claim_paths = claim_paths;
    return AURORA_OK;
}

int aurora_set_render_tree(aurora_ctx* c, int tree) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (tree < 0 || tree > 2)
        return fail(c, AURORA_ERR_INVALID, "render tree must be 0 (binary), 1 (four-wide) or 2 (flat)");
    c->render_tree = (uint32_t)tree;
    return AURORA_OK;
}

int aurora_set_ray_sort(aurora_ctx* c, int mode) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (mode < 0 || mode > 2) return fail(c, AURORA_ERR_INVALID, "ray sort mode must be 0, 1 or 2");
    c->ray_sort = mode;
    return AURORA_OK;
}

int aurora_scene_upload(aurora_ctx* c, const aurora_scene_desc* d, aurora_scene** out) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!d || !out) return fail(c, AURORA_ERR_INVALID, "desc or out is NULL");
    *out = nullptr;
    const std::string e = check_scene(*d);
    if (!e.empty()) return fail(c, AURORA_ERR_INVALID, e);
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    auto* s = new aurora_scene();
    s->ctx = c;
    auto cleanup = [&](const Status& st) {
        aurora_scene_destroy(s);
        return fail(c, st);
    };
    const hipStream_t st = c->stream;
    // quads: (Q, material) (u, kind) (v, 0)
    std::vector<float4> q((size_t)d->n_quads * 3);
    for (uint32_t i = 0; i < d->n_quads; ++i) {
        const float* p = d->quads + 9 * (size_t)i;
        uint32_t kind = d->quad_kind ? d->quad_kind[i] : AURORA_QUAD_PARALLELOGRAM;
        uint32_t mat = d->quad_material[i];
        float fm, fk;
        memcpy(&fm, &mat, 4);
        memcpy(&fk, &kind, 4);
        q[3 * (size_t)i] = make_float4(p[0], p[1], p[2], fm);
        q[3 * (size_t)i + 1] = make_float4(p[3], p[4], p[5], fk);
        q[3 * (size_t)i + 2] = make_float4(p[6], p[7], p[8], 0.0f);
    }
    std::vector<aurora_material> mats(d->materials, d->materials + d->n_materials);
    for (auto& m : mats)
        if (m.type == AURORA_MAT_METAL && m.fuzz > 1.0f) m.fuzz = 1.0f;
    std::vector<aurora_texture> texs(d->textures, d->textures + d->n_textures);
    uint32_t slots = 0;
    for (auto& t : texs)
        if (t.type == AURORA_TEX_NOISE) t.even = slots++;
    std::vector<float> pv((size_t)std::max(1u, slots) * 768);
    std::vector<int> pp((size_t)std::max(1u, slots) * 768);
    for (auto& t : texs)
        if (t.type == AURORA_TEX_NOISE) perlin_fill(t.seed, &pv[(size_t)t.even * 768], &pp[(size_t)t.even * 768]);
    Status r = upload(s->spheres, d->spheres, (size_t)d->n_spheres * 16, st);
    if (!r.bad()) r = upload(s->sphere_mat, d->sphere_material, (size_t)d->n_spheres * 4, st);
    if (!r.bad()) r = upload(s->quads, q.data(), q.size() * 16, st);
    if (!r.bad()) r = s->quad_n.alloc((size_t)std::max(1u, d->n_quads) * 16);
    if (!r.bad() && d->n_quads) {
        k_quad_normals<<<(d->n_quads + 255) / 256, 256, 0, st>>>(s->quads.as<float4>(), d->n_quads,
                                                                  s->quad_n.as<float4>());
        r = hip_status(hipGetLastError(), "k_quad_normals");
    }
    if (!r.bad()) r = upload(s->mats, mats.data(), mats.size() * sizeof(aurora_material), st);
    if (!r.bad()) r = upload(s->texs, texs.data(), texs.size() * sizeof(aurora_texture), st);
    if (!r.bad()) r = upload(s->texels, d->texels, (size_t)d->n_texels * 12, st);
    if (!r.bad()) r = upload(s->perlin_vec, pv.data(), pv.size() * 4, st);
    if (!r.bad()) r = upload(s->perlin_perm, pp.data(), pp.size() * 4, st);
    if (!r.bad()) r = hip_status(hipStreamSynchronize(st), "hipStreamSynchronize");
    if (r.bad()) return cleanup(r);
    DevScene& v = s->view;
    v.spheres = s->spheres.as<float4>();
    v.sphere_mat = s->sphere_mat.as<uint32_t>();
    v.quads = s->quads.as<float4>();
    v.quad_n = s->quad_n.as<float4>();
    v.mats = s->mats.as<aurora_material>();
    v.texs = s->texs.as<aurora_texture>();
    v.texels = s->texels.as<float>();
    v.perlin_vec = s->perlin_vec.as<float>();
    v.perlin_perm = s->perlin_perm.as<int>();
    v.nodes = nullptr;
    v.n_spheres = d->n_spheres;
    v.n_quads = d->n_quads;
    v.n_nodes = 0;
    s->n_prims = d->n_spheres + d->n_quads;
    *out = s;
    return AURORA_OK;
}

void aurora_scene_destroy(aurora_scene* s) {
    if (!s) return;
    if (s->ctx && s->ctx->stream) (void)hipStreamSynchronize(s->ctx->stream);
    DevBuf* bufs[] = {&s->spheres, &s->sphere_mat, &s->quads, &s->quad_n, &s->mats, &s->texs, &s->texels, &s->perlin_vec,
                      &s->perlin_perm, &s->nodes, &s->wide, &s->morton, &s->order, &s->flat, &s->rec1, &s->tr_o, &s->tr_d,
                      &s->tr_hit, &s->tr_spill, &s->tr_pool, &s->totals, &s->tr_key[0], &s->tr_key[1],
                      &s->tr_val[0], &s->tr_val[1], &s->tr_hist};
    for (DevBuf* b : bufs) b->release();
    s->scr.release();
    delete s;
}

int aurora_bvh_build(aurora_scene* s, aurora_bvh_stats* stats) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    aurora_bvh_stats st{};
    CK(c, lbvh_build(s->view, s->n_prims, s->nodes, s->wide, s->morton, s->order, s->scr, c->stream, &st));
    s->view.nodes = s->nodes.as<float4>();
    s->view.n_nodes = st.n_nodes;
    s->view.wide = s->wide.as<float4>();
    s->view.n_wide = st.n_nodes ? st.n_wide : 0u;
    CK(c, flat_build(s->n_prims, s->nodes, s->scr, s->flat, c->stream));
    s->view.flat = s->flat.as<float4>();
    s->view.n_flat = s->n_prims <= kFlatMaxPrims ? s->n_prims : 0u;
    s->stats = st;
    s->built = true;
    if (stats) *stats = st;
    return AURORA_OK;
}

int aurora_bvh_download(aurora_scene* s, aurora_bvh_node* nodes, uint32_t capacity) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (!s->built) return fail(c, AURORA_ERR_STATE, "BVH not built");
    if (capacity < s->stats.n_nodes) return fail(c, AURORA_ERR_INVALID, "capacity smaller than the node count");
    if (s->stats.n_nodes == 0) return AURORA_OK;
    if (!nodes) return fail(c, AURORA_ERR_INVALID, "nodes is NULL");
    CK(c, hip_status(hipMemcpyAsync(nodes, s->nodes.p, (size_t)s->stats.n_nodes * 32, hipMemcpyDeviceToHost,
                                    c->stream),
                     "hipMemcpyAsync"));
    CK(c, hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return AURORA_OK;
}

int aurora_bvh_download_wide(aurora_scene* s, aurora_bvh4_node* nodes, uint32_t capacity) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (!s->built) return fail(c, AURORA_ERR_STATE, "BVH not built");
    const uint32_t n = s->view.n_wide;
    if (capacity < n) return fail(c, AURORA_ERR_INVALID, "capacity smaller than the wide node count");
    if (n == 0) return AURORA_OK;
    if (!nodes) return fail(c, AURORA_ERR_INVALID, "nodes is NULL");
    CK(c, hip_status(hipMemcpyAsync(nodes, s->wide.p, (size_t)n * 128, hipMemcpyDeviceToHost, c->stream),
                     "hipMemcpyAsync"));
    CK(c, hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return AURORA_OK;
}

int aurora_bvh_download_order(aurora_scene* s, uint32_t* morton, uint32_t* order, uint32_t capacity) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (!s->built) return fail(c, AURORA_ERR_STATE, "BVH not built");
    if (capacity < s->n_prims) return fail(c, AURORA_ERR_INVALID, "capacity smaller than the primitive count");
    if (s->n_prims == 0) return AURORA_OK;
    if (!morton || !order) return fail(c, AURORA_ERR_INVALID, "output is NULL");
    CK(c, hip_status(hipMemcpyAsync(morton, s->morton.p, (size_t)s->n_prims * 4, hipMemcpyDeviceToHost, c->stream),
                     "hipMemcpyAsync"));
    CK(c, hip_status(hipMemcpyAsync(order, s->order.p, (size_t)s->n_prims * 4, hipMemcpyDeviceToHost, c->stream),
                     "hipMemcpyAsync"));
    CK(c, hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return AURORA_OK;
}

int aurora_trace_closest_device(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                                int32_t* prim, float* t, uint32_t* counts) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    if (n && (!o || !d || !prim || !t)) return fail(s->ctx, AURORA_ERR_INVALID, "NULL ray or output buffer");
    CK(s->ctx, hip_status(hipSetDevice(s->ctx->device), "hipSetDevice"));
    CK(s->ctx, trace_device(s, o, d, n, tmin, tmax, prim, t, counts, false));
    return AURORA_OK;
}

int aurora_trace_closest_device_stats(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin,
                                      float tmax, int32_t* prim, float* t, int count_fetches,
                                      aurora_trace_stats* stats) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    if (n && (!o || !d || !prim || !t)) return fail(s->ctx, AURORA_ERR_INVALID, "NULL ray or output buffer");
    aurora_trace_stats local{};
    CK(s->ctx, hip_status(hipSetDevice(s->ctx->device), "hipSetDevice"));
    CK(s->ctx, trace_device(s, o, d, n, tmin, tmax, prim, t, nullptr, false, &local, count_fetches != 0));
    CK(s->ctx, hip_status(hipStreamSynchronize(s->ctx->stream), "hipStreamSynchronize"));
    if (stats) *stats = local;
    return AURORA_OK;
}

int aurora_trace_closest_device_noalloc(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin,
                                        float tmax, int32_t* prim, float* t, uint32_t* counts) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (n > kNoAllocMaxRays)
        return fail(c, AURORA_ERR_INVALID,
                    "aurora_trace_closest_device_noalloc: " + std::to_string(n) +
                        " rays exceed the allocation-free maximum of " + std::to_string(kNoAllocMaxRays) +
                        " per call");
    if (!s->built) return fail(c, AURORA_ERR_STATE, "BVH not built: call aurora_bvh_build first");
    if (n == 0) return AURORA_OK;
    if (!o || !d || !prim || !t) return fail(c, AURORA_ERR_INVALID, "NULL buffer");
    if (!c->scratch) return fail(c, AURORA_ERR_STATE, "no scratch address for this context's device");
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    unsigned char* base = c->scratch;
    float4* ro = reinterpret_cast<float4*>(base);
    float4* rd = reinterpret_cast<float4*>(base + kNaOffDir);
    float2* hit = reinterpret_cast<float2*>(base + kNaOffHit);
    uint32_t* pool = reinterpret_cast<uint32_t*>(base + kNaOffPool);
    SceneRec* rec = reinterpret_cast<SceneRec*>(base + kNaOffRec);
    uint32_t* spill = reinterpret_cast<uint32_t*>(base + kNaOffSpill);
    const uint32_t blocks = (n + 255) / 256;
    k_pack3<<<blocks, 256, 0, c->stream>>>(o, ro, n);
    k_pack3<<<blocks, 256, 0, c->stream>>>(d, rd, n);
    SceneRec r{};
    r.sc = s->view;
    r.tmin = tmin;
    k_noalloc_setup<<<1, kPools * kPoolStride, 0, c->stream>>>(pool, rec, r);
    CK(c, hip_status(hipGetLastError(), "allocation-free query setup"));
    ExtendArgs a{};
    a.ray_o = ro;
    a.ray_d = rd;
    a.n = n;
    a.pool = pool;
    a.tmax = tmax;
    a.hit = hit;
    a.counts = counts;
    a.spill = spill;
    a.spill_stride = (uint32_t)kNoAllocGrid * kExtendBlock;
    a.recs = rec;
    a.one_scene = 1;
    a.wide = 1;
    CK(c, launch_extend(a, std::min(grid_for_extend(c, n), kNoAllocGrid), c->stream));
    k_unpack_hits<<<blocks, 256, 0, c->stream>>>(hit, prim, t, n);
    CK(c, hip_status(hipGetLastError(), "allocation-free query unpack"));
    return AURORA_OK;
}

int aurora_trace_brute_device(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                              int32_t* prim, float* t) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    if (n && (!o || !d || !prim || !t)) return fail(s->ctx, AURORA_ERR_INVALID, "NULL ray or output buffer");
    CK(s->ctx, hip_status(hipSetDevice(s->ctx->device), "hipSetDevice"));
    CK(s->ctx, trace_device(s, o, d, n, tmin, tmax, prim, t, nullptr, true));
    return AURORA_OK;
}

int aurora_trace_closest(aurora_scene* s, const float* o, const float* d, uint32_t n, float tmin, float tmax,
                         int32_t* prim, float* t) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (n && (!o || !d || !prim || !t)) return fail(c, AURORA_ERR_INVALID, "NULL ray or output buffer");
    if (!s->built) return fail(c, AURORA_ERR_STATE, "BVH not built: call aurora_bvh_build first");
    if (n == 0) return AURORA_OK;
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    DevBuf bo, bd, bp, bt;
    Status r = bo.alloc((size_t)n * 12);
    if (!r.bad()) r = bd.alloc((size_t)n * 12);
    if (!r.bad()) r = bp.alloc((size_t)n * 4);
    if (!r.bad()) r = bt.alloc((size_t)n * 4);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(bo.p, o, (size_t)n * 12, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipMemcpyAsync(bd.p, d, (size_t)n * 12, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad()) r = trace_device(s, bo.as<float>(), bd.as<float>(), n, tmin, tmax, bp.as<int32_t>(), bt.as<float>(),
                                   nullptr, false);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(prim, bp.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipMemcpyAsync(t, bt.p, (size_t)n * 4, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(c->stream);
    bo.release();
    bd.release();
    bp.release();
    bt.release();
    if (r.bad()) return fail(c, r);
    return AURORA_OK;
}

int aurora_render_device(aurora_scene* s, const aurora_camera_desc* cam, const aurora_render_desc* rd,
                         float* rgb, float* alb, float* nrm, aurora_render_stats* stats) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    CK(s->ctx, hip_status(hipSetDevice(s->ctx->device), "hipSetDevice"));
    const bool count = rd && (rd->flags & AURORA_RENDER_COUNT_FETCHES);
    CK(s->ctx, render_device(s, cam, rd, rgb, alb, nrm, stats, count));
    return AURORA_OK;
}

int aurora_render_batch_device(aurora_ctx* c, const aurora_batch_item* items, uint32_t n, uint32_t flags,
                               aurora_render_stats* stats) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (n == 0) {
        if (stats) *stats = aurora_render_stats{};
        return AURORA_OK;
    }
    if (!items) return fail(c, AURORA_ERR_INVALID, "items is NULL");
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    std::vector<Job> jobs(n);
    for (uint32_t k = 0; k < n; ++k) {
        const aurora_batch_item& it = items[k];
        if (it.scene && it.scene->ctx != c) return fail(c, AURORA_ERR_INVALID, "batch item scene belongs to another context");
        Status s = validate_job(it.scene, &it.camera, &it.render, it.d_rgb_sum, it.d_albedo_sum, it.d_normal_sum,
                                jobs[k]);
        if (s.bad()) return fail(c, s.code, "batch item " + std::to_string(k) + ": " + s.msg);
    }
    CK(c, render_jobs(c, jobs, stats, false, (flags & AURORA_RENDER_GRAPH) != 0));
    return AURORA_OK;
}

int aurora_render(aurora_scene* s, const aurora_camera_desc* cam, const aurora_render_desc* rd, float* rgb,
                  float* alb, float* nrm, aurora_render_stats* stats) {
    if (!s) return fail(nullptr, AURORA_ERR_INVALID, "scene is NULL");
    aurora_ctx* c = s->ctx;
    if (!cam || !rd || !rgb) return fail(c, AURORA_ERR_INVALID, "null camera, descriptor or rgb_sum");
    const bool aov = (rd->flags & AURORA_RENDER_AOV) != 0;
    if (aov && (!alb || !nrm)) return fail(c, AURORA_ERR_INVALID, "AOV flag set but albedo/normal buffers are NULL");
    if (cam->image_width == 0 || cam->image_height == 0)
        return fail(c, AURORA_ERR_INVALID, "image width and height must be positive");
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    const size_t bytes = (size_t)cam->image_width * cam->image_height * 12;
    DevBuf b_rgb, b_a, b_n;
    Status r = b_rgb.alloc(bytes);
    if (!r.bad() && aov) r = b_a.alloc(bytes);
    if (!r.bad() && aov) r = b_n.alloc(bytes);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(b_rgb.p, rgb, bytes, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad() && aov) r = hip_status(hipMemcpyAsync(b_a.p, alb, bytes, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad() && aov) r = hip_status(hipMemcpyAsync(b_n.p, nrm, bytes, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad())
        r = render_device(s, cam, rd, b_rgb.as<float>(), b_a.as<float>(), b_n.as<float>(), stats,
                          (rd->flags & AURORA_RENDER_COUNT_FETCHES) != 0);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(rgb, b_rgb.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad() && aov) r = hip_status(hipMemcpyAsync(alb, b_a.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad() && aov) r = hip_status(hipMemcpyAsync(nrm, b_n.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(c->stream);
    b_rgb.release();
    b_a.release();
    b_n.release();
    if (r.bad()) return fail(c, r);
    return AURORA_OK;
}

int aurora_camera_rays_device(aurora_ctx* c, const aurora_camera_desc* cam, uint32_t seed, const uint32_t* pixels,
                              const uint32_t* samples, uint32_t n, float* org, float* dir) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!cam) return fail(c, AURORA_ERR_INVALID, "camera is NULL");
    if (n && (!pixels || !samples || !org || !dir)) return fail(c, AURORA_ERR_INVALID, "NULL buffer");
    CamF f;
    const std::string e = camera_frame(*cam, f);
    if (!e.empty()) return fail(c, AURORA_ERR_INVALID, e);
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    CK(c, launch_camera_rays(f, seed, pixels, samples, n, org, dir, c->stream));
    return AURORA_OK;
}

int aurora_camera_rays(aurora_ctx* c, const aurora_camera_desc* cam, uint32_t seed, const uint32_t* pixels,
                       const uint32_t* samples, uint32_t n, float* org, float* dir) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!cam) return fail(c, AURORA_ERR_INVALID, "camera is NULL");
    if (n == 0) return AURORA_OK;
    if (!pixels || !samples || !org || !dir) return fail(c, AURORA_ERR_INVALID, "NULL buffer");
    const uint64_t wh = (uint64_t)cam->image_width * cam->image_height;
    for (uint32_t k = 0; k < n; ++k)
        if (pixels[k] >= wh) return fail(c, AURORA_ERR_INVALID, "pixel index beyond width*height");
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    DevBuf b_pix, b_smp, b_o, b_d;
    Status r = b_pix.alloc((size_t)n * 4);
    if (!r.bad()) r = b_smp.alloc((size_t)n * 4);
    if (!r.bad()) r = b_o.alloc((size_t)n * 12);
    if (!r.bad()) r = b_d.alloc((size_t)n * 12);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(b_pix.p, pixels, (size_t)n * 4, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipMemcpyAsync(b_smp.p, samples, (size_t)n * 4, hipMemcpyHostToDevice, c->stream), "copy");
    CamF f;
    if (!r.bad()) {
        const std::string e = camera_frame(*cam, f);
        if (!e.empty()) r = Status{AURORA_ERR_INVALID, e};
    }
    if (!r.bad()) r = launch_camera_rays(f, seed, b_pix.as<uint32_t>(), b_smp.as<uint32_t>(), n, b_o.as<float>(),
                                         b_d.as<float>(), c->stream);
    if (!r.bad()) r = hip_status(hipMemcpyAsync(org, b_o.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipMemcpyAsync(dir, b_d.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(c->stream);
    if (r.bad()) return fail(c, r);
    return AURORA_OK;
}


int aurora_paste_device(aurora_ctx* c, const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw,
                        uint32_t dh, const int32_t* corners) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    CK(c, check_paste(src, sw, sh, dst, dw, dh, corners));
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    bool painted = false;
    CK(c, paste_device(src, sw, sh, dst, dw, dh, corners, c->stream, &painted));
    return AURORA_OK;
}

int aurora_paste(aurora_ctx* c, const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                 const int32_t* corners) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    CK(c, check_paste(src, sw, sh, dst, dw, dh, corners));
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    const size_t sb = (size_t)sw * sh * 12, db = (size_t)dw * dh * 12;
    DevBuf b_src, b_dst;
    Status r = b_src.alloc(sb ? sb : 16);
    if (!r.bad()) r = b_dst.alloc(db);
    if (!r.bad() && sb) r = hip_status(hipMemcpyAsync(b_src.p, src, sb, hipMemcpyHostToDevice, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipMemcpyAsync(b_dst.p, dst, db, hipMemcpyHostToDevice, c->stream), "copy");
    bool painted = false;
    if (!r.bad()) r = paste_device(b_src.as<float>(), sw, sh, b_dst.as<float>(), dw, dh, corners, c->stream, &painted);
    if (!r.bad() && painted)
        r = hip_status(hipMemcpyAsync(dst, b_dst.p, db, hipMemcpyDeviceToHost, c->stream), "copy");
    if (!r.bad()) r = hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipStreamSynchronize(c->stream);
    if (r.bad()) return fail(c, r);
    return AURORA_OK;
}

int aurora_patch_render_device(aurora_ctx* c, const aurora_patch_triangle* tris, uint32_t n_tris,
                               const aurora_patch_material* mats, uint32_t n_mats, const aurora_patch_camera* cam,
                               const aurora_patch_config* cfg, double* d_rgb, aurora_patch_stats* stats) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!d_rgb) return fail(c, AURORA_ERR_INVALID, "patch render: output is NULL");
    CK(c, patch_check(tris, n_tris, mats, n_mats, cam, cfg));
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    CK(c, patch_render_device(c->patch, tris, n_tris, mats, cam, cfg, d_rgb, c->stream, stats));
    return AURORA_OK;
}

int aurora_patch_render(aurora_ctx* c, const aurora_patch_triangle* tris, uint32_t n_tris,
                        const aurora_patch_material* mats, uint32_t n_mats, const aurora_patch_camera* cam,
                        const aurora_patch_config* cfg, double* rgb, aurora_patch_stats* stats) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!rgb) return fail(c, AURORA_ERR_INVALID, "patch render: output is NULL");
    CK(c, patch_check(tris, n_tris, mats, n_mats, cam, cfg));
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    const size_t bytes = size_t(cam->width) * cam->height * 3 * sizeof(double);
    CK(c, c->patch.out.alloc(bytes));
    CK(c, patch_render_device(c->patch, tris, n_tris, mats, cam, cfg, c->patch.out.as<double>(), c->stream, stats));
    CK(c, hip_status(hipMemcpyAsync(rgb, c->patch.out.p, bytes, hipMemcpyDeviceToHost, c->stream), "copy"));
    CK(c, hip_status(hipStreamSynchronize(c->stream), "hipStreamSynchronize"));
    return AURORA_OK;
}

int aurora_resolve_rgb8(const float* rgb, uint32_t w, uint32_t h, uint32_t spp, uint32_t transfer, uint8_t* out) {
    if (!rgb || !out || spp == 0) return fail(nullptr, AURORA_ERR_INVALID, "NULL buffer or spp == 0");
    if (transfer > AURORA_TRANSFER_LINEAR) return fail(nullptr, AURORA_ERR_INVALID, "unknown transfer curve");
    const float fs = (float)spp;
    const size_t n = (size_t)w * h * 3;
    for (size_t i = 0; i < n; ++i) out[i] = resolve_byte(rgb[i], fs, transfer);
    return AURORA_OK;
}

int aurora_resolve_rgb8_device(aurora_ctx* c, const float* rgb, uint32_t w, uint32_t h, uint32_t spp,
                               uint32_t transfer, uint8_t* out) {
    if (!c) return fail(nullptr, AURORA_ERR_INVALID, "ctx is NULL");
    if (!rgb || !out || spp == 0) return fail(c, AURORA_ERR_INVALID, "NULL buffer or spp == 0");
    if (transfer > AURORA_TRANSFER_LINEAR) return fail(c, AURORA_ERR_INVALID, "unknown transfer curve");
    if (((uintptr_t)rgb & 15) || ((uintptr_t)out & 3))
        return fail(c, AURORA_ERR_INVALID, "rgb_sum must be 16-byte and out 4-byte aligned");
    CK(c, hip_status(hipSetDevice(c->device), "hipSetDevice"));
    CK(c, launch_resolve_rgb8(rgb, (size_t)w * h * 3, spp, transfer, out, c->stream));
    return AURORA_OK;
}

}  // extern "C"
