// render.hip — wavefront path tracing on gfx950: raygen, shade (+ queue
// compaction) and accumulation.  Extension (closest hit) is k_extend in
// trace.hip.
//
// A pass renders S samples of every pixel of one or more scenes (a batch):
// scene k owns path ids [path_off_k, path_off_k + W_k*H_k*S_k), sample-major
// (id = path_off + s*W*H + pixel), so the final accumulation reads each
// sample plane coalesced and adds samples in order, exactly as a sequential
// per-pixel loop would.  A path's scene index travels in its queue entry
// (direction.w), and every kernel reads the scene's record from a table.  Path state lives in HBM as structure-of-arrays
// float4 queues (origin+id, direction, throughput); the only radiance a path
// ever gains comes at its last segment (a light or the background), so the
// shade kernel writes it once to L[id] and carries no radiance in the queue.
//
// Reference counterparts: the per-pixel loop of experiments/rt.cpp:337-374
// (camera basis :339-343, pixel mapping :364-366) and the bounce loop of
// rt.cpp:251-334; reflect/refract from src/basic/vec3.cpp:182-199.
#include "internal.hpp"
#include "traverse.hpp"

#include <algorithm>

namespace aur {

namespace {

constexpr int kBlock = 256;

// The number of lanes below this one whose bit is set in mask (v_mbcnt:
// a ballot rank without a per-lane mask register).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// The scene record of a path.  A single-scene launch reads recs[0] through
// the constant address space: nothing writes the table while a render
// kernel runs, so its fields arrive by scalar loads and the scene's pointers
// and counts stay in scalar registers.
template <bool ONE>
__device__ __forceinline__ SceneRec scene_rec(const SceneRec* recs, uint32_t sid) {
    if constexpr (ONE) {
        static_assert(sizeof(SceneRec) % 4 == 0, "dword copy");
        typedef const __attribute__((address_space(4))) uint32_t* cu32p;
        SceneRec out;
        uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
        const cu32p src = (cu32p)recs;
#pragma unroll
        for (size_t i = 0; i < sizeof(SceneRec) / 4; ++i) dst[i] = src[i];
        return out;
    } else {
        return recs[sid];
    }
}

// last record whose offset (path_off or pix_off) is <= x
template <bool PIX>
__device__ __forceinline__ uint32_t find_rec(const SceneRec* recs, uint32_t n, uint32_t x) {
    uint32_t lo = 0, hi = n;  // invariant: answer in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        const uint32_t off = PIX ? recs[mid].pix_off : recs[mid].path_off;
        if (off <= x) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Camera ray of one sample through pixel (i, j) (pixel = j*W + i): the pixel
// centre p00 + i*du + j*dv jittered by U[-1/2, 1/2)^2 from Philox(pixel,
// sample, 0, 0), and, with depth of field, the origin moved to a point of the
// lens disk found by rejection among up to 32 draws (pixel, sample, 0, 1+k).
// The role of experiments/rt.cpp:364-366 (pixel -> direction), with the
// "Ray Tracing in One Weekend" viewport the host computes in camera_frame.
__device__ __forceinline__ void camera_ray(const CamF& cam, uint32_t pixel, uint32_t i, uint32_t j,
                                           uint32_t sample, uint32_t seed, v3& org, v3& dir) {
    const u4 q = philox(u4{pixel, sample, 0u, 0u}, seed, STREAM_RENDER);
    const float ox = u01(q.x) - 0.5f, oy = u01(q.y) - 0.5f;
    const v3 ps = (cam.p00 + ((float)i + ox) * cam.du) + ((float)j + oy) * cam.dv;
    org = cam.center;
    if (cam.defocus) {
        float px = 0.0f, py = 0.0f;
        for (uint32_t k = 0; k < 32; ++k) {
            const u4 c = philox(u4{pixel, sample, 0u, 1u + k}, seed, STREAM_RENDER);
            const float ax = 2.0f * u01(c.x) - 1.0f, ay = 2.0f * u01(c.y) - 1.0f;
            if (ax * ax + ay * ay < 1.0f) {
                px = ax;
                py = ay;
                break;
            }
        }
        org = (org + px * cam.disk_u) + py * cam.disk_v;
    }
    dir = unitv(ps - org);
}

// ONE (in every kernel below): the pass holds a single scene, so its record
// is read at a wave-uniform address (scalar loads) instead of per path.
template <bool ONE>
__global__ __launch_bounds__(kBlock) void k_raygen(Pass pa, float4* qo, float4* qd, float4* qb, float4* L,
                                                   uint32_t* counters) {
    const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
    const uint32_t n = pa.n_paths;
    if (p == 0) {
        counters[0] = n;   // queue 0 count
        counters[1] = 0u;  // queue 1 count (the segment counter spans the call)
    }
    if (p < kPools) counters[kPoolBase + kPoolStride * p] = 0u;  // extend work pools
    if (p >= n) return;
    const uint32_t sid = ONE ? 0u : find_rec<false>(pa.recs, pa.n_recs, p);
    const SceneRec r = scene_rec<ONE>(pa.recs, sid);
    const uint32_t local = p - r.path_off;
    const uint32_t s_local = local / r.WH;
    const uint32_t pixel = local - s_local * r.WH;
    const uint32_t i = pixel % r.cam.W, j = pixel / r.cam.W;
    v3 org, dir;
    camera_ray(r.cam, pixel, i, j, r.sample_base + s_local, r.seed, org, dir);
    qd[p] = make_float4(dir.x, dir.y, dir.z, __uint_as_float(sid));
    // for the fused tail a queue position is its path, and a pinhole
    // camera's rays all leave its centre, which the tail reads from the
    // scene record: only a thin lens needs its origins written
    if (pa.lean_queue && !r.cam.defocus) return;
    qo[p] = make_float4(org.x, org.y, org.z, __uint_as_float(p));
    if (pa.lean_queue) return;  // the fused tail sets throughput 1 and writes every L
    qb[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    L[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Camera rays of n (pixel, sample) pairs (the parity entry
// aurora_camera_rays_device): the same camera_ray as k_raygen.
__global__ __launch_bounds__(kBlock) void k_camera_rays(CamF cam, uint32_t seed, const uint32_t* pixels,
                                                        const uint32_t* samples, uint32_t n, float* org,
                                                        float* dir) {
    const uint32_t k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const uint32_t pixel = pixels[k];
    v3 o, d;
    camera_ray(cam, pixel, pixel % cam.W, pixel / cam.W, samples[k], seed, o, d);
    org[3 * (size_t)k] = o.x;
    org[3 * (size_t)k + 1] = o.y;
    org[3 * (size_t)k + 2] = o.z;
    dir[3 * (size_t)k] = d.x;
    dir[3 * (size_t)k + 1] = d.y;
    dir[3 * (size_t)k + 2] = d.z;
}

// sin and cos of 2*pi*u for u in [0, 1) (u a multiple of 2^-24): exact
// reduction to an eighth of a turn (8u, its integer and fractional parts and
// 1 - f are exact), then Taylor polynomials on [0, pi/4] (truncation below
// 2e-9), folded back by octant symmetry
__device__ __forceinline__ void sincos_turn(float u, float& s, float& c) {
    const float t = 8.0f * u;
    const float fk = floorf(t);
    const int k = (int)fk;
    const float f = t - fk;
    const bool odd = (k & 1) != 0;
    const float a = (odd ? 1.0f - f : f) * 0.785398163f;
    const float a2 = a * a;
    const float sp = a + (a * a2) * (-0.166666672f + a2 * (0.00833333377f + a2 * (-0.000198412701f + a2 * 2.75573188e-6f)));
    const float cp =
        1.0f + a2 * (-0.5f + a2 * (0.0416666679f + a2 * (-0.00138888892f + a2 * (2.48015876e-5f + a2 * -2.75573188e-7f))));
    const float sa = odd ? cp : sp, ca = odd ? sp : cp;  // sin, cos of the angle within the quadrant
    const int q = k >> 1;
    s = q == 0 ? sa : q == 1 ? ca : q == 2 ? -sa : -ca;
    c = q == 0 ? ca : q == 1 ? -sa : q == 2 ? -ca : sa;
}

// uniform direction on the unit sphere from one Philox draw, without
// rejection: z uniform in (-1, 1] (exact), azimuth uniform (Archimedes)
__device__ __forceinline__ v3 random_unit(uint32_t pixel, uint32_t sample, uint32_t bounce, uint32_t seed) {
    const u4 r = philox(u4{pixel, sample, bounce, 0u}, seed, STREAM_RENDER);
    const float z = 1.0f - 2.0f * u01(r.x);
    const float rho = sqrtf(fmaxf(0.0f, 1.0f - z * z));
    float sn, cs;
    sincos_turn(u01(r.y), sn, cs);
    return V(rho * cs, rho * sn, z);
}

__device__ __forceinline__ float schlick(float cosine, float ri) {
    float r0 = (1.0f - ri) / (1.0f + ri);
    r0 = r0 * r0;
    const float x = 1.0f - cosine;
    const float x5 = ((x * x) * (x * x)) * x;
    return r0 + (1.0f - r0) * x5;
}

// first-hit albedo and normal planes: 3 floats per path (12 bytes, a quarter
// less traffic than a float4 per path)
__device__ __forceinline__ void store3(float* p, uint32_t path, v3 v) {
    float* q = p + 3 * (size_t)path;
    q[0] = v.x;
    q[1] = v.y;
    q[2] = v.z;
}
__device__ __forceinline__ v3 load3(const float* p, size_t i) {
    const float* q = p + 3 * i;
    return V(q[0], q[1], q[2]);
}

struct ShadeBufs {
    const float4 *qo, *qd, *qb;
    const float2* hit;
    float4 *no, *nd, *nb;
    float4* L;
    float *aov_a, *aov_n;  // 3 floats per path
    uint32_t* counters;
    int cur;
};

struct ShadeOut {
    v3 p, nd, beta;
    bool alive;
    bool emitted;  // L[path] written (background or emitter)
};

// The pixel and sample of path id `path` (sample-major within its scene).
__device__ __forceinline__ void path_pixel(const SceneRec& r, uint32_t path, uint32_t& pixel, uint32_t& sample) {
    const uint32_t local = path - r.path_off;
    const uint32_t s_local = local / r.WH;
    pixel = local - s_local * r.WH;
    sample = r.sample_base + s_local;
}

// One shading event of path `path` (at `pixel`, `sample`) at segment
// `depth`: background or light emission on termination (written once to
// L[path]), else scatter.
__device__ __forceinline__ ShadeOut shade_core(const SceneRec& r, const ShadeBufs& b, v3 org, v3 dir, v3 beta,
                                               float t, int prim, uint32_t path, uint32_t pixel, uint32_t sample,
                                               uint32_t depth) {
    const DevScene& sc = r.sc;
    const bool aov = r.aov && depth == 0;
    ShadeOut out;
    out.alive = false;
    out.emitted = false;
    const v3 bg = V(r.bg[0], r.bg[1], r.bg[2]);
    if (prim < 0) {
        const v3 e = beta * bg;
        b.L[path] = make_float4(e.x, e.y, e.z, 0.0f);
        out.emitted = true;
        if (aov) {
            const v3 a = clamp01(bg);
            store3(b.aov_a, path, a);
            store3(b.aov_n, path, V(0.0f, 0.0f, 0.0f));
        }
        return out;
    }
    const uint32_t id = (uint32_t)prim;
    const v3 p = org + t * dir;
    v3 outward;
    uint32_t mi;
    const bool is_sphere = id < sc.n_spheres;
    float4 q0, q1, q2;
    if (is_sphere) {
        const float4 s = sc.spheres[id];
        outward = (p - xyz(s)) / s.w;
        mi = sc.sphere_mat[id];
    } else {
        const float4* q = sc.quads + 3 * (size_t)(id - sc.n_spheres);
        q0 = q[0];
        q1 = q[1];
        q2 = q[2];
        outward = xyz(sc.quad_n[id - sc.n_spheres]);  // unitv(cross(u, v)), precomputed
        mi = __float_as_uint(q0.w);
    }
    const aurora_material m = sc.mats[mi];
    float tu = 0.0f, tv = 0.0f;
    if (m.type != AURORA_MAT_DIELECTRIC && tex_needs_uv(sc, m.texture)) {
        if (is_sphere) {
            sphere_uv(outward, tu, tv);
        } else {
            float tt;
            hit_quad(q0, q1, q2, org, dir, r.tmin, INFINITY, tt, tu, tv);
        }
    }
    const bool front = dot(dir, outward) < 0.0f;
    const v3 n = front ? outward : -outward;
    v3 att = V(1.0f, 1.0f, 1.0f);
    v3 nd = V(0.0f, 0.0f, 0.0f);
    bool scatter = true;
    if (m.type == AURORA_MAT_DIFFUSE_LIGHT) {
        const v3 e0 = texture_value(sc, m.texture, tu, tv, p);
        const v3 e = beta * e0;
        b.L[path] = make_float4(e.x, e.y, e.z, 0.0f);
        out.emitted = true;
        att = clamp01(e0);
        scatter = false;
    } else if (m.type == AURORA_MAT_LAMBERTIAN) {
        att = texture_value(sc, m.texture, tu, tv, p);
        v3 sd = n + random_unit(pixel, sample, depth + 1, r.seed);
        if (near_zero(sd)) sd = n;
        nd = unitv(sd);
    } else if (m.type == AURORA_MAT_METAL) {
        att = texture_value(sc, m.texture, tu, tv, p);
        v3 rf = unitv(reflect(dir, n));
        if (m.fuzz > 0.0f) rf = rf + m.fuzz * random_unit(pixel, sample, depth + 1, r.seed);
        if (dot(rf, n) <= 0.0f) scatter = false;  // absorbed
        else nd = unitv(rf);
    } else {
        const float ri = front ? rcp_rn(m.ior) : m.ior;
        const float cos_t = fminf(dot(-dir, n), 1.0f);
        const float sin_t = sqrtf(1.0f - cos_t * cos_t);
        const bool cannot = ri * sin_t > 1.0f;
        const u4 c = philox(u4{pixel, sample, depth + 1, 0u}, r.seed, STREAM_RENDER);
        const float uu = u01(c.x);
        const v3 d2 = (cannot || schlick(cos_t, ri) > uu) ? reflect(dir, n) : refract(dir, n, ri);
        nd = unitv(d2);
    }
    if (aov) {
        const v3 a = clamp01(att);
        store3(b.aov_a, path, a);
        store3(b.aov_n, path, n);
    }
    if (scatter) {
        beta = beta * att;
        out.alive = depth + 1 < r.max_depth;
    }
    out.p = p;
    out.nd = nd;
    out.beta = beta;
    return out;
}

// Grid-stride over the live queue, one path per thread.  Survivors are
// appended to the next queue with one atomic per block and iteration: counts
// are scanned within each wave (ballot popcounts) and across the 4 waves
// (LDS).  The segments traced in the pass are counted with one atomic per
// block.
template <bool ONE>
__global__ __launch_bounds__(kBlock) void k_shade(Pass pa, ShadeBufs b, uint32_t depth) {
    __shared__ uint32_t s_wave[kBlock / 64];
    __shared__ uint32_t s_base;
    const uint32_t n = b.counters[b.cur];
    uint32_t* next_count = &b.counters[b.cur ^ 1];
    if (blockIdx.x == 0 && threadIdx.x < kPools) b.counters[kPoolBase + kPoolStride * threadIdx.x] = 0u;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    uint32_t segs = 0;  // segments this lane traced (a few hundred per launch)
    for (uint32_t base = blockIdx.x * kBlock; base < n; base += gridDim.x * kBlock) {
        const uint32_t i = base + tid;
        ShadeOut o;
        o.alive = false;
        uint32_t path = 0, sid = 0;
        if (i < n) {
            const float4 O = b.qo[i], D = b.qd[i], B = b.qb[i];
            const float2 H = b.hit[i];
            path = __float_as_uint(O.w);
            sid = ONE ? 0u : __float_as_uint(D.w);
            const SceneRec r = scene_rec<ONE>(pa.recs, sid);
            uint32_t pixel, sample;
            path_pixel(r, path, pixel, sample);
            o = shade_core(r, b, xyz(O), xyz(D), xyz(B), H.x, __float_as_int(H.y), path, pixel, sample, depth);
            ++segs;
        }
        const uint64_t m = __ballot(o.alive);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t before = 0, total = 0;
#pragma unroll
        for (int w = 0; w < kBlock / 64; ++w) {
            const uint32_t c = s_wave[w];
            before += w < wave ? c : 0u;
            total += c;
        }
        if (tid == 0) s_base = total ? atomicAdd(next_count, total) : 0u;
        __syncthreads();
        if (o.alive) {
            const uint32_t slot = s_base + before + lanes_below(m);
            b.no[slot] = make_float4(o.p.x, o.p.y, o.p.z, __uint_as_float(path));
            b.nd[slot] = make_float4(o.nd.x, o.nd.y, o.nd.z, __uint_as_float(sid));
            b.nb[slot] = make_float4(o.beta.x, o.beta.y, o.beta.z, 0.0f);
        }
        __syncthreads();  // s_wave / s_base are reused by the next iteration
    }
    unsigned long long wsegs = segs;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wsegs += __shfl_xor(wsegs, off);
    if (lane == 0 && wsegs) atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kTailSegs), wsegs);
}

// Fused tail: once the live queue is short, each lane carries one path to
// termination (traverse -> shade -> ...), so the last bounces cost one
// launch instead of two per bounce.  The grid is persistent: a lane whose
// path ends takes the next one from the queue (paths claimed per wave in
// batches from one of the per-XCD pools, dealt out by ballot rank), so a
// wave is not held by its longest path.  Same arithmetic per path as the
// extend/shade pair, so results are unchanged.
constexpr int kTailRefill = 4;
// the four-wide tail walks until this many lanes of a wave have ended their
// walk, then shades those while the others resume walking
constexpr int kWalkShadeBatch = 32;
// Paths a wave claims per pool atomic (unless the context sets a size):
// about a sixteenth of a wave's share of the pass, within [64, 128].  Large passes take 128 (Cornell, 33.5 M
// paths: 1.5% faster than 64, fewer claims on the shared counters; 32 was 7%
// slower); small passes of long paths keep 64, which balances the waves'
// last claims (a 2 M-path stress render lost 5-15% at 128).
constexpr uint32_t kTailBatchMin = 64, kTailBatchMax = 128;

// Five waves per SIMD for the tree walks: the register budget this leaves
// holds them without spills; the resident grid (five blocks per CU) matches
// it.  The flat walk needs no LDS stack and fewer registers, and runs eight
// waves per SIMD (64 registers, a few spilled: 2.54 G paths/s at seven waves,
// 2.58 at eight on the Cornell box).  TREE: 1 traverse the four-wide tree, 0
// the binary one, 2 (single-scene passes of small scenes) the flat leaf array
// (the faster form depends on the scene, so the launch shape chooses).
template <bool ONE, int TREE>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(TREE == 2 ? 8 : (ONE && TREE == 1) ? 6 : 5))) void k_tail(
    Pass pa, ShadeBufs b, uint32_t depth0, uint32_t* spill,
                                                 uint32_t spill_stride) {
    __shared__ uint32_t stk[kTravLds * kBlock];
    const uint32_t n = b.counters[b.cur];
    uint32_t* my_spill = spill + (size_t)blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & 63;
    // the queue is split into one contiguous range per XCD; a wave whose
    // range is spent moves on to the next ones, so uneven ranges (a batch's
    // scenes differ in cost per path) still finish together
    const uint32_t parts = min((uint32_t)kPools, gridDim.x);
    uint32_t part = blockIdx.x % parts, spent = 0;
    uint32_t r_lo = (uint32_t)(((uint64_t)n * part) / parts);
    uint32_t r_n = (uint32_t)(((uint64_t)n * (part + 1)) / parts) - r_lo;
    uint32_t* pool = b.counters + kPoolBase + kPoolStride * part;
    uint32_t b_next = 0, b_end = 0;  // wave-uniform, offsets within the range
    const uint32_t batch = __builtin_amdgcn_readfirstlane(
        pa.claim ? pa.claim : min(kTailBatchMax, max(kTailBatchMin, n / (gridDim.x * (kBlock / 64) * 16u))));
    bool has = false, done = false;
    uint32_t path = 0, sid = 0, depth = 0, pixel = 0, sample = 0;  // pixel, sample: once per path
    v3 org = V(0, 0, 0), dir = V(0, 0, 0), beta = V(0, 0, 0);
    uint32_t segs = 0;  // segments this lane traced (a few hundred per launch)
    WideWalk walk{};       // TREE 1: the lane's walk of its current segment
    bool walking = false;  // TREE 1: that walk has not ended
    for (;;) {
        const bool idle = !has && !done;
        const uint64_t im = __ballot(idle);
        const uint64_t am = __ballot(has);
        if (im != 0 && (am == 0 || __popcll(im) >= kTailRefill)) {
            // at most `parts` ranges are tried, each once after it is spent
            while (b_next >= b_end && spent < parts) {
                uint32_t base = r_n;
                if (lane == 0) {
                    if (__hip_atomic_load(pool, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < r_n)
                        base = atomicAdd(pool, batch);
                }
                // every lane of the wave is active here (no lane leaves the
                // loop before the whole wave does), so lane 0 is the first
                // active lane; a scalar copy keeps the range state in SGPRs
                base = __builtin_amdgcn_readfirstlane(base);
                if (base < r_n) {
                    b_next = base;
                    b_end = min(r_n, base + batch);
                } else if (++spent < parts) {
                    part = part + 1 == parts ? 0u : part + 1;
                    r_lo = (uint32_t)(((uint64_t)n * part) / parts);
                    r_n = (uint32_t)(((uint64_t)n * (part + 1)) / parts) - r_lo;
                    pool = b.counters + kPoolBase + kPoolStride * part;
                }
            }
            const uint32_t avail = b_next < b_end ? b_end - b_next : 0u;
            const uint32_t rank = lanes_below(im);
            if (idle) {
                if (avail == 0) {
                    done = true;
                } else if (rank < avail) {
                    const uint32_t i = r_lo + b_next + rank;
                    const float4 D = b.qd[i];
                    sid = ONE ? 0u : __float_as_uint(D.w);
                    dir = xyz(D);
                    const SceneRec rc = scene_rec<ONE>(pa.recs, sid);
                    if (pa.lean_queue) {
                        // the raygen queue: position i is path i; a pinhole
                        // camera's origin is its centre (k_raygen)
                        path = i;
                        org = rc.cam.defocus ? xyz(b.qo[i]) : rc.cam.center;
                        beta = V(1.0f, 1.0f, 1.0f);
                    } else {
                        const float4 O = b.qo[i];
                        path = __float_as_uint(O.w);
                        org = xyz(O);
                        beta = xyz(b.qb[i]);
                    }
                    depth = depth0;
                    path_pixel(rc, path, pixel, sample);
                    has = true;
                    if constexpr (TREE == 1) {
                        wide_walk_start(walk, dir, INFINITY);
                        walking = true;
                    }
                }
            }
            b_next += min(avail, (uint32_t)__popcll(im));
        }
        if (__ballot(has || !done) == 0) break;
        if (!has) continue;
        // one segment of this lane's path
        const SceneRec r = scene_rec<ONE>(pa.recs, sid);
        float t;
        int prim;
        if constexpr (TREE == 1) {
            // node steps until kWalkShadeBatch lanes have ended their walks
            // (or none walks any more); those shade below, the others walk on
            // after the refill
            for (;;) {
                if (walking && wide_walk_step(walk, r.sc, org, dir, r.tmin, INFINITY, stk + threadIdx.x, kBlock,
                                              my_spill, spill_stride))
                    walking = false;
                const uint64_t still = __ballot(walking);
                if (still == 0 || __popcll(__ballot(!walking)) >= kWalkShadeBatch) break;
            }
            if (walking) continue;
            wide_walk_result(walk, INFINITY, t, prim);
        } else if constexpr (TREE == 2) {
            traverse_closest_flat(r.sc, org, dir, r.tmin, INFINITY, t, prim);
        } else {
            traverse_closest_binary(r.sc, org, dir, r.tmin, INFINITY, stk + threadIdx.x, kBlock, my_spill,
                                    spill_stride, t, prim);
        }
        ++segs;
        const ShadeOut o = shade_core(r, b, org, dir, beta, t, prim, path, pixel, sample, depth);
        if (!o.alive) {
            // a path absorbed or stopped by the depth limit adds nothing (the
            // raygen kernel zeroes L unless the queue is lean)
            if (pa.lean_queue && !o.emitted) b.L[path] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            has = false;
        } else {
            org = o.p;
            dir = o.nd;
            beta = o.beta;
            ++depth;
            if constexpr (TREE == 1) {
                wide_walk_start(walk, dir, INFINITY);
                walking = true;
            }
        }
    }
    unsigned long long wsegs = segs;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wsegs += __shfl_xor(wsegs, off);
    if (lane == 0 && wsegs) atomicAdd(reinterpret_cast<unsigned long long*>(b.counters + kTailSegs), wsegs);
}

// per pixel: add the pass's samples in sample order to the scene's sums
template <bool ONE>
__global__ __launch_bounds__(kBlock) void k_accumulate(Pass pa, const float4* L, const float* A, const float* Nn) {
    const uint32_t q = blockIdx.x * kBlock + threadIdx.x;
    if (q >= pa.n_pixels) return;
    const SceneRec& r = pa.recs[ONE ? 0u : find_rec<true>(pa.recs, pa.n_recs, q)];
    const uint32_t px = q - r.pix_off;
    const uint32_t WH = r.WH, S = r.spp_pass;
    const float4* Ls = L + r.path_off;
    float* rgb = r.rgb;
    v3 acc = V(rgb[3 * (size_t)px], rgb[3 * (size_t)px + 1], rgb[3 * (size_t)px + 2]);
    for (uint32_t s = 0; s < S; ++s) acc = acc + xyz(Ls[(size_t)s * WH + px]);
    rgb[3 * (size_t)px] = acc.x;
    rgb[3 * (size_t)px + 1] = acc.y;
    rgb[3 * (size_t)px + 2] = acc.z;
    if (r.aov) {
        const float* As = A + 3 * (size_t)r.path_off;
        const float* Ns = Nn + 3 * (size_t)r.path_off;
        float* alb = r.alb;
        float* nrm = r.nrm;
        v3 a = V(alb[3 * (size_t)px], alb[3 * (size_t)px + 1], alb[3 * (size_t)px + 2]);
        v3 c = V(nrm[3 * (size_t)px], nrm[3 * (size_t)px + 1], nrm[3 * (size_t)px + 2]);
        for (uint32_t s = 0; s < S; ++s) {
            a = a + load3(As, (size_t)s * WH + px);
            c = c + load3(Ns, (size_t)s * WH + px);
        }
        alb[3 * (size_t)px] = a.x;
        alb[3 * (size_t)px + 1] = a.y;
        alb[3 * (size_t)px + 2] = a.z;
        nrm[3 * (size_t)px] = c.x;
        nrm[3 * (size_t)px + 1] = c.y;
        nrm[3 * (size_t)px + 2] = c.z;
    }
}

__global__ void k_advance(SceneRec* recs, uint32_t n) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) recs[k].sample_base += recs[k].spp_pass;
}

ShadeBufs shade_bufs(Workspace& ws, int cur) {
    ShadeBufs b;
    b.qo = ws.q_o[cur].as<float4>();
    b.qd = ws.q_d[cur].as<float4>();
    b.qb = ws.q_b[cur].as<float4>();
    b.hit = ws.hit.as<float2>();
    b.no = ws.q_o[cur ^ 1].as<float4>();
    b.nd = ws.q_d[cur ^ 1].as<float4>();
    b.nb = ws.q_b[cur ^ 1].as<float4>();
    b.L = ws.L.as<float4>();
    b.aov_a = ws.aov_a.as<float>();
    b.aov_n = ws.aov_n.as<float>();
    b.counters = ws.counters.as<uint32_t>();
    b.cur = cur;
    return b;
}

}  // namespace

Status launch_raygen(const Pass& p, Workspace& ws, hipStream_t s) {
    const uint32_t blocks = std::max<uint32_t>(1, (p.n_paths + kBlock - 1) / kBlock);
    auto k = p.n_recs == 1 ? k_raygen<true> : k_raygen<false>;
    k<<<blocks, kBlock, 0, s>>>(p, ws.q_o[0].as<float4>(), ws.q_d[0].as<float4>(), ws.q_b[0].as<float4>(),
                                ws.L.as<float4>(), ws.counters.as<uint32_t>());
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_camera_rays(const CamF& cam, uint32_t seed, const uint32_t* pixels, const uint32_t* samples,
                          uint32_t n, float* org, float* dir, hipStream_t s) {
    if (n == 0) return Status::ok();
    k_camera_rays<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(cam, seed, pixels, samples, n, org, dir);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

// 8-bit resolve of n floats (three per pixel), four per thread
__global__ __launch_bounds__(kBlock) void k_resolve_rgb8(const float* sum, size_t n, float spp, uint32_t transfer,
                                                         uint8_t* out) {
    const size_t k = ((size_t)blockIdx.x * kBlock + threadIdx.x) * 4;
    if (k + 4 <= n) {
        const float4 v = *reinterpret_cast<const float4*>(sum + k);
        const uint32_t b = (uint32_t)resolve_byte(v.x, spp, transfer) | ((uint32_t)resolve_byte(v.y, spp, transfer) << 8) |
                           ((uint32_t)resolve_byte(v.z, spp, transfer) << 16) |
                           ((uint32_t)resolve_byte(v.w, spp, transfer) << 24);
        *reinterpret_cast<uint32_t*>(out + k) = b;
    } else {
        for (size_t i = k; i < n; ++i) out[i] = resolve_byte(sum[i], spp, transfer);
    }
}

Status launch_resolve_rgb8(const float* sum, size_t n, uint32_t spp, uint32_t transfer, uint8_t* out, hipStream_t s) {
    if (n == 0) return Status::ok();
    const size_t blocks = (n + 4 * kBlock - 1) / (4 * kBlock);
    k_resolve_rgb8<<<(uint32_t)blocks, kBlock, 0, s>>>(sum, n, (float)spp, transfer, out);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_shade(const Pass& p, Workspace& ws, int cur, uint32_t depth, int grid, hipStream_t s) {
    auto k = p.n_recs == 1 ? k_shade<true> : k_shade<false>;
    k<<<grid, kBlock, 0, s>>>(p, shade_bufs(ws, cur), depth);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_tail(const Pass& p, Workspace& ws, int cur, uint32_t depth, int grid, hipStream_t s) {
    ShadeBufs b = shade_bufs(ws, cur);
    b.hit = nullptr;
    b.no = b.nd = b.nb = nullptr;
    const size_t need = (size_t)grid * kBlock * (kTravMaxStack - kTravLds) * 4;
    if (need > ws.spill.bytes) return Status{AURORA_ERR_RUNTIME, "tail grid exceeds the stack spill buffer"};
    auto k = p.n_recs == 1 ? (p.flat ? k_tail<true, 2> : p.wide ? k_tail<true, 1> : k_tail<true, 0>)
                           : (p.wide ? k_tail<false, 1> : k_tail<false, 0>);
    k<<<grid, kBlock, 0, s>>>(p, b, depth, ws.spill.as<uint32_t>(), (uint32_t)(grid * kBlock));
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_accumulate(const Pass& p, Workspace& ws, hipStream_t s) {
    if (p.n_pixels == 0) return Status::ok();
    auto k = p.n_recs == 1 ? k_accumulate<true> : k_accumulate<false>;
    k<<<(p.n_pixels + kBlock - 1) / kBlock, kBlock, 0, s>>>(p, ws.L.as<float4>(), ws.aov_a.as<float>(),
                                                            ws.aov_n.as<float>());
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_advance(SceneRec* recs, uint32_t n, hipStream_t s) {
    k_advance<<<(n + 63) / 64, 64, 0, s>>>(recs, n);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

}  // namespace aur
