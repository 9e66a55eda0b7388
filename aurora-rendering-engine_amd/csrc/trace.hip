// trace.hip — closest-hit traversal of the LBVH on gfx950 (the extend phase
// of the wavefront path tracer and the parity entry aurora_trace_closest).
//
// k_extend is persistent: a grid sized to the resident capacity of the chip
// pulls rays from a global pool.  Each lane owns one ray; when enough lanes
// of a wave64 have finished, the wave claims that many new rays with one
// atomic and hands them out by ballot prefix (refilling idle lanes).  The
// rays walk the four-wide tree collapsed from the LBVH: one step fetches a
// 128-byte node (one cache line, four child boxes), tests the four slabs,
// intersects leaf primitives at once and descends to the nearest hit child,
// pushing the others farthest first.  The stack lives in LDS (kLds entries
// per lane, laid out [level][thread] so a wave's pushes hit 64 consecutive
// banks) and spills deeper levels to a per-thread global array.
//
// Reference counterpart: the per-ray linear scan over all triangles,
// experiments/rt.cpp:209-218, with are::Triangle::intersect_ray
// (src/object/triangle.cpp:82-121) as the primitive test.
#include "internal.hpp"
#include "traverse.hpp"

namespace aur {

namespace {

constexpr int kBlock = 256;
constexpr int kLds = 16;
constexpr int kMaxStack = 96;  // three pushes per level of a tree of depth <= 64
constexpr int kRefill = 8;
constexpr int kParts = 8;
constexpr uint32_t kNone = 0xffffffffu;
constexpr uint32_t kNoPrim = 0x7fffffffu;

// ONE: every ray belongs to the first scene record, whose fields are then
// wave-uniform (scalar registers, loaded once); otherwise each claimed ray
// loads its scene's record.
//
// WIDE: walk the four-wide tree (ray queries, memory bound); otherwise the
// binary tree, one 64-byte child pair per step (render bounces, issue bound).
template <bool COUNT, bool ONE, bool WIDE>
__global__ __launch_bounds__(kBlock) void k_extend(ExtendArgs a) {
    __shared__ uint32_t stk[kLds][kBlock];
    const uint32_t n = a.count ? *a.count : a.n;
    if (a.zero_after && blockIdx.x == 0 && threadIdx.x == 0) *a.zero_after = 0u;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    uint32_t* spill = a.spill + (size_t)blockIdx.x * kBlock + tid;
    const size_t sstride = a.spill_stride;
    const float tmax = a.tmax;
    float tmin = 0.0f;  // the scene's ray offset
    // the scene of the lane's current ray
    const float4* N = nullptr;  // the wide nodes (WIDE) or the binary nodes
    const float4* S = nullptr;
    const float4* Qd = nullptr;
    uint32_t n_spheres = 0, n_nodes = 0, n_prims = 0;
    if constexpr (ONE) {
        const SceneRec& rec = a.recs[0];
        tmin = rec.tmin;
        N = WIDE ? rec.sc.wide : rec.sc.nodes;
        S = rec.sc.spheres;
        Qd = rec.sc.quads;
        n_spheres = rec.sc.n_spheres;
        n_nodes = WIDE ? rec.sc.n_wide : rec.sc.n_nodes;
        n_prims = rec.sc.n_spheres + rec.sc.n_quads;
    }

    uint32_t ri = kNone;
    bool done = false;
    v3 o = V(0, 0, 0), d = V(0, 0, 0), inv = V(0, 0, 0);
    float best_t = tmax;
    uint32_t best_id = kNoPrim;
    uint32_t cur = 0;
    int sp = 0;
    uint32_t cn = 0, cs = 0, cq = 0;                  // per-ray counts
    unsigned long long tn = 0, ts = 0, tq = 0;        // per-lane totals

    // te: the entry distance of the primitive's box; a hit before it does not
    // count (this makes the closest hit independent of the visit order)
    // (an id past the primitives, which no well-formed tree holds, reads
    // primitive 0 and is neither counted nor taken)
    auto test_prim = [&](uint32_t id, float te) {
        const bool valid = id < n_prims;
        const uint32_t k = valid ? id : 0u;
        float t = 0.0f;
        bool h;
        if (k < n_spheres) {
            if (COUNT) cs += valid ? 1u : 0u;
            h = hit_sphere_t(ldg4(S, k), o, d, tmin, tmax, t);
        } else {
            if (COUNT) cq += valid ? 1u : 0u;
            const size_t q = 3 * (size_t)(k - n_spheres);
            h = hit_quad_t(ldg4(Qd, q), ldg4(Qd, q + 1), ldg4(Qd, q + 2), o, d, tmin, tmax, t);
        }
        const bool take = valid & h & (t >= te) & ((t < best_t) | ((t == best_t) & (id < best_id)));
        best_t = take ? t : best_t;
        best_id = take ? id : best_id;
    };
    auto finish = [&]() {
        const bool miss = best_id == kNoPrim;
        a.hit[ri] = make_float2(miss ? tmax : best_t, __int_as_float(miss ? AURORA_MISS : (int)best_id));
        if (COUNT) {
            if (a.counts) {
                a.counts[3 * (size_t)ri] = cn;
                a.counts[3 * (size_t)ri + 1] = cs;
                a.counts[3 * (size_t)ri + 2] = cq;
            }
            tn += cn;
            ts += cs;
            tq += cq;
        }
        ri = kNone;
    };

    // Work distribution: the rays are split into up to kParts contiguous
    // ranges, one per XCD (blocks are dealt to the 8 XCDs round-robin), each
    // with its own pool counter on its own 128-byte line.  A wave claims a
    // batch of its range with one atomic and deals it out to idle lanes by
    // ballot rank; a plain load skips the atomic once the range is spent.
    // A wave whose range is spent moves on to the next ones, so uneven
    // ranges still finish together.
    const uint32_t parts = min((uint32_t)kParts, gridDim.x);
    uint32_t part = blockIdx.x % parts, spent = 0;
    uint32_t r_lo = (uint32_t)(((uint64_t)n * part) / parts);
    uint32_t r_n = (uint32_t)(((uint64_t)n * (part + 1)) / parts) - r_lo;
    uint32_t* pool = a.pool + kPoolStride * part;
    const uint32_t waves = ((gridDim.x - part + parts - 1) / parts) * (kBlock / 64);
    // about an eighth of a wave's share of its range per claim: half a share
    // left the last claims unbalanced (2^22 sorted stress queries: 146 ms
    // against 126 ms with eighths; 32nds measured the same as eighths)
    const uint32_t batch = min(4096u, max(64u, r_n / (waves * 8u) + 1u));
    uint32_t b_next = 0, b_end = 0;  // wave-uniform, offsets within the range
    for (;;) {
        const bool idle = (ri == kNone) && !done;
        const uint64_t im = __ballot(idle);
        const uint64_t am = __ballot(ri != kNone);
        if (im != 0 && (am == 0 || __popcll(im) >= kRefill)) {
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
                    pool = a.pool + kPoolStride * part;
                }
            }
            const uint32_t avail = b_next < b_end ? b_end - b_next : 0u;
            const uint32_t rank = (uint32_t)__popcll(im & lt);
            if (idle) {
                if (avail == 0) {
                    done = true;
                } else if (rank < avail) {
                    ri = r_lo + b_next + rank;
                    if (a.order) ri = a.order[ri];
                    const float4 ro = a.ray_o[ri], rd = a.ray_d[ri];
                    o = xyz(ro);
                    d = xyz(rd);
                    if constexpr (!ONE) {
                        const SceneRec& rec = a.recs[__float_as_uint(rd.w)];
                        tmin = rec.tmin;
                        N = WIDE ? rec.sc.wide : rec.sc.nodes;
                        S = rec.sc.spheres;
                        Qd = rec.sc.quads;
                        n_spheres = rec.sc.n_spheres;
                        n_nodes = WIDE ? rec.sc.n_wide : rec.sc.n_nodes;
                        n_prims = rec.sc.n_spheres + rec.sc.n_quads;
                    }
                    inv = V(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
                    best_t = tmax;
                    best_id = kNoPrim;
                    sp = 0;
                    cn = cs = cq = 0;
                    if (n_nodes == 0) {
                        finish();
                    } else if constexpr (WIDE) {
                        cur = 0;  // the root of the four-wide tree
                    } else {
                        const float4 r0 = ldg4(N, 0), r1 = ldg4(N, 1);
                        if (COUNT) ++cn;
                        float tr;
                        const uint32_t rc = __float_as_uint(r0.w);
                        if (!slab(r0, r1, o, inv, tmin, best_t, tr)) {
                            finish();
                        } else if (rc & AURORA_NODE_LEAF) {
                            test_prim(rc & ~AURORA_NODE_LEAF, tr);
                            finish();
                        } else {
                            cur = rc;
                        }
                    }
                }
            }
            b_next += min(avail, (uint32_t)__popcll(im));
        }
        if (__ballot((ri != kNone) || !done) == 0) break;
        if (ri == kNone) continue;

        auto push = [&](uint32_t v) {
            // the tree is shallow enough that the bound only keeps every
            // access inside the stack arrays
            if (sp < kLds) stk[sp][tid] = v;
            else if (sp < kMaxStack) spill[(size_t)(sp - kLds) * sstride] = v;
            if (sp < kMaxStack) ++sp;
        };
        auto pop_or_finish = [&]() {
            if (sp == 0) {
                finish();
            } else {
                --sp;
                if (sp < kLds) cur = stk[sp][tid];
                else cur = ldg1(spill, (size_t)(sp - kLds) * sstride);
            }
        };
        if constexpr (WIDE) {
            // a well-formed tree only refers to nodes below n_wide; any other
            // value ends the ray instead of reading out of bounds
            if (cur >= n_nodes) {
                finish();
                continue;
            }
            // one step: the four slots of wide node cur
            uint32_t r0, r1, r2, r3;
            const int m = wide_step(N, cur, o, inv, tmin, best_t, test_prim, r0, r1, r2, r3);
            if (COUNT) cn += 4;
            if (m == 0) {
                pop_or_finish();
                continue;
            }
            // the others farthest first
            if (m > 3) push(r3);
            if (m > 2) push(r2);
            if (m > 1) push(r1);
            cur = r0;
        } else {
            // a well-formed tree only holds even pair indices below n_nodes;
            // any other value ends the ray instead of reading out of bounds
            if ((cur & 1u) || cur + 1 >= n_nodes) {
                finish();
                continue;
            }
            // one step: the child pair at nodes[cur], nodes[cur+1]
            const float4 a0 = ldg4(N, 2 * (size_t)cur), a1 = ldg4(N, 2 * (size_t)cur + 1);
            const float4 b0 = ldg4(N, 2 * (size_t)cur + 2), b1 = ldg4(N, 2 * (size_t)cur + 3);
            if (COUNT) cn += 2;
            float ta, tb;
            bool ha = slab(a0, a1, o, inv, tmin, best_t, ta);
            bool hb = slab(b0, b1, o, inv, tmin, best_t, tb);
            const uint32_t ca = __float_as_uint(a0.w), cb = __float_as_uint(b0.w);
            if (ha && (ca & AURORA_NODE_LEAF)) {
                test_prim(ca & ~AURORA_NODE_LEAF, ta);
                ha = false;
            }
            if (hb && (cb & AURORA_NODE_LEAF)) {
                test_prim(cb & ~AURORA_NODE_LEAF, tb);
                hb = false;
            }
            if (ha && hb) {
                const bool bnear = tb < ta;
                push(bnear ? ca : cb);
                cur = bnear ? cb : ca;
            } else if (ha) {
                cur = ca;
            } else if (hb) {
                cur = cb;
            } else {
                pop_or_finish();
            }
        }
    }
    if (COUNT && a.totals) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            tn += __shfl_xor(tn, off);
            ts += __shfl_xor(ts, off);
            tq += __shfl_xor(tq, off);
        }
        if (lane == 0) {
            atomicAdd(&a.totals[0], tn);
            atomicAdd(&a.totals[1], ts);
            atomicAdd(&a.totals[2], tq);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_brute(DevScene sc, const float4* ro, const float4* rd, uint32_t n,
                                                  float tmin, float tmax, float2* hit) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const v3 o = xyz(ro[i]), d = xyz(rd[i]);
    const v3 inv = V(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
    float best_t = tmax;
    uint32_t best_id = kNoPrim;
    const uint32_t np = sc.n_spheres + sc.n_quads;
    for (uint32_t id = 0; id < np; ++id) {
        // the rule of the tree traversals: the ray must enter the
        // primitive's box, and a hit before the box's entry does not count
        v3 lo, hi;
        prim_box(sc, id, lo, hi);
        float te;
        if (!slab(make_float4(lo.x, lo.y, lo.z, 0.0f), make_float4(hi.x, hi.y, hi.z, 0.0f), o, inv, tmin, tmax, te))
            continue;
        float t;
        bool h;
        if (id < sc.n_spheres) {
            h = hit_sphere(sc.spheres[id], o, d, tmin, tmax, t);
        } else {
            const float4* q = sc.quads + 3 * (size_t)(id - sc.n_spheres);
            float al, be;
            h = hit_quad(q[0], q[1], q[2], o, d, tmin, tmax, t, al, be);
        }
        if (h && t >= te && (t < best_t || (t == best_t && id < best_id))) {
            best_t = t;
            best_id = id;
        }
    }
    const bool miss = best_id == kNoPrim;
    hit[i] = make_float2(miss ? tmax : best_t, __int_as_float(miss ? AURORA_MISS : (int)best_id));
}

__device__ __forceinline__ uint32_t spread9(uint32_t v) {
    // bits 0..8 of v to every third bit
    v &= 0x1ffu;
    v = (v | (v << 16)) & 0x030000ffu;
    v = (v | (v << 8)) & 0x0300f00fu;
    v = (v | (v << 4)) & 0x030c30c3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__device__ __forceinline__ uint32_t cell9(float x, float lo, float hi) {
    const float e = hi - lo;
    const float u = e > 0.0f ? (x - lo) / e : 0.0f;
    return (uint32_t)fminf(fmaxf(u * 512.0f, 0.0f), 511.0f);
}

__global__ __launch_bounds__(kBlock) void k_ray_keys(const float4* ro, const float4* rd, uint32_t n,
                                                     const float4* nodes, uint32_t* keys, uint32_t* vals) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float4 lo = nodes[0], hi = nodes[1];
    const float4 o = ro[i], d = rd[i];
    const uint32_t m = (spread9(cell9(o.x, lo.x, hi.x)) << 2) | (spread9(cell9(o.y, lo.y, hi.y)) << 1) |
                       spread9(cell9(o.z, lo.z, hi.z));
    const uint32_t oct = (d.x < 0.0f ? 4u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 1u : 0u);
    keys[i] = (m << 3) | oct;
    vals[i] = i;
}

}  // namespace

Status launch_ray_keys(const float4* ro, const float4* rd, uint32_t n, const float4* nodes, uint32_t* keys,
                       uint32_t* vals, hipStream_t s) {
    if (n == 0) return Status::ok();
    k_ray_keys<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(ro, rd, n, nodes, keys, vals);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

static_assert(kExtendBlock == kBlock && kExtendSpillLevels == kMaxStack - kLds, "extend shape constants");
int extend_block() { return kBlock; }
int extend_spill_levels() { return kMaxStack - kLds; }

Status launch_extend(const ExtendArgs& a, int grid, hipStream_t s) {
    if ((size_t)grid * kBlock > a.spill_stride) return Status{AURORA_ERR_RUNTIME, "extend grid exceeds spill stride"};
    const bool count = a.counts || a.totals;
    const int v = (count ? 4 : 0) | (a.one_scene ? 2 : 0) | (a.wide ? 1 : 0);
    void (*const ks[8])(ExtendArgs) = {k_extend<false, false, false>, k_extend<false, false, true>,
                                       k_extend<false, true, false>,  k_extend<false, true, true>,
                                       k_extend<true, false, false>,  k_extend<true, false, true>,
                                       k_extend<true, true, false>,   k_extend<true, true, true>};
    ks[v]<<<grid, kBlock, 0, s>>>(a);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status launch_brute(const DevScene& sc, const float4* o, const float4* d, uint32_t n, float tmin, float tmax,
                    float2* hit, hipStream_t s) {
    if (n == 0) return Status::ok();
    k_brute<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(sc, o, d, n, tmin, tmax, hit);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

}  // namespace aur
