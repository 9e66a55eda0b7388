// traverse.hpp — closest-hit traversal over the LBVH, in two forms.
// Four-wide: the 128-byte node (one cache line) holds four child boxes; per
// step the four slab tests against the current best t, the hit leaf
// primitives in slot order, then descend to the nearest hit internal slot
// and push the others farthest first.  It halves the dependent fetches and
// uses whole lines, so it wins where traversal is memory bound (large
// incoherent ray queries).  Binary: the 64-byte child pair per step, a
// cheaper step that wins where traversal is issue bound (render bounces,
// whose nodes sit in cache).
//
// Order independence.  A primitive's hit counts only if its t is at least
// the entry distance of the primitive's own (leaf) box, and ties in t go to
// the lower primitive id.  Slab arithmetic is monotone in the box bounds
// (correctly rounded subtract and multiply, min and max), and every
// ancestor's bounds contain its children's as floats, so an ancestor's entry
// distance is never above its leaf's, hence never above a counted hit's t:
// no ancestor of the closest counted hit is culled while the best t is at
// or above it, and every form and visit order returns the same hit.
#pragma once

#include "dev_common.hpp"

namespace aur {

constexpr int kTravLds = 16;       // stack levels kept in LDS
constexpr int kTravMaxStack = 96;  // a four-wide step pushes at most 3 per level (depth <= 64)

// Compare-exchange of (key, ref) pairs, swapping only when the second key
// is strictly smaller.  Keys are the unsigned bits of entry t + 0 (-0 taken
// as +0): for t >= +0 that is the float order, and one 32-bit compare orders
// the pair and carries the ref along.  Slots that are not hit are keyed all
// ones and sort last.
__device__ __forceinline__ void wide_cs(uint32_t& ka, uint32_t& ra, uint32_t& kb, uint32_t& rb) {
    const bool sw = kb < ka;
    const uint32_t k = sw ? kb : ka, r = sw ? rb : ra;
    kb = sw ? ka : kb;
    rb = sw ? ra : rb;
    ka = k;
    ra = r;
}

// One step of the four-wide traversal at node `cur` (8 float4: per slot
// (lo, ref), (hi, 0)): slab tests of the four slots against tcur; the hit
// internal slots ordered nearest first by the comparators (0,1) (2,3) (0,2)
// (1,3) (1,2) (equal distances stay where the network leaves them); the hit
// leaf slots' primitives tested in slot order (test_prim(id, entry t)).
// Returns the count of hit internal slots; ref[0..count) holds them in that
// order.
template <class TestPrim>
__device__ __forceinline__ int wide_step(const float4* W, uint32_t cur, v3 o, v3 inv, float tmin, float tcur,
                                         TestPrim&& test_prim, uint32_t& r0, uint32_t& r1, uint32_t& r2,
                                         uint32_t& r3) {
    const size_t b = 8 * (size_t)cur;
    const float4 l0 = ldg4(W, b), h0 = ldg4(W, b + 1), l1 = ldg4(W, b + 2), h1 = ldg4(W, b + 3);
    const float4 l2 = ldg4(W, b + 4), h2 = ldg4(W, b + 5), l3 = ldg4(W, b + 6), h3 = ldg4(W, b + 7);
    const uint32_t f0 = __float_as_uint(l0.w), f1 = __float_as_uint(l1.w);
    const uint32_t f2 = __float_as_uint(l2.w), f3 = __float_as_uint(l3.w);
    float t0 = 0.0f, t1 = 0.0f, t2 = 0.0f, t3 = 0.0f;
    // all four slabs evaluated (an empty slot's box is zeros), no branches
    bool e0 = slab(l0, h0, o, inv, tmin, tcur, t0) & (f0 != AURORA_BVH4_EMPTY);
    bool e1 = slab(l1, h1, o, inv, tmin, tcur, t1) & (f1 != AURORA_BVH4_EMPTY);
    bool e2 = slab(l2, h2, o, inv, tmin, tcur, t2) & (f2 != AURORA_BVH4_EMPTY);
    bool e3 = slab(l3, h3, o, inv, tmin, tcur, t3) & (f3 != AURORA_BVH4_EMPTY);
    // the hit leaf slots, compacted in slot order with their entry
    // distances: the wave runs the j-th primitive test only when some lane
    // has more than j leaves, instead of once per slot that any lane hit
    uint32_t p0 = 0, p1 = 0, p2 = 0, p3 = 0;
    float n0 = 0.0f, n1 = 0.0f, n2 = 0.0f, n3 = 0.0f;
    int np = 0;
    auto add = [&](bool& e, uint32_t f, float tn) {
        const bool lf = e & ((f & AURORA_NODE_LEAF) != 0u);
        const uint32_t id = f & ~AURORA_NODE_LEAF;
        p0 = (lf & (np == 0)) ? id : p0;
        p1 = (lf & (np == 1)) ? id : p1;
        p2 = (lf & (np == 2)) ? id : p2;
        p3 = (lf & (np == 3)) ? id : p3;
        n0 = (lf & (np == 0)) ? tn : n0;
        n1 = (lf & (np == 1)) ? tn : n1;
        n2 = (lf & (np == 2)) ? tn : n2;
        n3 = (lf & (np == 3)) ? tn : n3;
        np += lf ? 1 : 0;
        e = e & !lf;
    };
    add(e0, f0, t0);
    add(e1, f1, t1);
    add(e2, f2, t2);
    add(e3, f3, t3);
    // order the hit internal slots first (their entry distances do not
    // depend on the leaf tests), so the slot words die before the tests
    uint32_t k0 = e0 ? __float_as_uint(t0 + 0.0f) : ~0u, k1 = e1 ? __float_as_uint(t1 + 0.0f) : ~0u;
    uint32_t k2 = e2 ? __float_as_uint(t2 + 0.0f) : ~0u, k3 = e3 ? __float_as_uint(t3 + 0.0f) : ~0u;
    r0 = f0;
    r1 = f1;
    r2 = f2;
    r3 = f3;
    wide_cs(k0, r0, k1, r1);
    wide_cs(k2, r2, k3, r3);
    wide_cs(k0, r0, k2, r2);
    wide_cs(k1, r1, k3, r3);
    wide_cs(k1, r1, k2, r2);
    const int m = (int)e0 + (int)e1 + (int)e2 + (int)e3;
    if (np > 0) test_prim(p0, n0);
    if (np > 1) test_prim(p1, n1);
    if (np > 2) test_prim(p2, n2);
    if (np > 3) test_prim(p3, n3);
    return m;
}

// The four-wide walk of one ray, resumable: its state (reciprocal
// direction, best hit, next node, stack depth) lives in the caller's
// registers and the thread's stack column (level k at stk[k * col_stride]
// in LDS, at spill[(k - kTravLds) * sstride] from kTravLds on), and
// wide_walk_step runs one node step: wide_step at the current node, then
// the nearest hit child next and the others pushed farthest first, or a pop
// when no child is hit.  A persistent kernel can so let the lanes whose walk
// ended shade and start their next segment while the other lanes of the
// wave keep walking.
struct WideWalk {
    v3 inv;
    float best_t;
    uint32_t best_id;
    uint32_t cur;  // the next node; kWalkDone once the walk has ended
    int sp;
};
constexpr uint32_t kWalkDone = 0xffffffffu;

__device__ __forceinline__ void wide_walk_start(WideWalk& w, v3 d, float tmax) {
    w.inv = V(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
    w.best_t = tmax;
    w.best_id = 0x7fffffffu;
    w.cur = 0;
    w.sp = 0;
}

// one node step of the walk; true once it has ended
__device__ __forceinline__ bool wide_walk_step(WideWalk& w, const DevScene& sc, v3 o, v3 d, float tmin, float tmax,
                                               uint32_t* stk, int col_stride, uint32_t* spill, size_t sstride) {
    const uint32_t n_spheres = sc.n_spheres;
    const uint32_t n_prims = sc.n_spheres + sc.n_quads;
    // a well-formed tree only refers to nodes below n_wide; any other value
    // ends the ray instead of reading out of bounds
    if (w.cur >= sc.n_wide) {
        w.cur = kWalkDone;
        return true;
    }
    auto test_prim = [&](uint32_t id, float tn) {
        const bool valid = id < n_prims;
        const uint32_t k = valid ? id : 0u;
        float t = 0.0f;
        bool h;
        if (k < n_spheres) {
            h = hit_sphere_t(ldg4(sc.spheres, k), o, d, tmin, tmax, t);
        } else {
            const size_t q = 3 * (size_t)(k - n_spheres);
            h = hit_quad_t(ldg4(sc.quads, q), ldg4(sc.quads, q + 1), ldg4(sc.quads, q + 2), o, d, tmin, tmax, t);
        }
        const bool take = valid & h & (t >= tn) & ((t < w.best_t) | ((t == w.best_t) & (id < w.best_id)));
        w.best_t = take ? t : w.best_t;
        w.best_id = take ? id : w.best_id;
    };
    uint32_t r0, r1, r2, r3;
    const int m = wide_step(sc.wide, w.cur, o, w.inv, tmin, w.best_t, test_prim, r0, r1, r2, r3);
    if (m == 0) {
        if (w.sp == 0) {
            w.cur = kWalkDone;
            return true;
        }
        --w.sp;
        w.cur = w.sp < kTravLds ? stk[w.sp * col_stride] : ldg1(spill, (size_t)(w.sp - kTravLds) * sstride);
        return false;
    }
    auto push = [&](uint32_t v) {
        if (w.sp < kTravLds) stk[w.sp * col_stride] = v;
        else if (w.sp < kTravMaxStack) spill[(size_t)(w.sp - kTravLds) * sstride] = v;
        if (w.sp < kTravMaxStack) ++w.sp;
    };
    if (m > 3) push(r3);
    if (m > 2) push(r2);
    if (m > 1) push(r1);
    w.cur = r0;
    return false;
}

__device__ __forceinline__ void wide_walk_result(const WideWalk& w, float tmax, float& t_out, int& prim_out) {
    const bool miss = w.best_id == 0x7fffffffu;
    t_out = miss ? tmax : w.best_t;
    prim_out = miss ? AURORA_MISS : (int)w.best_id;
}

// Flat form, for scenes of a few primitives: every leaf record in the flat
// array's order (largest boxes first, so the best t falls early and more
// boxes are culled), its box's slab test against the current best t and its
// primitive test, the same sequence for every ray; a leaf whose box no lane
// of the wave reaches is skipped at once.  Within a single-scene launch the
// records sit at wave-uniform addresses (scalar loads) and no lane waits for
// another's traversal: no stack, no divergence.  A primitive is taken under
// the same conditions as in the tree walks (its leaf box passes the slab
// test at the current best t, the hit lies at or beyond the box's entry
// distance, nearer or an equal t with a lower id), so the closest hit is the
// tree walks' one.  tmax must not be NaN (the render passes +inf), so the
// best t never is; tmin is the launch's one scene's, so wave-uniform.
__device__ __forceinline__ void traverse_closest_flat(const DevScene& sc, v3 o, v3 d, float tmin, float tmax,
                                                      float& t_out, int& prim_out) {
    // one scene per launch: the scene's fields are wave-uniform
    const float4* const F = uniform_ptr(sc.flat);
    const float4* const S = uniform_ptr(sc.spheres);
    const float4* const Qd = uniform_ptr(sc.quads);
    const uint32_t n_spheres = __builtin_amdgcn_readfirstlane(sc.n_spheres);
    const uint32_t n_flat = __builtin_amdgcn_readfirstlane(sc.n_flat);
    const uint32_t n_prims = n_spheres + __builtin_amdgcn_readfirstlane(sc.n_quads);
    const float tmin_s = __uint_as_float(__builtin_amdgcn_readfirstlane(__float_as_uint(tmin)));
    const v3 inv = V(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
    float best_t = tmax;
    uint32_t best_id = 0x7fffffffu;
    for (uint32_t k = 0; k < n_flat; ++k) {
        const float4 lo = ldu4(F, 2 * (size_t)k), hi = ldu4(F, 2 * (size_t)k + 1);
        float tn = 0.0f;
        const bool hb = slab_tcur(lo, hi, o, inv, tmin_s, best_t, tn);
        const uint32_t id = __float_as_uint(lo.w) & ~AURORA_NODE_LEAF;
        const bool valid = id < n_prims;
        const uint32_t q = valid ? id : 0u;
        if (__ballot(hb) == 0) continue;  // no lane reaches this leaf's box: a wave-uniform skip
        float t = 0.0f;
        bool h;
        if (q < n_spheres) {
            h = hit_sphere_t(ldu4(S, q), o, d, tmin, tmax, t);
        } else {
            const size_t b = 3 * (size_t)(q - n_spheres);
            h = hit_quad_t(ldu4(Qd, b), ldu4(Qd, b + 1), ldu4(Qd, b + 2), o, d, tmin, tmax, t);
        }
        const bool take = hb & valid & h & (t >= tn) & ((t < best_t) | ((t == best_t) & (id < best_id)));
        best_t = take ? t : best_t;
        best_id = take ? id : best_id;
    }
    const bool miss = best_id == 0x7fffffffu;
    t_out = miss ? tmax : best_t;
    prim_out = miss ? AURORA_MISS : (int)best_id;
}

// Binary form: one step fetches the aligned child pair at `cur` (two node
// records, 64 bytes), tests both slabs, tests leaf primitives at once and
// descends to the nearer child, pushing the farther.  Used where traversal
// is issue bound (render bounces), where its cheaper step wins.
// stk: this thread's LDS column (level k at stk[k * col_stride]); spill: its
// global column for levels >= kTravLds (level k at spill[(k-kTravLds)*sstride]).
__device__ __forceinline__ void traverse_closest_binary(const DevScene& sc, v3 o, v3 d, float tmin, float tmax,
                                                 uint32_t* stk, int col_stride, uint32_t* spill, size_t sstride,
                                                 float& t_out, int& prim_out) {
    // scene fields in registers: the spill stores below could otherwise
    // force a reload of every field read through `sc` after them
    const float4* const N = sc.nodes;
    const float4* const S = sc.spheres;
    const float4* const Qd = sc.quads;
    const uint32_t n_spheres = sc.n_spheres, n_nodes = sc.n_nodes;
    const uint32_t n_prims = sc.n_spheres + sc.n_quads;
    const v3 inv = V(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
    float best_t = tmax;
    uint32_t best_id = 0x7fffffffu;
    // tn: the entry distance of the primitive's box; hits before it do not count
    // (an id past the primitives, which no well-formed tree holds, reads
    // primitive 0 and is never taken)
    auto test_prim = [&](uint32_t id, float tn) {
        const bool valid = id < n_prims;
        const uint32_t k = valid ? id : 0u;
        float t = 0.0f;
        bool h;
        if (k < n_spheres) {
            h = hit_sphere_t(ldg4(S, k), o, d, tmin, tmax, t);
        } else {
            const size_t q = 3 * (size_t)(k - n_spheres);
            h = hit_quad_t(ldg4(Qd, q), ldg4(Qd, q + 1), ldg4(Qd, q + 2), o, d, tmin, tmax, t);
        }
        const bool take = valid & h & (t >= tn) & ((t < best_t) | ((t == best_t) & (id < best_id)));
        best_t = take ? t : best_t;
        best_id = take ? id : best_id;
    };
    if (n_nodes > 0) {
        const float4 r0 = ldg4(N, 0), r1 = ldg4(N, 1);
        float tr;
        const uint32_t rc = __float_as_uint(r0.w);
        if (slab(r0, r1, o, inv, tmin, best_t, tr)) {
            if (rc & AURORA_NODE_LEAF) {
                test_prim(rc & ~AURORA_NODE_LEAF, tr);
            } else {
                uint32_t cur = rc;
                int sp = 0;
                for (;;) {
                    if ((cur & 1u) || cur + 1 >= n_nodes) break;
                    const float4 a0 = ldg4(N, 2 * (size_t)cur), a1 = ldg4(N, 2 * (size_t)cur + 1);
                    const float4 b0 = ldg4(N, 2 * (size_t)cur + 2), b1 = ldg4(N, 2 * (size_t)cur + 3);
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
                        const uint32_t far_ = bnear ? ca : cb;
                        cur = bnear ? cb : ca;
                        if (sp < kTravLds) stk[sp * col_stride] = far_;
                        else if (sp < kTravMaxStack) spill[(size_t)(sp - kTravLds) * sstride] = far_;
                        if (sp < kTravMaxStack) ++sp;
                    } else if (ha) {
                        cur = ca;
                    } else if (hb) {
                        cur = cb;
                    } else if (sp == 0) {
                        break;
                    } else {
                        --sp;
                        if (sp < kTravLds) cur = stk[sp * col_stride];
                        else cur = ldg1(spill, (size_t)(sp - kTravLds) * sstride);
                    }
                }
            }
        }
    }
    const bool miss = best_id == 0x7fffffffu;
    t_out = miss ? tmax : best_t;
    prim_out = miss ? AURORA_MISS : (int)best_id;
}

}  // namespace aur
